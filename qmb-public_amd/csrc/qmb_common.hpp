// qmb_common.hpp -- shared host/device helpers for the qmb_hip library (gfx950).
//
// Configurations are nq bytes, site s = bit (s % 8) of byte s / 8 (LSB first, qmb/bitspack.py:50,
// qmb/_hamiltonian_cpu.cpp:36-46).  Loaded little-endian, the nq bytes form W = ceil(nq / 8) 64-bit
// words in which site s is bit (s % 64) of word s / 64, so every bit operation on a config is a
// word operation.  Byte-lexicographic order (byte 0 most significant, qmb/_hamiltonian_cpu.cpp:7-20)
// is the order of the byte-swapped words compared most-significant word first.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmb {

constexpr int kMaxOps = 4;       // operators per term (qmb/_hamiltonian.cpp:64, prepare<4>)
constexpr int kMaxWords = 4;     // configs up to 32 bytes = 256 sites
constexpr int kMaxQubytes = 8 * kMaxWords;
constexpr uint64_t kEmptySlot = ~0ull;

// ---------------------------------------------------------------------------------------------
// Compiled term record.
//
// A term is a product of at most four a / a^dagger operators applied right to left
// (qmb/_hamiltonian_cpu.cpp:59-76).  Whether it succeeds on a config depends only on the initial
// bits of the (at most four) distinct sites it touches: the first operator applied on a site needs
// that site's bit to be the opposite of its kind, later operators on the same site are then fully
// determined.  So a term compiles to, per touched site ("slot"): the required initial bit, the final
// bit, and whether the site occurs an odd number of times (which decides its contribution to the
// Jordan-Wigner string).  The fermionic parity of the reference's bit loop (:71-75) is
//     popcount(config & P & ~T) xor pconst
// with P = xor of lowmask(s) over all operator sites s, T the touched-site mask and pconst the part
// of the loop that falls on touched sites, which is static once the required pattern holds.
//
// Bit layout of the 64-bit record:
//   [ 0,32)  site of slot k in bits [8k, 8k+8)
//   [32,36)  valid slots (a prefix mask)
//   [36,40)  required initial bit per slot
//   [40,44)  final bit per slot
//   [44,48)  odd-multiplicity flag per slot (always 0 for hard-core bosons)
//   48       pconst
// ---------------------------------------------------------------------------------------------
enum CompileStatus : int32_t {
    kCompileOk = 0,
    kCompileNeverSucceeds = 1,  // e.g. a_p a_p: statically zero
    kCompileBadSite = 2,        // site < 0 or >= 8 * n_qubytes
    kCompileBadKind = 3,        // kind not in {0, 1, 2}
};

struct CompiledTerm {
    uint64_t rec;
    int32_t status;
    int32_t n_required;    // number of sites required occupied
    int32_t required[4];   // those sites, ascending
    bool diagonal;         // no bit changes: the term maps a config to itself
};

__host__ __device__ inline CompiledTerm compile_term(const int16_t* site, const uint8_t* kind, int particle_cut, int n_bits) {
    CompiledTerm out{};
    int ts[kMaxOps];
    int init[kMaxOps];
    int cur[kMaxOps];
    int mult[kMaxOps];
    int nt = 0;
    for (int op = kMaxOps; op-- > 0;) {
        int k = kind[op];
        if (k == 2) {
            continue;
        }
        if (k > 2) {
            out.status = kCompileBadKind;
            return out;
        }
        int s = site[op];
        if (s < 0 || s >= n_bits) {
            out.status = kCompileBadSite;
            return out;
        }
        int slot = -1;
        for (int u = 0; u < nt; ++u) {
            if (ts[u] == s) {
                slot = u;
            }
        }
        if (slot < 0) {
            slot = nt++;
            ts[slot] = s;
            init[slot] = 1 - k;
            cur[slot] = init[slot];
            mult[slot] = 0;
        }
        if (cur[slot] == k) {
            out.status = kCompileNeverSucceeds;
            return out;
        }
        cur[slot] = k;
        mult[slot] += 1;
    }
    // Static part of the parity: replay the reference loop on a config holding only the required
    // pattern; untouched sites contribute through the runtime mask instead.
    int pconst = 0;
    if (particle_cut == 1) {
        int cv[kMaxOps];
        for (int u = 0; u < nt; ++u) {
            cv[u] = init[u];
        }
        for (int op = kMaxOps; op-- > 0;) {
            int k = kind[op];
            if (k == 2) {
                continue;
            }
            int s = site[op];
            int slot = 0;
            for (int u = 0; u < nt; ++u) {
                if (ts[u] == s) {
                    slot = u;
                }
            }
            cv[slot] = k;
            for (int u = 0; u < nt; ++u) {
                if (ts[u] < s && cv[u]) {
                    pconst ^= 1;
                }
            }
        }
    }
    uint64_t rec = 0;
    bool diag = true;
    int nr = 0;
    for (int u = 0; u < nt; ++u) {
        rec |= static_cast<uint64_t>(ts[u]) << (8 * u);
        rec |= 1ull << (32 + u);
        rec |= static_cast<uint64_t>(init[u]) << (36 + u);
        rec |= static_cast<uint64_t>(cur[u]) << (40 + u);
        if (particle_cut == 1 && (mult[u] & 1)) {
            rec |= 1ull << (44 + u);
        }
        if (init[u] != cur[u]) {
            diag = false;
        }
        if (init[u]) {
            out.required[nr++] = ts[u];
        }
    }
    rec |= static_cast<uint64_t>(pconst) << 48;
    // ascending required sites
    for (int a = 1; a < nr; ++a) {
        for (int b = a; b > 0 && out.required[b - 1] > out.required[b]; --b) {
            int tmp = out.required[b];
            out.required[b] = out.required[b - 1];
            out.required[b - 1] = tmp;
        }
    }
    out.rec = rec;
    out.status = kCompileOk;
    out.n_required = nr;
    out.diagonal = diag;
    return out;
}

// Buckets: terms are grouped by the set R of sites they require occupied.  |R| = 0 -> bucket 0,
// |R| = 1 -> 1 + r, |R| = 2 -> 1 + n + r1 (r1 - 1) / 2 + r0 (r0 < r1), |R| >= 3 -> the residual
// bucket, which every config visits.  A config only visits buckets whose R it contains.
__host__ __device__ inline int64_t bucket_count(int n_bits) {
    return 2 + static_cast<int64_t>(n_bits) + static_cast<int64_t>(n_bits) * (n_bits - 1) / 2;
}

__host__ __device__ inline int64_t bucket_of(const CompiledTerm& c, int n_bits) {
    if (c.n_required == 0) {
        return 0;
    }
    if (c.n_required == 1) {
        return 1 + c.required[0];
    }
    if (c.n_required == 2) {
        int64_t r0 = c.required[0], r1 = c.required[1];
        return 1 + n_bits + r1 * (r1 - 1) / 2 + r0;
    }
    return bucket_count(n_bits) - 1;
}

// Inverse of bucket_of for pair buckets: index k = r1 (r1 - 1) / 2 + r0.
__host__ __device__ inline void decode_pair(int64_t k, int* r0, int* r1) {
    int64_t a = static_cast<int64_t>((1.0 + __builtin_sqrt(1.0 + 8.0 * static_cast<double>(k))) * 0.5);
    while (a * (a - 1) / 2 > k) {
        --a;
    }
    while ((a + 1) * a / 2 <= k) {
        ++a;
    }
    *r1 = static_cast<int>(a);
    *r0 = static_cast<int>(k - a * (a - 1) / 2);
}

// ---------------------------------------------------------------------------------------------
// Pair excitations: terms on four distinct sites that annihilate two (r0 < r1) and create two
// (c0 < c1), with no condition on any other site.  Every operator order of such a term succeeds
// exactly on the configs holding r0, r1 and not c0, c1, maps them to the same target, and has the
// Jordan-Wigner parity  pair_parity(c) ^ delta  with
//     pair_parity(c) = popcount(c & (L(r0) ^ L(r1) ^ L(c0) ^ L(c1))) mod 2,  L(s) = sites below s,
// where delta depends on the term only (its parity mask is the same XOR of low masks minus the
// four sites).  So all such terms with the same (r0, r1, c0, c1) merge into one coefficient
// sum_t (-1)^delta_t coef_t, applied with sign (-1)^pair_parity(source config).
// ---------------------------------------------------------------------------------------------
struct PairExcitation {
    int r0, r1;  // annihilated sites, r0 < r1
    int c0, c1;  // created sites, c0 < c1
    int delta;   // parity offset of this operator order
};

// Index of the unordered site pair (a < b) among all pairs of n sites: b (b - 1) / 2 + a.
__host__ __device__ __forceinline__ int64_t pair_index(int a, int b) { return static_cast<int64_t>(b) * (b - 1) / 2 + a; }

// The same pairs of n_bits sites (n_bits even) in blocks by site parity: even-even pairs, then
// odd-odd, then even-odd.  With sites interleaved by spin (even = alpha, odd = beta) the blocks are
// the alpha-alpha, beta-beta and alpha-beta pairs, so the coefficients a spin-conserving excitation
// can meet sit in one block.  A permutation of pair_index, so any table is merged the same way.
__host__ __device__ __forceinline__ int64_t pair_slot(int a, int b, int n_bits) {
    const int64_t half = n_bits >> 1;
    const int64_t same = half * (half - 1) / 2;
    if (((a ^ b) & 1) == 0) {
        return ((a & 1) ? same : 0) + pair_index(a >> 1, b >> 1);
    }
    const int ev = (a & 1) ? b : a;
    const int od = (a & 1) ? a : b;
    return 2 * same + static_cast<int64_t>(ev >> 1) * half + (od >> 1);
}

__host__ __device__ inline bool pair_excitation(uint64_t rec, int particle_cut, PairExcitation* out) {
    if ((rec >> 63) & 1ull) {
        return false;
    }
    const uint32_t valid = static_cast<uint32_t>(rec >> 32) & 0xf;
    if (valid != 0xf) {
        return false;
    }
    const uint32_t req = static_cast<uint32_t>(rec >> 36) & 0xf;
    const uint32_t fin = static_cast<uint32_t>(rec >> 40) & 0xf;
    int ann[2], cre[2], na = 0, nc = 0;
    int sites[4];
    for (int k = 0; k < 4; ++k) {
        sites[k] = static_cast<int>((rec >> (8 * k)) & 0xff);
        for (int q = 0; q < k; ++q) {
            if (sites[q] == sites[k]) {
                return false;
            }
        }
        const uint32_t r = (req >> k) & 1u, f = (fin >> k) & 1u;
        if (r == 1u && f == 0u) {
            ann[na++ & 1] = sites[k];
        } else if (r == 0u && f == 1u) {
            cre[nc++ & 1] = sites[k];
        } else {
            return false;
        }
    }
    if (na != 2 || nc != 2) {
        return false;
    }
    PairExcitation e;
    e.r0 = ann[0] < ann[1] ? ann[0] : ann[1];
    e.r1 = ann[0] < ann[1] ? ann[1] : ann[0];
    e.c0 = cre[0] < cre[1] ? cre[0] : cre[1];
    e.c1 = cre[0] < cre[1] ? cre[1] : cre[0];
    // pair_parity of the smallest source config {r0, r1}: r0 is below r1, plus the four
    // comparisons with the created sites; the term's own parity there is its constant bit.
    const int canon = (1 + (e.r0 < e.c0) + (e.r1 < e.c0) + (e.r0 < e.c1) + (e.r1 < e.c1)) & 1;
    const int pconst = static_cast<int>((rec >> 48) & 1);
    e.delta = particle_cut == 1 ? (pconst ^ canon) : 0;
    *out = e;
    return true;
}

// Single excitations: terms that annihilate one site r and create one site p, conditioned on at
// most one other site k (occupied or empty) -- the one-body terms and the number-dressed hops
// a^+_p n_k a_r of a molecular Hamiltonian.  Their parity is  single_parity(c) ^ delta  with
//     single_parity(c) = popcount(c & (L(r) ^ L(p))) mod 2,
// so per (r, p) they merge into a constant part A (valid with k empty) and a per-k change K:
// the matrix element on a source config c is (-1)^single_parity(c) (A + sum_{k in c} K[k]).
struct SingleExcitation {
    int r, p;   // annihilated, created
    int k;      // conditioned site, or -1
    int kocc;   // 1: k must be occupied, 0: empty
    int delta;  // parity offset of this operator order
};

__host__ __device__ inline bool single_excitation(uint64_t rec, int particle_cut, SingleExcitation* out) {
    if ((rec >> 63) & 1ull) {
        return false;
    }
    const uint32_t valid = static_cast<uint32_t>(rec >> 32) & 0xf;
    if (valid != 0x3u && valid != 0x7u) {
        return false;
    }
    const uint32_t req = static_cast<uint32_t>(rec >> 36) & 0xf;
    const uint32_t fin = static_cast<uint32_t>(rec >> 40) & 0xf;
    SingleExcitation e{-1, -1, -1, 0, 0};
    for (int u = 0; u < 4; ++u) {
        if (!((valid >> u) & 1u)) {
            continue;
        }
        const int s = static_cast<int>((rec >> (8 * u)) & 0xff);
        const uint32_t r = (req >> u) & 1u, f = (fin >> u) & 1u;
        if (r == 1u && f == 0u) {
            if (e.r >= 0) {
                return false;
            }
            e.r = s;
        } else if (r == 0u && f == 1u) {
            if (e.p >= 0) {
                return false;
            }
            e.p = s;
        } else {
            if (e.k >= 0) {
                return false;
            }
            e.k = s;
            e.kocc = static_cast<int>(r);
        }
    }
    if (e.r < 0 || e.p < 0) {
        return false;
    }
    // single_parity of the required pattern v = {r} (+ k when it must be occupied)
    const int canon = (e.r < e.p ? 1 : 0) ^ ((e.k >= 0 && e.kocc && ((e.k < e.r) != (e.k < e.p))) ? 1 : 0);
    const int pconst = static_cast<int>((rec >> 48) & 1);
    e.delta = particle_cut == 1 ? (pconst ^ canon) : 0;
    *out = e;
    return true;
}

// ---------------------------------------------------------------------------------------------
// Config words.
// ---------------------------------------------------------------------------------------------
template <int W>
struct Cfg {
    uint64_t w[W];
};

__host__ __device__ inline uint64_t low_mask_word(int s, int word) {
    int lo = s - 64 * word;
    if (lo >= 64) {
        return ~0ull;
    }
    if (lo <= 0) {
        return 0ull;
    }
    return (1ull << lo) - 1ull;
}

// Applies a compiled record.  Returns true on success and writes the resulting config and the
// parity bit (0 -> +1, 1 -> -1).
template <int W>
__host__ __device__ __forceinline__ bool apply_rec(uint64_t rec, const Cfg<W>& c, Cfg<W>& out, int& parity) {
    uint64_t touched[W];
    uint64_t pm[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        out.w[w] = c.w[w];
        touched[w] = 0;
        pm[w] = 0;
    }
    const uint32_t valid = static_cast<uint32_t>(rec >> 32) & 0xf;
    const uint32_t req = static_cast<uint32_t>(rec >> 36) & 0xf;
    const uint32_t fin = static_cast<uint32_t>(rec >> 40) & 0xf;
    const uint32_t odd = static_cast<uint32_t>(rec >> 44) & 0xf;
#pragma unroll
    for (int k = 0; k < kMaxOps; ++k) {
        if (!((valid >> k) & 1u)) {
            break;
        }
        const int s = static_cast<int>((rec >> (8 * k)) & 0xff);
        const int wi = s >> 6;
        const uint64_t bit = 1ull << (s & 63);
        uint64_t word = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            word = (w == wi) ? c.w[w] : word;
        }
        const uint32_t have = (word & bit) ? 1u : 0u;
        if (have != ((req >> k) & 1u)) {
            return false;
        }
        const bool f = (fin >> k) & 1u;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (w == wi) {
                out.w[w] = f ? (out.w[w] | bit) : (out.w[w] & ~bit);
                touched[w] |= bit;
            }
            if ((odd >> k) & 1u) {
                pm[w] ^= low_mask_word(s, w);
            }
        }
    }
    int p = static_cast<int>((rec >> 48) & 1);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        p ^= __builtin_popcountll(c.w[w] & pm[w] & ~touched[w]) & 1;
    }
    parity = p;
    return true;
}

// A compiled record expanded to word masks: the term succeeds on c iff (c & m) == v for every word;
// the target is c ^ f and the parity popcount(c & p) ^ pconst (p excludes the touched sites).
template <int W>
struct Masks {
    uint64_t m[W], v[W], f[W], p[W];
    uint64_t hf;  // index hash of the flip mask f
    int pconst;
};

template <int W>
__host__ __device__ __forceinline__ Masks<W> expand_rec(uint64_t rec, const uint64_t* __restrict__ cols) {
    Masks<W> x;
    x.hf = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        x.m[w] = 0;
        x.v[w] = 0;
        x.f[w] = 0;
        x.p[w] = 0;
    }
    const uint32_t valid = static_cast<uint32_t>(rec >> 32) & 0xf;
    const uint32_t req = static_cast<uint32_t>(rec >> 36) & 0xf;
    const uint32_t fin = static_cast<uint32_t>(rec >> 40) & 0xf;
    const uint32_t odd = static_cast<uint32_t>(rec >> 44) & 0xf;
#pragma unroll
    for (int k = 0; k < kMaxOps; ++k) {
        if ((valid >> k) & 1u) {
            const int s = static_cast<int>((rec >> (8 * k)) & 0xff);
            const uint64_t bit = 1ull << (s & 63);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (w == (s >> 6)) {
                    x.m[w] |= bit;
                    x.v[w] |= ((req >> k) & 1u) ? bit : 0ull;
                    x.f[w] |= (((req ^ fin) >> k) & 1u) ? bit : 0ull;
                }
            }
            if (cols != nullptr && (((req ^ fin) >> k) & 1u)) {
                x.hf ^= cols[s];
            }
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if ((odd >> k) & 1u) {
                    x.p[w] ^= low_mask_word(s, w);
                }
            }
        }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        x.p[w] &= ~x.m[w];
    }
    x.pconst = static_cast<int>((rec >> 48) & 1);
    return x;
}

// ---------------------------------------------------------------------------------------------
// Hashing (MurmurHash3 64-bit finaliser, a bijection on 64-bit words).
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// The index hash is linear over GF(2): H(x) = XOR of cols[s] over the occupied sites s, with cols
// pseudorandom 64-bit vectors.  Random linear maps are a 2-universal family, and linearity gives
// H(c ^ f) = H(c) ^ H(f): a connection's target hash is the source's hash XOR the term's flip hash.
__host__ __device__ __forceinline__ uint64_t hash_column(int site) {
    return fmix64(0x6a09e667f3bcc909ull + 0x9e3779b97f4a7c15ull * static_cast<uint64_t>(site + 1));
}

template <int W>
__host__ __device__ __forceinline__ uint64_t hash_cfg(const Cfg<W>& c, const uint64_t* __restrict__ cols) {
    uint64_t h = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t x = c.w[w];
        while (x) {
            h ^= cols[64 * w + __builtin_ctzll(x)];
            x &= x - 1;
        }
    }
    return h;
}

// Blocked Bloom filter over the same hash: one 64-bit word per key (chosen by the high half of the
// hash, like the table slot) with three bits chosen by the low half.
__host__ __device__ __forceinline__ uint64_t filter_word(uint64_t h, uint64_t fmask) { return (h >> 32) & fmask; }

__host__ __device__ __forceinline__ uint64_t filter_bits(uint64_t h) {
    return (1ull << (h & 63)) | (1ull << ((h >> 6) & 63)) | (1ull << ((h >> 12) & 63));
}

template <int W>
__host__ __device__ __forceinline__ bool cfg_equal(const Cfg<W>& a, const Cfg<W>& b) {
    bool eq = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        eq = eq && (a.w[w] == b.w[w]);
    }
    return eq;
}

__host__ __device__ __forceinline__ uint64_t bswap64(uint64_t x) {
    return __builtin_bswap64(x);
}

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85.
// ---------------------------------------------------------------------------------------------
struct U32x4 {
    uint32_t v[4];
};

__host__ __device__ inline U32x4 philox4x32_10(U32x4 ctr, uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * ctr.v[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * ctr.v[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        U32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ctr;
}

// Uniform double in (0, 1] from counter `index` under a 64-bit seed: the top 53 bits of the first
// two output words, plus one, times 2^-53.
__host__ __device__ inline double philox_uniform(uint64_t seed, uint64_t index) {
    U32x4 ctr{{static_cast<uint32_t>(index), static_cast<uint32_t>(index >> 32), 0u, 0u}};
    U32x4 r = philox4x32_10(ctr, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const uint64_t bits = ((static_cast<uint64_t>(r.v[0]) << 32) | r.v[1]) >> 11;
    return static_cast<double>(bits + 1) * (1.0 / 9007199254740992.0);
}

}  // namespace qmb
