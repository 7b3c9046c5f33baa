// qmb_pairs.hip -- the pair-excitation path of apply_within on gfx950.
//
// Two configs a, b of one set are connected by a pair excitation (annihilate the pair Pa of a,
// create the pair Pb) exactly when a \ Pa == b \ Pb ("core") and Pa, Pb are disjoint.  So instead
// of generating every (config, term) target and probing the index for it -- the reference's
// scheme, qmb/_hamiltonian_cuda.cu:86-135, which for a molecular Hamiltonian spends ~90% of its
// work on double-excitation targets that are not in the set -- every config emits one entry per
// occupied pair (its cores), the entries are radix-sorted by core hash, and a wave walks each group
// of equal cores, combining its rows pairwise with the merged coefficient table of qmb_pairs.
// Only connections whose target is in the set are visited; each output row sums in registers and
// reaches HBM with one FP64 atomic per group it belongs to.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "qmb_internal.hpp"

namespace qmb {

namespace {

constexpr int kBlock = 256;
// Sort key of a core entry, 32 bits: the core class (the number of even sites of the core mod 3)
// times kHashRange, plus the 64-bit core hash mod kHashRange.  The class is a function of the core,
// so a group never splits; in a set where every config has the same numbers of even and odd
// occupied sites (spin sectors) it tells the pair class of all the group's entries, so groups come
// in class order and waves running side by side read one block of the coefficient table.  Cores
// with equal keys share a group ("mixed"), where the kernels compare cores exactly; the modulus
// breaks the hash's linearity, so such merges are chance events (about 1 group in 100 at 12M cores
// per class) with no structure.  Unused entries get kKeySentinel and sort last.  A 32-bit key
// sorts in four 8-bit radix passes, with the packed (row, removed sites) value riding along; the
// sort is stable, so inside a group the entries keep the order they were emitted in: by the rank
// of the removed pair (order 1: site) among the row's own, then by row.
constexpr uint32_t kHashRange = 0x55555555u;  // 3 * kHashRange = 2^32 - 1
constexpr uint32_t kKeySentinel = 0xFFFFFFFFu;
constexpr int kKeyBits = 32;
// A packed core entry is one u64: key << kPackedBits | row << sub_bits | sub, sorted on bits
// [kPackedBits, kPackedBits + 32) alone.  The range ends at bit 63, not 64: rocPRIM's radix sort
// takes a merge path for small inputs whose comparison mask is (1 << end_bit) - 1 in the key type,
// which a 64-bit end_bit overflows (it then compared the low bits).
constexpr int kPackedBits = 31;
constexpr int kMaxPairBytes = 16;            // the dense table is n_pairs^2 entries

inline unsigned blocks_for(int64_t n, int block = kBlock) {
    int64_t b = (n + block - 1) / block;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

// Index of the calling wave among the grid's persistent waves, ordered so that each XCD holds a
// contiguous eighth of [0, waves).  Workgroups go to the 8 XCDs round-robin by blockIdx.x, so with
// a grid that divides evenly, waves walking units w, w + waves, ... keep each XCD's L2 on the
// coefficient rows of an eighth of every window of units instead of rows from the whole window.
// Measured: the batch kernel gains (138.9 -> 134 ms on 64 molecular instances); the single-set
// pair and single kernels lose (8.07 -> 8.46 ms at 1M configs), so they keep the plain order.
__device__ __forceinline__ int64_t xcd_wave() {
    constexpr int kXcds = 8;
    constexpr int64_t kWavesPerBlock = kBlock / 64;
    const int64_t in_block = threadIdx.x >> 6;
    if (gridDim.x % kXcds != 0) {
        return static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + in_block;
    }
    const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    return static_cast<int64_t>(blockIdx.x % kXcds) * (waves / kXcds) + static_cast<int64_t>(blockIdx.x / kXcds) * kWavesPerBlock +
           in_block;
}

__global__ void zero_counts_kernel(int64_t* counts) {
    if (threadIdx.x < 4) {
        counts[threadIdx.x] = 0;
    }
}

// ------------------------------------------------------------------------------ coefficients --
// A diagonal term needs every site it touches back in its first state, so each of its sites carries
// an even number of its (at most four) operators: it conditions at most two sites, and its sign is
// the constant (-1)^pconst.  Its indicator prod_s (v_s ? n_s : 1 - n_s) expands into a constant,
// per-site and per-site-pair part: diag[0], diag[1 + p], diag[1 + n_bits + pair_index(p, q)].
__device__ __forceinline__ void add2(double2* d, double sgn, const double2& v) {
    unsafeAtomicAdd(&d->x, sgn * v.x);
    unsafeAtomicAdd(&d->y, sgn * v.y);
}

__device__ void fold_diagonal(uint64_t rec, int particle_cut, int n_bits, const double2& coef, double2* diag) {
    const uint32_t valid = static_cast<uint32_t>(rec >> 32) & 0xf;
    const uint32_t req = static_cast<uint32_t>(rec >> 36) & 0xf;
    const int pconst = particle_cut == 1 ? static_cast<int>((rec >> 48) & 1) : 0;
    const double2 v = pconst ? make_double2(-coef.x, -coef.y) : coef;
    int sites[2], occ[2], ns = 0;
    for (int u = 0; u < 4; ++u) {
        if ((valid >> u) & 1u) {
            if (ns < 2) {
                sites[ns] = static_cast<int>((rec >> (8 * u)) & 0xff);
                occ[ns] = static_cast<int>((req >> u) & 1u);
            }
            ++ns;
        }
    }
    double2* e1 = diag + 1;
    double2* e2 = diag + 1 + n_bits;
    if (ns == 0) {
        add2(diag, 1.0, v);
    } else if (ns == 1) {
        if (occ[0]) {
            add2(e1 + sites[0], 1.0, v);
        } else {
            add2(diag, 1.0, v);
            add2(e1 + sites[0], -1.0, v);
        }
    } else if (ns == 2) {
        const int p = sites[0] < sites[1] ? sites[0] : sites[1];
        const int q = sites[0] < sites[1] ? sites[1] : sites[0];
        const int op = sites[0] < sites[1] ? occ[0] : occ[1];
        const int oq = sites[0] < sites[1] ? occ[1] : occ[0];
        double2* pq = e2 + pair_index(p, q);
        // (op ? n_p : 1 - n_p) (oq ? n_q : 1 - n_q)
        add2(pq, (op == oq) ? 1.0 : -1.0, v);
        if (!oq) {
            add2(e1 + p, op ? 1.0 : -1.0, v);
        }
        if (!op) {
            add2(e1 + q, oq ? 1.0 : -1.0, v);
        }
        if (!op && !oq) {
            add2(diag, 1.0, v);
        }
    }
}

// *counter += the lanes of the wave where `add` holds, as one atomic from the wave's first such lane
// (a same-address atomic per thread serialises: 600k of them took 7 ms).  Every lane must call it.
__device__ __forceinline__ void wave_count(int64_t* counter, bool add) {
    const uint64_t m = __ballot(add);
    if (m != 0ull && static_cast<int>(threadIdx.x & 63u) == __builtin_ctzll(m)) {
        atomicAdd(reinterpret_cast<unsigned long long*>(counter), static_cast<unsigned long long>(__popcll(m)));
    }
}

// Launched with whole waves: threads past n_terms take part in the counts only.
__global__ void pairs_fill_kernel(const int16_t* __restrict__ site, const uint8_t* __restrict__ kind,
                                  const double2* __restrict__ coef, int64_t n_terms, int particle_cut, int n_bits,
                                  int64_t n_pairs, double2* __restrict__ table, double2* __restrict__ diag,
                                  int64_t* __restrict__ counts) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    bool is_diag = false, is_pair = false;
    if (t < n_terms) {
        const CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, particle_cut, n_bits);
        PairExcitation e;
        if (c.status == kCompileOk && c.diagonal) {
            fold_diagonal(c.rec, particle_cut, n_bits, coef[t], diag);
            is_diag = true;
        } else if (c.status == kCompileOk && pair_excitation(c.rec, particle_cut, &e)) {
            const double2 v = coef[t];
            double2* dst = table + pair_slot(e.r0, e.r1, n_bits) * n_pairs + pair_slot(e.c0, e.c1, n_bits);
            unsafeAtomicAdd(&dst->x, e.delta ? -v.x : v.x);
            unsafeAtomicAdd(&dst->y, e.delta ? -v.y : v.y);
            is_pair = true;
        }
    }
    wave_count(&counts[0], is_pair);
    wave_count(&counts[1], is_diag);
}

int pairs_layout(Carver& cv, int n_qubytes, qmb_pairs* out) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n_qubytes > kMaxPairBytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "the pair-excitation table supports n_qubytes <= %d, got %d", kMaxPairBytes,
                         n_qubytes);
    }
    qmb_pairs p{};
    p.n_qubytes = n_qubytes;
    p.n_bits = 8 * n_qubytes;
    p.n_pairs = static_cast<int64_t>(p.n_bits) * (p.n_bits - 1) / 2;
    p.buffer = cv.base;
    p.counts = cv.take<int64_t>(4);
    p.coef = cv.take<double>(2 * static_cast<size_t>(p.n_pairs) * static_cast<size_t>(p.n_pairs));
    p.coef_re = cv.take<double>(static_cast<size_t>(p.n_pairs) * static_cast<size_t>(p.n_pairs));
    p.diag = cv.take<double>(2 * (1 + static_cast<size_t>(p.n_bits) + static_cast<size_t>(p.n_pairs)));
    p.bytes = cv.off;
    if (out) {
        *out = p;
    }
    return QMB_OK;
}

// Real parts of the merged coefficients; counts[2] += entries with a nonzero imaginary part.
__global__ void pairs_real_kernel(const double2* __restrict__ coef, int64_t n, double* __restrict__ re, int64_t* __restrict__ counts) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) {
        return;
    }
    const double2 v = coef[k];
    re[k] = v.x;
    if (v.y != 0.0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&counts[2]), 1ull);
    }
}

// ----------------------------------------------------------------------------------- cores --
template <int W>
__device__ __forceinline__ Cfg<W> load_key(const uint64_t* __restrict__ keys, int64_t i) {
    Cfg<W> c;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        c.w[w] = keys[i * W + w];
    }
    return c;
}

template <int W>
__device__ __forceinline__ void flip_site(Cfg<W>& c, int s) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if (w == (s >> 6)) {
            c.w[w] ^= 1ull << (s & 63);
        }
    }
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int lane) {
    const uint32_t lo = __shfl(static_cast<uint32_t>(v), lane);
    const uint32_t hi = __shfl(static_cast<uint32_t>(v >> 32), lane);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t core_key(uint64_t h, uint32_t cls) {
    return cls * kHashRange + static_cast<uint32_t>(h % kHashRange);
}

// One thread per config: every occupied pair (x < y) -- ORDER 2 -- or every occupied site x --
// ORDER 1 -- gives the entry (key of the config without those sites, row << 16 | y << 8 | x); unused
// slots of a row with fewer occupied sites get the sentinel key.  Entry k of row i sits at k * n + i,
// so the lanes of a wave store to consecutive addresses.  A row's pairs are emitted in pair_slot
// order (even-even pairs, then odd-odd, each by the higher site first; then even-odd, by the even
// site first) and its sites in ascending order.  Rows of one group share the core, so a pair's rank
// among its row's pairs grows with its slot (ties where no core site lies between two slots), and
// the stable sort leaves each group's entries close to ascending slots: the lanes of a wave then
// read nearby columns of a coefficient table row.
template <int W, int ORDER, bool PACKED>
__global__ void core_emit_kernel(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ cols, int64_t n, int n_bits,
                                 int per_config, int sub_bits, uint32_t* __restrict__ okeys, void* __restrict__ ovals,
                                 int64_t* __restrict__ counts) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    constexpr uint64_t kEven = 0x5555555555555555ull;
    const Cfg<W> c = load_key<W>(keys, i);
    const uint64_t h = hash_cfg<W>(c, cols);
    int even = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        even += __popcll(c.w[w] & kEven);
    }
    int64_t k = 0;
    // the value: row << 16 | y << 8 | x, or packed into 32 bits as row << sub_bits | (pair slot of
    // (x, y), or the site x); the packed value halves what the sort moves per value
    auto emit = [&](uint64_t key, uint64_t val) {
        if (k < per_config) {
            if (PACKED) {
                const uint32_t sub = ORDER == 2 ? static_cast<uint32_t>(pair_slot(static_cast<int>(val & 0xff), static_cast<int>(val >> 8), n_bits))
                                                : static_cast<uint32_t>(val);
                static_cast<uint64_t*>(ovals)[k * n + i] = (key << kPackedBits) | ((static_cast<uint32_t>(i) << sub_bits) | sub);
            } else {
                okeys[k * n + i] = static_cast<uint32_t>(key);
                static_cast<uint64_t*>(ovals)[k * n + i] = (static_cast<uint64_t>(i) << 16) | val;
            }
        }
        ++k;
    };
    if (ORDER == 1) {
#pragma unroll
        for (int wx = 0; wx < W; ++wx) {
            uint64_t mx = c.w[wx];
            while (mx) {
                const int x = 64 * wx + __builtin_ctzll(mx);
                mx &= mx - 1;
                emit(core_key(h ^ cols[x], 0u), static_cast<uint64_t>(x));
            }
        }
    } else {
        // same-parity pairs: the core loses two even sites (parity 0) or none (parity 1)
        for (int par = 0; par < 2; ++par) {
            const uint64_t pm = par ? ~kEven : kEven;
            const uint32_t cls = static_cast<uint32_t>((even - (par ? 0 : 2)) % 3);
#pragma unroll
            for (int wy = 0; wy < W; ++wy) {
                uint64_t my = c.w[wy] & pm;
                while (my) {
                    const int y = 64 * wy + __builtin_ctzll(my);
                    my &= my - 1;
                    const uint64_t hy = h ^ cols[y];
#pragma unroll
                    for (int wx = 0; wx < W; ++wx) {
                        if (wx > wy) {
                            continue;
                        }
                        uint64_t mx = c.w[wx] & pm & (wx == wy ? (1ull << (y & 63)) - 1ull : ~0ull);
                        while (mx) {
                            const int x = 64 * wx + __builtin_ctzll(mx);
                            mx &= mx - 1;
                            emit(core_key(hy ^ cols[x], cls), (static_cast<uint64_t>(y) << 8) | static_cast<uint64_t>(x));
                        }
                    }
                }
            }
        }
        // even-odd pairs: the core loses one even site
        const uint32_t cls = static_cast<uint32_t>((even - 1) % 3);
#pragma unroll
        for (int we = 0; we < W; ++we) {
            uint64_t me = c.w[we] & kEven;
            while (me) {
                const int ev = 64 * we + __builtin_ctzll(me);
                me &= me - 1;
                const uint64_t he = h ^ cols[ev];
#pragma unroll
                for (int wo = 0; wo < W; ++wo) {
                    uint64_t mo = c.w[wo] & ~kEven;
                    while (mo) {
                        const int od = 64 * wo + __builtin_ctzll(mo);
                        mo &= mo - 1;
                        const int x = ev < od ? ev : od;
                        const int y = ev < od ? od : ev;
                        emit(core_key(he ^ cols[od], cls), (static_cast<uint64_t>(y) << 8) | static_cast<uint64_t>(x));
                    }
                }
            }
        }
    }
    if (k > per_config) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&counts[2]), 1ull);
    }
    for (; k < per_config; ++k) {
        if (PACKED) {
            static_cast<uint64_t*>(ovals)[k * n + i] = static_cast<uint64_t>(kKeySentinel) << kPackedBits;
        } else {
            okeys[k * n + i] = kKeySentinel;
            static_cast<uint64_t*>(ovals)[k * n + i] = 0ull;
        }
    }
}

// inv[pair_slot(x, y)] = y << 8 | x for every pair x < y of n_bits sites.
__global__ void pair_inverse_kernel(int n_bits, uint32_t* __restrict__ inv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_bits * n_bits) {
        return;
    }
    const int x = t % n_bits, y = t / n_bits;
    if (x < y) {
        inv[pair_slot(x, y, n_bits)] = (static_cast<uint32_t>(y) << 8) | static_cast<uint32_t>(x);
    }
}

// Flags of a used sorted core entry: the first entry of its key (a group head, bit 0) and an entry
// whose core differs from the previous entry's under the same key (two cores merged by equal keys,
// bit 31).  Unused slots (the sentinel keys, sorted last) get 0; the flag pass writes the number of
// used entries to counts[0] and each wave's number of heads to its chunk's counter; a head's group
// is the heads before its chunk (an exclusive scan of the counters) plus those before it in the chunk.
constexpr uint32_t kFlagHead = 1u;
constexpr uint32_t kFlagBreak = 0x80000000u;
// Bit 31 of starts[g] (g < G): group g holds more than one core (its rows' cores are compared exactly).
constexpr uint32_t kGroupMixed = 0x80000000u;
constexpr uint32_t kGroupStart = 0x7FFFFFFFu;

// The core of an entry (row << 16 | y << 8 | x; order 1: row << 16 | x): its row's config without
// the entry's sites.
template <int W, int ORDER>
__device__ __forceinline__ Cfg<W> entry_core(const uint64_t* __restrict__ ckeys, uint64_t v) {
    Cfg<W> c = load_key<W>(ckeys, static_cast<int64_t>(v >> 16));
    flip_site<W>(c, static_cast<int>(v & 0xff));
    if (ORDER == 2) {
        flip_site<W>(c, static_cast<int>((v >> 8) & 0xff));
    }
    return c;
}

template <int ORDER>
__device__ __forceinline__ uint64_t unpack_entry(uint64_t x, int sub_bits, const uint32_t* __restrict__ inv) {
    const uint32_t pv = static_cast<uint32_t>(x & ((uint64_t{1} << kPackedBits) - 1u));
    const uint32_t sub = pv & ((1u << sub_bits) - 1u);
    return (static_cast<uint64_t>(pv >> sub_bits) << 16) | (ORDER == 2 ? inv[sub] : sub);
}

// Flags of kFlagUnroll * 64 consecutive sorted entries per wave: entry base + 64 t + lane in slot t of
// each lane.  A wave loads its entries once (the neighbours' keys come from the lanes beside, across
// slots from lane 63 / lane 0 of the slot beside, and the wave's outer neighbours are read by lanes 0
// and 63).  Only entries of groups of two or more read their config: each compares its core with
// the previous entry's, taken the same way; the slots' config loads are independent, so a lane keeps
// kFlagUnroll of them in flight.  Every lane runs every shuffle.  load(j, k, v) reads entry j < cap.
constexpr int kFlagUnroll = 4;

template <int W, int ORDER, typename Load>
__device__ __forceinline__ void core_flags(const uint64_t* __restrict__ ckeys, int64_t cap, Load load, uint32_t (&f)[kFlagUnroll],
                                           uint32_t (&k)[kFlagUnroll], uint64_t (&v)[kFlagUnroll], int64_t base, int lane,
                                           int64_t* __restrict__ counts) {
    constexpr int U = kFlagUnroll;
#pragma unroll
    for (int t = 0; t < U; ++t) {
        k[t] = kKeySentinel;
        v[t] = 0ull;
        const int64_t e = base + 64 * t + lane;
        if (e < cap) {
            load(e, k[t], v[t]);
        }
    }
    // the wave's outer neighbours
    uint32_t k_before = kKeySentinel, k_after = kKeySentinel;
    uint64_t v_before = 0ull;
    if (lane == 0 && base > 0 && base <= cap) {
        load(base - 1, k_before, v_before);
    }
    if (lane == 63 && base + 64 * U < cap) {
        uint64_t v_after;
        load(base + 64 * U, k_after, v_after);
    }
    uint32_t kp[U], kn[U];
#pragma unroll
    for (int t = 0; t < U; ++t) {
        const uint32_t below = __shfl(k[t], (lane + 63) & 63);
        const uint32_t above = __shfl(k[t], (lane + 1) & 63);
        const uint32_t prev_slot = __shfl(k[t > 0 ? t - 1 : 0], 63);
        const uint32_t next_slot = __shfl(k[t < U - 1 ? t + 1 : U - 1], 0);
        kp[t] = lane != 0 ? below : (t > 0 ? prev_slot : k_before);
        kn[t] = lane != 63 ? above : (t < U - 1 ? next_slot : k_after);
    }
    Cfg<W> core[U];
#pragma unroll
    for (int t = 0; t < U; ++t) {
        const bool used = k[t] != kKeySentinel;
        const bool multi = used && (kp[t] == k[t] || kn[t] == k[t]);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            core[t].w[w] = 0ull;
        }
        if (multi) {
            core[t] = entry_core<W, ORDER>(ckeys, v[t]);
        }
    }
    Cfg<W> core_before{};
    if (lane == 0 && k_before != kKeySentinel && k_before == k[0]) {
        core_before = entry_core<W, ORDER>(ckeys, v_before);
    }
#pragma unroll
    for (int t = 0; t < U; ++t) {
        bool same = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t below = shfl_u64(core[t].w[w], (lane + 63) & 63);
            const uint64_t prev_slot = shfl_u64(core[t > 0 ? t - 1 : 0].w[w], 63);
            const uint64_t prev = lane != 0 ? below : (t > 0 ? prev_slot : core_before.w[w]);
            same = same & (prev == core[t].w[w]);
        }
        const bool used = k[t] != kKeySentinel;
        const bool head = used && kp[t] != k[t];
        f[t] = !used ? 0u : (head ? kFlagHead : (same ? 0u : kFlagBreak));
        if (used && kn[t] == kKeySentinel) {
            counts[0] = base + 64 * t + lane + 1;  // the last used entry
        }
    }
}

// Entries per flag chunk: one wave's kFlagUnroll * 64 consecutive entries.
constexpr int kFlagChunk = 64 * kFlagUnroll;

inline int64_t flag_chunks(int64_t cap) { return static_cast<int64_t>(blocks_for(cap, kBlock * kFlagUnroll)) * (kBlock / 64); }

// chunk_heads[c] = number of group heads among the entries of chunk c (the wave's own).  Every wave
// of the grid writes its chunk, also past the entries.
__device__ __forceinline__ void store_chunk_heads(const uint32_t (&f)[kFlagUnroll], int64_t base, int lane,
                                                  uint32_t* __restrict__ chunk_heads) {
    uint32_t heads = 0;
#pragma unroll
    for (int t = 0; t < kFlagUnroll; ++t) {
        heads += static_cast<uint32_t>(__popcll(__ballot((f[t] & kFlagHead) != 0u)));
    }
    if (lane == 0) {
        chunk_heads[base / kFlagChunk] = heads;
    }
}

// Flags of packed entries (key << 31 | row << sub_bits | sub, sorted on the key); also splits each
// entry into its key and its value in the kernels' form (row << 16 | y << 8 | x), and counts each
// chunk's heads.  Grid: flag_chunks(cap) waves of kFlagChunk entries.
template <int W, int ORDER>
__global__ __launch_bounds__(kBlock) void core_flag_unpack_kernel(const uint64_t* __restrict__ entries, int64_t cap,
                                                                  uint32_t* __restrict__ flag, uint32_t* __restrict__ chunk_heads,
                                                                  int sub_bits, const uint32_t* __restrict__ inv,
                                                                  const uint64_t* __restrict__ ckeys, uint32_t* __restrict__ keys,
                                                                  uint64_t* __restrict__ vals, int64_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t base = (blockIdx.x * static_cast<int64_t>(blockDim.x) + (threadIdx.x & ~63)) * kFlagUnroll;
    auto load = [&](int64_t j, uint32_t& kj, uint64_t& vj) {
        const uint64_t xj = entries[j];
        kj = static_cast<uint32_t>(xj >> kPackedBits);
        vj = kj != kKeySentinel ? unpack_entry<ORDER>(xj, sub_bits, inv) : 0ull;
    };
    uint32_t f[kFlagUnroll], k[kFlagUnroll];
    uint64_t v[kFlagUnroll];
    core_flags<W, ORDER>(ckeys, cap, load, f, k, v, base, lane, counts);
    store_chunk_heads(f, base, lane, chunk_heads);
#pragma unroll
    for (int t = 0; t < kFlagUnroll; ++t) {
        const int64_t e = base + 64 * t + lane;
        if (e < cap) {
            flag[e] = f[t];
            keys[e] = k[t];
            vals[e] = v[t];
        }
    }
}

// Flags of sorted (key, value) pairs and each chunk's heads, as core_flag_unpack_kernel.
template <int W, int ORDER>
__global__ __launch_bounds__(kBlock) void core_flag_kernel(const uint32_t* __restrict__ keys, const uint64_t* __restrict__ vals, int64_t cap,
                                                           const uint64_t* __restrict__ ckeys, uint32_t* __restrict__ flag,
                                                           uint32_t* __restrict__ chunk_heads, int64_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t base = (blockIdx.x * static_cast<int64_t>(blockDim.x) + (threadIdx.x & ~63)) * kFlagUnroll;
    auto load = [&](int64_t j, uint32_t& kj, uint64_t& vj) {
        kj = keys[j];
        vj = vals[j];
    };
    uint32_t f[kFlagUnroll], k[kFlagUnroll];
    uint64_t v[kFlagUnroll];
    core_flags<W, ORDER>(ckeys, cap, load, f, k, v, base, lane, counts);
    store_chunk_heads(f, base, lane, chunk_heads);
#pragma unroll
    for (int t = 0; t < kFlagUnroll; ++t) {
        const int64_t e = base + 64 * t + lane;
        if (e < cap) {
            flag[e] = f[t];
        }
    }
}

// flag[e] = 1 where a sorted key differs from the previous one (the strings index).
__global__ void key_flag_kernel(const uint32_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ flag) {
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= n) {
        return;
    }
    const uint32_t k = keys[e];
    flag[e] = (k != kKeySentinel && (e == 0 || keys[e - 1] != k)) ? 1u : 0u;
}

// starts[g] = first entry of group g, with kGroupMixed when a break lies inside the group;
// starts[G] = number of used entries; counts {entries, G}.  One wave per flag chunk, 64 flags at a
// time: chunk_off[c] = heads before chunk c (the exclusive scan of the chunks' head counts), so a
// head's group is that plus the heads before it in the chunk.  A head finds its group's end (the
// next head or the end of the used entries) and the breaks before it from the wave's ballots; the
// last head of the 64, whose group may run past them, reads on 64 flags at a time with the whole
// wave.  Grid: flag_chunks(capacity) waves.
__global__ __launch_bounds__(kBlock) void core_start_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ chunk_off,
                                                            uint32_t* __restrict__ starts, int64_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t used = counts[0];
    const int64_t chunk = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) >> 6;
    const uint64_t below = (1ull << lane) - 1ull;                    // lanes below this one
    const uint64_t after = lane == 63 ? 0ull : ~0ull << (lane + 1);  // lanes above this one
    if (chunk * kFlagChunk >= used) {
        return;  // the whole wave
    }
    uint32_t g0 = chunk_off[chunk];
    for (int t = 0; t < kFlagUnroll; ++t) {
        const int64_t base = chunk * kFlagChunk + 64 * t;
        if (base >= used) {
            break;  // the whole wave
        }
        const int64_t e = base + lane;
        const uint32_t f = e < used ? flag[e] : 0u;
        const bool head = (f & kFlagHead) != 0u;
        const uint64_t heads = __ballot(head);
        const uint64_t stop = __ballot(e >= used || head);
        const uint64_t brk = __ballot((f & kFlagBreak) != 0u);
        const uint64_t next = stop & after;
        // the group of the 64's last head runs on past them when no stop follows it
        const bool open = head && next == 0ull;
        const uint64_t open_mask = __ballot(open);
        bool tail_break = false;
        int64_t tail_end = 0;
        if (open_mask != 0ull) {
            for (int64_t j0 = base + 64;; j0 += 64) {
                const int64_t j = j0 + lane;
                const uint32_t fj = j < used ? flag[j] : 0u;
                const uint64_t st = __ballot(j >= used || (fj & kFlagHead) != 0u);
                const uint64_t bk = __ballot((fj & kFlagBreak) != 0u);
                if (st != 0ull) {
                    const int nx = __builtin_ctzll(st);
                    tail_break = tail_break || (bk & ((1ull << nx) - 1ull)) != 0ull;
                    tail_end = j0 + nx;
                    break;
                }
                tail_break = tail_break || bk != 0ull;
            }
        }
        if (head) {
            bool mixed;
            int64_t end;
            if (open) {
                mixed = (brk & after) != 0ull || tail_break;
                end = tail_end;
            } else {
                const int nx = __builtin_ctzll(next);
                mixed = (brk & after & ((1ull << nx) - 1ull)) != 0ull;
                end = base + nx;
            }
            const uint32_t g = g0 + static_cast<uint32_t>(__popcll(heads & below));
            starts[g] = static_cast<uint32_t>(e) | (mixed ? kGroupMixed : 0u);
            if (end == used) {
                starts[g + 1] = static_cast<uint32_t>(end);
                counts[1] = static_cast<int64_t>(g) + 1;
            }
        }
        g0 += static_cast<uint32_t>(__popcll(heads));
    }
}

// Bits of the sub field of a packed core value (order 2: pair slots, order 1: sites) and of the
// row; the 32-bit value applies when both fit.
inline int bits_for(int64_t range) {
    int b = 0;
    while ((int64_t{1} << b) < range) {
        ++b;
    }
    return b;
}
inline int core_sub_bits(int n_bits, int order) {
    return bits_for(order == 2 ? static_cast<int64_t>(n_bits) * (n_bits - 1) / 2 : n_bits);
}
inline bool core_packed(int64_t n, int n_bits, int order) { return bits_for(n) + core_sub_bits(n_bits, order) <= kPackedBits; }

int cores_layout(Carver& cv, int n_qubytes, int64_t n, int max_occupied, int order, qmb_cores* out, void** scratch,
                 size_t* scratch_bytes) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "n must be >= 0");
    }
    if (order != 1 && order != 2) {
        return set_error(QMB_ERR_ARGUMENT, "order must be 1 or 2, got %d", order);
    }
    if (max_occupied < 0 || max_occupied > 8 * n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "max_occupied must be in [0, %d], got %d", 8 * n_qubytes, max_occupied);
    }
    qmb_cores c{};
    c.n_qubytes = n_qubytes;
    c.n_words = n_words_of(n_qubytes);
    c.max_occupied = max_occupied;
    c.order = order;
    c.per_config = order == 2 ? max_occupied * (max_occupied - 1) / 2 : max_occupied;
    c.n = n;
    c.capacity = n * c.per_config;
    if (c.capacity >= (1ll << 31)) {
        return set_error(QMB_ERR_UNSUPPORTED, "core index of %lld entries exceeds 2^31 - 1", static_cast<long long>(c.capacity));
    }
    const size_t cap = static_cast<size_t>(c.capacity > 0 ? c.capacity : 1);
    c.buffer = cv.base;
    c.counts = cv.take<int64_t>(4);
    c.keys = cv.take<uint32_t>(cap);
    c.vals = cv.take<uint64_t>(cap);
    c.starts = cv.take<uint32_t>(cap + 1);
    size_t tmp = 0;
    {
        uint64_t* v = nullptr;
        uint32_t* u = nullptr;
        if (rocprim::radix_sort_pairs(nullptr, tmp, u, u, v, v, static_cast<unsigned>(cap), 0, kKeyBits) != hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim radix_sort_pairs size query failed");
        }
        size_t packed = 0;
        if (rocprim::radix_sort_keys(nullptr, packed, v, v, static_cast<unsigned>(cap), kPackedBits, kPackedBits + kKeyBits) !=
            hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim radix_sort_keys size query failed");
        }
        tmp = tmp > packed ? tmp : packed;
        size_t scan = 0;
        if (rocprim::exclusive_scan(nullptr, scan, u, u, 0u, cap, rocprim::plus<uint32_t>()) != hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim exclusive_scan size query failed");
        }
        tmp = tmp > scan ? tmp : scan;
    }
    const size_t s0 = cv.off;
    const size_t inv_bytes = static_cast<size_t>(8 * n_qubytes) * (8 * n_qubytes) / 2 * sizeof(uint32_t) + 256;
    const size_t chunk_bytes = 2 * static_cast<size_t>(flag_chunks(static_cast<int64_t>(cap))) * sizeof(uint32_t) + 512;
    void* s = cv.take_bytes(2 * cap * sizeof(uint64_t) + cap * sizeof(uint32_t) + chunk_bytes + inv_bytes + tmp + 4096);
    if (scratch) {
        *scratch = s;
    }
    if (scratch_bytes) {
        *scratch_bytes = cv.off - s0;
    }
    c.bytes = cv.off;
    if (out) {
        *out = c;
    }
    return QMB_OK;
}

// ------------------------------------------------------------------------------ group kernel --
struct PairParams {
    const double2* coef;    // merged pair-excitation coefficients, [pair_slot][pair_slot]
    const double* coef_re;  // their real parts when every coefficient is real, else null
    int64_t n_pairs;
    int n_bits;
    int particle_cut;
    const uint64_t* ckeys;  // config words of the set, [n][W]
    const uint64_t* vals;   // sorted entries
    const uint32_t* starts;
    const int64_t* counts;
    const double2* in;
    double2* out;
    const double2* in2;  // second input/output of a two-vector pass (NV = 2), else null
    double2* out2;
};

// Sites in [a, b), a <= b: L(a) ^ L(b) with L(s) the sites below s.
template <int W>
__device__ __forceinline__ Cfg<W> range_mask(int a, int b) {
    Cfg<W> m;
    if (W == 1) {
        m.w[0] = (1ull << b) - (1ull << a);  // sites of a one-word config are < 64
    } else {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int lo = a - 64 * w, hi = b - 64 * w;
            const uint64_t mh = hi <= 0 ? 0ull : (hi >= 64 ? ~0ull : (1ull << hi) - 1ull);
            const uint64_t ml = lo <= 0 ? 0ull : (lo >= 64 ? ~0ull : (1ull << lo) - 1ull);
            m.w[w] = mh & ~ml;
        }
    }
    return m;
}

template <int W>
__device__ __forceinline__ int masked_parity(const Cfg<W>& c, const Cfg<W>& m) {
    int par = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        par ^= __popcll(c.w[w] & m.w[w]);
    }
    return par & 1;
}

// pair_parity(c) = popcount(c & (L(a) ^ L(b) ^ L(x) ^ L(y))) mod 2 for a < b, x < y.
template <int W>
__device__ __forceinline__ int pair_parity(const Cfg<W>& c, int a, int b, int x, int y) {
    const Cfg<W> m1 = range_mask<W>(a, b);
    const Cfg<W> m2 = range_mask<W>(x, y);
    int par = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        par ^= __popcll(c.w[w] & (m1.w[w] ^ m2.w[w]));
    }
    return par & 1;
}

// One staged group entry (b: a config of the group, removed sites q0 < q1), shared by the wave
// through LDS: what every row a of the group needs of it, so the per-pair work is a few mask
// operations, one coefficient load and the multiply-adds.
template <int W>
struct PairStage {
    double2 g;          // in[row]
    uint64_t bits[W];   // the sites q0 and q1 (all ones for an empty slot: it meets every row)
    uint32_t part;      // the entry's part of the coefficient index: pair_slot(q0, q1) (ADJ), times n_pairs (!ADJ)
    uint32_t ebit;      // parity of the group core's sites in [q0, q1)
    uint64_t m[W];      // sites in [q0, q1)
    uint32_t row;       // config row of the entry
    uint32_t pad;
    double2 g2;         // in2[row] of a two-vector pass
};
constexpr int kWavesPerBlock = kBlock / 64;
constexpr size_t kPairStageBytes = kWavesPerBlock * 64 * sizeof(PairStage<2>);
static_assert(sizeof(PairStage<1>) <= sizeof(PairStage<2>) && sizeof(PairStage<2>) % 16 == 0, "stage slot layout");

__device__ __forceinline__ double2 shfl_d2(const double2& v, int lane) {
    return make_double2(__shfl(v.x, lane), __shfl(v.y, lane));
}

// Coefficient gathers kept in flight per lane in pair_group: 4 in the one-vector pair kernel; 3 in
// the two-vector pair kernel and the batched kernel, whose accumulators and other paths hold more
// registers (4 there spills).
constexpr int kPairBatch = 4;
constexpr int kPairBatchInBatch = 3;

template <int W>
__device__ __forceinline__ Cfg<W> site_bits(int a, int b) {
    Cfg<W> m;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        m.w[w] = ((a >> 6) == w ? (1ull << (a & 63)) : 0ull) | ((b >> 6) == w ? (1ull << (b & 63)) : 0ull);
    }
    return m;
}

// Stages entry b (value vb) of a group whose first core is core0.  An inactive slot gets all-ones
// site bits, so every row sees it as overlapping and skips it.
template <int W, bool ADJ, int NV>
__device__ __forceinline__ PairStage<W> stage_entry(const PairParams& p, bool act, uint64_t vb, const Cfg<W>& core0, bool fermi) {
    PairStage<W> st;
    if (act) {
        const int q0 = static_cast<int>(vb & 0xff);
        const int q1 = static_cast<int>((vb >> 8) & 0xff);
        st.row = static_cast<uint32_t>(vb >> 16);
        st.g = p.in[st.row];
        st.g2 = NV == 2 ? p.in2[st.row] : make_double2(0.0, 0.0);
        const Cfg<W> mq = range_mask<W>(q0, q1);
        const Cfg<W> bq = site_bits<W>(q0, q1);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            st.m[w] = fermi ? mq.w[w] : 0ull;  // zero for hard-core bosons: no sign
            st.bits[w] = bq.w[w];
        }
        const int64_t slot = pair_slot(q0, q1, p.n_bits);
        st.part = static_cast<uint32_t>(ADJ ? slot : slot * p.n_pairs);
        st.ebit = fermi ? static_cast<uint32_t>(masked_parity<W>(core0, mq)) : 0u;
    } else {
        st.row = 0;
        st.g = make_double2(0.0, 0.0);
        st.g2 = make_double2(0.0, 0.0);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            st.m[w] = 0;
            st.bits[w] = ~0ull;
        }
        st.part = 0;
        st.ebit = 0;
    }
    st.pad = 0;
    return st;
}

// What row a (removed sites r0 < r1, core `core`) keeps while it walks the staged entries.
template <int W>
struct PairRow {
    Cfg<W> core;   // the row's config without r0, r1
    Cfg<W> bits;   // r0 and r1
    Cfg<W> mr;     // sites in [r0, r1)
    int64_t part;  // the row's part of the coefficient index
    int rbit;      // parity(core & [r0, r1)) ^ 1
};

// Contribution of staged entry st to row ra, or false when the pair is not a connection: the
// removed pairs overlap (which includes the row itself), or -- in a group where different cores
// share the hash (mixed) -- the entry's core differs.  Signs (fermions):
//   ADJ:  A_ba = (-1)^(rbit ^ ebit ^ parity(bits_a & [q0, q1))) T[Pa][Pb], applied conjugated
//   !ADJ: A_ab = (-1)^(rbit ^ ebit ^ parity(bits_b & [r0, r1))) T[Pb][Pa]
// with ebit taken from the row's own core in a mixed group.
template <int W, bool ADJ, bool REAL>
__device__ __forceinline__ bool pair_term(const PairParams& p, const PairRow<W>& ra, const PairStage<W>& st, bool mixed, bool en,
                                          double2& x) {
    uint64_t meet = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        meet |= ra.bits.w[w] & st.bits[w];
    }
    bool ok = en && meet == 0ull;
    int ebit = static_cast<int>(st.ebit);
    if (mixed) {
        if (ok) {
            const Cfg<W> cb = load_key<W>(p.ckeys, static_cast<int64_t>(st.row));
#pragma unroll
            for (int w = 0; w < W; ++w) {
                ok = ok && ((cb.w[w] ^ st.bits[w]) == ra.core.w[w]);
            }
        }
        Cfg<W> mq;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            mq.w[w] = st.m[w];
        }
        ebit = masked_parity<W>(ra.core, mq);
    }
    // only the lanes with a connection load a coefficient
    x = make_double2(0.0, 0.0);
    if (ok) {
        const int64_t idx = ra.part + static_cast<int64_t>(st.part);
        x = REAL ? make_double2(p.coef_re[idx], 0.0) : p.coef[idx];
    }
    // The sign's parity in bit 0: rbit ^ ebit plus the other pair's sites inside this pair's range
    // (the range masks, rbit and ebit are zero for hard-core bosons).  Adding par << 31 to the high
    // word flips the sign bit by bit 0 of par alone, without a branch.
    uint32_t par = static_cast<uint32_t>(ra.rbit ^ ebit);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t m = ADJ ? (ra.bits.w[w] & st.m[w]) : (st.bits[w] & ra.mr.w[w]);
        par += static_cast<uint32_t>(__popc(static_cast<uint32_t>(m))) + static_cast<uint32_t>(__popc(static_cast<uint32_t>(m >> 32)));
    }
    const int flip = static_cast<int>(par << 31);
    x.x = __hiloint2double(__double2hiint(x.x) + flip, __double2loint(x.x));
    if (!REAL) {
        x.y = __hiloint2double(__double2hiint(x.y) + flip, __double2loint(x.y));
    }
    return ok;
}

template <int W>
__device__ __forceinline__ PairRow<W> pair_row(const PairParams& p, const Cfg<W>& core, int r0, int r1, bool adj, bool fermi) {
    PairRow<W> ra;
    ra.core = core;
    ra.bits = site_bits<W>(r0, r1);
    ra.mr = range_mask<W>(r0, r1);
    const int64_t slot = pair_slot(r0, r1, p.n_bits);
    ra.part = adj ? slot * p.n_pairs : slot;
    ra.rbit = fermi ? (masked_parity<W>(ra.core, ra.mr) ^ 1) : 0;
    if (!fermi) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            ra.mr.w[w] = 0ull;  // no sign for hard-core bosons
        }
    }
    return ra;
}

// acc += x g (!ADJ) or conj(x) g (ADJ); a real table skips the imaginary part of x.
template <bool ADJ, bool REAL>
__device__ __forceinline__ void pair_madd(const double2& x, const double2& g, double& re, double& im) {
    if (REAL) {
        re = __fma_rn(x.x, g.x, re);
        im = __fma_rn(x.x, g.y, im);
    } else if (ADJ) {
        re = __fma_rn(x.y, g.y, __fma_rn(x.x, g.x, re));
        im = __fma_rn(-x.y, g.x, __fma_rn(x.x, g.y, im));
    } else {
        re = __fma_rn(-x.y, g.y, __fma_rn(x.x, g.x, re));
        im = __fma_rn(x.y, g.x, __fma_rn(x.x, g.y, im));
    }
}

// One wave, one group of entries [s, e) (wave-uniform).  Lane = row a; the group's entries b are
// staged 64 at a time in this wave's LDS slots and read back as broadcasts.
//   ADJ:  out[a] += sum_b conj(A_ba) in[b]
//   !ADJ: out[a] += sum_b A_ab in[b]
// !ADJ reads T[Pb][Pa]: one table row per entry, at the columns of the lanes' rows.
// s_raw: the group's start with its kGroupMixed bit.  In a group of one core every row's core is
// the first entry's, so only that entry's config is read; a mixed group (different cores with
// equal keys) reads every row's own config and checks every pair exactly.
template <int W, bool ADJ, int B, int NV, bool REAL>
__device__ void pair_group(const PairParams& p, uint32_t s_raw, uint32_t e, int lane, PairStage<W>* stage) {
    const uint32_t s = s_raw & kGroupStart;
    const bool mixed = (s_raw & kGroupMixed) != 0u;
    const uint64_t v0 = p.vals[s];
    const Cfg<W> core0 = entry_core<W, 2>(p.ckeys, v0);
    const bool fermi = p.particle_cut == 1;
    for (uint32_t a0 = s; a0 < e; a0 += 64) {
        const uint32_t a = a0 + lane;
        const bool act = a < e;
        const uint64_t va = act ? p.vals[a] : v0;
        const PairRow<W> ra = pair_row<W>(p, mixed ? entry_core<W, 2>(p.ckeys, va) : core0, static_cast<int>(va & 0xff),
                                          static_cast<int>((va >> 8) & 0xff), ADJ, fermi);
        const int64_t ia = static_cast<int64_t>(va >> 16);
        double acc_re[NV] = {}, acc_im[NV] = {};
        for (uint32_t b0 = s; b0 < e; b0 += 64) {
            const uint32_t bl = b0 + lane;
            stage[lane] = stage_entry<W, ADJ, NV>(p, bl < e, bl < e ? p.vals[bl] : 0ull, core0, fermi);
            __builtin_amdgcn_wave_barrier();
            const int nb = static_cast<int>(e - b0 < 64u ? e - b0 : 64u);
            for (int u = 0; u < nb; u += B) {
                double2 x[B], g[B], g2[B];
                bool ok[B];
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    const PairStage<W>& st = stage[(u + k) & 63];
                    g[k] = st.g;
                    g2[k] = NV == 2 ? st.g2 : make_double2(0.0, 0.0);
                    // slots past nb hold entries of the previous chunk or nothing: bounded by u + k < nb
                    ok[k] = pair_term<W, ADJ, REAL>(p, ra, st, mixed, act && (u + k < nb), x[k]);
                }
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    // x[k] is zero where there is no connection, so the multiply-add needs no select
                    // (a non-finite input amplitude then also reaches the rows of its group it is
                    // not connected to)
                    (void)ok[k];
                    pair_madd<ADJ, REAL>(x[k], g[k], acc_re[0], acc_im[0]);
                    if (NV == 2) {
                        pair_madd<ADJ, REAL>(x[k], g2[k], acc_re[NV - 1], acc_im[NV - 1]);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

// A pass over groups of 9..32 entries, lane = row: whole groups are taken from `mask` in lane order
// while their rows fit the 64 lanes, a group's rows consecutive from `lead`.  nmax: the largest
// group of the pass.  Lanes past the pass's groups get s = e = 0.
struct RowLane {
    uint32_t s, e;  // the lane's group (s with its kGroupMixed bit)
    int lead;       // lane of the group's first row
    int nmax;
};

__device__ __forceinline__ RowLane row_pass(uint64_t& mask, uint32_t gs_l, uint32_t ge_l, int lane) {
    RowLane f{0u, 0u, 0, 0};
    int used = 0;
    while (mask) {
        const int l = __builtin_ctzll(mask);
        const uint32_t sj = __builtin_amdgcn_readlane(gs_l, l);
        const uint32_t ej = __builtin_amdgcn_readlane(ge_l, l);
        const int nj = static_cast<int>(ej - (sj & kGroupStart));
        if (used + nj > 64) {
            break;
        }
        mask &= mask - 1;
        if (lane >= used && lane < used + nj) {
            f.s = sj;
            f.e = ej;
            f.lead = used;
        }
        used += nj;
        f.nmax = nj > f.nmax ? nj : f.nmax;
    }
    return f;
}

// Groups of 9..32 entries laid out by row_pass (a 9-entry group takes 9 lanes, so 7 of them share a
// pass).  A group's entries are its rows, so each lane stages its own entry in its LDS slot and the
// rows of a group walk the slots of that group, as in pair_group.
template <int W, bool ADJ, int B, int NV, bool REAL>
__device__ void pair_group_rows(const PairParams& p, uint64_t mask, uint32_t gs_l, uint32_t ge_l, int lane, PairStage<W>* stage) {
    const bool fermi = p.particle_cut == 1;
    while (mask) {
        const RowLane f = row_pass(mask, gs_l, ge_l, lane);
        const bool mixed = (f.s & kGroupMixed) != 0u;
        const uint32_t s = f.s & kGroupStart;
        const int g = static_cast<int>(f.e - s);  // 0 for a lane without a group
        const int ri = lane - f.lead;
        const bool act = ri < g;
        const uint64_t va = act ? p.vals[s + static_cast<uint32_t>(ri)] : 0ull;
        const int64_t ia = static_cast<int64_t>(va >> 16);
        const int r0 = static_cast<int>(va & 0xff);
        const int r1 = static_cast<int>((va >> 8) & 0xff);
        // the group's first row reads its config for the group (a mixed group: every row its own)
        Cfg<W> own;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            own.w[w] = act && (mixed || ri == 0) ? p.ckeys[ia * W + w] : 0ull;
        }
        if (act) {
            flip_site<W>(own, r0);
            flip_site<W>(own, r1);
        }
        // the group's first core, from the lane holding its first row (every lane runs every shuffle)
        Cfg<W> core0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            core0.w[w] = shfl_u64(own.w[w], f.lead);
        }
        const PairRow<W> ra = pair_row<W>(p, mixed ? own : core0, r0, r1, ADJ, fermi);
        // the exact per-pair check runs wave-wide when any group of the pass is mixed: a row of a
        // group of one core holds its true core (core0) and passes it for its connections
        const bool wmixed = __ballot(mixed) != 0ull;
        // this lane's entry (its own row), staged for the rows of its group
        stage[lane] = stage_entry<W, ADJ, NV>(p, act, va, core0, fermi);
        double acc_re[NV] = {}, acc_im[NV] = {};
        __builtin_amdgcn_wave_barrier();
        for (int u = 0; u < f.nmax; u += B) {
            double2 x[B], gv[B], gv2[B];
            bool ok[B];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const int uu = u + k;
                // slots past the group belong to other groups or to no group: bounded by uu < g
                const PairStage<W>& st = stage[(f.lead + uu) & 63];
                gv[k] = st.g;
                gv2[k] = NV == 2 ? st.g2 : make_double2(0.0, 0.0);
                ok[k] = pair_term<W, ADJ, REAL>(p, ra, st, wmixed, act && uu < g, x[k]);
            }
#pragma unroll
            for (int k = 0; k < B; ++k) {
                (void)ok[k];  // x[k] is zero where there is no connection
                pair_madd<ADJ, REAL>(x[k], gv[k], acc_re[0], acc_im[0]);
                if (NV == 2) {
                    pair_madd<ADJ, REAL>(x[k], gv2[k], acc_re[NV - 1], acc_im[NV - 1]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

// A pass over groups of 2..8 entries with their ordered (row, entry) pairs side by side: whole
// groups are taken from `mask` in lane order while their pairs fit the 64 lanes.  A group of g
// entries takes g (g - 1) lanes from `lead`, row a's g - 1 pairs consecutive (entry b = j, or j + 1
// from j = a on), so no lane pairs a row with itself; a 3-entry group takes 6 lanes instead of a
// 4 x 4 block.  Lanes past the pass's groups get s = e = 0.
struct FlatLane {
    uint32_t s, e;  // the lane's group (s with its kGroupMixed bit)
    int ai, bi;     // row and entry in the group
    int j, d;       // position in row ai's run of d = g - 1 lanes
    int lead;       // lane of the group's first pair (row 0, j = 0)
    int rowb;       // lane of row bi's first pair
};

__device__ __forceinline__ FlatLane flat_pass(uint64_t& mask, uint32_t gs_l, uint32_t ge_l, int lane) {
    FlatLane f{0u, 0u, 0, 0, 0, 1, 0, 0};
    int used = 0;
    while (mask) {
        const int l = __builtin_ctzll(mask);
        const uint32_t sj = __builtin_amdgcn_readlane(gs_l, l);
        const uint32_t ej = __builtin_amdgcn_readlane(ge_l, l);
        const int nj = static_cast<int>(ej - (sj & kGroupStart));
        const int len = nj * (nj - 1);
        if (used + len > 64) {
            break;
        }
        mask &= mask - 1;
        if (lane >= used && lane < used + len) {
            f.s = sj;
            f.e = ej;
            f.lead = used;
        }
        used += len;
    }
    const int g = static_cast<int>(f.e - (f.s & kGroupStart));
    f.d = g > 1 ? g - 1 : 1;
    const uint32_t k = static_cast<uint32_t>(lane - f.lead);
    f.ai = static_cast<int>(k / static_cast<uint32_t>(f.d));
    f.j = static_cast<int>(k) - f.ai * f.d;
    f.bi = f.j + (f.j >= f.ai ? 1 : 0);
    f.rowb = f.lead + f.bi * f.d;
    return f;
}

// Sums each row's run of d lanes into its first lane (j = 0); d <= 7.
__device__ __forceinline__ void flat_row_sum(double& re, double& im, int j, int d) {
    for (int m = 1; m < 8; m <<= 1) {
        const double ore = __shfl_down(re, m);
        const double oim = __shfl_down(im, m);
        if (j + m < d) {
            re += ore;
            im += oim;
        }
    }
}

// Groups of 2..8 entries laid out by flat_pass: one lane per ordered (row a, entry b) pair.  Each
// lane checks its pair exactly (cores compared through a shuffle from the lane holding row b), the
// rows are summed over their runs, and the first lane of each row adds the sum.  `mask` holds the
// lanes of the 64-group chunk whose group has 2..8 entries; (gs_l, ge_l) is that lane's group.
template <int W, bool ADJ, int NV>
__device__ void pair_flat_class(const PairParams& p, uint64_t mask, uint32_t gs_l, uint32_t ge_l, int lane) {
    const bool fermi = p.particle_cut == 1;
    while (mask) {
        const FlatLane f = flat_pass(mask, gs_l, ge_l, lane);
        const bool mixed = (f.s & kGroupMixed) != 0u;
        const uint32_t s = f.s & kGroupStart;
        const bool act = f.e > s;
        const uint64_t va = act ? p.vals[s + f.ai] : 0ull;
        const uint64_t vb = act ? p.vals[s + f.bi] : 0ull;
        const int64_t ia = static_cast<int64_t>(va >> 16);
        const int r0 = static_cast<int>(va & 0xff), r1 = static_cast<int>((va >> 8) & 0xff);
        const int q0 = static_cast<int>(vb & 0xff), q1 = static_cast<int>((vb >> 8) & 0xff);
        // the coefficient and the input row need only the entries: loaded beside the stored configs
        // and dropped below when the cores differ (entry 0 of the table and row 0 of the input are
        // always there)
        const bool pre = act && q0 != r0 && q0 != r1 && q1 != r0 && q1 != r1;
        const int64_t idx = !pre ? 0
                                 : (ADJ ? pair_slot(r0, r1, p.n_bits) * p.n_pairs + pair_slot(q0, q1, p.n_bits)
                                        : pair_slot(q0, q1, p.n_bits) * p.n_pairs + pair_slot(r0, r1, p.n_bits));
        const double2 x = p.coef_re ? make_double2(p.coef_re[idx], 0.0) : p.coef[idx];
        const int64_t ib = pre ? static_cast<int64_t>(vb >> 16) : 0;
        double2 gv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gv[v] = (v == 0 ? p.in : p.in2)[ib];
        }
        // the lanes of the group's first row read its config for the group (a mixed group: every
        // row its own)
        Cfg<W> own;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            own.w[w] = act && (mixed || f.ai == 0) ? p.ckeys[ia * W + w] : 0ull;
        }
        if (act) {
            flip_site<W>(own, r0);
            flip_site<W>(own, r1);
        }
        // every lane runs every shuffle (a short-circuit would leave lanes out)
        Cfg<W> core;
        bool same = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t c0 = shfl_u64(own.w[w], f.lead);
            core.w[w] = mixed ? own.w[w] : c0;
            const bool eq = shfl_u64(own.w[w], f.rowb) == own.w[w];
            same = same & eq;
        }
        const bool ok = pre && (same || !mixed);
        double acc_re[NV] = {}, acc_im[NV] = {};
        if (ok) {
            int sg;
            if (ADJ) {
                Cfg<W> ca = core;
                flip_site<W>(ca, r0);
                flip_site<W>(ca, r1);
                sg = fermi ? pair_parity<W>(ca, r0, r1, q0, q1) : 0;
            } else {
                Cfg<W> cb = core;
                flip_site<W>(cb, q0);
                flip_site<W>(cb, q1);
                sg = fermi ? pair_parity<W>(cb, r0, r1, q0, q1) : 0;
            }
            const double xr = sg ? -x.x : x.x;
            const double xi = sg ? -x.y : x.y;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (ADJ) {
                    acc_re[v] = xr * gv[v].x + xi * gv[v].y;
                    acc_im[v] = xr * gv[v].y - xi * gv[v].x;
                } else {
                    acc_re[v] = xr * gv[v].x - xi * gv[v].y;
                    acc_im[v] = xr * gv[v].y + xi * gv[v].x;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            flat_row_sum(acc_re[v], acc_im[v], f.j, f.d);
            if (f.j == 0 && act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

// The groups above 8 entries of a 64-group chunk: packed several to a wave up to 32 entries (PACK),
// then one wave per group.
template <int W, bool ADJ, int B, int NV, bool PACK, bool REAL>
__device__ __forceinline__ void pair_chunk_groups(const PairParams& p, uint32_t gs, uint32_t ge, uint32_t sz, int lane,
                                                  PairStage<W>* stage) {
    if (PACK) {
        pair_group_rows<W, ADJ, B, NV, REAL>(p, __ballot(sz > 8u && sz <= 32u), gs, ge, lane, stage);
    }
    uint64_t big = __ballot(sz > (PACK ? 32u : 8u));
    while (big) {
        const int l = __builtin_ctzll(big);
        big &= big - 1;
        pair_group<W, ADJ, B, NV, REAL>(p, __shfl(gs, l), __shfl(ge, l), lane, stage);
    }
}

// The pair work of the groups [g0, g0 + 64): groups of up to 8 entries with their ordered pairs
// side by side (pair_flat_class), larger ones packed or a row block at a time (pair_group).
template <int W, bool ADJ, int B, int NV, bool PACK = true>
__device__ __forceinline__ void pair_chunk(const PairParams& p, int64_t groups, int64_t g0, int lane, char* stage_bytes) {
    const int64_t gg = g0 + lane;
    uint32_t gs = 0, ge = 0;  // gs keeps the group's kGroupMixed bit for the group functions
    if (gg < groups) {
        gs = p.starts[gg];
        ge = p.starts[gg + 1] & kGroupStart;
    }
    const uint32_t sz = ge - (gs & kGroupStart);
    pair_flat_class<W, ADJ, NV>(p, __ballot(sz >= 2u && sz <= 8u), gs, ge, lane);
    // each wave owns 64 slots of the largest stage size: waves of one block may run different W
    PairStage<W>* stage = reinterpret_cast<PairStage<W>*>(stage_bytes + (threadIdx.x >> 6) * 64 * sizeof(PairStage<2>));
    // PACK: groups of 9..32 entries several to a wave (else one group per wave)
    // a real table (wave-uniform) takes the loops specialised for 8-byte coefficients
    if (p.coef_re != nullptr) {
        pair_chunk_groups<W, ADJ, B, NV, PACK, true>(p, gs, ge, sz, lane, stage);
    } else {
        pair_chunk_groups<W, ADJ, B, NV, PACK, false>(p, gs, ge, sz, lane, stage);
    }
}

__device__ __forceinline__ void mark_overflow(double2* out) {
    // a row had more occupied sites than the core index was built for
    out[0] = make_double2(__longlong_as_double(0x7ff8000000000000ll), __longlong_as_double(0x7ff8000000000000ll));
}

// Persistent waves over the groups, 64 groups per step.  Every wave leaves when the group count
// is done.
template <int W, bool ADJ, int NV = 1>
__global__ __launch_bounds__(kBlock) void pair_kernel(PairParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    if (p.counts[2] != 0) {
        if (wave == 0 && lane == 0) {
            mark_overflow(p.out);
        }
        return;
    }
    __shared__ __align__(16) char stage[kPairStageBytes];
    const int64_t groups = p.counts[1];
    for (int64_t g0 = wave * 64; g0 < groups; g0 += nwaves * 64) {
        // two vectors hold twice the accumulators: fewer coefficient gathers in flight per lane
        pair_chunk<W, ADJ, NV == 1 ? kPairBatch : kPairBatchInBatch, NV>(p, groups, g0, lane, stage);
    }
}

// ------------------------------------------------------------------------- single excitations --
// Launched with whole waves: threads past n_terms take part in the counts only.
__global__ void singles_fill_kernel(const int16_t* __restrict__ site, const uint8_t* __restrict__ kind,
                                    const double2* __restrict__ coef, int64_t n_terms, int particle_cut, int n_bits,
                                    double2* __restrict__ a, double2* __restrict__ kd, int64_t* __restrict__ counts) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    bool is_single = false, is_dressed = false;
    if (t < n_terms) {
        const CompiledTerm c = compile_term(site + 4 * t, kind + 4 * t, particle_cut, n_bits);
        SingleExcitation e;
        if (c.status == kCompileOk && single_excitation(c.rec, particle_cut, &e)) {
            const double2 v0 = coef[t];
            const double2 v = e.delta ? make_double2(-v0.x, -v0.y) : v0;
            const int64_t rp = static_cast<int64_t>(e.r) * n_bits + e.p;
            if (e.k < 0 || !e.kocc) {
                // holds whenever k is empty
                unsafeAtomicAdd(&a[rp].x, v.x);
                unsafeAtomicAdd(&a[rp].y, v.y);
            }
            if (e.k >= 0) {
                double2* dk = kd + rp * n_bits + e.k;
                unsafeAtomicAdd(&dk->x, e.kocc ? v.x : -v.x);
                unsafeAtomicAdd(&dk->y, e.kocc ? v.y : -v.y);
                is_dressed = true;
            }
            is_single = true;
        }
    }
    wave_count(&counts[0], is_single);
    wave_count(&counts[1], is_dressed);
}

int singles_layout(Carver& cv, int n_qubytes, qmb_singles* out) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n_qubytes > kMaxPairBytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "the single-excitation table supports n_qubytes <= %d, got %d", kMaxPairBytes,
                         n_qubytes);
    }
    qmb_singles g{};
    g.n_qubytes = n_qubytes;
    g.n_bits = 8 * n_qubytes;
    const size_t nb = static_cast<size_t>(g.n_bits);
    g.buffer = cv.base;
    g.counts = cv.take<int64_t>(4);
    g.a = cv.take<double>(2 * nb * nb);
    g.kd = cv.take<double>(2 * nb * nb * nb);
    g.a_re = cv.take<double>(nb * nb);
    g.kd_re = cv.take<double>(nb * nb * nb);
    g.bytes = cv.off;
    if (out) {
        *out = g;
    }
    return QMB_OK;
}

struct SingleParams {
    const double2* a;
    const double2* kd;
    const double* a_re;   // their real parts when every coefficient is real, else null
    const double* kd_re;
    int n_bits;
    int particle_cut;
    const uint64_t* ckeys;
    const uint64_t* vals;
    const uint32_t* starts;
    const int64_t* counts;
    const double2* in;
    double2* out;
    const double2* in2;  // second input/output of a two-vector pass (NV = 2), else null
    double2* out2;
};

template <int W>
__device__ __forceinline__ int single_parity(const Cfg<W>& c, int r, int p) {
    return masked_parity<W>(c, range_mask<W>(r < p ? r : p, r < p ? p : r));
}

// Sum of row[k] over the set bits k of c, ascending k.  Up to CH indices are taken from the mask
// before their loads are issued, so the loads of one element overlap instead of each waiting for the
// last add (same order of additions, same result as a bit-by-bit walk).  The single-set kernels use
// CH = kSingleLoads (13: the 13 core sites of a 14-electron config in one batch; single kernel 1.68
// -> 1.40 ms at 52 spin orbitals and 7a+7b against 8, while 16 ran 1.67 ms); the batch kernel, held
// to 128 VGPRs, uses 4 (measured on 64 molecular instances: 126.8 ms
// per step against 128.8 ms with 2 and 130.7 ms with the plain walk).  row must hold at least
// 64 * (W - 1) + 1 entries (a lane without an index reads row[64 w] and drops it).
constexpr int kSingleLoads = 13;
__device__ __forceinline__ double zero_like(double) { return 0.0; }
__device__ __forceinline__ double2 zero_like(double2) { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double add_to(double a, double b) { return a + b; }
__device__ __forceinline__ double2 add_to(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

template <int W, typename T, int CH>
__device__ __forceinline__ T masked_row_sum(const T* __restrict__ row, const Cfg<W>& c, T acc) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t m = c.w[w];
        while (m) {
            T v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const bool have = m != 0ull;
                const int k = 64 * w + (have ? __builtin_ctzll(m) : 0);
                m = have ? m & (m - 1ull) : 0ull;
                const T x = row[k];
                v[u] = have ? x : zero_like(x);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                acc = add_to(acc, v[u]);
            }
        }
    }
    return acc;
}

// Element of the single excitation (annihilate r, create p) on a source config whose other
// occupied sites are `core`.
template <int W, int CH>
__device__ __forceinline__ double2 single_element(const SingleParams& p, const Cfg<W>& core, int r, int q) {
    const int64_t rq = static_cast<int64_t>(r) * p.n_bits + q;
    if (p.kd_re) {
        // real table: 8-byte reads, the sum kept in one register
        return make_double2(masked_row_sum<W, double, CH>(p.kd_re + rq * p.n_bits, core, p.a_re[rq]), 0.0);
    }
    return masked_row_sum<W, double2, CH>(p.kd + rq * p.n_bits, core, p.a[rq]);
}

// One wave, one group of order-1 entries [s, e): rows a (config = core + r) against entries b
// (config = core + q), q != r:
//   ADJ:  out[a] += conj(M_ba) in[b],  M_ba = (-1)^single_parity(c_a; r, q) (A[r][q] + sum_{k in core} K[r][q][k])
//   !ADJ: out[a] += M_ab in[b],        M_ab = (-1)^single_parity(c_b; q, r) (A[q][r] + sum_{k in core} K[q][r][k])
// s_raw: the group's start with its kGroupMixed bit (see pair_group).
template <int W, bool ADJ, int NV, int CH>
__device__ void single_group(const SingleParams& p, uint32_t s_raw, uint32_t e, int lane) {
    const uint32_t s = s_raw & kGroupStart;
    const bool mixed = (s_raw & kGroupMixed) != 0u;
    const uint64_t v0 = p.vals[s];
    const Cfg<W> core0 = entry_core<W, 1>(p.ckeys, v0);
    const bool fermi = p.particle_cut == 1;
    for (uint32_t a0 = s; a0 < e; a0 += 64) {
        const uint32_t a = a0 + lane;
        const bool act = a < e;
        const uint64_t va = act ? p.vals[a] : v0;
        const int64_t ia = static_cast<int64_t>(va >> 16);
        const int r = static_cast<int>(va & 0xff);
        const Cfg<W> core = mixed ? entry_core<W, 1>(p.ckeys, va) : core0;
        Cfg<W> ca = core;
        flip_site<W>(ca, r);
        double acc_re[NV] = {}, acc_im[NV] = {};
        for (uint32_t b0 = s; b0 < e; b0 += 64) {
            const uint32_t bl = b0 + lane;
            const bool bin = bl < e;
            const uint64_t vbl = bin ? p.vals[bl] : 0ull;
            const double2 gbl = bin ? p.in[vbl >> 16] : make_double2(0.0, 0.0);
            const double2 gbl2 = (NV == 2 && bin) ? p.in2[vbl >> 16] : make_double2(0.0, 0.0);
            const int nb = static_cast<int>(e - b0 < 64u ? e - b0 : 64u);
            for (int u = 0; u < nb; ++u) {
                const uint64_t vb = shfl_u64(vbl, u);
                const double2 g = shfl_d2(gbl, u);
                const double2 g2 = NV == 2 ? shfl_d2(gbl2, u) : make_double2(0.0, 0.0);
                const int q = static_cast<int>(vb & 0xff);
                bool ok = act && (b0 + static_cast<uint32_t>(u) != a) && q != r;
                if (mixed && ok) {
                    Cfg<W> cb = load_key<W>(p.ckeys, static_cast<int64_t>(vb >> 16));
                    flip_site<W>(cb, q);
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        ok = ok && cb.w[w] == core.w[w];
                    }
                }
                if (!ok) {
                    continue;
                }
                double2 x;
                int sg;
                if (ADJ) {
                    x = single_element<W, CH>(p, core, r, q);
                    sg = fermi ? single_parity<W>(ca, r, q) : 0;
                } else {
                    Cfg<W> cb = core;
                    flip_site<W>(cb, q);
                    x = single_element<W, CH>(p, core, q, r);
                    sg = fermi ? single_parity<W>(cb, q, r) : 0;
                }
                const double xr = sg ? -x.x : x.x;
                const double xi = sg ? -x.y : x.y;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const double2 gg = v == 0 ? g : g2;
                    if (ADJ) {
                        acc_re[v] += xr * gg.x + xi * gg.y;
                        acc_im[v] += xr * gg.y - xi * gg.x;
                    } else {
                        acc_re[v] += xr * gg.x - xi * gg.y;
                        acc_im[v] += xr * gg.y + xi * gg.x;
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

// Single-excitation groups of 9..32 entries laid out by row_pass; a group's entries are its rows,
// read from the lanes of the group by shuffles (as single_group reads its staged entries), so
// several groups share a wave.
template <int W, bool ADJ, int NV, int CH>
__device__ void single_group_rows(const SingleParams& p, uint64_t mask, uint32_t gs_l, uint32_t ge_l, int lane) {
    const bool fermi = p.particle_cut == 1;
    while (mask) {
        const RowLane f = row_pass(mask, gs_l, ge_l, lane);
        const bool mixed = (f.s & kGroupMixed) != 0u;
        const uint32_t s = f.s & kGroupStart;
        const int g = static_cast<int>(f.e - s);  // 0 for a lane without a group
        const int ri = lane - f.lead;
        const bool act = ri < g;
        const uint64_t va = act ? p.vals[s + static_cast<uint32_t>(ri)] : 0ull;
        const int64_t ia = static_cast<int64_t>(va >> 16);
        const int r = static_cast<int>(va & 0xff);
        // the group's first row reads its config for the group (a mixed group: every row its own)
        Cfg<W> own;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            own.w[w] = act && (mixed || ri == 0) ? p.ckeys[ia * W + w] : 0ull;
        }
        if (act) {
            flip_site<W>(own, r);
        }
        Cfg<W> core;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t c0 = shfl_u64(own.w[w], f.lead);  // every lane runs every shuffle
            core.w[w] = mixed ? own.w[w] : c0;
        }
        // the exact per-pair check runs wave-wide when any group of the pass is mixed (a row of a
        // group of one core holds its true core and passes it)
        const bool wmixed = __ballot(mixed) != 0ull;
        Cfg<W> ca = core;
        flip_site<W>(ca, r);
        const double2 gin = act ? p.in[ia] : make_double2(0.0, 0.0);
        const double2 gin2 = (NV == 2 && act) ? p.in2[ia] : make_double2(0.0, 0.0);
        double acc_re[NV] = {}, acc_im[NV] = {};
        for (int u = 0; u < f.nmax; ++u) {
            // lanes past the group belong to other groups or to no group: bounded by u < g
            const int src = (f.lead + u) & 63;
            const uint64_t vb = shfl_u64(va, src);
            const double2 gb = shfl_d2(gin, src);
            const double2 gb2 = NV == 2 ? shfl_d2(gin2, src) : make_double2(0.0, 0.0);
            const int q = static_cast<int>(vb & 0xff);
            bool ok = act && u < g && u != ri && q != r;
            if (wmixed && ok) {
                Cfg<W> cb = load_key<W>(p.ckeys, static_cast<int64_t>(vb >> 16));
                flip_site<W>(cb, q);
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    ok = ok && cb.w[w] == core.w[w];
                }
            }
            if (!ok) {
                continue;
            }
            double2 x;
            int sg;
            if (ADJ) {
                x = single_element<W, CH>(p, core, r, q);
                sg = fermi ? single_parity<W>(ca, r, q) : 0;
            } else {
                Cfg<W> cb = core;
                flip_site<W>(cb, q);
                x = single_element<W, CH>(p, core, q, r);
                sg = fermi ? single_parity<W>(cb, q, r) : 0;
            }
            const double xr = sg ? -x.x : x.x;
            const double xi = sg ? -x.y : x.y;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double2 gg = v == 0 ? gb : gb2;
                if (ADJ) {
                    acc_re[v] += xr * gg.x + xi * gg.y;
                    acc_im[v] += xr * gg.y - xi * gg.x;
                } else {
                    acc_re[v] += xr * gg.x - xi * gg.y;
                    acc_im[v] += xr * gg.y + xi * gg.x;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

// Single-excitation groups of 2..8 entries laid out by flat_pass, as pair_flat_class.
template <int W, bool ADJ, int NV, int CH>
__device__ void single_flat_class(const SingleParams& p, uint64_t mask, uint32_t gs_l, uint32_t ge_l, int lane) {
    const bool fermi = p.particle_cut == 1;
    while (mask) {
        const FlatLane f = flat_pass(mask, gs_l, ge_l, lane);
        const bool mixed = (f.s & kGroupMixed) != 0u;
        const uint32_t s = f.s & kGroupStart;
        const bool act = f.e > s;
        const uint64_t va = act ? p.vals[s + f.ai] : 0ull;
        const uint64_t vb = act ? p.vals[s + f.bi] : 0ull;
        const int64_t ia = static_cast<int64_t>(va >> 16);
        const int r = static_cast<int>(va & 0xff);
        const int q = static_cast<int>(vb & 0xff);
        // the input row needs only the entry: loaded beside the stored configs (row 0 when unused)
        const bool pre = act && q != r;
        const int64_t ib = pre ? static_cast<int64_t>(vb >> 16) : 0;
        double2 gv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gv[v] = (v == 0 ? p.in : p.in2)[ib];
        }
        // the lanes of the group's first row read its config for the group (a mixed group: every
        // row its own)
        Cfg<W> own;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            own.w[w] = act && (mixed || f.ai == 0) ? p.ckeys[ia * W + w] : 0ull;
        }
        if (act) {
            flip_site<W>(own, r);
        }
        // every lane runs every shuffle (a short-circuit would leave lanes out)
        Cfg<W> core;
        bool same = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t c0 = shfl_u64(own.w[w], f.lead);
            core.w[w] = mixed ? own.w[w] : c0;
            const bool eq = shfl_u64(own.w[w], f.rowb) == own.w[w];
            same = same & eq;
        }
        const bool ok = pre && (same || !mixed);
        double acc_re[NV] = {}, acc_im[NV] = {};
        if (ok) {
            double2 x;
            int sg;
            if (ADJ) {
                Cfg<W> ca = core;
                flip_site<W>(ca, r);
                x = single_element<W, CH>(p, core, r, q);
                sg = fermi ? single_parity<W>(ca, r, q) : 0;
            } else {
                Cfg<W> cb = core;
                flip_site<W>(cb, q);
                x = single_element<W, CH>(p, core, q, r);
                sg = fermi ? single_parity<W>(cb, q, r) : 0;
            }
            const double xr = sg ? -x.x : x.x;
            const double xi = sg ? -x.y : x.y;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (ADJ) {
                    acc_re[v] = xr * gv[v].x + xi * gv[v].y;
                    acc_im[v] = xr * gv[v].y - xi * gv[v].x;
                } else {
                    acc_re[v] = xr * gv[v].x - xi * gv[v].y;
                    acc_im[v] = xr * gv[v].y + xi * gv[v].x;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            flat_row_sum(acc_re[v], acc_im[v], f.j, f.d);
            if (f.j == 0 && act && (acc_re[v] != 0.0 || acc_im[v] != 0.0)) {
                double2* o = v == 0 ? p.out : p.out2;
                unsafeAtomicAdd(&o[ia].x, acc_re[v]);
                unsafeAtomicAdd(&o[ia].y, acc_im[v]);
            }
        }
    }
}

template <int W, bool ADJ, int NV, int CH, bool PACK = true>
__device__ __forceinline__ void single_chunk(const SingleParams& p, int64_t groups, int64_t g0, int lane) {
    const int64_t gg = g0 + lane;
    uint32_t gs = 0, ge = 0;  // gs keeps the group's kGroupMixed bit for the group functions
    if (gg < groups) {
        gs = p.starts[gg];
        ge = p.starts[gg + 1] & kGroupStart;
    }
    const uint32_t sz = ge - (gs & kGroupStart);
    single_flat_class<W, ADJ, NV, CH>(p, __ballot(sz >= 2u && sz <= 8u), gs, ge, lane);
    // PACK: groups of 9..32 entries several to a wave (else one group per wave)
    if (PACK) {
        single_group_rows<W, ADJ, NV, CH>(p, __ballot(sz > 8u && sz <= 32u), gs, ge, lane);
    }
    uint64_t big = __ballot(sz > (PACK ? 32u : 8u));
    while (big) {
        const int l = __builtin_ctzll(big);
        big &= big - 1;
        single_group<W, ADJ, NV, CH>(p, __shfl(gs, l), __shfl(ge, l), lane);
    }
}

template <int W, bool ADJ, int NV = 1>
__global__ __launch_bounds__(kBlock) void single_kernel(SingleParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    if (p.counts[2] != 0) {
        if (wave == 0 && lane == 0) {
            mark_overflow(p.out);
        }
        return;
    }
    const int64_t groups = p.counts[1];
    for (int64_t g0 = wave * 64; g0 < groups; g0 += nwaves * 64) {
        single_chunk<W, ADJ, NV, kSingleLoads>(p, groups, g0, lane);
    }
}

// ----------------------------------------------------------------------------------- hops --
// A single excitation between two sites of one parity (a spin-conserving hop when sites interleave
// spin: even = up, odd = down) leaves the string on the other parity's sites unchanged, so the two
// configs it connects hold equal strings there.  The strings index of parity s groups a set's rows
// by their string on the sites of parity 1 - s: one entry per row, sorted by a 32-bit key of that
// string (a hash of it mod 2^32 - 1).  Inside a group two rows are connected by such a hop iff their
// configs differ in exactly two sites, both of parity s, one occupied in each; rows of different
// strings merged by equal keys fail that test.  Building it sorts n keys, where the order-1 core
// index sorts n per occupied site.
constexpr uint64_t kEvenSites = 0x5555555555555555ull;

__host__ __device__ __forceinline__ uint64_t parity_sites(int parity) { return parity ? ~kEvenSites : kEvenSites; }

// The sites of one parity of a config word, packed into 32 bits (site 2j + parity -> bit j).
__device__ __forceinline__ uint64_t parity_bits(uint64_t x, int parity) {
    x = (parity ? x >> 1 : x) & kEvenSites;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return x;
}

// The key of a row: its other-parity string packed (bit 32 w + j = site 64 w + 2 j + 1 - parity),
// mixed by the 64-bit finaliser, mod 2^32 - 1 (the key no entry holds marks nothing here).
template <int W>
__global__ void strings_emit_kernel(const uint64_t* __restrict__ keys, int64_t n, int parity, uint32_t* __restrict__ okeys,
                                    uint32_t* __restrict__ orows) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const Cfg<W> c = load_key<W>(keys, i);
    uint64_t other = 0ull;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        other |= parity_bits(c.w[w], parity ^ 1) << (32 * w);
    }
    okeys[i] = static_cast<uint32_t>(fmix64(other ^ 0x2545f4914f6cdd1dull) % 0xFFFFFFFFull);
    orows[i] = static_cast<uint32_t>(i);
}

// group[e] holds the inclusive count of group starts up to entry e; turned into the entry's group.
__global__ void strings_start_kernel(const uint32_t* __restrict__ flag, uint32_t* __restrict__ group, int64_t n,
                                     uint32_t* __restrict__ starts, int64_t* __restrict__ counts) {
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= n) {
        return;
    }
    const uint32_t g = group[e] - 1u;
    group[e] = g;
    if (flag[e]) {
        starts[g] = static_cast<uint32_t>(e);
    }
    if (e == n - 1) {
        starts[g + 1] = static_cast<uint32_t>(n);
        counts[0] = n;
        counts[1] = static_cast<int64_t>(g) + 1;
    }
}

// counts[2] = sum of size^2 over the groups of two or more rows (the pairs the hop kernel tests),
// counts[3] = the largest group: a grid-stride pass, one atomic per block.
__global__ __launch_bounds__(kBlock) void strings_stats_kernel(const uint32_t* __restrict__ starts, int64_t* __restrict__ counts) {
    __shared__ unsigned long long s_sq[kBlock / 64], s_mx[kBlock / 64];
    const int64_t groups = counts[1];
    unsigned long long sq = 0ull, mx = 0ull;
    for (int64_t g = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; g < groups; g += static_cast<int64_t>(gridDim.x) * kBlock) {
        const unsigned long long sz = starts[g + 1] - starts[g];
        sq += sz >= 2ull ? sz * sz : 0ull;
        mx = sz > mx ? sz : mx;
    }
    for (int d = 32; d > 0; d >>= 1) {
        sq += __shfl_down(sq, d);
        const unsigned long long o = __shfl_down(mx, d);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        s_sq[threadIdx.x >> 6] = sq;
        s_mx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            sq += s_sq[w];
            mx = s_mx[w] > mx ? s_mx[w] : mx;
        }
        if (mx != 0ull) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&counts[2]), sq);
            atomicMax(reinterpret_cast<unsigned long long*>(&counts[3]), mx);
        }
    }
}

// mflag[e] = 1 when entry e's group holds two or more rows.
__global__ void strings_multi_kernel(const uint32_t* __restrict__ group, const uint32_t* __restrict__ starts, int64_t n,
                                     uint32_t* __restrict__ mflag) {
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= n) {
        return;
    }
    const uint32_t g = group[e];
    mflag[e] = starts[g + 1] - starts[g] >= 2u ? 1u : 0u;
}

// The entries of groups of two or more rows, in index order, with their config words and hop-parity
// string (site 2j + parity -> bit j) gathered and each entry's group as a range of the compacted
// order (mpos: exclusive scan of mflag); counts[4] = their number.
template <int W>
__global__ void strings_compact_kernel(const uint64_t* __restrict__ ckeys, const uint32_t* __restrict__ rows,
                                       const uint32_t* __restrict__ group, const uint32_t* __restrict__ starts,
                                       const uint32_t* __restrict__ mflag, const uint32_t* __restrict__ mpos, int64_t n, int parity,
                                       uint32_t* __restrict__ mrows, uint2* __restrict__ mrange, uint64_t* __restrict__ mcfg,
                                       uint64_t* __restrict__ mstr, int64_t* __restrict__ counts) {
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= n) {
        return;
    }
    if (e == n - 1) {
        counts[4] = static_cast<int64_t>(mpos[e]) + mflag[e];
    }
    if (!mflag[e]) {
        return;
    }
    const uint32_t p = mpos[e];
    const uint32_t g = group[e];
    const uint32_t lo = mpos[starts[g]];
    const uint32_t row = rows[e];
    mrows[p] = row;
    mrange[p] = make_uint2(lo, lo + (starts[g + 1] - starts[g]));
    uint64_t str = 0ull;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t c = ckeys[static_cast<int64_t>(row) * W + w];
        mcfg[static_cast<int64_t>(p) * W + w] = c;
        str |= parity_bits(c, parity) << (32 * w);
    }
    mstr[p] = str;
}

// A compacted group whose rows hold more than one other-parity string (strings merged by equal
// keys): mix[first entry] = 1.
template <int W>
__global__ void strings_mixed_kernel(const uint2* __restrict__ mrange, const uint64_t* __restrict__ mcfg, const int64_t* __restrict__ counts,
                                     int parity, uint32_t* __restrict__ mix) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p < 1 || p >= counts[4]) {
        return;
    }
    const uint32_t lo = mrange[p].x;
    if (lo == p) {
        return;
    }
    const uint64_t other = parity_sites(parity ^ 1);
    bool differ = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        differ = differ || ((mcfg[p * W + w] ^ mcfg[(p - 1) * W + w]) & other) != 0ull;
    }
    if (differ) {
        mix[lo] = 1u;
    }
}

// Marks the entries of merged groups: bit 31 of their range's end.
__global__ void strings_mark_kernel(const uint32_t* __restrict__ mix, const int64_t* __restrict__ counts, uint2* __restrict__ mrange) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= counts[4]) {
        return;
    }
    if (mix[mrange[p].x]) {
        mrange[p].y |= 0x80000000u;
    }
}

template <int W>
struct HopStage {
    uint64_t c[W];  // config of the entry's row
    double2 g;      // in[row]
};

struct HopParams {
    SingleParams s;          // the single-excitation tables, in, out and the set's config words
    const int64_t* scounts;  // the tables' counts: [1] = terms with a conditioned site
    uint64_t sites;          // the hop parity's sites, per word
    int parity;
    const uint32_t* mrows;   // the compacted entries (groups of two or more rows): row,
    const uint2* mrange;     // group as a range of the compacted order (bit 31 of the end: merged group),
    const uint64_t* mcfg;    // config words,
    const uint64_t* mstr;    // hop-parity string
    const int64_t* counts;   // [4] = number of compacted entries
};

// One wave per 64 consecutive compacted entries (rows of groups of two or more): lane = row a.  The
// wave walks the entries of the union of its rows' groups (consecutive in the compacted order),
// staged 64 at a time in LDS (amplitudes gathered) and read back as broadcasts, and tests each
// pair: in a wave of unmerged groups only the hop-parity strings of the rows of the lane's own
// group (the other parity's strings are equal there), else the whole configs.  The connections a
// chunk finds (a row may have many: the seed of a star of hops) are then listed in LDS and shared
// out over the lanes, one (row, entry) per lane per round, so their elements and coefficient loads
// run side by side; each row's sum collects in LDS and is added to out[a] without atomics (each row
// has one entry).
//   ADJ:  out[a] += conj(M_ba) in[b],  M_ba = (-1)^single_parity(c_a; r, q) (A[r][q] + sum_{k in core} K[r][q][k])
//   !ADJ: out[a] += M_ab in[b],        M_ab = (-1)^single_parity(c_b; q, r) (A[q][r] + sum_{k in core} K[q][r][k])
// with r the site of a, q the site of b (c_a = core + r, c_b = core + q), as in single_group.
constexpr int kHopList = 256;  // connections listed per round, per wave

template <int W>
struct HopWaveLds {
    HopStage<W> stage[64];
    uint64_t rowc[64][W];      // the lanes' row configs
    double2 acc[64];           // the lanes' row sums
    uint16_t list[kHopList];   // lane << 6 | entry slot
};

template <bool ADJ>
__device__ __forceinline__ void hop_madd(double xr, double xi, const double2& g, double& re, double& im) {
    if (ADJ) {
        re += xr * g.x + xi * g.y;
        im += xr * g.y - xi * g.x;
    } else {
        re += xr * g.x - xi * g.y;
        im += xr * g.y + xi * g.x;
    }
}

template <int W, bool ADJ>
__global__ __launch_bounds__(kBlock) void hop_kernel(HopParams p) {
    __shared__ HopWaveLds<W> lds_all[kBlock / 64];
    const int lane = threadIdx.x & 63;
    HopWaveLds<W>& L = lds_all[threadIdx.x >> 6];
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    const int64_t entries = p.counts[4];
    const bool fermi = p.s.particle_cut == 1;
    const bool plain = p.scounts[1] == 0;
    for (int64_t base = wave * 64; base < entries; base += nwaves * 64) {
        const int64_t e = base + lane;
        const bool act = e < entries;
        const uint2 range = act ? p.mrange[e] : make_uint2(0xFFFFFFFFu, 0u);
        const uint32_t gs = range.x, ge = range.y & 0x7FFFFFFFu;
        // a wave meeting a merged group compares whole configs; the others compare the hop-parity
        // strings of the rows of their own groups
        const bool merged = __ballot(act && (range.y >> 31) != 0u) != 0ull;
        // the entries any row of the wave can meet: the union of its rows' groups
        uint32_t lo = gs, hi = ge;
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t ol = __shfl_xor(lo, d);
            const uint32_t oh = __shfl_xor(hi, d);
            lo = ol < lo ? ol : lo;
            hi = oh > hi ? oh : hi;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        hi = __builtin_amdgcn_readfirstlane(hi);
        const int64_t ia = act ? static_cast<int64_t>(p.mrows[e]) : 0;
        Cfg<W> ca;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            ca.w[w] = act ? p.mcfg[e * W + w] : 0ull;
            L.rowc[lane][w] = ca.w[w];
        }
        const uint64_t sa = act ? p.mstr[e] : 0ull;
        // the particle count of the wave's first row's string, and whether every row of the wave
        // holds as many (lane 0 always has a row)
        const int pc_ref = __builtin_amdgcn_readfirstlane(__popcll(sa));
        const bool rows_even = __ballot(act && __popcll(sa) != pc_ref) == 0ull;
        L.acc[lane] = make_double2(0.0, 0.0);
        for (uint32_t c0 = lo; c0 < hi; c0 += 64u) {
            const uint32_t b = c0 + static_cast<uint32_t>(lane);
            bool uneven = false;
            if (b < hi) {
                HopStage<W> st;
                if (merged) {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        st.c[w] = p.mcfg[static_cast<int64_t>(b) * W + w];
                    }
                } else {
                    st.c[0] = p.mstr[b];
                    uneven = __popcll(st.c[0]) != pc_ref;
                }
                st.g = p.s.in[p.mrows[b]];
                L.stage[lane] = st;
            }
            // every string the chunk's pair tests meet holds pc_ref particles (every lane runs the
            // ballot)
            const bool staged_even = __ballot(uneven) == 0ull;
            const bool even = rows_even && staged_even;
            __builtin_amdgcn_wave_barrier();
            const int nb = static_cast<int>(hi - c0 < 64u ? hi - c0 : 64u);
            // the pair tests: bit u of `hit` marks a connection of this lane's row to entry u.  A row
            // meets only the entries of its own group, so each lane walks its group's part of the
            // chunk: the wave's walk is the longest such part, not the whole union.
            const int ulo = static_cast<int>(gs) - static_cast<int>(c0) > 0 ? static_cast<int>(gs) - static_cast<int>(c0) : 0;
            const int uend = static_cast<int>(ge) - static_cast<int>(c0);
            const int uhi = act ? (uend < nb ? uend : nb) : 0;
            uint64_t hit = 0ull;
            if (merged) {
                for (int u = ulo; u < uhi; ++u) {
                    const HopStage<W>& st = L.stage[u];
                    int pc = 0, pa = 0;
                    uint64_t off = 0ull;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        const uint64_t d = ca.w[w] ^ st.c[w];
                        pc += __popcll(d);
                        pa += __popcll(d & ca.w[w]);
                        off |= d & ~p.sites;
                    }
                    // two sites differ, both of the hop parity, one occupied in each config
                    hit |= (pc == 2 && pa == 1 && off == 0ull) ? (1ull << u) : 0ull;
                }
            } else if (even) {
                // rows of one (unmerged) group share the other parity's string, and strings of
                // equal particle counts that differ in two sites hold one of them each: the count
                // of differing sites decides
                for (int u = ulo; u < uhi; ++u) {
                    hit |= static_cast<uint64_t>(__popcll(sa ^ L.stage[u].c[0]) == 2) << u;
                }
            } else {
                // rows of one (unmerged) group share the other parity's string: two hop-parity
                // sites differ, one occupied in each
                for (int u = ulo; u < uhi; ++u) {
                    const uint64_t d = sa ^ L.stage[u].c[0];
                    const bool ok = __popcll(d) == 2 && __popcll(d & sa) == 1;
                    hit |= ok ? (1ull << u) : 0ull;
                }
            }
            hit = act ? hit : 0ull;
            // list the chunk's connections (lane-major) and share them out, kHopList per round
            const int cnt = __popcll(hit);
            int off = cnt;
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(off, d);
                off += lane >= d ? o : 0;
            }
            const int total = __shfl(off, 63);
            off -= cnt;  // exclusive
            for (int r0 = 0; r0 < total; r0 += kHopList) {
                {
                    uint64_t h = hit;
                    int k = off;
                    while (h) {
                        const int u = __builtin_ctzll(h);
                        h &= h - 1ull;
                        if (k >= r0 && k < r0 + kHopList) {
                            L.list[k - r0] = static_cast<uint16_t>((lane << 6) | u);
                        }
                        ++k;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const int m = total - r0 < kHopList ? total - r0 : kHopList;
                for (int i = lane; i < m; i += 64) {
                    const int item = L.list[i];
                    const int la = item >> 6, u = item & 63;
                    const HopStage<W>& st = L.stage[u];
                    Cfg<W> cx;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        cx.w[w] = L.rowc[la][w];
                    }
                    int r = 0, q = 0;
                    if (merged) {
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const uint64_t d = cx.w[w] ^ st.c[w];
                            const uint64_t da = d & cx.w[w], db = d & st.c[w];
                            r = da ? 64 * w + __builtin_ctzll(da) : r;
                            q = db ? 64 * w + __builtin_ctzll(db) : q;
                        }
                    } else {
                        // the row's hop-parity string from its config words
                        uint64_t sx = 0ull;
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            sx |= parity_bits(cx.w[w], p.parity) << (32 * w);
                        }
                        const uint64_t d = sx ^ st.c[0];
                        r = 2 * __builtin_ctzll(d & sx) + p.parity;
                        q = 2 * __builtin_ctzll(d & st.c[0]) + p.parity;
                    }
                    Cfg<W> core = cx;
                    flip_site<W>(core, r);
                    Cfg<W> cb = core;
                    flip_site<W>(cb, q);
                    const int sg = fermi ? (ADJ ? single_parity<W>(cx, r, q) : single_parity<W>(cb, q, r)) : 0;
                    double2 x;
                    if (plain) {
                        const int64_t rq = ADJ ? static_cast<int64_t>(r) * p.s.n_bits + q : static_cast<int64_t>(q) * p.s.n_bits + r;
                        x = p.s.a_re ? make_double2(p.s.a_re[rq], 0.0) : p.s.a[rq];
                    } else {
                        x = ADJ ? single_element<W, 8>(p.s, core, r, q) : single_element<W, 8>(p.s, core, q, r);
                    }
                    double re = 0.0, im = 0.0;
                    hop_madd<ADJ>(sg ? -x.x : x.x, sg ? -x.y : x.y, st.g, re, im);
                    if (re != 0.0 || im != 0.0) {
                        atomicAdd(&L.acc[la].x, re);
                        atomicAdd(&L.acc[la].y, im);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
        const double2 acc = L.acc[lane];
        if (act && (acc.x != 0.0 || acc.y != 0.0)) {
            p.s.out[ia].x += acc.x;
            p.s.out[ia].y += acc.y;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------------------- diagonal --
// d(c) = diag[0] + sum_{p in c} diag[1 + p] + sum_{p < q in c} diag[1 + n_bits + pair_index(p, q)]
template <int W, int CH>
__device__ __forceinline__ double2 diag_value(const Cfg<W>& c, const double2* __restrict__ diag, int n_bits) {
    double2 d = diag[0];
    const double2* e1 = diag + 1;
    const double2* e2 = diag + 1 + n_bits;
#pragma unroll
    for (int wp = 0; wp < W; ++wp) {
        uint64_t mp = c.w[wp];
        while (mp) {
            const int p = 64 * wp + __builtin_ctzll(mp);
            mp &= mp - 1;
            const double2 a = e1[p];
            d.x += a.x;
            d.y += a.y;
#pragma unroll
            for (int wq = 0; wq < W; ++wq) {
                if (wq < wp) {
                    continue;
                }
                uint64_t mq = wq == wp ? mp : c.w[wq];
                // up to CH partners q at a time: their loads go out together (same order of additions)
                while (mq) {
                    double2 v[CH];
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        const bool have = mq != 0ull;
                        const int q = 64 * wq + (have ? __builtin_ctzll(mq) : 0);
                        mq = have ? mq & (mq - 1ull) : 0ull;
                        const double2 b = e2[have ? pair_index(p, q) : 0];
                        v[u] = have ? b : make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        d.x += v[u].x;
                        d.y += v[u].y;
                    }
                }
            }
        }
    }
    return d;
}

// out[i] += d_i in[i] (conj(d_i) for ADJ)
template <int W, bool ADJ>
__global__ __launch_bounds__(kBlock) void diag_apply_kernel(const uint64_t* __restrict__ ckeys, int64_t n,
                                                            const double2* __restrict__ diag, int n_bits,
                                                            const double2* __restrict__ in, double2* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double2 d = diag_value<W, 8>(load_key<W>(ckeys, i), diag, n_bits);
    const double im = ADJ ? -d.y : d.y;
    if (d.x != 0.0 || im != 0.0) {
        const double2 g = in[i];
        out[i].x += d.x * g.x - im * g.y;
        out[i].y += d.x * g.y + im * g.x;
    }
}

// The same diagonal on two vectors in one pass (d, or conj(d) for the adjoint): the config words
// and the folded table are read once for both.
template <int W, bool ADJ>
__global__ __launch_bounds__(kBlock) void diag_apply2_kernel(const uint64_t* __restrict__ ckeys, int64_t n,
                                                             const double2* __restrict__ diag, int n_bits,
                                                             const double2* __restrict__ in, double2* __restrict__ out,
                                                             const double2* __restrict__ in2, double2* __restrict__ out2) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double2 d = diag_value<W, 8>(load_key<W>(ckeys, i), diag, n_bits);
    const double im = ADJ ? -d.y : d.y;
    if (d.x != 0.0 || im != 0.0) {
        const double2 g = in[i];
        const double2 g2 = in2[i];
        out[i].x += d.x * g.x - im * g.y;
        out[i].y += d.x * g.y + im * g.x;
        out2[i].x += d.x * g2.x - im * g2.y;
        out2[i].y += d.x * g2.y + im * g2.x;
    }
}

// --------------------------------------------------------------------- batched persistent launch --
// Many instances (term tables + config sets) in one launch: the work of every instance is cut into
// units -- 64 pair-core groups, 64 single-core groups, 64 diagonal rows -- numbered type-major
// (all pair units, then singles, then diagonal, so the heaviest work is taken first), and
// persistent waves take units round-robin until all are done.
struct BatchItemDev {
    int32_t n_words, particle_cut, has_pairs, has_singles;
    int32_t n_bits, reserved;
    int64_t n;
    const uint64_t* ckeys;
    const double2* diag;
    const double2* pcoef;
    const double* pcoef_re;
    int64_t n_pairs;
    const uint64_t* pvals;
    const uint32_t* pstarts;
    const int64_t* pcounts;
    const double2* sa;
    const double2* skd;
    const double* sa_re;
    const double* skd_re;
    const uint64_t* svals;
    const uint32_t* sstarts;
    const int64_t* scounts;
    const double2* in;
    double2* out;
    const double2* in2;  // two-vector batches: the second input and output
    double2* out2;
};

constexpr int kBatchMaxItems = 4096;

template <int W, bool ADJ, int NV>
__device__ void diag_rows(const BatchItemDev& it, int64_t r0, int lane) {
    const int64_t i = r0 + lane;
    if (i >= it.n) {
        return;
    }
    const double2 d = diag_value<W, 4>(load_key<W>(it.ckeys, i), it.diag, it.n_bits);
    const double im = ADJ ? -d.y : d.y;
    if (d.x != 0.0 || im != 0.0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double2 g = (v == 0 ? it.in : it.in2)[i];
            double2* o = v == 0 ? it.out : it.out2;
            // pair and single units of the same launch add into these rows concurrently
            unsafeAtomicAdd(&o[i].x, d.x * g.x - im * g.y);
            unsafeAtomicAdd(&o[i].y, d.x * g.y + im * g.x);
        }
    }
}

// PART 1: the pair units (type 0); PART 2: the single-excitation (type 1) and diagonal (type 2)
// units.  Both pack the groups of 9..32 entries several to a wave.
template <int W, bool ADJ, int NV, int PART>
__device__ void batch_unit(const BatchItemDev& it, int type, int64_t local, int lane, char* stage) {
    if (PART == 1 && type == 0) {
        PairParams p{};
        p.coef = it.pcoef;
        p.coef_re = it.pcoef_re;
        p.n_pairs = it.n_pairs;
        p.n_bits = it.n_bits;
        p.particle_cut = it.particle_cut;
        p.ckeys = it.ckeys;
        p.vals = it.pvals;
        p.starts = it.pstarts;
        p.counts = it.pcounts;
        p.in = it.in;
        p.out = it.out;
        p.in2 = it.in2;
        p.out2 = it.out2;
        const int64_t groups = it.pcounts[1];
        if (local * 64 < groups && it.pcounts[2] == 0) {
            pair_chunk<W, ADJ, kPairBatchInBatch, NV, true>(p, groups, local * 64, lane, stage);
        }
    } else if (PART == 2 && type == 1) {
        SingleParams p{};
        p.a = it.sa;
        p.kd = it.skd;
        p.a_re = it.sa_re;
        p.kd_re = it.skd_re;
        p.n_bits = it.n_bits;
        p.particle_cut = it.particle_cut;
        p.ckeys = it.ckeys;
        p.vals = it.svals;
        p.starts = it.sstarts;
        p.counts = it.scounts;
        p.in = it.in;
        p.out = it.out;
        p.in2 = it.in2;
        p.out2 = it.out2;
        const int64_t groups = it.scounts[1];
        if (local * 64 < groups && it.scounts[2] == 0) {
            single_chunk<W, ADJ, NV, 4, true>(p, groups, local * 64, lane);
        }
    } else if (PART == 2 && type == 2) {
        diag_rows<W, ADJ, NV>(it, local * 64, lane);
        if (local == 0 && lane == 0 &&
            ((it.has_pairs && it.pcounts[2] != 0) || (it.has_singles && it.scounts[2] != 0))) {
            mark_overflow(it.out);
            if (NV == 2) {
                mark_overflow(it.out2);
            }
        }
    }
}

// One batch application is two persistent launches over the units, the pair units (PART 1), then the
// single and diagonal units (PART 2): each kernel holds only its own unit code, so both keep the
// packed groups within 128 VGPRs (4 waves per SIMD; 3 waves measured ~20% slower).  As one launch
// over all unit types the kernel spilled at that limit and had to walk the 9..32-entry groups one
// per wave (64 molecular instances, forward + backward: 115.0 ms against 108.2 ms split).
template <bool ADJ, int NV, int PART>
__global__ __launch_bounds__(kBlock, 4) void batch_kernel(const BatchItemDev* __restrict__ items, const int64_t* __restrict__ offsets,
                                                       int n_items, int64_t total) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = xcd_wave();
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    __shared__ __align__(16) char stage[kPairStageBytes];
    const int slots = 3 * n_items;
    const int64_t first = PART == 2 ? offsets[n_items] : 0;
    const int64_t last = PART == 1 ? offsets[n_items] : total;
    for (int64_t u = first + wave; u < last; u += nwaves) {
        // last slot whose first unit is <= u (offsets[slots] == total > u)
        int lo = 0, hi = slots;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (offsets[mid] <= u) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const int type = lo / n_items;
        const BatchItemDev& it = items[lo - type * n_items];
        const int64_t local = u - offsets[lo];
        if (it.n_words == 1) {
            batch_unit<1, ADJ, NV, PART>(it, type, local, lane, stage);
        } else {
            batch_unit<2, ADJ, NV, PART>(it, type, local, lane, stage);
        }
    }
}

}  // namespace

size_t pairs_bytes(int n_qubytes) {
    Carver cv{nullptr, 0};
    return pairs_layout(cv, n_qubytes, nullptr) == QMB_OK ? cv.off : 0;
}

int pairs_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                void* buffer, size_t bytes, qmb_pairs* out, hipStream_t stream) {
    if (particle_cut != 1 && particle_cut != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "particle_cut must be 1 or 2, got %d", particle_cut);
    }
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_pairs p{};
    QMB_TRY(pairs_layout(cv, n_qubytes, &p));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "pair table buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    p.particle_cut = particle_cut;
    p.bytes = bytes;
    zero_counts_kernel<<<1, 64, 0, stream>>>(p.counts);
    QMB_CHECK_LAUNCH();
    QMB_TRY(launch_zero_f64(p.coef, 2 * p.n_pairs * p.n_pairs, stream));
    QMB_TRY(launch_zero_f64(p.diag, 2 * (1 + p.n_bits + p.n_pairs), stream));
    if (n_terms > 0) {
        pairs_fill_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(site, kind, reinterpret_cast<const double2*>(coef), n_terms,
                                                                      particle_cut, p.n_bits, p.n_pairs,
                                                                      reinterpret_cast<double2*>(p.coef),
                                                                      reinterpret_cast<double2*>(p.diag), p.counts);
        QMB_CHECK_LAUNCH();
    }
    *out = p;
    return QMB_OK;
}

int singles_set_real(qmb_singles* singles, int real, hipStream_t stream) {
    if (!singles) {
        return set_error(QMB_ERR_ARGUMENT, "singles must be non-null");
    }
    singles->real = 0;
    if (real) {
        const int64_t nb = singles->n_bits;
        pairs_real_kernel<<<blocks_for(nb * nb), kBlock, 0, stream>>>(reinterpret_cast<const double2*>(singles->a), nb * nb,
                                                                      singles->a_re, singles->counts);
        QMB_CHECK_LAUNCH();
        pairs_real_kernel<<<blocks_for(nb * nb * nb), kBlock, 0, stream>>>(reinterpret_cast<const double2*>(singles->kd),
                                                                           nb * nb * nb, singles->kd_re, singles->counts);
        QMB_CHECK_LAUNCH();
        singles->real = 1;
    }
    return QMB_OK;
}

int pairs_set_real(qmb_pairs* pairs, int real, hipStream_t stream) {
    if (!pairs) {
        return set_error(QMB_ERR_ARGUMENT, "pairs must be non-null");
    }
    pairs->real = 0;
    if (real) {
        const int64_t n = pairs->n_pairs * pairs->n_pairs;
        pairs_real_kernel<<<blocks_for(n), kBlock, 0, stream>>>(reinterpret_cast<const double2*>(pairs->coef), n, pairs->coef_re,
                                                                pairs->counts);
        QMB_CHECK_LAUNCH();
        pairs->real = 1;
    }
    return QMB_OK;
}

size_t cores_bytes(int n_qubytes, int64_t n, int max_occupied, int order) {
    Carver cv{nullptr, 0};
    return cores_layout(cv, n_qubytes, n, max_occupied, order, nullptr, nullptr, nullptr) == QMB_OK ? cv.off : 0;
}

int cores_build(const qmb_targets* targets, int max_occupied, int order, void* buffer, size_t bytes, qmb_cores* out,
                hipStream_t stream) {
    if (!targets) {
        return set_error(QMB_ERR_ARGUMENT, "targets must be non-null");
    }
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_cores c{};
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    QMB_TRY(cores_layout(cv, targets->n_qubytes, targets->n, max_occupied, order, &c, &scratch, &scratch_bytes));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "core index buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    c.bytes = bytes;
    zero_counts_kernel<<<1, 64, 0, stream>>>(c.counts);
    QMB_CHECK_LAUNCH();
    if (c.capacity > 0) {
        const size_t cap = static_cast<size_t>(c.capacity);
        Carver sc{static_cast<char*>(scratch), 0};
        // a u64 per entry: the emitted keys (u32) or, packed, the emitted entries
        uint64_t* entries_in = sc.take<uint64_t>(cap);
        uint32_t* keys_in = reinterpret_cast<uint32_t*>(entries_in);
        uint64_t* vals_in = sc.take<uint64_t>(cap);
        uint32_t* flag = sc.take<uint32_t>(cap);
        const int64_t chunks = flag_chunks(c.capacity);
        uint32_t* chunk_heads = sc.take<uint32_t>(static_cast<size_t>(chunks));
        uint32_t* chunk_off = sc.take<uint32_t>(static_cast<size_t>(chunks));
        const int n_bits = 8 * c.n_qubytes;
        uint32_t* inv = sc.take<uint32_t>(static_cast<size_t>(n_bits) * n_bits / 2);
        size_t tmp_bytes = scratch_bytes > sc.off + 512 ? scratch_bytes - sc.off - 512 : 0;
        void* tmp = sc.take_bytes(tmp_bytes);
        const unsigned nb = blocks_for(c.n);
        // packed entries (key << 31 | row and pair slot) when row and slot fit in 31 bits: a
        // keys-only sort of one u64 per entry, from entries_in into vals_in
        const bool packed = core_packed(c.n, n_bits, order);
        const int sub_bits = core_sub_bits(n_bits, order);
        auto emit = [&](auto kernel) {
            kernel<<<nb, kBlock, 0, stream>>>(targets->keys, targets->cols, c.n, n_bits, c.per_config, sub_bits, keys_in,
                                              packed ? static_cast<void*>(entries_in) : static_cast<void*>(vals_in), c.counts);
        };
        const int shape = c.n_words * 2 + (order - 1);
        if (packed) {
            switch (shape) {
                case 2: emit(&core_emit_kernel<1, 1, true>); break;
                case 3: emit(&core_emit_kernel<1, 2, true>); break;
                case 4: emit(&core_emit_kernel<2, 1, true>); break;
                case 5: emit(&core_emit_kernel<2, 2, true>); break;
                case 6: emit(&core_emit_kernel<3, 1, true>); break;
                case 7: emit(&core_emit_kernel<3, 2, true>); break;
                case 8: emit(&core_emit_kernel<4, 1, true>); break;
                default: emit(&core_emit_kernel<4, 2, true>); break;
            }
        } else {
            switch (shape) {
                case 2: emit(&core_emit_kernel<1, 1, false>); break;
                case 3: emit(&core_emit_kernel<1, 2, false>); break;
                case 4: emit(&core_emit_kernel<2, 1, false>); break;
                case 5: emit(&core_emit_kernel<2, 2, false>); break;
                case 6: emit(&core_emit_kernel<3, 1, false>); break;
                case 7: emit(&core_emit_kernel<3, 2, false>); break;
                case 8: emit(&core_emit_kernel<4, 1, false>); break;
                default: emit(&core_emit_kernel<4, 2, false>); break;
            }
        }
        QMB_CHECK_LAUNCH();
        size_t need = tmp_bytes;
        if (packed) {
            QMB_HIP_TRY(rocprim::radix_sort_keys(tmp, need, entries_in, vals_in, static_cast<unsigned>(cap), kPackedBits,
                                                 kPackedBits + kKeyBits, stream));
            if (order == 2) {
                pair_inverse_kernel<<<blocks_for(static_cast<int64_t>(n_bits) * n_bits), kBlock, 0, stream>>>(n_bits, inv);
                QMB_CHECK_LAUNCH();
            }
            auto unpack = [&](auto kernel) {
                kernel<<<blocks_for(c.capacity, kBlock * kFlagUnroll), kBlock, 0, stream>>>(vals_in, c.capacity, flag, chunk_heads, sub_bits,
                                                                                         inv, targets->keys, c.keys, c.vals, c.counts);
            };
            switch (shape) {
                case 2: unpack(&core_flag_unpack_kernel<1, 1>); break;
                case 3: unpack(&core_flag_unpack_kernel<1, 2>); break;
                case 4: unpack(&core_flag_unpack_kernel<2, 1>); break;
                case 5: unpack(&core_flag_unpack_kernel<2, 2>); break;
                case 6: unpack(&core_flag_unpack_kernel<3, 1>); break;
                case 7: unpack(&core_flag_unpack_kernel<3, 2>); break;
                case 8: unpack(&core_flag_unpack_kernel<4, 1>); break;
                default: unpack(&core_flag_unpack_kernel<4, 2>); break;
            }
        } else {
            QMB_HIP_TRY(rocprim::radix_sort_pairs(tmp, need, keys_in, c.keys, vals_in, c.vals, static_cast<unsigned>(cap), 0, kKeyBits,
                                                  stream));
            auto flags = [&](auto kernel) {
                kernel<<<blocks_for(c.capacity, kBlock * kFlagUnroll), kBlock, 0, stream>>>(c.keys, c.vals, c.capacity, targets->keys,
                                                                                         flag, chunk_heads, c.counts);
            };
            switch (shape) {
                case 2: flags(&core_flag_kernel<1, 1>); break;
                case 3: flags(&core_flag_kernel<1, 2>); break;
                case 4: flags(&core_flag_kernel<2, 1>); break;
                case 5: flags(&core_flag_kernel<2, 2>); break;
                case 6: flags(&core_flag_kernel<3, 1>); break;
                case 7: flags(&core_flag_kernel<3, 2>); break;
                case 8: flags(&core_flag_kernel<4, 1>); break;
                default: flags(&core_flag_kernel<4, 2>); break;
            }
        }
        QMB_CHECK_LAUNCH();
        need = tmp_bytes;
        QMB_HIP_TRY(rocprim::exclusive_scan(tmp, need, chunk_heads, chunk_off, 0u, static_cast<size_t>(chunks), rocprim::plus<uint32_t>(),
                                            stream));
        // one wave per flag chunk, as the flag kernel
        core_start_kernel<<<blocks_for(c.capacity, kBlock * kFlagUnroll), kBlock, 0, stream>>>(flag, chunk_off, c.starts, c.counts);
        c.key_bits = kKeyBits;
        QMB_CHECK_LAUNCH();
    }
    *out = c;
    return QMB_OK;
}

int apply_pairs(const qmb_pairs* pairs, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                int mode, hipStream_t stream) {
    if (!pairs || !cores || !targets) {
        return set_error(QMB_ERR_ARGUMENT, "pairs, cores and targets must be non-null");
    }
    if (pairs->n_qubytes != cores->n_qubytes || cores->n_qubytes != targets->n_qubytes || cores->n != targets->n) {
        return set_error(QMB_ERR_ARGUMENT, "pairs, cores and targets describe different shapes");
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (cores->order != 2) {
        return set_error(QMB_ERR_ARGUMENT, "the pair path needs a core index of order 2");
    }
    if (cores->capacity == 0) {
        return QMB_OK;
    }
    PairParams p{};
    p.coef = reinterpret_cast<const double2*>(pairs->coef);
    p.coef_re = pairs->real ? pairs->coef_re : nullptr;
    p.n_pairs = pairs->n_pairs;
    p.n_bits = pairs->n_bits;
    p.particle_cut = pairs->particle_cut;
    p.ckeys = targets->keys;
    p.vals = cores->vals;
    p.starts = cores->starts;
    p.counts = cores->counts;
    p.in = reinterpret_cast<const double2*>(in);
    p.out = reinterpret_cast<double2*>(out);
    // persistent waves: enough to fill every CU several times over, never more than the groups need
    const int64_t want = (cores->capacity + 64 * 4 - 1) / (64 * 4);
    const unsigned grid = static_cast<unsigned>(want < 4096 ? (want < 1 ? 1 : want) : 4096);
    auto go = [&](auto kernel) -> int {
        kernel<<<grid, kBlock, 0, stream>>>(p);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    };
    if (cores->n_words == 1) {
        return mode == 1 ? go(&pair_kernel<1, true>) : go(&pair_kernel<1, false>);
    }
    if (cores->n_words == 2) {
        return mode == 1 ? go(&pair_kernel<2, true>) : go(&pair_kernel<2, false>);
    }
    return set_error(QMB_ERR_UNSUPPORTED, "pair path supports n_qubytes <= %d", kMaxPairBytes);
}

size_t singles_bytes(int n_qubytes) {
    Carver cv{nullptr, 0};
    return singles_layout(cv, n_qubytes, nullptr) == QMB_OK ? cv.off : 0;
}

int singles_build(int n_qubytes, int particle_cut, const int16_t* site, const uint8_t* kind, const double* coef, int64_t n_terms,
                  void* buffer, size_t bytes, qmb_singles* out, hipStream_t stream) {
    if (particle_cut != 1 && particle_cut != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "particle_cut must be 1 or 2, got %d", particle_cut);
    }
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_singles g{};
    QMB_TRY(singles_layout(cv, n_qubytes, &g));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "single-excitation buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    g.particle_cut = particle_cut;
    g.bytes = bytes;
    const int64_t nb = g.n_bits;
    zero_counts_kernel<<<1, 64, 0, stream>>>(g.counts);
    QMB_CHECK_LAUNCH();
    QMB_TRY(launch_zero_f64(g.a, 2 * nb * nb, stream));
    QMB_TRY(launch_zero_f64(g.kd, 2 * nb * nb * nb, stream));
    if (n_terms > 0) {
        singles_fill_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(site, kind, reinterpret_cast<const double2*>(coef), n_terms,
                                                                        particle_cut, g.n_bits, reinterpret_cast<double2*>(g.a),
                                                                        reinterpret_cast<double2*>(g.kd), g.counts);
        QMB_CHECK_LAUNCH();
    }
    *out = g;
    return QMB_OK;
}

int apply_singles(const qmb_singles* singles, const qmb_cores* cores, const qmb_targets* targets, const double* in, double* out,
                  int mode, hipStream_t stream) {
    if (!singles || !cores || !targets) {
        return set_error(QMB_ERR_ARGUMENT, "singles, cores and targets must be non-null");
    }
    if (singles->n_qubytes != cores->n_qubytes || cores->n_qubytes != targets->n_qubytes || cores->n != targets->n) {
        return set_error(QMB_ERR_ARGUMENT, "singles, cores and targets describe different shapes");
    }
    if (cores->order != 1) {
        return set_error(QMB_ERR_ARGUMENT, "the single-excitation path needs a core index of order 1");
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (cores->capacity == 0) {
        return QMB_OK;
    }
    SingleParams p{};
    p.a = reinterpret_cast<const double2*>(singles->a);
    p.kd = reinterpret_cast<const double2*>(singles->kd);
    p.a_re = singles->real ? singles->a_re : nullptr;
    p.kd_re = singles->real ? singles->kd_re : nullptr;
    p.n_bits = singles->n_bits;
    p.particle_cut = singles->particle_cut;
    p.ckeys = targets->keys;
    p.vals = cores->vals;
    p.starts = cores->starts;
    p.counts = cores->counts;
    p.in = reinterpret_cast<const double2*>(in);
    p.out = reinterpret_cast<double2*>(out);
    const int64_t want = (cores->capacity + 64 * 4 - 1) / (64 * 4);
    const unsigned grid = static_cast<unsigned>(want < 4096 ? (want < 1 ? 1 : want) : 4096);
    auto go = [&](auto kernel) -> int {
        kernel<<<grid, kBlock, 0, stream>>>(p);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    };
    if (cores->n_words == 1) {
        return mode == 1 ? go(&single_kernel<1, true>) : go(&single_kernel<1, false>);
    }
    if (cores->n_words == 2) {
        return mode == 1 ? go(&single_kernel<2, true>) : go(&single_kernel<2, false>);
    }
    return set_error(QMB_ERR_UNSUPPORTED, "single-excitation path supports n_qubytes <= %d", kMaxPairBytes);
}

int apply_diagonal(const qmb_pairs* pairs, const qmb_targets* targets, const double* in, double* out, int mode,
                   hipStream_t stream) {
    if (!pairs || !targets) {
        return set_error(QMB_ERR_ARGUMENT, "pairs and targets must be non-null");
    }
    if (pairs->n_qubytes != targets->n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "pairs built for %d bytes, targets for %d", pairs->n_qubytes, targets->n_qubytes);
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (targets->n == 0) {
        return QMB_OK;
    }
    const unsigned grid = blocks_for(targets->n);
    const auto* dg = reinterpret_cast<const double2*>(pairs->diag);
    const auto* vi = reinterpret_cast<const double2*>(in);
    auto* vo = reinterpret_cast<double2*>(out);
    const int nb = pairs->n_bits;
    switch (targets->n_words * 2 + mode) {
        case 2: diag_apply_kernel<1, false><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, nb, vi, vo); break;
        case 3: diag_apply_kernel<1, true><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, nb, vi, vo); break;
        case 4: diag_apply_kernel<2, false><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, nb, vi, vo); break;
        default: diag_apply_kernel<2, true><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, nb, vi, vo); break;
    }
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

int strings_layout(Carver& cv, int n_qubytes, int64_t n, qmb_strings* out, void** scratch, size_t* scratch_bytes) {
    QMB_TRY(check_qubytes(n_qubytes));
    if (n_qubytes > kMaxPairBytes) {
        return set_error(QMB_ERR_UNSUPPORTED, "the strings index supports n_qubytes <= %d, got %d", kMaxPairBytes, n_qubytes);
    }
    if (n < 0 || n >= (1ll << 31)) {
        return set_error(QMB_ERR_ARGUMENT, "n must be in [0, 2^31), got %lld", static_cast<long long>(n));
    }
    qmb_strings t{};
    t.n_qubytes = n_qubytes;
    t.n_words = n_words_of(n_qubytes);
    t.n = n;
    const size_t cap = static_cast<size_t>(n > 0 ? n : 1);
    t.buffer = cv.base;
    t.counts = cv.take<int64_t>(8);
    t.keys = cv.take<uint32_t>(cap);
    t.rows = cv.take<uint32_t>(cap);
    t.group = cv.take<uint32_t>(cap);
    t.starts = cv.take<uint32_t>(cap + 1);
    t.mrows = cv.take<uint32_t>(cap);
    t.mrange = cv.take<uint32_t>(2 * cap);
    t.mcfg = cv.take<uint64_t>(cap * static_cast<size_t>(t.n_words));
    t.mstr = cv.take<uint64_t>(cap);
    size_t tmp = 0;
    {
        uint32_t* u = nullptr;
        if (rocprim::radix_sort_pairs(nullptr, tmp, u, u, u, u, static_cast<unsigned>(cap), 0, 32) != hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim radix_sort_pairs size query failed");
        }
        size_t scan = 0;
        if (rocprim::inclusive_scan(nullptr, scan, u, u, cap, rocprim::plus<uint32_t>()) != hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim inclusive_scan size query failed");
        }
        tmp = tmp > scan ? tmp : scan;
        if (rocprim::exclusive_scan(nullptr, scan, u, u, 0u, cap, rocprim::plus<uint32_t>()) != hipSuccess) {
            return set_error(QMB_ERR_HIP, "rocprim exclusive_scan size query failed");
        }
        tmp = tmp > scan ? tmp : scan;
    }
    const size_t s0 = cv.off;
    void* s = cv.take_bytes(4 * cap * sizeof(uint32_t) + tmp + 4096);
    if (scratch) {
        *scratch = s;
    }
    if (scratch_bytes) {
        *scratch_bytes = cv.off - s0;
    }
    t.bytes = cv.off;
    if (out) {
        *out = t;
    }
    return QMB_OK;
}

size_t strings_bytes(int n_qubytes, int64_t n) {
    Carver cv{nullptr, 0};
    return strings_layout(cv, n_qubytes, n, nullptr, nullptr, nullptr) == QMB_OK ? cv.off : 0;
}

int strings_build(const qmb_targets* targets, int parity, void* buffer, size_t bytes, qmb_strings* out, hipStream_t stream) {
    if (!targets || !out) {
        return set_error(QMB_ERR_ARGUMENT, "targets and out must be non-null");
    }
    if (parity != 0 && parity != 1) {
        return set_error(QMB_ERR_ARGUMENT, "parity must be 0 (hops between even sites) or 1 (odd sites), got %d", parity);
    }
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_strings t{};
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    QMB_TRY(strings_layout(cv, targets->n_qubytes, targets->n, &t, &scratch, &scratch_bytes));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "strings index buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    t.bytes = bytes;
    t.parity = parity;
    QMB_TRY(launch_zero_f64(reinterpret_cast<double*>(t.counts), 8, stream));
    const int64_t n = t.n;
    if (n > 0) {
        const size_t cap = static_cast<size_t>(n);
        Carver sc{static_cast<char*>(scratch), 0};
        uint32_t* keys_in = sc.take<uint32_t>(cap);
        uint32_t* rows_in = sc.take<uint32_t>(cap);
        uint32_t* flag = sc.take<uint32_t>(cap);
        uint32_t* mpos = sc.take<uint32_t>(cap);
        size_t tmp_bytes = scratch_bytes > sc.off + 512 ? scratch_bytes - sc.off - 512 : 0;
        void* tmp = sc.take_bytes(tmp_bytes);
        const unsigned nb = blocks_for(n);
        if (t.n_words == 1) {
            strings_emit_kernel<1><<<nb, kBlock, 0, stream>>>(targets->keys, n, parity, keys_in, rows_in);
        } else {
            strings_emit_kernel<2><<<nb, kBlock, 0, stream>>>(targets->keys, n, parity, keys_in, rows_in);
        }
        QMB_CHECK_LAUNCH();
        size_t need = tmp_bytes;
        QMB_HIP_TRY(rocprim::radix_sort_pairs(tmp, need, keys_in, t.keys, rows_in, t.rows, static_cast<unsigned>(cap), 0, 32, stream));
        key_flag_kernel<<<nb, kBlock, 0, stream>>>(t.keys, n, flag);
        QMB_CHECK_LAUNCH();
        need = tmp_bytes;
        QMB_HIP_TRY(rocprim::inclusive_scan(tmp, need, flag, t.group, cap, rocprim::plus<uint32_t>(), stream));
        strings_start_kernel<<<nb, kBlock, 0, stream>>>(flag, t.group, n, t.starts, t.counts);
        QMB_CHECK_LAUNCH();
        strings_stats_kernel<<<nb < 512u ? nb : 512u, kBlock, 0, stream>>>(t.starts, t.counts);
        QMB_CHECK_LAUNCH();
        strings_multi_kernel<<<nb, kBlock, 0, stream>>>(t.group, t.starts, n, flag);
        QMB_CHECK_LAUNCH();
        need = tmp_bytes;
        QMB_HIP_TRY(rocprim::exclusive_scan(tmp, need, flag, mpos, 0u, cap, rocprim::plus<uint32_t>(), stream));
        auto* mrange = reinterpret_cast<uint2*>(t.mrange);
        if (t.n_words == 1) {
            strings_compact_kernel<1><<<nb, kBlock, 0, stream>>>(targets->keys, t.rows, t.group, t.starts, flag, mpos, n, parity,
                                                                 t.mrows, mrange, t.mcfg, t.mstr, t.counts);
        } else {
            strings_compact_kernel<2><<<nb, kBlock, 0, stream>>>(targets->keys, t.rows, t.group, t.starts, flag, mpos, n, parity,
                                                                 t.mrows, mrange, t.mcfg, t.mstr, t.counts);
        }
        QMB_CHECK_LAUNCH();
        // merged groups (flag is free again: it marks their first entries)
        QMB_TRY(launch_zero_f64(reinterpret_cast<double*>(flag), static_cast<int64_t>((cap + 1) / 2), stream));
        if (t.n_words == 1) {
            strings_mixed_kernel<1><<<nb, kBlock, 0, stream>>>(mrange, t.mcfg, t.counts, parity, flag);
        } else {
            strings_mixed_kernel<2><<<nb, kBlock, 0, stream>>>(mrange, t.mcfg, t.counts, parity, flag);
        }
        QMB_CHECK_LAUNCH();
        strings_mark_kernel<<<nb, kBlock, 0, stream>>>(flag, t.counts, mrange);
        QMB_CHECK_LAUNCH();
    }
    *out = t;
    return QMB_OK;
}

int apply_hops(const qmb_singles* singles, const qmb_strings* strings, const qmb_targets* targets, const double* in, double* out,
               int mode, hipStream_t stream) {
    if (!singles || !strings || !targets || !in || !out) {
        return set_error(QMB_ERR_ARGUMENT, "singles, strings, targets, in and out must be non-null");
    }
    if (singles->n_qubytes != strings->n_qubytes || strings->n_qubytes != targets->n_qubytes || strings->n != targets->n) {
        return set_error(QMB_ERR_ARGUMENT, "singles, strings and targets describe different shapes");
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (strings->n == 0) {
        return QMB_OK;
    }
    HopParams p{};
    p.s.a = reinterpret_cast<const double2*>(singles->a);
    p.s.kd = reinterpret_cast<const double2*>(singles->kd);
    p.s.a_re = singles->real ? singles->a_re : nullptr;
    p.s.kd_re = singles->real ? singles->kd_re : nullptr;
    p.s.n_bits = singles->n_bits;
    p.s.particle_cut = singles->particle_cut;
    p.s.ckeys = targets->keys;
    p.s.in = reinterpret_cast<const double2*>(in);
    p.s.out = reinterpret_cast<double2*>(out);
    p.scounts = singles->counts;
    p.sites = parity_sites(strings->parity);
    p.parity = strings->parity;
    p.mrows = strings->mrows;
    p.mrange = reinterpret_cast<const uint2*>(strings->mrange);
    p.mcfg = strings->mcfg;
    p.mstr = strings->mstr;
    p.counts = strings->counts;
    const int64_t want = (strings->n + 64 * 4 - 1) / (64 * 4);
    const unsigned grid = static_cast<unsigned>(want < 8192 ? (want < 1 ? 1 : want) : 8192);
    auto go = [&](auto kernel) -> int {
        kernel<<<grid, kBlock, 0, stream>>>(p);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    };
    if (strings->n_words == 1) {
        return mode == 1 ? go(&hop_kernel<1, true>) : go(&hop_kernel<1, false>);
    }
    return mode == 1 ? go(&hop_kernel<2, true>) : go(&hop_kernel<2, false>);
}

int apply_structured_dual(const qmb_pairs* pairs, const qmb_singles* singles, const qmb_cores* cores2, const qmb_cores* cores1,
                          const qmb_targets* targets, const double* in, double* out, const double* in2, double* out2, int mode,
                          hipStream_t stream) {
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (!pairs || !cores2 || !targets || !in || !out || !in2 || !out2 || (singles && !cores1)) {
        return set_error(QMB_ERR_ARGUMENT, "pairs, cores2, targets, the four vectors (and cores1 with singles) must be non-null");
    }
    if (pairs->n_qubytes != targets->n_qubytes || cores2->n != targets->n || cores2->order != 2 ||
        (singles && (cores1->n != targets->n || cores1->order != 1 || singles->n_qubytes != targets->n_qubytes))) {
        return set_error(QMB_ERR_ARGUMENT, "pairs, singles, cores and targets describe different shapes");
    }
    if (targets->n == 0) {
        return QMB_OK;
    }
    const int W = targets->n_words;
    if (W != 1 && W != 2) {
        return set_error(QMB_ERR_UNSUPPORTED, "the structured path supports n_qubytes <= %d", kMaxPairBytes);
    }
    const auto* vi = reinterpret_cast<const double2*>(in);
    const auto* vi2 = reinterpret_cast<const double2*>(in2);
    auto* vo = reinterpret_cast<double2*>(out);
    auto* vo2 = reinterpret_cast<double2*>(out2);
    const auto* dg = reinterpret_cast<const double2*>(pairs->diag);
    const unsigned dgrid = blocks_for(targets->n);
    if (W == 1) {
        if (mode == 1) {
            diag_apply2_kernel<1, true><<<dgrid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, pairs->n_bits, vi, vo, vi2, vo2);
        } else {
            diag_apply2_kernel<1, false><<<dgrid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, pairs->n_bits, vi, vo, vi2, vo2);
        }
    } else if (mode == 1) {
        diag_apply2_kernel<2, true><<<dgrid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, pairs->n_bits, vi, vo, vi2, vo2);
    } else {
        diag_apply2_kernel<2, false><<<dgrid, kBlock, 0, stream>>>(targets->keys, targets->n, dg, pairs->n_bits, vi, vo, vi2, vo2);
    }
    QMB_CHECK_LAUNCH();
    if (singles && cores1->capacity > 0) {
        SingleParams p{};
        p.a = reinterpret_cast<const double2*>(singles->a);
        p.kd = reinterpret_cast<const double2*>(singles->kd);
        p.a_re = singles->real ? singles->a_re : nullptr;
        p.kd_re = singles->real ? singles->kd_re : nullptr;
        p.n_bits = singles->n_bits;
        p.particle_cut = singles->particle_cut;
        p.ckeys = targets->keys;
        p.vals = cores1->vals;
        p.starts = cores1->starts;
        p.counts = cores1->counts;
        p.in = vi;
        p.out = vo;
        p.in2 = vi2;
        p.out2 = vo2;
        const int64_t want = (cores1->capacity + 64 * 4 - 1) / (64 * 4);
        const unsigned grid = static_cast<unsigned>(want < 4096 ? (want < 1 ? 1 : want) : 4096);
        if (W == 1) {
            mode == 1 ? single_kernel<1, true, 2><<<grid, kBlock, 0, stream>>>(p) : single_kernel<1, false, 2><<<grid, kBlock, 0, stream>>>(p);
        } else {
            mode == 1 ? single_kernel<2, true, 2><<<grid, kBlock, 0, stream>>>(p) : single_kernel<2, false, 2><<<grid, kBlock, 0, stream>>>(p);
        }
        QMB_CHECK_LAUNCH();
    }
    if (cores2->capacity > 0) {
        PairParams p{};
        p.coef = reinterpret_cast<const double2*>(pairs->coef);
        p.coef_re = pairs->real ? pairs->coef_re : nullptr;
        p.n_pairs = pairs->n_pairs;
        p.n_bits = pairs->n_bits;
        p.particle_cut = pairs->particle_cut;
        p.ckeys = targets->keys;
        p.vals = cores2->vals;
        p.starts = cores2->starts;
        p.counts = cores2->counts;
        p.in = vi;
        p.out = vo;
        p.in2 = vi2;
        p.out2 = vo2;
        const int64_t want = (cores2->capacity + 64 * 4 - 1) / (64 * 4);
        const unsigned grid = static_cast<unsigned>(want < 4096 ? (want < 1 ? 1 : want) : 4096);
        if (W == 1) {
            mode == 1 ? pair_kernel<1, true, 2><<<grid, kBlock, 0, stream>>>(p) : pair_kernel<1, false, 2><<<grid, kBlock, 0, stream>>>(p);
        } else {
            mode == 1 ? pair_kernel<2, true, 2><<<grid, kBlock, 0, stream>>>(p) : pair_kernel<2, false, 2><<<grid, kBlock, 0, stream>>>(p);
        }
        QMB_CHECK_LAUNCH();
    }
    return QMB_OK;
}

int batch_layout(Carver& cv, int n_items, qmb_batch* out) {
    if (n_items < 1 || n_items > kBatchMaxItems) {
        return set_error(QMB_ERR_ARGUMENT, "n_items must be in [1, %d], got %d", kBatchMaxItems, n_items);
    }
    qmb_batch b{};
    b.n_items = n_items;
    b.buffer = cv.base;
    b.items = cv.take<BatchItemDev>(static_cast<size_t>(n_items));
    b.offsets = cv.take<int64_t>(3 * static_cast<size_t>(n_items) + 1);
    b.bytes = cv.off;
    if (out) {
        *out = b;
    }
    return QMB_OK;
}

size_t batch_bytes(int n_items) {
    Carver cv{nullptr, 0};
    return batch_layout(cv, n_items, nullptr) == QMB_OK ? cv.off : 0;
}

int batch_plan(const qmb_batch_item* items, int n_items, void* buffer, size_t bytes, qmb_batch* out, hipStream_t stream) {
    Carver cv{static_cast<char*>(buffer), 0};
    qmb_batch b{};
    QMB_TRY(batch_layout(cv, n_items, &b));
    if (cv.off > bytes) {
        return set_error(QMB_ERR_WORKSPACE, "batch buffer too small: need %zu bytes, have %zu", cv.off, bytes);
    }
    b.bytes = bytes;
    std::vector<BatchItemDev> dev(static_cast<size_t>(n_items));
    int dual = 0;
    std::vector<int64_t> units(3 * static_cast<size_t>(n_items), 0);
    for (int k = 0; k < n_items; ++k) {
        const qmb_batch_item& x = items[k];
        if (!x.pairs || !x.targets) {
            return set_error(QMB_ERR_ARGUMENT, "item %d: pairs and targets must be non-null", k);
        }
        const qmb_targets& tg = *x.targets;
        if (x.pairs->n_qubytes != tg.n_qubytes || tg.n_words > 2) {
            return set_error(QMB_ERR_ARGUMENT, "item %d: pairs / targets shapes differ or n_qubytes > 16", k);
        }
        if (tg.n > 0 && (!x.in || !x.out)) {
            return set_error(QMB_ERR_ARGUMENT, "item %d: in and out must be non-null", k);
        }
        if ((x.pairs && (!x.cores2 || x.cores2->order != 2 || x.cores2->n != tg.n || x.pairs->n_qubytes != tg.n_qubytes)) ||
            (x.singles && (!x.cores1 || x.cores1->order != 1 || x.cores1->n != tg.n || x.singles->n_qubytes != tg.n_qubytes))) {
            return set_error(QMB_ERR_ARGUMENT, "item %d: pair / single tables need matching core indexes of order 2 / 1", k);
        }
        BatchItemDev d{};
        d.n_words = tg.n_words;
        d.particle_cut = x.pairs->particle_cut;
        d.n_bits = 8 * tg.n_qubytes;
        d.n = tg.n;
        d.ckeys = tg.keys;
        d.diag = reinterpret_cast<const double2*>(x.pairs->diag);
        if (x.pairs) {
            d.has_pairs = 1;
            d.pcoef = reinterpret_cast<const double2*>(x.pairs->coef);
            d.pcoef_re = x.pairs->real ? x.pairs->coef_re : nullptr;
            d.n_pairs = x.pairs->n_pairs;
            d.pvals = x.cores2->vals;
            d.pstarts = x.cores2->starts;
            d.pcounts = x.cores2->counts;
            units[k] = (x.cores2->capacity + 63) / 64;
        }
        if (x.singles) {
            d.has_singles = 1;
            d.sa = reinterpret_cast<const double2*>(x.singles->a);
            d.skd = reinterpret_cast<const double2*>(x.singles->kd);
            d.sa_re = x.singles->real ? x.singles->a_re : nullptr;
            d.skd_re = x.singles->real ? x.singles->kd_re : nullptr;
            d.svals = x.cores1->vals;
            d.sstarts = x.cores1->starts;
            d.scounts = x.cores1->counts;
            units[n_items + k] = (x.cores1->capacity + 63) / 64;
        }
        units[2 * n_items + k] = (tg.n + 63) / 64;
        d.in = reinterpret_cast<const double2*>(x.in);
        d.out = reinterpret_cast<double2*>(x.out);
        d.in2 = reinterpret_cast<const double2*>(x.in2);
        d.out2 = reinterpret_cast<double2*>(x.out2);
        const int two = (x.in2 != nullptr && x.out2 != nullptr) ? 1 : 0;
        if ((x.in2 != nullptr) != (x.out2 != nullptr) || (k > 0 && two != dual)) {
            return set_error(QMB_ERR_ARGUMENT, "item %d: in2 and out2 must be given together, for every item or none", k);
        }
        dual = two;
        dev[static_cast<size_t>(k)] = d;
    }
    std::vector<int64_t> offsets(units.size() + 1, 0);
    for (size_t q = 0; q < units.size(); ++q) {
        offsets[q + 1] = offsets[q] + units[q];
    }
    b.total_units = offsets.back();
    b.dual = dual;
    QMB_HIP_TRY(hipMemcpyAsync(b.items, dev.data(), dev.size() * sizeof(BatchItemDev), hipMemcpyHostToDevice, stream));
    QMB_HIP_TRY(hipMemcpyAsync(b.offsets, offsets.data(), offsets.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    // the host tables above go out of scope on return
    QMB_HIP_TRY(hipStreamSynchronize(stream));
    *out = b;
    return QMB_OK;
}

int apply_batch(const qmb_batch* batch, int mode, int waves, hipStream_t stream) {
    if (!batch) {
        return set_error(QMB_ERR_ARGUMENT, "batch must be non-null");
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (batch->total_units == 0) {
        return QMB_OK;
    }
    int64_t w = waves > 0 ? waves : 8192;
    w = w < 4 ? 4 : w;
    int64_t blocks = (w + 3) / 4;
    const int64_t need = (batch->total_units + 3) / 4;
    blocks = blocks < need ? blocks : need;
    const auto* items = static_cast<const BatchItemDev*>(batch->items);
    const unsigned g = static_cast<unsigned>(blocks);
    if (batch->dual) {
        if (mode == 1) {
            batch_kernel<true, 2, 1><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
            batch_kernel<true, 2, 2><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
        } else {
            batch_kernel<false, 2, 1><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
            batch_kernel<false, 2, 2><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
        }
    } else if (mode == 1) {
        batch_kernel<true, 1, 1><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
        batch_kernel<true, 1, 2><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
    } else {
        batch_kernel<false, 1, 1><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
        batch_kernel<false, 1, 2><<<g, kBlock, 0, stream>>>(items, batch->offsets, batch->n_items, batch->total_units);
    }
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

}  // namespace qmb
