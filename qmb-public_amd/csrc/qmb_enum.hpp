// qmb_enum.hpp -- the tiled, bucketed connection enumeration shared by every gfx950 kernel that
// walks the (config, term) pairs of a Hamiltonian: apply_within forward and adjoint,
// find_relative's selection passes and the batched persistent kernel.
//
// A workgroup owns a TILE of C configs and one GROUP of term buckets:
//   1. the tile's configs are read once (row bytes -> 64-bit words) into LDS;
//   2. one wave ballot per (site, 64 configs) builds per-site occupancy bitsets of the tile;
//   3. for each bucket of the group (terms sharing the set R of required-occupied sites), the
//      member configs containing R are one AND of <= 2 bitset rows, compacted with popcount and a
//      wave prefix sum, so only (member, term) pairs of that bucket are evaluated (the reference
//      evaluates the full term x config grid, qmb/_hamiltonian_cuda.cu:137-169);
//   4. each success is handed to an Action with the target config and its parity bit.
// Group g of a tile is processed by workgroup (tile * G + g); consecutive workgroups land on the
// eight XCDs in turn, so with G = 8 every XCD's L2 holds only its own eighth of the term table.
#pragma once

#include "qmb_internal.hpp"

namespace qmb {

template <int W>
__device__ __forceinline__ Cfg<W> load_row(const uint8_t* __restrict__ row, int nq) {
    Cfg<W> c;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        c.w[k] = 0;
    }
    for (int b = 0; b < nq; ++b) {
        const uint64_t byte = row[b];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (k == (b >> 3)) {
                c.w[k] |= byte << (8 * (b & 7));
            }
        }
    }
    return c;
}

template <int W>
__device__ __forceinline__ void store_row(uint8_t* __restrict__ row, int nq, const Cfg<W>& c) {
    for (int b = 0; b < nq; ++b) {
        uint64_t word = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            word = (k == (b >> 3)) ? c.w[k] : word;
        }
        row[b] = static_cast<uint8_t>(word >> (8 * (b & 7)));
    }
}

// Hash-index lookup: row of the indexed set equal to c, or -1.  The table is at most 1/4 full,
// so the probe sequence reaches an empty slot; the loop is additionally capped at the capacity.
template <int W>
__device__ __forceinline__ int64_t lookup(const uint64_t* __restrict__ table, uint64_t mask, const uint64_t* __restrict__ keys,
                                          const Cfg<W>& c) {
    const uint64_t h = hash_cfg<W>(c);
    const uint32_t fp = static_cast<uint32_t>(h);
    uint64_t slot = (h >> 32) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const uint64_t e = table[slot];
        if (e == kEmptySlot) {
            return -1;
        }
        if (static_cast<uint32_t>(e) == fp) {
            const uint64_t j = e >> 32;
            bool eq = true;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                eq = eq && (keys[j * W + k] == c.w[k]);
            }
            if (eq) {
                return static_cast<int64_t>(j);
            }
        }
        slot = (slot + 1) & mask;
    }
    return -1;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ a, int64_t n, uint64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return static_cast<uint32_t>(lo);
}

// Everything the enumeration needs about one (term table, config tile source) pair.
struct EnumSource {
    const uint8_t* configs;  // rows of nq bytes
    int64_t n_i;
    int nq;
    int n_bits;
    const uint64_t* recs;
    const uint32_t* bucket_start;
    int64_t n_buckets;
    const int64_t* counts;
};

constexpr int kEnumBlock = 256;

// LDS carve-up for a tile of C = 1 << tile_log2 configs.
struct EnumLds {
    uint64_t* cfg;    // [C][W]
    uint64_t* bits;   // [n_bits][C / 64]
    uint64_t* valid;  // [C / 64]
    uint16_t* mem;    // [C]
};

template <int W>
__device__ __forceinline__ EnumLds carve_lds(uint64_t* smem, int tile_log2, int n_bits) {
    const int C = 1 << tile_log2;
    const int CW = C >> 6;
    EnumLds l;
    l.cfg = smem;
    l.bits = l.cfg + C * W;
    l.valid = l.bits + n_bits * CW;
    l.mem = reinterpret_cast<uint16_t*>(l.valid + CW);
    return l;
}

// A compiled record expanded to word masks: the term succeeds on c iff (c & m) == v for every word;
// the target is c ^ f and the parity popcount(c & p) ^ pconst (p excludes the touched sites).
template <int W>
struct Masks {
    uint64_t m[W], v[W], f[W], p[W];
    int pconst;
};

template <int W>
__device__ __forceinline__ Masks<W> expand_rec(uint64_t rec) {
    Masks<W> x;
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

template <int W>
__device__ __forceinline__ bool apply_masks(const Masks<W>& x, const Cfg<W>& c, Cfg<W>& o, int& parity) {
    bool ok = true;
    int par = x.pconst;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        ok = ok && ((c.w[w] & x.m[w]) == x.v[w]);
        o.w[w] = c.w[w] ^ x.f[w];
        par ^= __popcll(c.w[w] & x.p[w]) & 1;
    }
    parity = par;
    return ok;
}

// Processes tile `tile` x bucket group `g` (of `groups`): buckets b with b % groups == g, so the
// heavy buckets (pairs of low sites that nearly every config occupies) spread over all groups.
// Action provides on_success(i, t, target, parity) for every term position t that succeeds on
// config row i.  Every thread of the workgroup must call this (it synchronises the workgroup).
template <int W, typename Action>
__device__ void enumerate_tile(const EnumSource& p, int64_t tile, int tile_log2, int g, int groups, uint64_t* smem, Action& act) {
    __shared__ int s_count;
    const int C = 1 << tile_log2;
    const int CW = C >> 6;
    const EnumLds L = carve_lds<W>(smem, tile_log2, p.n_bits);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t i0 = tile << tile_log2;
    const int64_t left = p.n_i - i0;
    const int nc = left < C ? static_cast<int>(left) : C;

    for (int r = tid; r < C; r += kEnumBlock) {
        Cfg<W> c;
        if (r < nc) {
            c = load_row<W>(p.configs + (i0 + r) * p.nq, p.nq);
        } else {
#pragma unroll
            for (int k = 0; k < W; ++k) {
                c.w[k] = 0;
            }
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            L.cfg[r * W + k] = c.w[k];
        }
    }
    __syncthreads();
    for (int q = wave; q < CW; q += kEnumBlock / 64) {
        const int r = q * 64 + lane;
        const uint64_t v = __ballot(r < nc);
        if (lane == 0) {
            L.valid[q] = v;
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const uint64_t word = L.cfg[r * W + k];
            for (int bb = 0; bb < 64; ++bb) {
                const int b = 64 * k + bb;
                if (b >= p.n_bits) {
                    break;
                }
                const uint64_t m = __ballot((word >> bb) & 1ull);
                if (lane == 0) {
                    L.bits[b * CW + q] = m;
                }
            }
        }
    }
    __syncthreads();

    const int64_t single_end = 1 + p.n_bits;
    const int64_t residual = p.n_buckets - 1;
    for (int64_t b = g; b < p.n_buckets; b += groups) {
        const uint32_t ts = p.bucket_start[b];
        const uint32_t te = p.bucket_start[b + 1];
        if (ts == te) {
            continue;  // uniform across the workgroup
        }
        if (wave == 0) {
            int r0 = -1, r1 = -1;
            if (b >= 1 && b < single_end) {
                r0 = static_cast<int>(b) - 1;
            } else if (b >= single_end && b < residual) {
                decode_pair(b - single_end, &r0, &r1);
            }
            uint64_t m = 0;
            if (lane < CW) {
                m = L.valid[lane];
                if (r0 >= 0) {
                    m &= L.bits[r0 * CW + lane];
                }
                if (r1 >= 0) {
                    m &= L.bits[r1 * CW + lane];
                }
            }
            const int cnt = __popcll(m);
            int incl = cnt;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(incl, d);
                if (lane >= d) {
                    incl += y;
                }
            }
            int pos = incl - cnt;
            while (m) {
                const int bit = __builtin_ctzll(m);
                L.mem[pos++] = static_cast<uint16_t>(lane * 64 + bit);
                m &= m - 1;
            }
            if (lane == 63) {
                s_count = incl;
            }
        }
        __syncthreads();
        const int mcount = s_count;
        const int k = static_cast<int>(te - ts);
        // 2-D split of the (term, member) pairs: TX threads along the bucket's terms, each holding
        // its term expanded to masks while it walks TY-strided members (configs are LDS broadcasts).
        const int kk = k < kEnumBlock ? k : kEnumBlock;
        const int lg = kk <= 1 ? 0 : 32 - __clz(kk - 1);
        const int TX = 1 << lg;
        const int TY = kEnumBlock >> lg;
        const int tx = tid & (TX - 1);
        const int ty = tid >> lg;
        for (uint32_t t = ts + tx; t < te; t += TX) {
            const Masks<W> x = expand_rec<W>(p.recs[t]);
            for (int q = ty; q < mcount; q += TY) {
                const int r = L.mem[q];
                Cfg<W> c;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    c.w[w] = L.cfg[r * W + w];
                }
                Cfg<W> o;
                int par;
                if (apply_masks<W>(x, c, o, par)) {
                    act.on_success(i0 + r, t, o, par);
                }
            }
        }
        __syncthreads();
    }
}

// Direct enumeration for small term tables: each thread owns one config in registers and walks
// every term, the terms staged through LDS (expanded to masks) in chunks shared by the workgroup.
constexpr int kDirectChunk = 256;

template <int W>
struct DirectLds {
    uint64_t m[kDirectChunk * W];
    uint64_t v[kDirectChunk * W];
    uint64_t f[kDirectChunk * W];
    uint64_t p[kDirectChunk * W];
    int32_t pc[kDirectChunk];
};

template <int W, typename Action>
__device__ void enumerate_direct(const EnumSource& p, int64_t block, DirectLds<W>& L, Action& act) {
    const int tid = threadIdx.x;
    const int64_t i = block * kEnumBlock + tid;
    const bool valid = i < p.n_i;
    Cfg<W> c;
    if (valid) {
        c = load_row<W>(p.configs + i * p.nq, p.nq);
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c.w[k] = 0;
        }
    }
    const int64_t tb = p.counts[0];
    for (int64_t base = 0; base < tb; base += kDirectChunk) {
        const int cnt = static_cast<int>(tb - base < kDirectChunk ? tb - base : kDirectChunk);
        __syncthreads();
        for (int u = tid; u < cnt; u += kEnumBlock) {
            const Masks<W> x = expand_rec<W>(p.recs[base + u]);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                L.m[u * W + w] = x.m[w];
                L.v[u * W + w] = x.v[w];
                L.f[u * W + w] = x.f[w];
                L.p[u * W + w] = x.p[w];
            }
            L.pc[u] = x.pconst;
        }
        __syncthreads();
        if (valid) {
            for (int u = 0; u < cnt; ++u) {
                bool ok = true;
                int par = L.pc[u];
                Cfg<W> o;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    ok = ok && ((c.w[w] & L.m[u * W + w]) == L.v[u * W + w]);
                    o.w[w] = c.w[w] ^ L.f[u * W + w];
                    par ^= __popcll(c.w[w] & L.p[u * W + w]) & 1;
                }
                if (ok) {
                    act.on_success(i, static_cast<uint32_t>(base + u), o, par);
                }
            }
        }
    }
}

inline size_t enum_lds_bytes_of(int n_words, int n_bits, int tile_log2) {
    const size_t C = static_cast<size_t>(1) << tile_log2;
    const size_t CW = C / 64;
    return C * n_words * 8 + static_cast<size_t>(n_bits) * CW * 8 + CW * 8 + C * 2;
}

template <typename F>
int dispatch_words(int W, F&& f) {
    switch (W) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        default: return set_error(QMB_ERR_UNSUPPORTED, "unsupported word count %d", W);
    }
}

inline EnumSource make_source(const qmb_terms* terms, const uint8_t* configs, int64_t n_i) {
    EnumSource s{};
    s.configs = configs;
    s.n_i = n_i;
    s.nq = terms->n_qubytes;
    s.n_bits = terms->n_bits;
    s.recs = terms->recs;
    s.bucket_start = terms->bucket_start;
    s.n_buckets = terms->n_buckets;
    s.counts = terms->counts;
    return s;
}

}  // namespace qmb
