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

// Hash-index lookup: row of the indexed set equal to c, or -1.  The table is at most half full,
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

// Processes tile `tile` x bucket group `g` (of `groups`).  Action provides
//   begin_member(i)                   before the terms of member config i (global row)
//   on_success(i, t, target, parity)  for every term position t that succeeds on config i
//   end_member(i)                     after them
// Every thread of the workgroup must call this (it synchronises the workgroup).
template <int W, typename Action>
__device__ void enumerate_tile(const EnumSource& p, int64_t tile, int tile_log2, int g, int groups, uint64_t* smem, Action& act) {
    __shared__ int s_count;
    __shared__ uint32_t s_range[2];
    const int C = 1 << tile_log2;
    const int CW = C >> 6;
    const EnumLds L = carve_lds<W>(smem, tile_log2, p.n_bits);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t i0 = tile << tile_log2;
    const int64_t left = p.n_i - i0;
    const int nc = left < C ? static_cast<int>(left) : C;

    if (tid == 0) {
        const uint64_t tb = static_cast<uint64_t>(p.counts[0]);
        const uint64_t lo_t = tb * g / groups;
        const uint64_t hi_t = tb * (g + 1) / groups;
        s_range[0] = lower_bound_u32(p.bucket_start, p.n_buckets, lo_t);
        s_range[1] = (g == groups - 1) ? static_cast<uint32_t>(p.n_buckets) : lower_bound_u32(p.bucket_start, p.n_buckets, hi_t);
    }
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

    const uint32_t b_lo = s_range[0];
    const uint32_t b_hi = s_range[1];
    const int64_t single_end = 1 + p.n_bits;
    const int64_t residual = p.n_buckets - 1;
    for (uint32_t b = b_lo; b < b_hi; ++b) {
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
                decode_pair(static_cast<int64_t>(b) - single_end, &r0, &r1);
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
        // 2-D split of the (member, term) pairs: TX threads along the bucket's terms.
        const int kk = k < kEnumBlock ? k : kEnumBlock;
        const int lg = kk <= 1 ? 0 : 32 - __clz(kk - 1);
        const int TX = 1 << lg;
        const int TY = kEnumBlock >> lg;
        const int tx = tid & (TX - 1);
        const int ty = tid >> lg;
        for (int q = ty; q < mcount; q += TY) {
            const int r = L.mem[q];
            Cfg<W> c;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                c.w[w] = L.cfg[r * W + w];
            }
            const int64_t i = i0 + r;
            act.begin_member(i);
            for (uint32_t t = ts + tx; t < te; t += TX) {
                Cfg<W> o;
                int par;
                if (apply_rec<W>(p.recs[t], c, o, par)) {
                    act.on_success(i, t, o, par);
                }
            }
            act.end_member(i);
        }
        __syncthreads();
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
