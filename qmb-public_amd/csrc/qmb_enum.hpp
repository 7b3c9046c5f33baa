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
                                          const uint64_t* __restrict__ cols, const Cfg<W>& c) {
    const uint64_t h = hash_cfg<W>(c, cols);
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

// Optional hooks of the Action protocol (no-ops here; an Action overrides what it needs):
//   tile_begin(i0, nc, acc)  every thread, after the tile of nc configs from row i0 is loaded and
//                            before the barrier that precedes the first post(); acc is the tile's
//                            LDS accumulator row per config (double2[C]);
//   tile_end(i0, nc, acc)    every thread, after the barrier that follows the last post();
//   direct_end(i, valid)     the direct enumeration's thread for config i, after its last post().
struct ActionHooks {
    __device__ __forceinline__ void tile_begin(int64_t, int, double2*) {}
    __device__ __forceinline__ void tile_end(int64_t, int, double2*) {}
    __device__ __forceinline__ void direct_end(int64_t, bool) {}
};

// A hash index over a config set as the kernels see it.
struct HashIndex {
    const uint64_t* table;
    const uint64_t* keys;
    uint64_t mask;
    const uint64_t* filter;
    uint64_t fmask;
    const uint64_t* cols;
    // The index was built from the source configs themselves, rows distinct: a term that flips no
    // site targets the source's own row, which is answered without a lookup.
    bool self_rows;
};

// Linear probe for target o (index hash h) from table slot `slot` until its key or an empty slot.
// Bounded by the capacity: the table is at most a quarter full.
template <int W>
__device__ __forceinline__ int64_t lookup_from(const HashIndex& ix, const Cfg<W>& o, uint64_t h, uint64_t slot) {
    const uint32_t fp = static_cast<uint32_t>(h);
    for (uint64_t probe = 0; probe <= ix.mask; ++probe) {
        const uint64_t ee = ix.table[slot];
        if (ee == kEmptySlot) {
            return -1;
        }
        if (static_cast<uint32_t>(ee) == fp) {
            const uint64_t jj = ee >> 32;
            bool same = true;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                same = same && (ix.keys[jj * W + k] == o.w[k]);
            }
            if (same) {
                return static_cast<int64_t>(jj);
            }
        }
        slot = (slot + 1) & ix.mask;
    }
    return -1;
}

// Batched lookups of U targets.  Action protocol (hs, ht: index hashes of source and target; ht == hs
// iff the term flips nothing):
//   bool pre(i, t, target, parity, hs, ht)       -> true if the target must be looked up
//   void post(i, t, target, parity, j, distinct) -> j = row of the target in the index, or -1;
//                                                   distinct = (ht != hs)
// stay[u]: the term flips no site and the index holds the source rows (ix.self_rows): the target is
// row i[u] itself and nothing is loaded.
// The filter words of all U targets are loaded first.  PHASED then also loads the first table slot
// of every target that passed, then the stored keys of every slot whose fingerprint matched, and
// calls post() last; otherwise each target's probe chain and post() run in turn.  The direct
// enumeration (few lookups per term row, 2 in flight at 8 waves) gains from the phases (a 64-site
// Hubbard chain with 4M configs: forward 5.25 -> 4.2 ms with them); the tile enumeration, whose posts
// do more work between lookups, loses (find_relative at 1M configs 73.6 -> 90.7 ms), so it keeps the
// chains.
template <int W, int U, bool PHASED, typename Action>
__device__ __forceinline__ void lookup_batch(const HashIndex& ix, Action& act, const int64_t (&i)[U], const uint32_t (&t)[U],
                                             const Cfg<W> (&o)[U], const int (&par)[U], const bool (&want)[U],
                                             const bool (&stay)[U], const uint64_t (&h)[U], const uint64_t (&hs)[U]) {
    uint64_t e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        e[u] = (want[u] && !stay[u]) ? ix.filter[filter_word(h[u], ix.fmask)] : 0ull;
    }
    if (!PHASED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!want[u]) {
                continue;
            }
            int64_t j = -1;
            const uint64_t bits = filter_bits(h[u]);
            if (stay[u]) {
                j = i[u];
            } else if ((e[u] & bits) == bits) {
                j = lookup_from<W>(ix, o[u], h[u], (h[u] >> 32) & ix.mask);
            }
            act.post(i[u], t[u], o[u], par[u], j, h[u] != hs[u]);
        }
        return;
    }
    // e[u]: the first table slot of a target that passed the filter, else kEmptySlot
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t bits = filter_bits(h[u]);
        const bool pass = want[u] && !stay[u] && (e[u] & bits) == bits;
        e[u] = pass ? ix.table[(h[u] >> 32) & ix.mask] : kEmptySlot;
    }
    uint64_t key[U][W];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const bool fp = e[u] != kEmptySlot && static_cast<uint32_t>(e[u]) == static_cast<uint32_t>(h[u]);
        const uint64_t jj = fp ? e[u] >> 32 : 0ull;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            key[u][k] = fp ? ix.keys[jj * W + k] : ~o[u].w[k];
        }
    }
    int64_t j[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        bool eq = true;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            eq = eq && key[u][k] == o[u].w[k];
        }
        j[u] = -1;
        if (stay[u]) {
            j[u] = i[u];
        } else if (eq) {
            j[u] = static_cast<int64_t>(e[u] >> 32);
        } else if (e[u] != kEmptySlot) {
            // the first slot held another key: continue the linear probe from the next slot
            j[u] = lookup_from<W>(ix, o[u], h[u], (((h[u] >> 32) & ix.mask) + 1) & ix.mask);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (want[u]) {
            act.post(i[u], t[u], o[u], par[u], j[u], h[u] != hs[u]);
        }
    }
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
// Lookups kept in flight per thread.
constexpr int kBatch = 2;

// LDS carve-up for a tile of C = 1 << tile_log2 configs.
struct EnumLds {
    double2* acc;     // [C] per-config accumulator for gathering actions
    uint64_t* cfg;    // [C][W] tile configs
    uint64_t* hc;     // [C] their index hashes
    uint64_t* bits;   // [n_bits][C / 64] per-site occupancy bitsets
    uint64_t* valid;  // [C / 64]
    uint64_t* mcfg;   // [C][W] members of the current bucket (compacted copies)
    uint64_t* mh;     // [C] their hashes
    uint16_t* mem;    // [C] their tile rows
};

template <int W>
__device__ __forceinline__ EnumLds carve_lds(uint64_t* smem, int tile_log2, int n_bits) {
    const int C = 1 << tile_log2;
    const int CW = C >> 6;
    EnumLds l;
    l.acc = reinterpret_cast<double2*>(smem);
    l.cfg = smem + 2 * C;
    l.hc = l.cfg + C * W;
    l.bits = l.hc + C;
    l.valid = l.bits + n_bits * CW;
    l.mcfg = l.valid + CW;
    l.mh = l.mcfg + C * W;
    l.mem = reinterpret_cast<uint16_t*>(l.mh + C);
    return l;
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
// Every success goes through the Action protocol of lookup_batch against `ix`.  Every thread of the
// workgroup must call this (it synchronises the workgroup).
template <int W, typename Action>
__device__ void enumerate_tile(const EnumSource& p, const HashIndex& ix, int64_t tile, int tile_log2, int g, int groups, uint64_t* smem,
                               Action& act) {
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
        L.hc[r] = hash_cfg<W>(c, ix.cols);
    }
    act.tile_begin(i0, nc, L.acc);
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
                const int r = lane * 64 + __builtin_ctzll(m);
                L.mem[pos] = static_cast<uint16_t>(r);
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    L.mcfg[pos * W + w] = L.cfg[r * W + w];
                }
                L.mh[pos] = L.hc[r];
                ++pos;
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
            const Masks<W> x = expand_rec<W>(p.recs[t], ix.cols);
            bool no_flip = true;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                no_flip = no_flip && x.f[w] == 0ull;
            }
            for (int q0 = ty; q0 < mcount; q0 += TY * kBatch) {
                int64_t ii[kBatch];
                uint32_t tt[kBatch];
                Cfg<W> o[kBatch];
                int par[kBatch];
                bool want[kBatch];
                bool stay[kBatch];
                uint64_t h[kBatch];
                uint64_t hs[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int q = q0 + u * TY;
                    const bool in = q < mcount;
                    const int qq = in ? q : 0;
                    Cfg<W> c;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        c.w[w] = L.mcfg[qq * W + w];
                    }
                    hs[u] = L.mh[qq];
                    h[u] = hs[u] ^ x.hf;
                    tt[u] = t;
                    const bool ok = apply_masks<W>(x, c, o[u], par[u]) && in;
                    ii[u] = ok ? i0 + L.mem[qq] : i0;
                    want[u] = ok && act.pre(ii[u], t, o[u], par[u], hs[u], h[u]);
                    stay[u] = ix.self_rows && no_flip;
                }
                lookup_batch<W, kBatch, false>(ix, act, ii, tt, o, par, want, stay, h, hs);
            }
        }
        __syncthreads();
    }
    act.tile_end(i0, nc, L.acc);
}

// Direct enumeration for small term tables: each thread owns one config in registers and walks
// every term, the terms staged through LDS (expanded to masks) in chunks shared by the workgroup.
constexpr int kDirectChunk = 256;
// Lookups kept in flight per thread by the direct enumeration: 2 keeps the kernel at 60 VGPRs and 8
// waves per SIMD, which hides more latency than 3 (74 VGPRs, 6 waves) or 4 (92 VGPRs, 5 waves).
constexpr int kDirectBatch = 2;

template <int W>
struct DirectLds {
    uint64_t m[kDirectChunk * W];
    uint64_t v[kDirectChunk * W];
    uint64_t f[kDirectChunk * W];
    uint64_t p[kDirectChunk * W];
    uint64_t hf[kDirectChunk];
    int32_t pc[kDirectChunk];
};

template <int W, typename Action>
__device__ void enumerate_direct(const EnumSource& p, const HashIndex& ix, int64_t block, DirectLds<W>& L, Action& act) {
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
    const uint64_t hc = hash_cfg<W>(c, ix.cols);
    const int64_t tb = p.counts[0];
    for (int64_t base = 0; base < tb; base += kDirectChunk) {
        const int cnt = static_cast<int>(tb - base < kDirectChunk ? tb - base : kDirectChunk);
        __syncthreads();
        for (int u = tid; u < cnt; u += kEnumBlock) {
            const Masks<W> x = expand_rec<W>(p.recs[base + u], ix.cols);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                L.m[u * W + w] = x.m[w];
                L.v[u * W + w] = x.v[w];
                L.f[u * W + w] = x.f[w];
                L.p[u * W + w] = x.p[w];
            }
            L.pc[u] = x.pconst;
            L.hf[u] = x.hf;
        }
        __syncthreads();
        if (valid) {
            // Per 32 terms: a mask of those that apply to this config, then lookups of kDirectBatch of
            // them at a time (in term order), so the loads in flight follow the successful terms
            // rather than the terms tried.
            for (int t0 = 0; t0 < cnt; t0 += 32) {
                const int lim = cnt - t0 < 32 ? cnt - t0 : 32;
                uint32_t todo = 0;
                for (int k = 0; k < lim; ++k) {
                    const int u = t0 + k;
                    bool ok = true;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        ok = ok && ((c.w[w] & L.m[u * W + w]) == L.v[u * W + w]);
                    }
                    todo |= ok ? (1u << k) : 0u;
                }
                while (todo != 0u) {
                    int64_t ii[kDirectBatch];
                    uint32_t tt[kDirectBatch];
                    Cfg<W> o[kDirectBatch];
                    int par[kDirectBatch];
                    bool want[kDirectBatch];
                    bool stay[kDirectBatch];
                    uint64_t h[kDirectBatch];
                    uint64_t hs[kDirectBatch];
#pragma unroll
                    for (int b = 0; b < kDirectBatch; ++b) {
                        const bool have = todo != 0u;
                        const int u = t0 + (have ? __builtin_ctz(todo) : 0);
                        todo = have ? todo & (todo - 1u) : 0u;
                        int pr = L.pc[u];
                        bool no_flip = true;
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            o[b].w[w] = c.w[w] ^ L.f[u * W + w];
                            pr ^= __popcll(c.w[w] & L.p[u * W + w]) & 1;
                            no_flip = no_flip && L.f[u * W + w] == 0ull;
                        }
                        stay[b] = ix.self_rows && no_flip;
                        par[b] = pr;
                        ii[b] = i;
                        tt[b] = static_cast<uint32_t>(base + u);
                        hs[b] = hc;
                        h[b] = hc ^ L.hf[u];
                        want[b] = have && act.pre(i, tt[b], o[b], pr, hc, h[b]);
                    }
                    lookup_batch<W, kDirectBatch, true>(ix, act, ii, tt, o, par, want, stay, h, hs);
                }
            }
        }
    }
    act.direct_end(i, valid);
}

inline size_t enum_lds_bytes_of(int n_words, int n_bits, int tile_log2) {
    const size_t C = static_cast<size_t>(1) << tile_log2;
    const size_t CW = C / 64;
    return C * 16 + 2 * (C * n_words * 8 + C * 8) + static_cast<size_t>(n_bits) * CW * 8 + CW * 8 + C * 2;
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
