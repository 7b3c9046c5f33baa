// qmb_apply.hip -- the hot path: Hamiltonian application on bit-packed configs for gfx950.
//
// apply_within forward (scatter) and adjoint (gather) run on the tiled, bucketed enumeration of
// qmb_enum.hpp; each success is looked up in the hash index of configs_j and accumulated in FP64.
// Also here: the reference-scheme comparator kernel, diagonal_term and single_relative.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "qmb_enum.hpp"

namespace qmb {

namespace {

constexpr int kBlock = kEnumBlock;

inline unsigned blocks_for(int64_t n, int block = kBlock) {
    int64_t b = (n + block - 1) / block;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

struct ApplyParams {
    EnumSource src;
    const double2* in;
    double2* out;
    const double2* coefs;
    const uint64_t* tkeys;
    const uint64_t* table;
    uint64_t mask;
    const uint64_t* filter;
    uint64_t fmask;
    const uint64_t* cols;
    int tile_log2;
    int groups;
    int self_rows;  // symmetric launches: configs_i is the indexed set, rows distinct
};

// out[j] += sign * coef * in[i]: scatter with hardware FP64 atomics.
// SYM: the operator is Hermitian and the index covers the source set itself, so each connected pair
// (i, j) is handled once, from the end with the smaller index hash, applying both A_ji and
// A_ij = conj(A_ji); connections seen from the other end are skipped before their lookup.
template <int W, bool SYM>
struct ForwardAction : ActionHooks {
    const ApplyParams& p;
    __device__ explicit ForwardAction(const ApplyParams& pp) : p(pp) {}
    __device__ __forceinline__ bool pre(int64_t, uint32_t, const Cfg<W>&, int, uint64_t hs, uint64_t ht) {
        return !SYM || ht >= hs;
    }
    __device__ __forceinline__ void post(int64_t i, uint32_t t, const Cfg<W>&, int par, int64_t j, bool distinct) {
        if (j < 0) {
            return;
        }
        const double2 cf = p.coefs[t];
        const double2 v = p.in[i];
        const double sc_re = par ? -cf.x : cf.x;
        const double sc_im = par ? -cf.y : cf.y;
        unsafeAtomicAdd(&p.out[j].x, sc_re * v.x - sc_im * v.y);
        unsafeAtomicAdd(&p.out[j].y, sc_re * v.y + sc_im * v.x);
        if (SYM && distinct) {
            const double2 g = p.in[j];
            unsafeAtomicAdd(&p.out[i].x, sc_re * g.x + sc_im * g.y);
            unsafeAtomicAdd(&p.out[i].y, sc_re * g.y - sc_im * g.x);
        }
    }
};

// conj(sign * coef) * in[j]: the adjoint's contribution to source config i from target row j.
__device__ __forceinline__ double2 adjoint_term(const double2& cf, const double2& g, int par) {
    double re = cf.x * g.x + cf.y * g.y;
    double im = cf.x * g.y - cf.y * g.x;
    return par ? make_double2(-re, -im) : make_double2(re, im);
}

// out[i] += conj(sign * coef) * in[j], gathered per source config.  Tiled enumeration: the tile's
// sums live in LDS (ds_add_f64, contended only by the lanes that hit the same config in one step)
// and reach HBM once per (tile, bucket group) with one FP64 atomic per component, instead of one
// global atomic per contribution.
template <int W>
struct AdjointTileAction : ActionHooks {
    const ApplyParams& p;
    double2* acc;
    int64_t base;
    __device__ explicit AdjointTileAction(const ApplyParams& pp) : p(pp), acc(nullptr), base(0) {}
    __device__ __forceinline__ void tile_begin(int64_t i0, int, double2* a) {
        acc = a;
        base = i0;
        for (int r = threadIdx.x; r < (1 << p.tile_log2); r += kEnumBlock) {
            a[r] = make_double2(0.0, 0.0);
        }
    }
    __device__ __forceinline__ void tile_end(int64_t i0, int nc, double2* a) {
        for (int r = threadIdx.x; r < nc; r += kEnumBlock) {
            const double2 v = a[r];
            if (v.x != 0.0 || v.y != 0.0) {
                if (p.groups == 1) {
                    p.out[i0 + r].x += v.x;
                    p.out[i0 + r].y += v.y;
                } else {
                    unsafeAtomicAdd(&p.out[i0 + r].x, v.x);
                    unsafeAtomicAdd(&p.out[i0 + r].y, v.y);
                }
            }
        }
    }
    __device__ __forceinline__ bool pre(int64_t, uint32_t, const Cfg<W>&, int, uint64_t, uint64_t) { return true; }
    __device__ __forceinline__ void post(int64_t i, uint32_t t, const Cfg<W>&, int par, int64_t j, bool) {
        if (j < 0) {
            return;
        }
        const double2 v = adjoint_term(p.coefs[t], p.in[j], par);
        double2* a = acc + (i - base);
        atomicAdd(&a->x, v.x);
        atomicAdd(&a->y, v.y);
    }
};

// The same gather for the direct enumeration, where one thread owns each source config: the sum
// stays in registers and is added to out[i] once, without atomics.
template <int W>
struct AdjointDirectAction : ActionHooks {
    const ApplyParams& p;
    double re, im;
    __device__ explicit AdjointDirectAction(const ApplyParams& pp) : p(pp), re(0.0), im(0.0) {}
    __device__ __forceinline__ bool pre(int64_t, uint32_t, const Cfg<W>&, int, uint64_t, uint64_t) { return true; }
    __device__ __forceinline__ void post(int64_t, uint32_t t, const Cfg<W>&, int par, int64_t j, bool) {
        if (j < 0) {
            return;
        }
        const double2 v = adjoint_term(p.coefs[t], p.in[j], par);
        re += v.x;
        im += v.y;
    }
    __device__ __forceinline__ void direct_end(int64_t i, bool valid) {
        if (valid && (re != 0.0 || im != 0.0)) {
            p.out[i].x += re;
            p.out[i].y += im;
        }
    }
};

// Counts connections (successful applications) and contributions (connections whose target is in
// the index) for the traffic model; `out` is reinterpreted as two unsigned 64-bit counters.
template <int W>
struct CountAction : ActionHooks {
    const ApplyParams& p;
    unsigned long long gen, hit;
    __device__ explicit CountAction(const ApplyParams& pp) : p(pp), gen(0), hit(0) {}
    __device__ __forceinline__ bool pre(int64_t, uint32_t, const Cfg<W>&, int, uint64_t, uint64_t) {
        ++gen;
        return true;
    }
    __device__ __forceinline__ void post(int64_t, uint32_t, const Cfg<W>&, int, int64_t j, bool) { hit += j >= 0 ? 1 : 0; }
    __device__ __forceinline__ void flush() {
        unsigned long long g = gen, h = hit;
        for (int d = 32; d > 0; d >>= 1) {
            g += __shfl_down(g, d);
            h += __shfl_down(h, d);
        }
        if ((threadIdx.x & 63) == 0) {
            unsigned long long* c = reinterpret_cast<unsigned long long*>(p.out);
            atomicAdd(&c[0], g);
            atomicAdd(&c[1], h);
        }
    }
};

template <int W, int MODE, int DIRECT>
__global__ __launch_bounds__(kBlock) void apply_kernel(ApplyParams p) {
    extern __shared__ uint64_t smem[];
    const HashIndex ix{p.table, p.tkeys, p.mask, p.filter, p.fmask, p.cols, p.self_rows != 0};
    auto run = [&](auto& act) {
        if (DIRECT) {
            enumerate_direct<W>(p.src, ix, blockIdx.x, *reinterpret_cast<DirectLds<W>*>(smem), act);
        } else {
            enumerate_tile<W>(p.src, ix, blockIdx.x / p.groups, p.tile_log2, blockIdx.x % p.groups, p.groups, smem, act);
        }
    };
    if (MODE == 0) {
        ForwardAction<W, false> act(p);
        run(act);
    } else if (MODE == 1) {
        if (DIRECT) {
            AdjointDirectAction<W> act(p);
            run(act);
        } else {
            AdjointTileAction<W> act(p);
            run(act);
        }
    } else if (MODE == 3) {
        ForwardAction<W, true> act(p);
        run(act);
    } else {
        CountAction<W> act(p);
        run(act);
        act.flush();
    }
}

// ------------------------------------------------------------------------------------------
// The reference's scheme on the GPU (comparator): thread (term, config) over the full grid,
// binary search in configs_j sorted by bytes, FP64 atomics (qmb/_hamiltonian_cuda.cu:86-169).
// ------------------------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void plain_kernel(const uint8_t* __restrict__ configs, const double2* __restrict__ psi,
                                                      int64_t n_i, int nq, int particle_cut, const int16_t* __restrict__ site,
                                                      const uint8_t* __restrict__ kind, const double2* __restrict__ coef,
                                                      const uint64_t* __restrict__ sorted_be, const uint32_t* __restrict__ perm,
                                                      int64_t n_j, double2* __restrict__ out) {
    const int64_t t = blockIdx.x;
    int16_t ts[kMaxOps];
    uint8_t tk[kMaxOps];
#pragma unroll
    for (int op = 0; op < kMaxOps; ++op) {
        ts[op] = site[4 * t + op];
        tk[op] = kind[4 * t + op];
    }
    const double2 cf = coef[t];
    for (int64_t i = blockIdx.y * static_cast<int64_t>(kBlock) + threadIdx.x; i < n_i;
         i += static_cast<int64_t>(gridDim.y) * kBlock) {
        Cfg<W> c = load_row<W>(configs + i * nq, nq);
        bool ok = true;
        int parity = 0;
#pragma unroll
        for (int op = kMaxOps; op-- > 0;) {
            const int k = tk[op];
            if (k == 2 || !ok) {
                continue;
            }
            const int s = ts[op];
            const int wi = s >> 6;
            const uint64_t bit = 1ull << (s & 63);
            uint64_t word = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                word = (w == wi) ? c.w[w] : word;
            }
            if (((word & bit) ? 1 : 0) == k) {
                ok = false;
                continue;
            }
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (w == wi) {
                    c.w[w] = k ? (c.w[w] | bit) : (c.w[w] & ~bit);
                }
                if (particle_cut == 1) {
                    parity ^= __popcll(c.w[w] & low_mask_word(s, w)) & 1;
                }
            }
        }
        if (!ok) {
            continue;
        }
        uint64_t key[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            key[w] = bswap64(c.w[w]);
        }
        int64_t lo = 0, hi = n_j - 1, found = -1;
        while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            int cmp = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint64_t s = sorted_be[mid * W + w];
                if (cmp == 0) {
                    cmp = key[w] < s ? -1 : (key[w] > s ? 1 : 0);
                }
            }
            if (cmp < 0) {
                hi = mid - 1;
            } else if (cmp > 0) {
                lo = mid + 1;
            } else {
                found = mid;
                break;
            }
        }
        if (found < 0) {
            continue;
        }
        const int64_t j = perm[found];
        const double2 v = psi[i];
        double re = cf.x * v.x - cf.y * v.y;
        double im = cf.x * v.y + cf.y * v.x;
        if (parity) {
            re = -re;
            im = -im;
        }
        unsafeAtomicAdd(&out[j].x, re);
        unsafeAtomicAdd(&out[j].y, im);
    }
}

// ------------------------------------------------------------------------------------------
// diagonal_term: one thread per config, diagonal terms staged through LDS in original order,
// summed serially so the result does not depend on scheduling (qmb/_hamiltonian_cuda.cu:677-708).
// ------------------------------------------------------------------------------------------
constexpr int kDiagTile = 1024;

template <int W>
__global__ __launch_bounds__(kBlock) void diag_kernel(const uint8_t* __restrict__ configs, int64_t n, int nq,
                                                     const uint64_t* __restrict__ drecs, const double2* __restrict__ dcoefs,
                                                     const int64_t* __restrict__ counts, double2* __restrict__ out) {
    __shared__ uint64_t s_rec[kDiagTile];
    __shared__ double2 s_coef[kDiagTile];
    const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    const bool valid = i < n;
    Cfg<W> c;
    if (valid) {
        c = load_row<W>(configs + i * nq, nq);
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c.w[k] = 0;
        }
    }
    const int64_t td = counts[1];
    double re = 0.0, im = 0.0;
    for (int64_t base = 0; base < td; base += kDiagTile) {
        const int cnt = static_cast<int>(td - base < kDiagTile ? td - base : kDiagTile);
        __syncthreads();
        for (int u = threadIdx.x; u < cnt; u += kBlock) {
            s_rec[u] = drecs[base + u];
            s_coef[u] = dcoefs[base + u];
        }
        __syncthreads();
        for (int u = 0; u < cnt; ++u) {
            Cfg<W> o;
            int par;
            if (apply_rec<W>(s_rec[u], c, o, par)) {
                const double2 cf = s_coef[u];
                re += par ? -cf.x : cf.x;
                im += par ? -cf.y : cf.y;
            }
        }
    }
    if (valid) {
        out[i] = make_double2(re, im);
    }
}

// The diagonal terms applied on an indexed set: out[i] += d_i in[i] (mode 0) or conj(d_i) in[i]
// (mode 1), d_i the sum over the diagonal terms, in their original order, of the coefficients of
// those that succeed on config i (with their parity), as diag_kernel sums them.  Each tile of
// records is expanded to word masks in LDS once per block (succeeds iff (c & m) == v, parity
// popcount(c & p) folded with the record's constant into the coefficient's sign), so a row pays a
// few word operations per term.  For tables with few diagonal terms (a Hubbard U per site), where
// the folded pair table's sum over occupied pairs would read O(occupied^2) entries per row.
constexpr int kDiagMaskTile = 256;

template <int W>
struct DiagMask {
    uint64_t m[W], v[W], p[W];
    double2 coef;  // sign of the record's constant parity folded in
};

template <int W>
__device__ __forceinline__ void stage_diag_masks(const uint64_t* __restrict__ drecs, const double2* __restrict__ dcoefs, int64_t base,
                                                 int cnt, DiagMask<W>* s_mask) {
    for (int u = threadIdx.x; u < cnt; u += kBlock) {
        const Masks<W> x = expand_rec<W>(drecs[base + u], nullptr);
        DiagMask<W> d;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            d.m[w] = x.m[w];
            d.v[w] = x.v[w];
            d.p[w] = x.p[w];
        }
        const double2 cf = dcoefs[base + u];
        d.coef = x.pconst ? make_double2(-cf.x, -cf.y) : cf;
        s_mask[u] = d;
    }
}

template <int W>
__device__ __forceinline__ void sum_diag_masks(const Cfg<W>& c, const DiagMask<W>* s_mask, int cnt, double& re, double& im) {
    for (int u = 0; u < cnt; ++u) {
        const DiagMask<W>& d = s_mask[u];
        bool ok = true;
        int par = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            ok = ok && (c.w[w] & d.m[w]) == d.v[w];
            par ^= __popcll(c.w[w] & d.p[w]);
        }
        if (ok) {
            re += (par & 1) ? -d.coef.x : d.coef.x;
            im += (par & 1) ? -d.coef.y : d.coef.y;
        }
    }
}

template <bool ADJ>
__device__ __forceinline__ void add_diag(int64_t i, double re, double im, const double2* __restrict__ in, double2* __restrict__ out) {
    if (re != 0.0 || im != 0.0) {
        const double dim = ADJ ? -im : im;
        const double2 g = in[i];
        out[i].x += re * g.x - dim * g.y;
        out[i].y += re * g.y + dim * g.x;
    }
}

// The row's diagonal value: added times the input (VALUES false) or written to out (VALUES true).
template <bool ADJ, bool VALUES>
__device__ __forceinline__ void put_diag(int64_t i, double re, double im, const double2* __restrict__ in, double2* __restrict__ out) {
    if (VALUES) {
        out[i] = make_double2(re, im);
    } else {
        add_diag<ADJ>(i, re, im, in, out);
    }
}

// Grid-stride over rows.  Up to kDiagMaskTile diagonal terms are staged once per block; more are
// walked a tile at a time per block of rows.
template <int W, bool ADJ, bool VALUES>
__global__ __launch_bounds__(kBlock) void diag_terms_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                           const uint64_t* __restrict__ drecs, const double2* __restrict__ dcoefs,
                                                           const int64_t* __restrict__ counts, const double2* __restrict__ in,
                                                           double2* __restrict__ out) {
    __shared__ DiagMask<W> s_mask[kDiagMaskTile];
    const int64_t td = counts[1];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    if (td <= kDiagMaskTile) {
        stage_diag_masks<W>(drecs, dcoefs, 0, static_cast<int>(td), s_mask);
        __syncthreads();
        for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; i < n; i += stride) {
            Cfg<W> c;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                c.w[k] = keys[i * W + k];
            }
            double re = 0.0, im = 0.0;
            sum_diag_masks<W>(c, s_mask, static_cast<int>(td), re, im);
            put_diag<ADJ, VALUES>(i, re, im, in, out);
        }
        return;
    }
    for (int64_t i0 = blockIdx.x * static_cast<int64_t>(kBlock); i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        const bool valid = i < n;
        Cfg<W> c;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c.w[k] = valid ? keys[i * W + k] : 0ull;
        }
        double re = 0.0, im = 0.0;
        for (int64_t base = 0; base < td; base += kDiagMaskTile) {
            const int cnt = static_cast<int>(td - base < kDiagMaskTile ? td - base : kDiagMaskTile);
            __syncthreads();
            stage_diag_masks<W>(drecs, dcoefs, base, cnt, s_mask);
            __syncthreads();
            sum_diag_masks<W>(c, s_mask, cnt, re, im);
        }
        if (valid) {
            put_diag<ADJ, VALUES>(i, re, im, in, out);
        }
    }
}

// out[i] += d[i] in[i] (conj(d[i]) for the adjoint), d the rows' diagonal values.
template <bool ADJ>
__global__ __launch_bounds__(kBlock) void diag_scale_kernel(const double2* __restrict__ d, int64_t n, const double2* __restrict__ in,
                                                           double2* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double2 v = d[i];
    add_diag<ADJ>(i, v.x, v.y, in, out);
}

// ------------------------------------------------------------------------------------------
// single_relative: per-term Efraimidis-Spirakis keys, terms sorted by key (descending, ties by
// index), then each config takes the first term that succeeds with a result outside the input set.
// ------------------------------------------------------------------------------------------
__global__ void sr_keys_kernel(const double2* __restrict__ coef, int64_t n_terms, uint64_t seed, const double* __restrict__ keys_in,
                               uint64_t* __restrict__ keybits, uint32_t* __restrict__ vals, uint8_t* __restrict__ usable) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_terms) {
        return;
    }
    double key;
    if (keys_in) {
        key = keys_in[t];
    } else {
        const double2 cf = coef[t];
        const double weight = pow(cf.x * cf.x + cf.y * cf.y, 0.5);
        key = pow(philox_uniform(seed, static_cast<uint64_t>(t)), 1.0 / weight);
    }
    const bool ok = !(key != key) && key >= 0.0;  // a NaN key never wins (score < NaN is false)
    keybits[t] = ok ? static_cast<uint64_t>(__double_as_longlong(key)) : 0ull;
    usable[t] = ok ? 1 : 0;
    vals[t] = static_cast<uint32_t>(t);
}

template <int W>
__global__ __launch_bounds__(kBlock) void single_rel_kernel(const uint8_t* __restrict__ configs, int64_t n, int nq,
                                                           const uint32_t* __restrict__ order, const uint64_t* __restrict__ orecs,
                                                           const uint8_t* __restrict__ usable, int64_t n_terms,
                                                           const uint64_t* __restrict__ tkeys, const uint64_t* __restrict__ table,
                                                           uint64_t mask, const uint64_t* __restrict__ cols, uint8_t* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const Cfg<W> c = load_row<W>(configs + i * nq, nq);
    for (int64_t q = 0; q < n_terms; ++q) {
        const uint32_t t = order[q];
        const uint64_t rec = orecs[t];
        if ((rec & kNeverSucceeds) || !usable[t]) {
            continue;
        }
        Cfg<W> o;
        int par;
        if (!apply_rec<W>(rec, c, o, par)) {
            continue;
        }
        if (lookup<W>(table, mask, tkeys, cols, o) >= 0) {
            continue;
        }
        store_row<W>(out + i * nq, nq, o);
        return;
    }
    for (int b = 0; b < nq; ++b) {
        out[i * nq + b] = 0;
    }
}

}  // namespace

LaunchShape resolve_launch(const qmb_launch* launch, int n_words, int64_t n_i, int64_t n_terms) {
    LaunchShape s{};
    // 256-config tiles: the member lists of one bucket stay short enough that the compaction and the
    // barriers around it are amortised over ~10^2 terms per thread, and a 10^6-config set still gives
    // ~3900 tiles x 8 groups of work-groups.
    int tile = 256;
    int groups = 8;
    int strategy = launch ? launch->strategy : 0;
    if (strategy == 0) {
        // Small tables: every config walks every term from LDS.  Large tables: bucketed tiles.
        strategy = n_terms <= kDirectMaxTerms ? 2 : 1;
    }
    s.direct = strategy == 2;
    s.symmetric = launch != nullptr ? launch->symmetric : 0;
    if (launch && launch->tile > 0) {
        tile = launch->tile;
    } else {
        // small inputs: keep at least ~128 tiles in flight (tile >= 64)
        while (tile > 64 && (n_i + tile - 1) / tile < 128) {
            tile >>= 1;
        }
    }
    if (launch && launch->groups > 0) {
        groups = launch->groups;
    }
    if (tile < 64) {
        tile = 64;
    }
    if (tile > 4096) {
        tile = 4096;
    }
    int lg = 0;
    while ((1 << lg) < tile) {
        ++lg;
    }
    s.tile_log2 = lg;
    s.groups = groups < 1 ? 1 : (groups > 64 ? 64 : groups);
    return s;
}

size_t direct_lds_bytes(int n_words) {
    static_assert(sizeof(DirectLds<1>) <= static_cast<size_t>(kDirectChunk) * (32 * 1 + 12), "DirectLds<1> layout");
    static_assert(sizeof(DirectLds<2>) <= static_cast<size_t>(kDirectChunk) * (32 * 2 + 12), "DirectLds<2> layout");
    static_assert(sizeof(DirectLds<3>) <= static_cast<size_t>(kDirectChunk) * (32 * 3 + 12), "DirectLds<3> layout");
    static_assert(sizeof(DirectLds<4>) <= static_cast<size_t>(kDirectChunk) * (32 * 4 + 12), "DirectLds<4> layout");
    return static_cast<size_t>(kDirectChunk) * (32 * n_words + 12);
}

size_t enum_lds_bytes(int n_words, int n_bits, int tile_log2) { return enum_lds_bytes_of(n_words, n_bits, tile_log2); }

int apply_prepared(const qmb_terms* terms, const qmb_targets* targets, const uint8_t* configs_i, int64_t n_i, const double* in,
                   double* out, int mode, const qmb_launch* launch, hipStream_t stream) {
    if (!terms || !targets) {
        return set_error(QMB_ERR_ARGUMENT, "terms and targets must be non-null");
    }
    if (terms->n_qubytes != targets->n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "terms built for %d bytes, targets for %d", terms->n_qubytes, targets->n_qubytes);
    }
    if (mode != 0 && mode != 1 && mode != 2) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward), 1 (adjoint) or 2 (count)");
    }
    if (n_i < 0 || n_i >= (1ll << 40)) {
        return set_error(QMB_ERR_ARGUMENT, "n_i out of range");
    }
    if (n_i == 0 || terms->n_terms == 0) {
        return QMB_OK;
    }
    const LaunchShape ls = resolve_launch(launch, terms->n_words, n_i, terms->n_terms);
    ApplyParams p{};
    p.src = make_source(terms, configs_i, n_i);
    p.in = reinterpret_cast<const double2*>(in);
    p.out = reinterpret_cast<double2*>(out);
    p.coefs = reinterpret_cast<const double2*>(terms->coefs);
    p.tkeys = targets->keys;
    p.table = targets->table;
    p.mask = targets->capacity - 1;
    p.filter = targets->filter;
    p.fmask = targets->filter_words - 1;
    p.cols = targets->cols;
    p.tile_log2 = ls.tile_log2;
    p.groups = ls.groups;
    p.self_rows = ls.symmetric != 0 ? 1 : 0;
    int64_t grid;
    size_t lds;
    if (ls.direct) {
        grid = (n_i + kBlock - 1) / kBlock;
        lds = direct_lds_bytes(terms->n_words);
    } else {
        const int64_t tiles = (n_i + (1ll << ls.tile_log2) - 1) >> ls.tile_log2;
        grid = tiles * ls.groups;
        lds = enum_lds_bytes(terms->n_words, terms->n_bits, ls.tile_log2);
    }
    if (grid >= (1ll << 31)) {
        return set_error(QMB_ERR_ARGUMENT, "grid too large");
    }
    if (lds > 160 * 1024 - 1024) {
        return set_error(QMB_ERR_ARGUMENT, "tile of %d configs needs %zu bytes of LDS", 1 << ls.tile_log2, lds);
    }
    return dispatch_words(terms->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        auto launch_one = [&](auto kernel) -> int {
            QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(lds)));
            kernel<<<static_cast<unsigned>(grid), kBlock, lds, stream>>>(p);
            QMB_CHECK_LAUNCH();
            return QMB_OK;
        };
        // A Hermitian operator on its own index: forward and adjoint are the same map, computed as
        // the gather (symmetric 1) or as the half-lookup scatter (symmetric 2).
        const int m = (ls.symmetric != 0 && mode != 2) ? (ls.symmetric == 2 ? 3 : 1) : mode;
        if (ls.direct) {
            if (m == 0) return launch_one(&apply_kernel<W, 0, 1>);
            if (m == 1) return launch_one(&apply_kernel<W, 1, 1>);
            if (m == 3) return launch_one(&apply_kernel<W, 3, 1>);
            return launch_one(&apply_kernel<W, 2, 1>);
        }
        if (m == 0) return launch_one(&apply_kernel<W, 0, 0>);
        if (m == 1) return launch_one(&apply_kernel<W, 1, 0>);
        if (m == 3) return launch_one(&apply_kernel<W, 3, 0>);
        return launch_one(&apply_kernel<W, 2, 0>);
    });
}

int apply_plain(int n_qubytes, int particle_cut, const uint8_t* configs_i, const double* psi_i, int64_t n_i,
                const uint64_t* sorted_keys_be, const uint32_t* perm, int64_t n_j, const int16_t* site, const uint8_t* kind,
                const double* coef, int64_t n_terms, double* psi_j, hipStream_t stream) {
    if (n_i == 0 || n_terms == 0 || n_j == 0) {
        return QMB_OK;
    }
    if (n_terms >= (1ll << 31)) {
        return set_error(QMB_ERR_ARGUMENT, "too many terms for the plain grid");
    }
    int64_t gy = (n_i + kBlock - 1) / kBlock;
    if (gy > 65535) {
        gy = 65535;
    }
    dim3 grid(static_cast<unsigned>(n_terms), static_cast<unsigned>(gy));
    return dispatch_words(n_words_of(n_qubytes), [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        plain_kernel<W><<<grid, kBlock, 0, stream>>>(configs_i, reinterpret_cast<const double2*>(psi_i), n_i, n_qubytes,
                                                      particle_cut, site, kind, reinterpret_cast<const double2*>(coef),
                                                      sorted_keys_be, perm, n_j, reinterpret_cast<double2*>(psi_j));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int diagonal_prepared(const qmb_terms* terms, const uint8_t* configs, int64_t n, double* out, hipStream_t stream) {
    if (n == 0) {
        return QMB_OK;
    }
    return dispatch_words(terms->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        diag_kernel<W><<<blocks_for(n), kBlock, 0, stream>>>(configs, n, terms->n_qubytes, terms->drecs,
                                                             reinterpret_cast<const double2*>(terms->dcoefs), terms->counts,
                                                             reinterpret_cast<double2*>(out));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int apply_diagonal_terms(const qmb_terms* terms, const qmb_targets* targets, const double* in, double* out, int mode,
                         hipStream_t stream) {
    if (!terms || !targets || !in || !out) {
        return set_error(QMB_ERR_ARGUMENT, "terms, targets, in and out must be non-null");
    }
    if (terms->n_qubytes != targets->n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "terms built for %d bytes, targets for %d", terms->n_qubytes, targets->n_qubytes);
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (targets->n == 0) {
        return QMB_OK;
    }
    return dispatch_words(terms->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        const auto* dc = reinterpret_cast<const double2*>(terms->dcoefs);
        const auto* vi = reinterpret_cast<const double2*>(in);
        auto* vo = reinterpret_cast<double2*>(out);
        const unsigned grid = blocks_for(targets->n) < 2048u ? blocks_for(targets->n) : 2048u;
        if (mode == 1) {
            diag_terms_kernel<W, true, false><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, terms->drecs, dc, terms->counts, vi,
                                                                          vo);
        } else {
            diag_terms_kernel<W, false, false><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, terms->drecs, dc, terms->counts, vi,
                                                                           vo);
        }
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int diagonal_values(const qmb_terms* terms, const qmb_targets* targets, double* d, hipStream_t stream) {
    if (!terms || !targets || !d) {
        return set_error(QMB_ERR_ARGUMENT, "terms, targets and d must be non-null");
    }
    if (terms->n_qubytes != targets->n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "terms built for %d bytes, targets for %d", terms->n_qubytes, targets->n_qubytes);
    }
    if (targets->n == 0) {
        return QMB_OK;
    }
    return dispatch_words(terms->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        const unsigned grid = blocks_for(targets->n) < 2048u ? blocks_for(targets->n) : 2048u;
        diag_terms_kernel<W, false, true><<<grid, kBlock, 0, stream>>>(targets->keys, targets->n, terms->drecs,
                                                                      reinterpret_cast<const double2*>(terms->dcoefs), terms->counts,
                                                                      nullptr, reinterpret_cast<double2*>(d));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int apply_diagonal_values(int64_t n, const double* d, const double* in, double* out, int mode, hipStream_t stream) {
    if (n < 0 || (n > 0 && (!d || !in || !out))) {
        return set_error(QMB_ERR_ARGUMENT, "d, in and out must be non-null and n >= 0");
    }
    if (mode != 0 && mode != 1) {
        return set_error(QMB_ERR_ARGUMENT, "mode must be 0 (forward) or 1 (adjoint)");
    }
    if (n == 0) {
        return QMB_OK;
    }
    const auto* dv = reinterpret_cast<const double2*>(d);
    const auto* vi = reinterpret_cast<const double2*>(in);
    auto* vo = reinterpret_cast<double2*>(out);
    if (mode == 1) {
        diag_scale_kernel<true><<<blocks_for(n), kBlock, 0, stream>>>(dv, n, vi, vo);
    } else {
        diag_scale_kernel<false><<<blocks_for(n), kBlock, 0, stream>>>(dv, n, vi, vo);
    }
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

size_t single_relative_scratch_bytes(int64_t n_terms) {
    const size_t T = static_cast<size_t>(n_terms > 0 ? n_terms : 1);
    size_t sort_bytes = 0;
    uint64_t* k = nullptr;
    uint32_t* v = nullptr;
    (void)rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, k, k, v, v, static_cast<unsigned>(T));
    Carver cv;
    cv.take<uint64_t>(T);  // orecs
    cv.take<uint64_t>(T);  // keybits
    cv.take<uint64_t>(T);  // sorted keybits
    cv.take<uint32_t>(T);  // vals
    cv.take<uint32_t>(T);  // order
    cv.take<uint8_t>(T);   // usable
    cv.take_bytes(sort_bytes);
    return cv.off + 256;
}

int single_relative_prepared(int n_qubytes, int particle_cut, const uint8_t* configs, int64_t n, const int16_t* site,
                             const uint8_t* kind, const double* coef, int64_t n_terms, uint64_t seed, const double* keys,
                             const qmb_targets* exclude, uint8_t* out, void* scratch, size_t scratch_bytes, hipStream_t stream) {
    if (n == 0) {
        return QMB_OK;
    }
    const size_t T = static_cast<size_t>(n_terms > 0 ? n_terms : 1);
    Carver cv{static_cast<char*>(scratch), 0};
    uint64_t* orecs = cv.take<uint64_t>(T);
    uint64_t* keybits = cv.take<uint64_t>(T);
    uint64_t* skeybits = cv.take<uint64_t>(T);
    uint32_t* vals = cv.take<uint32_t>(T);
    uint32_t* order = cv.take<uint32_t>(T);
    uint8_t* usable = cv.take<uint8_t>(T);
    size_t sort_bytes = 0;
    (void)rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, keybits, skeybits, vals, order, static_cast<unsigned>(T));
    void* tmp = cv.take_bytes(sort_bytes);
    if (cv.off > scratch_bytes) {
        return set_error(QMB_ERR_WORKSPACE, "single_relative scratch too small");
    }
    if (n_terms > 0) {
        QMB_TRY(compile_records_original(n_qubytes, particle_cut, site, kind, n_terms, orecs, stream));
        sr_keys_kernel<<<blocks_for(n_terms), kBlock, 0, stream>>>(reinterpret_cast<const double2*>(coef), n_terms, seed, keys,
                                                                  keybits, vals, usable);
        QMB_CHECK_LAUNCH();
        QMB_HIP_TRY(rocprim::radix_sort_pairs_desc(tmp, sort_bytes, keybits, skeybits, vals, order,
                                                   static_cast<unsigned>(n_terms), 0, 64, stream));
    }
    return dispatch_words(n_words_of(n_qubytes), [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        single_rel_kernel<W><<<blocks_for(n), kBlock, 0, stream>>>(configs, n, n_qubytes, order, orecs, usable, n_terms,
                                                                   exclude->keys, exclude->table, exclude->capacity - 1, exclude->cols, out);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

// Row of each config in an indexed set, or -1 (a thread per config).
template <int W>
__global__ void targets_lookup_kernel(const uint64_t* __restrict__ table, uint64_t mask, const uint64_t* __restrict__ keys,
                                      const uint64_t* __restrict__ cols, const uint8_t* __restrict__ configs, int nq, int64_t n,
                                      int64_t* __restrict__ rows) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    rows[i] = lookup<W>(table, mask, keys, cols, load_row<W>(configs + i * nq, nq));
}

int targets_lookup(const qmb_targets* targets, const uint8_t* configs, int64_t n, int64_t* rows, hipStream_t stream) {
    if (!targets || n < 0 || (n > 0 && (!configs || !rows))) {
        return set_error(QMB_ERR_ARGUMENT, "targets, configs and rows must be non-null");
    }
    if (n == 0) {
        return QMB_OK;
    }
    return dispatch_words(targets->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
        targets_lookup_kernel<W><<<blocks, 256, 0, stream>>>(targets->table, targets->capacity - 1, targets->keys, targets->cols, configs,
                                                              targets->n_qubytes, n, rows);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

}  // namespace qmb
