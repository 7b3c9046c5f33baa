// qmb_find_relative.hip -- find_relative as a deterministic radix select on gfx950.
//
// The reference keeps a K-slot min-heap of (weight, config) entries, one entry per contribution
// |sign * coef * psi_i|^2 whose target is not excluded (qmb/_hamiltonian_cpu.cpp:341-405), guarded
// on the GPU by per-node spin locks with racy pre-reads (qmb/_hamiltonian_cuda.cu:276-427); then it
// drops zero weights and returns the distinct configs sorted by bytes (:537-545).
//
// Here the K largest contributions are found without locks.  Positive doubles order like their bit
// patterns, so the K-th largest weight is located digit by digit (12-bit digits of the weight bits):
// every pass re-runs the enumeration and histograms the next digit of the entries still tied with
// the threshold, and a single-thread kernel picks the digit at which the running count reaches K.
// Once the tied class fits the candidate buffer, one more pass collects every entry strictly above
// the threshold prefix and the tied candidates; the candidates are ordered by (weight desc, config
// asc) and the first K - above are kept.  All control lives in device memory, so the whole call is
// enqueued without host synchronisation.
//
// Large selections first try a sampled threshold, which needs one full pass instead of three: the
// contributions of an evenly spaced sample of n_s rows (with their exclusion lookups) are sorted,
// and the sample weights at ranks f K -+ 6 sqrt(f K) (f = n_s / n_i) bracket the K-th weight of the
// whole set with overwhelming probability.  One pass then collects every entry above the upper
// bound and every entry between the bounds as candidates.  The result is exact when fewer than K
// entries lie above and at least K lie above the lower bound, which the device checks; otherwise
// the state is reset and the digit-by-digit passes run as before.  Either way the candidates are
// finished by the same sort, so ties resolve identically.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "qmb_enum.hpp"

namespace qmb {

namespace {

constexpr int kHistBits = 12;
constexpr int kHistBins = 1 << kHistBits;
constexpr int kLevels = 6;  // digits [51,63) [39,51) [27,39) [15,27) [3,15) [0,3)

__host__ __device__ inline int fr_shift(int level) { return level < 5 ? 51 - 12 * level : 0; }
__host__ __device__ inline int fr_width(int level) { return level < 5 ? 12 : 3; }

inline unsigned blocks_for(int64_t n, int block = 256) {
    int64_t b = (n + block - 1) / block;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

struct FrState {
    int64_t k_rem;
    uint64_t prefix;
    int32_t level;
    int32_t shift;
    int32_t done;
    int32_t all;
    unsigned long long n_above;
    unsigned long long n_cand;
    int64_t cand_needed;
    uint64_t lo;   // collect: candidates are lo <= w < hi, entries with w >= hi are selected outright
    uint64_t hi;
    unsigned long long n_sample;
    int32_t fast_try;   // the sampled bounds are set
    int32_t collected;  // a collect pass produced the final selection
    int32_t failed;     // the sampled attempt did not bracket the K-th weight
    int32_t reserved;
};

struct FrParams {
    EnumSource src;
    const double2* psi;
    const double2* coefs;
    const uint64_t* xkeys;
    const uint64_t* xtable;
    uint64_t xmask;
    const uint64_t* xfilter;
    uint64_t xfmask;
    const uint64_t* xcols;
    FrState* state;
    unsigned long long* hist;
    uint64_t* sel_c;   // [K][W] selected configs ("above" entries, then the kept candidates)
    uint64_t* cand_w;  // [cap]
    uint64_t* cand_c;  // [cap][W]
    uint64_t* sample_w;  // [sample_cap] weights of the sampled rows' contributions
    int64_t sample_cap;
    int64_t k;
    int64_t cand_cap;
    int tile_log2;
    int groups;
    int direct;
};

// |sign * coef * psi|^2 with every product and sum rounded separately (no contraction), so the
// bits match a plain IEEE evaluation of the reference expression (qmb/_hamiltonian_cpu.cpp:388-392).
__device__ __forceinline__ uint64_t weight_bits(const double2 cf, const double2 v) {
    const double re = __dsub_rn(__dmul_rn(cf.x, v.x), __dmul_rn(cf.y, v.y));
    const double im = __dadd_rn(__dmul_rn(cf.x, v.y), __dmul_rn(cf.y, v.x));
    const double w = __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
    return static_cast<uint64_t>(__double_as_longlong(w));
}

template <int W>
struct HistAction : ActionHooks {
    const FrParams& p;
    uint32_t* hist;
    uint64_t prefix;
    int shift;
    int width;
    __device__ HistAction(const FrParams& pp, uint32_t* h, uint64_t pre_, int sh, int wd)
        : p(pp), hist(h), prefix(pre_), shift(sh), width(wd) {}
    // Only entries still tied with the threshold prefix need the (costly) exclusion lookup.
    __device__ __forceinline__ bool pre(int64_t i, uint32_t t, const Cfg<W>&, int, uint64_t, uint64_t) {
        const uint64_t w = weight_bits(p.coefs[t], p.psi[i]);
        return w != 0 && (w >> (shift + width)) == prefix;
    }
    __device__ __forceinline__ void post(int64_t i, uint32_t t, const Cfg<W>&, int, int64_t j, bool) {
        if (j >= 0) {
            return;  // excluded
        }
        const uint64_t w = weight_bits(p.coefs[t], p.psi[i]);
        atomicAdd(&hist[(w >> shift) & ((1u << width) - 1u)], 1u);
    }
};

// Appends to a device counter from the active lanes with one atomic per wave: returns this lane's slot.
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* counter) {
    const uint64_t active = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(active)) - 1;
    const int lane = static_cast<int>(threadIdx.x & 63);
    unsigned long long base = 0;
    if (lane == leader) {
        base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(active)));
    }
    base = __shfl(base, leader);
    return base + static_cast<unsigned long long>(__popcll(active & ((1ull << lane) - 1ull)));
}

__device__ __forceinline__ bool sample_full(const FrParams& p) {
    return *reinterpret_cast<const volatile unsigned long long*>(&p.state->n_sample) > static_cast<unsigned long long>(p.sample_cap);
}

// The sampled rows' contributions: every nonzero weight whose target is not excluded.  Threads
// that start after the buffer has overflowed (the attempt is then abandoned) skip their lookups.
template <int W>
struct SampleAction : ActionHooks {
    const FrParams& p;
    bool full;
    __device__ explicit SampleAction(const FrParams& pp) : p(pp), full(sample_full(pp)) {}
    __device__ __forceinline__ bool pre(int64_t i, uint32_t t, const Cfg<W>&, int, uint64_t, uint64_t) {
        return !full && weight_bits(p.coefs[t], p.psi[i]) != 0;
    }
    __device__ __forceinline__ void post(int64_t i, uint32_t t, const Cfg<W>&, int, int64_t j, bool) {
        if (j >= 0) {
            return;  // excluded
        }
        const unsigned long long s = wave_append(&p.state->n_sample);
        if (s < static_cast<unsigned long long>(p.sample_cap)) {
            p.sample_w[s] = weight_bits(p.coefs[t], p.psi[i]);
        }
    }
};

template <int W>
struct CollectAction : ActionHooks {
    const FrParams& p;
    uint64_t lo;
    uint64_t hi;
    __device__ CollectAction(const FrParams& pp, uint64_t lo_, uint64_t hi_) : p(pp), lo(lo_), hi(hi_) {}
    __device__ __forceinline__ bool pre(int64_t i, uint32_t t, const Cfg<W>&, int, uint64_t, uint64_t) {
        const uint64_t w = weight_bits(p.coefs[t], p.psi[i]);
        return w != 0 && w >= lo;
    }
    __device__ __forceinline__ void post(int64_t i, uint32_t t, const Cfg<W>& o, int, int64_t j, bool) {
        if (j >= 0) {
            return;  // excluded
        }
        const uint64_t w = weight_bits(p.coefs[t], p.psi[i]);
        if (w >= hi) {
            const unsigned long long s = atomicAdd(&p.state->n_above, 1ull);
            if (s < static_cast<unsigned long long>(p.k)) {
#pragma unroll
                for (int w2 = 0; w2 < W; ++w2) {
                    p.sel_c[s * W + w2] = o.w[w2];
                }
            }
        } else {
            const unsigned long long s = atomicAdd(&p.state->n_cand, 1ull);
            if (s < static_cast<unsigned long long>(p.cand_cap)) {
                p.cand_w[s] = w;
#pragma unroll
                for (int w2 = 0; w2 < W; ++w2) {
                    p.cand_c[s * W + w2] = o.w[w2];
                }
            }
        }
    }
};

template <int W, typename Action>
__device__ __forceinline__ void fr_enumerate(const FrParams& p, uint64_t* smem, Action& act) {
    const HashIndex ix{p.xtable, p.xkeys, p.xmask, p.xfilter, p.xfmask, p.xcols, false};
    if (p.direct) {
        enumerate_direct<W>(p.src, ix, blockIdx.x, *reinterpret_cast<DirectLds<W>*>(smem), act);
    } else {
        enumerate_tile<W>(p.src, ix, blockIdx.x / p.groups, p.tile_log2, blockIdx.x % p.groups, p.groups, smem, act);
    }
}

template <int W>
__global__ __launch_bounds__(kEnumBlock) void fr_hist_kernel(FrParams p) {
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_hist[kHistBins];
    if (p.state->done) {
        return;  // uniform: every thread reads the same flag
    }
    const int level = p.state->level;
    const int shift = fr_shift(level);
    const int width = fr_width(level);
    const uint64_t prefix = p.state->prefix;
    for (int b = threadIdx.x; b < kHistBins; b += kEnumBlock) {
        s_hist[b] = 0;
    }
    __syncthreads();
    HistAction<W> act(p, s_hist, prefix, shift, width);
    fr_enumerate<W>(p, smem, act);
    __syncthreads();
    for (int b = threadIdx.x; b < (1 << width); b += kEnumBlock) {
        if (s_hist[b]) {
            atomicAdd(&p.hist[b], static_cast<unsigned long long>(s_hist[b]));
        }
    }
}

// phase 0: the sampled attempt (runs when its bounds are set); phase 1: the digit-by-digit result
// (runs unless the sampled attempt already produced the selection).
template <int W>
__global__ __launch_bounds__(kEnumBlock) void fr_collect_kernel(FrParams p, int phase) {
    extern __shared__ uint64_t smem[];
    if (p.state->collected || (phase == 0 && !p.state->fast_try)) {
        return;  // uniform: every thread reads the same flags
    }
    CollectAction<W> act(p, p.state->lo, p.state->hi);
    fr_enumerate<W>(p, smem, act);
}

template <int W>
__global__ __launch_bounds__(kEnumBlock) void fr_sample_kernel(FrParams p) {
    extern __shared__ uint64_t smem[];
    SampleAction<W> act(p);
    fr_enumerate<W>(p, smem, act);
}

// Bounds from the sorted (descending) sample weights; single thread.
__global__ void fr_bounds_kernel(FrState* st, const uint64_t* __restrict__ sorted, int64_t sample_cap, double fraction, int64_t k) {
    if (threadIdx.x != 0) {
        return;
    }
    st->fast_try = 0;
    if (st->n_sample > static_cast<unsigned long long>(sample_cap)) {
        return;  // the sample overflowed its buffer: take the exact passes
    }
    const int64_t count = static_cast<int64_t>(st->n_sample);
    const double e = fraction * static_cast<double>(k);
    const double sd = sqrt(e);
    const int64_t k_hi = static_cast<int64_t>(floor(e - 6.0 * sd)) - 1;
    const int64_t k_lo = static_cast<int64_t>(ceil(e + 6.0 * sd)) + 1;
    st->hi = k_hi < 0 ? ~0ull : (k_hi < count ? sorted[k_hi] + 1ull : 1ull);
    st->lo = k_lo < count ? sorted[k_lo] : 1ull;
    if (st->lo < 1ull) {
        st->lo = 1ull;
    }
    if (st->hi < st->lo) {
        st->hi = st->lo;
    }
    st->fast_try = 1;
}

// Accepts the sampled attempt when the bounds bracket the K-th weight, else resets for the exact
// passes; single thread.
__global__ void fr_verify_kernel(FrState* st, int64_t k, int64_t cand_cap) {
    if (threadIdx.x != 0 || !st->fast_try) {
        return;
    }
    const unsigned long long kk = static_cast<unsigned long long>(k);
    // lo = 1 collects every nonzero entry, so fewer than K of them is a complete selection too
    const bool reached = st->n_above + st->n_cand >= kk || st->lo <= 1ull;
    if (st->n_above < kk && reached && st->n_cand <= static_cast<unsigned long long>(cand_cap)) {
        st->collected = 1;
        st->done = 1;
        st->cand_needed = static_cast<int64_t>(kk - st->n_above);
    } else {
        st->failed = 1;
        st->n_above = 0;
        st->n_cand = 0;
    }
}

// After a failed sampled attempt the candidate buffer must start empty again.
__global__ void fr_clear_failed_kernel(const FrState* __restrict__ st, uint64_t* __restrict__ cand_w, int64_t cap) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < cap && st->failed) {
        cand_w[i] = 0ull;
    }
}

// Every sample-th row of configs_i and psi_i, compacted.
__global__ void fr_gather_sample_kernel(const uint8_t* __restrict__ configs, const double2* __restrict__ psi, int64_t n_i, int nq,
                                        int64_t n_s, uint8_t* __restrict__ sample_configs, double2* __restrict__ sample_psi) {
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n_s) {
        return;
    }
    const int64_t r = j * n_i / n_s;
    for (int b = 0; b < nq; ++b) {
        sample_configs[j * nq + b] = configs[r * nq + b];
    }
    sample_psi[j] = psi[r];
}

__global__ void fr_init_kernel(FrState* st, unsigned long long* hist, int64_t k) {
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) {
        hist[b] = 0;
    }
    if (threadIdx.x == 0) {
        st->k_rem = k;
        st->prefix = 0;
        st->level = 0;
        st->shift = 0;
        st->done = 0;
        st->all = 0;
        st->n_above = 0;
        st->n_cand = 0;
        st->cand_needed = 0;
        st->lo = 1;
        st->hi = 1;
        st->n_sample = 0;
        st->fast_try = 0;
        st->collected = 0;
        st->failed = 0;
        st->reserved = 0;
    }
}

// Chooses the threshold digit of the current level (single thread; 4096 bins).
__global__ void fr_select_kernel(FrState* st, unsigned long long* hist, int64_t cand_cap) {
    if (threadIdx.x != 0 || st->done) {
        return;
    }
    const int level = st->level;
    const int width = fr_width(level);
    const int bins = 1 << width;
    unsigned long long total = 0;
    for (int b = 0; b < bins; ++b) {
        total += hist[b];
    }
    const unsigned long long need = static_cast<unsigned long long>(st->k_rem);
    if (level == 0 && total <= need) {
        st->all = 1;
        st->done = 1;
        st->cand_needed = 0;
        st->lo = 1;  // every nonzero entry is selected outright
        st->hi = 1;
    } else {
        unsigned long long cum = 0;
        int d = bins - 1;
        for (; d > 0; --d) {
            if (cum + hist[d] >= need) {
                break;
            }
            cum += hist[d];
        }
        st->k_rem = static_cast<int64_t>(need - cum);
        st->prefix = (st->prefix << width) | static_cast<uint64_t>(d);
        st->shift = fr_shift(level);
        if (hist[d] <= static_cast<unsigned long long>(cand_cap) || level == kLevels - 1) {
            st->done = 1;
            st->cand_needed = st->k_rem;
            st->lo = st->prefix << st->shift;         // the tied class
            st->hi = (st->prefix + 1ull) << st->shift;  // everything above it
        } else {
            st->level = level + 1;
        }
    }
    for (int b = 0; b < bins; ++b) {
        hist[b] = 0;
    }
}

template <int W>
__global__ void be_word_gather_kernel(const uint64_t* __restrict__ words, const uint32_t* __restrict__ perm, int64_t n, int w,
                                      uint64_t* __restrict__ out) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p < n) {
        out[p] = bswap64(words[static_cast<int64_t>(perm[p]) * W + w]);
    }
}

__global__ void u64_gather_kernel(const uint64_t* __restrict__ v, const uint32_t* __restrict__ perm, int64_t n,
                                  uint64_t* __restrict__ out) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p < n) {
        out[p] = v[perm[p]];
    }
}

__global__ void valid_key_kernel(const uint32_t* __restrict__ perm, int64_t n, const FrState* __restrict__ st, int64_t k,
                                 uint64_t* __restrict__ out) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p < n) {
        const unsigned long long above = st->n_above < static_cast<unsigned long long>(k) ? st->n_above : k;
        const int64_t n_sel = static_cast<int64_t>(above) + st->cand_needed;
        out[p] = static_cast<int64_t>(perm[p]) < n_sel ? 0ull : 1ull;
    }
}

__global__ void iota_u32_kernel(uint32_t* p, int64_t n) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        p[i] = static_cast<uint32_t>(i);
    }
}

// Appends the first cand_needed candidates (in sorted order) after the "above" entries.
template <int W>
__global__ void fr_take_kernel(const uint32_t* __restrict__ perm, int64_t cap, const uint64_t* __restrict__ cand_c,
                               const FrState* __restrict__ st, int64_t k, uint64_t* __restrict__ sel_c) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    const unsigned long long above = st->n_above < static_cast<unsigned long long>(k) ? st->n_above : k;
    const int64_t found = static_cast<int64_t>(st->n_cand < static_cast<unsigned long long>(cap) ? st->n_cand : cap);
    const int64_t need = st->cand_needed < found ? st->cand_needed : found;
    if (p < cap && p < need && static_cast<int64_t>(above) + p < k) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            sel_c[(above + p) * W + w] = cand_c[static_cast<int64_t>(perm[p]) * W + w];
        }
    }
}

template <int W>
__global__ void fr_unique_flag_kernel(const uint32_t* __restrict__ perm, int64_t n, const uint64_t* __restrict__ sel_c,
                                      const FrState* __restrict__ st, int64_t k, uint32_t* __restrict__ flag) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
    const unsigned long long above = st->n_above < static_cast<unsigned long long>(k) ? st->n_above : k;
    const int64_t found = static_cast<int64_t>(st->n_cand);
    const int64_t need = st->cand_needed < found ? st->cand_needed : found;
    int64_t n_sel = static_cast<int64_t>(above) + need;
    if (n_sel > k) {
        n_sel = k;
    }
    const int64_t a = perm[p];
    uint32_t f = a < n_sel ? 1u : 0u;
    if (f && p > 0) {
        const int64_t b = perm[p - 1];
        bool eq = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            eq = eq && (sel_c[a * W + w] == sel_c[b * W + w]);
        }
        if (eq) {
            f = 0u;
        }
    }
    flag[p] = f;
}

template <int W>
__global__ void fr_scatter_kernel(const uint32_t* __restrict__ perm, int64_t n, const uint64_t* __restrict__ sel_c,
                                  const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int nq,
                                  uint8_t* __restrict__ out, int64_t* __restrict__ out_rows) {
    const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
    if (flag[p]) {
        Cfg<W> c;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            c.w[w] = sel_c[static_cast<int64_t>(perm[p]) * W + w];
        }
        store_row<W>(out + static_cast<int64_t>(pos[p]) * nq, nq, c);
    }
    if (p == n - 1) {
        *out_rows = static_cast<int64_t>(pos[p]) + flag[p];
    }
}

int64_t cand_capacity(int64_t k) {
    int64_t cap = 2 * k;
    return cap < (1 << 20) ? (1 << 20) : cap;
}

// The sampled attempt: n_s rows (0 = no attempt) with f K >= 64 sampled ranks, and a weight
// buffer of at most 2^25 entries (n_s n_terms bounds the contributions of the sample).  At most
// 4096 rows are sampled, so an attempt whose sample overflows the buffer wastes little.
struct FrSamplePlan {
    int64_t n_s;
    int64_t cap;
};

FrSamplePlan fr_sample_plan(int64_t n_i, int64_t k, int64_t n_terms) {
    FrSamplePlan sp{0, 0};
    if (n_i < 4096 || k <= 0 || n_terms <= 0) {
        return sp;
    }
    int64_t n_s = (64 * n_i + k - 1) / k;
    n_s = n_s < 512 ? 512 : n_s;
    if (n_s > n_i / 4 || n_s > 4096) {
        return sp;  // a sample this large saves little over the exact passes
    }
    const int64_t limit = int64_t{1} << 25;
    sp.n_s = n_s;
    sp.cap = n_s > limit / n_terms ? limit : n_s * n_terms;
    return sp;
}

struct FrScratch {
    FrSamplePlan plan;
    uint8_t* sample_configs;
    double2* sample_psi;
    uint64_t* sample_w;
    uint64_t* sample_alt;
    FrState* state;
    unsigned long long* hist;
    uint64_t* sel_c;
    uint64_t* cand_w;
    uint64_t* cand_c;
    uint32_t* perm_a;
    uint32_t* perm_b;
    uint64_t* key_a;
    uint64_t* key_b;
    uint32_t* flag;
    uint32_t* pos;
    void* tmp;
    size_t tmp_bytes;
};

FrScratch fr_layout(Carver& cv, int W, int n_qubytes, int64_t n_i, int64_t k, int64_t n_terms) {
    const int64_t cap = cand_capacity(k);
    const size_t M = static_cast<size_t>(cap > k ? cap : k);
    FrScratch s{};
    s.plan = fr_sample_plan(n_i, k, n_terms);
    s.sample_configs = cv.take<uint8_t>(static_cast<size_t>(s.plan.n_s) * n_qubytes);
    s.sample_psi = cv.take<double2>(static_cast<size_t>(s.plan.n_s));
    s.sample_w = cv.take<uint64_t>(static_cast<size_t>(s.plan.cap));
    s.sample_alt = cv.take<uint64_t>(static_cast<size_t>(s.plan.cap));
    s.state = cv.take<FrState>(1);
    s.hist = cv.take<unsigned long long>(kHistBins);
    s.sel_c = cv.take<uint64_t>(static_cast<size_t>(k) * W);
    s.cand_w = cv.take<uint64_t>(static_cast<size_t>(cap));
    s.cand_c = cv.take<uint64_t>(static_cast<size_t>(cap) * W);
    s.perm_a = cv.take<uint32_t>(M);
    s.perm_b = cv.take<uint32_t>(M);
    s.key_a = cv.take<uint64_t>(M);
    s.key_b = cv.take<uint64_t>(M);
    s.flag = cv.take<uint32_t>(static_cast<size_t>(k));
    s.pos = cv.take<uint32_t>(static_cast<size_t>(k));
    size_t sort_bytes = 0, sort_desc_bytes = 0, scan_bytes = 0, sample_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, s.key_a, s.key_b, s.perm_a, s.perm_b, static_cast<unsigned>(M));
    (void)rocprim::radix_sort_pairs_desc(nullptr, sort_desc_bytes, s.key_a, s.key_b, s.perm_a, s.perm_b, static_cast<unsigned>(M));
    (void)rocprim::exclusive_scan(nullptr, scan_bytes, s.flag, s.pos, 0u, static_cast<size_t>(k), rocprim::plus<uint32_t>());
    if (s.plan.n_s > 0) {
        (void)rocprim::radix_sort_keys_desc(nullptr, sample_bytes, s.sample_w, s.sample_alt, static_cast<unsigned>(s.plan.cap));
    }
    size_t t = sort_bytes > sort_desc_bytes ? sort_bytes : sort_desc_bytes;
    t = t > scan_bytes ? t : scan_bytes;
    t = t > sample_bytes ? t : sample_bytes;
    s.tmp = cv.take_bytes(t);
    s.tmp_bytes = t;
    return s;
}

// Sorts perm (an index permutation over n entries) stably by config words ascending.
template <int W>
int sort_by_config(const uint64_t* words, int64_t n, FrScratch& s, hipStream_t stream) {
    for (int w = W - 1; w >= 0; --w) {
        be_word_gather_kernel<W><<<blocks_for(n), 256, 0, stream>>>(words, s.perm_a, n, w, s.key_a);
        QMB_CHECK_LAUNCH();
        size_t need = s.tmp_bytes;
        QMB_HIP_TRY(rocprim::radix_sort_pairs(s.tmp, need, s.key_a, s.key_b, s.perm_a, s.perm_b, static_cast<unsigned>(n), 0, 64,
                                              stream));
        QMB_HIP_TRY(hipMemcpyAsync(s.perm_a, s.perm_b, static_cast<size_t>(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    }
    return QMB_OK;
}

template <int W>
int find_relative_impl(const qmb_terms* terms, const qmb_targets* exclude, const uint8_t* configs_i, const double* psi_i,
                       int64_t n_i, int64_t k, uint8_t* out, int64_t* out_rows, void* scratch, size_t scratch_bytes,
                       hipStream_t stream) {
    Carver cv{static_cast<char*>(scratch), 0};
    FrScratch s = fr_layout(cv, W, terms->n_qubytes, n_i, k, terms->n_terms);
    if (cv.off > scratch_bytes) {
        return set_error(QMB_ERR_WORKSPACE, "find_relative scratch too small: need %zu, have %zu", cv.off, scratch_bytes);
    }
    const int64_t cap = cand_capacity(k);
    const LaunchShape ls = resolve_launch(nullptr, W, n_i, terms->n_terms);
    FrParams p{};
    p.src = make_source(terms, configs_i, n_i);
    p.psi = reinterpret_cast<const double2*>(psi_i);
    p.coefs = reinterpret_cast<const double2*>(terms->coefs);
    p.xkeys = exclude->keys;
    p.xtable = exclude->table;
    p.xmask = exclude->capacity - 1;
    p.xfilter = exclude->filter;
    p.xfmask = exclude->filter_words - 1;
    p.xcols = exclude->cols;
    p.state = s.state;
    p.hist = s.hist;
    p.sel_c = s.sel_c;
    p.cand_w = s.cand_w;
    p.cand_c = s.cand_c;
    p.sample_w = s.sample_w;
    p.sample_cap = s.plan.cap;
    p.k = k;
    p.cand_cap = cap;
    p.tile_log2 = ls.tile_log2;
    p.groups = ls.groups;
    p.direct = ls.direct ? 1 : 0;
    const int64_t tiles = (n_i + (1ll << ls.tile_log2) - 1) >> ls.tile_log2;
    const int64_t grid = n_i <= 0 ? 0 : (ls.direct ? (n_i + kEnumBlock - 1) / kEnumBlock : tiles * ls.groups);
    const size_t lds = ls.direct ? direct_lds_bytes(W) : enum_lds_bytes_of(W, terms->n_bits, ls.tile_log2);

    fr_init_kernel<<<1, 256, 0, stream>>>(s.state, s.hist, k);
    QMB_CHECK_LAUNCH();
    QMB_TRY(launch_fill_u64(s.cand_w, 0ull, cap, stream));
    if (grid > 0 && terms->n_terms > 0) {
        QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&fr_hist_kernel<W>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&fr_collect_kernel<W>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        if (s.plan.n_s > 0) {
            // the sampled attempt: sample contributions, bounds, one collect pass, check
            const int64_t n_s = s.plan.n_s;
            fr_gather_sample_kernel<<<blocks_for(n_s), 256, 0, stream>>>(configs_i, p.psi, n_i, terms->n_qubytes, n_s, s.sample_configs,
                                                                         s.sample_psi);
            QMB_CHECK_LAUNCH();
            QMB_TRY(launch_fill_u64(s.sample_w, 0ull, s.plan.cap, stream));
            FrParams ps = p;
            ps.src = make_source(terms, s.sample_configs, n_s);
            ps.psi = s.sample_psi;
            const LaunchShape lss = resolve_launch(nullptr, W, n_s, terms->n_terms);
            const int64_t tiles_s = (n_s + (1ll << lss.tile_log2) - 1) >> lss.tile_log2;
            const int64_t grid_s = lss.direct ? (n_s + kEnumBlock - 1) / kEnumBlock : tiles_s * lss.groups;
            const size_t lds_s = lss.direct ? direct_lds_bytes(W) : enum_lds_bytes_of(W, terms->n_bits, lss.tile_log2);
            ps.tile_log2 = lss.tile_log2;
            ps.groups = lss.groups;
            ps.direct = lss.direct ? 1 : 0;
            QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&fr_sample_kernel<W>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_s > lds ? lds_s : lds)));
            fr_sample_kernel<W><<<static_cast<unsigned>(grid_s), kEnumBlock, lds_s, stream>>>(ps);
            QMB_CHECK_LAUNCH();
            size_t need_s = s.tmp_bytes;
            QMB_HIP_TRY(rocprim::radix_sort_keys_desc(s.tmp, need_s, s.sample_w, s.sample_alt, static_cast<unsigned>(s.plan.cap), 0, 64,
                                                      stream));
            fr_bounds_kernel<<<1, 64, 0, stream>>>(s.state, s.sample_alt, s.plan.cap,
                                                   static_cast<double>(n_s) / static_cast<double>(n_i), k);
            QMB_CHECK_LAUNCH();
            fr_collect_kernel<W><<<static_cast<unsigned>(grid), kEnumBlock, lds, stream>>>(p, 0);
            QMB_CHECK_LAUNCH();
            fr_verify_kernel<<<1, 64, 0, stream>>>(s.state, k, cap);
            QMB_CHECK_LAUNCH();
            fr_clear_failed_kernel<<<blocks_for(cap), 256, 0, stream>>>(s.state, s.cand_w, cap);
            QMB_CHECK_LAUNCH();
        }
        for (int level = 0; level < kLevels; ++level) {
            fr_hist_kernel<W><<<static_cast<unsigned>(grid), kEnumBlock, lds, stream>>>(p);
            QMB_CHECK_LAUNCH();
            fr_select_kernel<<<1, 64, 0, stream>>>(s.state, s.hist, cap);
            QMB_CHECK_LAUNCH();
        }
        fr_collect_kernel<W><<<static_cast<unsigned>(grid), kEnumBlock, lds, stream>>>(p, 1);
        QMB_CHECK_LAUNCH();
    }
    // Candidates: order by (weight desc, config asc), keep the first cand_needed.
    iota_u32_kernel<<<blocks_for(cap), 256, 0, stream>>>(s.perm_a, cap);
    QMB_CHECK_LAUNCH();
    QMB_TRY(sort_by_config<W>(s.cand_c, cap, s, stream));
    u64_gather_kernel<<<blocks_for(cap), 256, 0, stream>>>(s.cand_w, s.perm_a, cap, s.key_a);
    QMB_CHECK_LAUNCH();
    size_t need = s.tmp_bytes;
    QMB_HIP_TRY(rocprim::radix_sort_pairs_desc(s.tmp, need, s.key_a, s.key_b, s.perm_a, s.perm_b, static_cast<unsigned>(cap), 0, 64,
                                               stream));
    fr_take_kernel<W><<<blocks_for(cap), 256, 0, stream>>>(s.perm_b, cap, s.cand_c, s.state, k, s.sel_c);
    QMB_CHECK_LAUNCH();
    // Selected entries: sort by config (valid entries first), drop repeats, write rows.
    iota_u32_kernel<<<blocks_for(k), 256, 0, stream>>>(s.perm_a, k);
    QMB_CHECK_LAUNCH();
    QMB_TRY(sort_by_config<W>(s.sel_c, k, s, stream));
    valid_key_kernel<<<blocks_for(k), 256, 0, stream>>>(s.perm_a, k, s.state, k, s.key_a);
    QMB_CHECK_LAUNCH();
    need = s.tmp_bytes;
    QMB_HIP_TRY(rocprim::radix_sort_pairs(s.tmp, need, s.key_a, s.key_b, s.perm_a, s.perm_b, static_cast<unsigned>(k), 0, 1, stream));
    fr_unique_flag_kernel<W><<<blocks_for(k), 256, 0, stream>>>(s.perm_b, k, s.sel_c, s.state, k, s.flag);
    QMB_CHECK_LAUNCH();
    need = s.tmp_bytes;
    QMB_HIP_TRY(rocprim::exclusive_scan(s.tmp, need, s.flag, s.pos, 0u, static_cast<size_t>(k), rocprim::plus<uint32_t>(), stream));
    fr_scatter_kernel<W><<<blocks_for(k), 256, 0, stream>>>(s.perm_b, k, s.sel_c, s.flag, s.pos, terms->n_qubytes, out, out_rows);
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

}  // namespace

size_t find_relative_scratch_bytes(int n_qubytes, int64_t n_i, int64_t count_selected, int64_t n_terms) {
    Carver cv;
    (void)fr_layout(cv, n_words_of(n_qubytes), n_qubytes, n_i, count_selected, n_terms);
    return cv.off + 256;
}

int find_relative_prepared(const qmb_terms* terms, const qmb_targets* exclude, const uint8_t* configs_i, const double* psi_i,
                           int64_t n_i, int64_t count_selected, uint8_t* out, int64_t* out_rows, void* scratch,
                           size_t scratch_bytes, hipStream_t stream) {
    if (count_selected <= 0) {
        return set_error(QMB_ERR_ARGUMENT, "count_selected must be positive, got %lld", static_cast<long long>(count_selected));
    }
    if (count_selected >= (1ll << 31)) {
        return set_error(QMB_ERR_ARGUMENT, "count_selected too large");
    }
    if (terms->n_qubytes != exclude->n_qubytes) {
        return set_error(QMB_ERR_ARGUMENT, "exclude set has a different config width");
    }
    return dispatch_words(terms->n_words, [&](auto wc) -> int {
        constexpr int W = decltype(wc)::value;
        return find_relative_impl<W>(terms, exclude, configs_i, psi_i, n_i, count_selected, out, out_rows, scratch, scratch_bytes,
                                     stream);
    });
}

}  // namespace qmb
