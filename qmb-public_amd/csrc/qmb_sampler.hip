// qmb_sampler.hip -- the element-wise and bit-level work of the autoregressive electron up/down
// network (qmb/mlp.py:61-314): its forward amplitude and its unique Gumbel-top-k sampler.
//
// The network sees a config as a sequence of sites in sampling order; site k of that sequence is
// physical site ordering[k], and its two bits (spin up, spin down) are config bits 2 ordering[k]
// and 2 ordering[k] + 1.  During sampling a prefix of i sites is kept as a bit string, two bits
// per site in sampling order (bit 2k: up, bit 2k + 1: down), in uint64 words.
//
// Per site the network emits four logits, child c = 2 up + down.  A child is allowed when the
// prefix can still reach the target electron counts (qmb/mlp.py:113-154); the allowed logits are
// normalised so that sum exp(2 x) = 1 (qmb/mlp.py:156-167).  The forward sums the selected
// normalised logit over the sites in double precision (qmb/mlp.py:169-205).  The sampler extends
// every kept prefix by its four children with Gumbel-perturbed log-probabilities conditioned on the
// parent's (qmb/mlp.py:247-270); the caller keeps the best `batch` children.
//
// Logit arrays use a site-major, row-minor layout ([S][4][B] for the forward, [4][n] per sampling
// step), the layout the transposed GEMMs of the network produce, so every access here is coalesced.
#include "qmb_internal.hpp"

namespace qmb {
namespace {

constexpr int kSamplerThreads = 256;

inline unsigned sampler_blocks(int64_t n) {
    const int64_t b = (n + kSamplerThreads - 1) / kSamplerThreads;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

__device__ __forceinline__ int prefix_bit(const uint64_t* row, int bit) { return static_cast<int>((row[bit >> 6] >> (bit & 63)) & 1ull); }

__device__ __forceinline__ int config_bit(const uint8_t* row, int bit) { return (row[bit >> 3] >> (bit & 7)) & 1; }

// Allowed children after a prefix of `done` sites holding `up` and `down` electrons.
__device__ __forceinline__ void allowed_children(int done, int up, int down, int n_sites, int spin_up, int spin_down, bool ok[4]) {
    const bool up_e = up < spin_up;
    const bool down_e = down < spin_down;
    const bool up_h = done - up < n_sites - spin_up;
    const bool down_h = done - down < n_sites - spin_down;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        ok[c] = ((c >> 1) ? up_e : up_h) && ((c & 1) ? down_e : down_h);
    }
}

// x[c] <- x[c] - log(sum_c exp(2 x[c])) / 2 with disallowed children at -inf.
template <typename T>
__device__ __forceinline__ void normalise(T x[4], const bool ok[4]) {
    T s = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        x[c] = ok[c] ? x[c] : -static_cast<T>(INFINITY);
        s += exp(2 * x[c]);
    }
    const T param = log(s) / 2;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        x[c] = x[c] - param;
    }
}

// Uniform in (0, 1) for (seed, step, index): Philox4x32-10 with the step in the third counter word.
__device__ __forceinline__ double sampler_uniform(uint64_t seed, uint32_t step, uint64_t index) {
    U32x4 ctr{{static_cast<uint32_t>(index), static_cast<uint32_t>(index >> 32), step, 0x5A4D504Cu}};
    const U32x4 r = philox4x32_10(ctr, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const uint64_t bits = ((static_cast<uint64_t>(r.v[0]) << 32) | r.v[1]) >> 11;
    return (static_cast<double>(bits) + 0.5) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ void __launch_bounds__(kSamplerThreads)
    children_kernel(int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* __restrict__ prefix, int words,
                    const T* __restrict__ logits, const T* __restrict__ unperturbed, const T* __restrict__ perturbed,
                    const T* __restrict__ uniform, uint64_t seed, T* __restrict__ child_unperturbed, T* __restrict__ child_perturbed) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (r >= n) {
        return;
    }
    int up = 0, down = 0;
    const uint64_t* row = prefix + r * words;
    for (int w = 0; w < words; ++w) {
        const int lo = 64 * w;
        if (lo >= 2 * site) {
            break;
        }
        const int width = 2 * site - lo;
        const uint64_t valid = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
        const uint64_t bits = row[w] & valid;
        up += __popcll(bits & 0x5555555555555555ull);
        down += __popcll(bits & 0xAAAAAAAAAAAAAAAAull);
    }
    bool ok[4];
    allowed_children(site, up, down, n_sites, spin_up, spin_down, ok);
    T x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        x[c] = logits[c * n + r];
    }
    normalise(x, ok);
    const T base = unperturbed[r];
    const T parent = perturbed[r];
    T l[4], big_l[4];
    T z = -static_cast<T>(INFINITY);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        l[c] = base + 2 * x[c];
        const T u = uniform ? uniform[c * n + r] : static_cast<T>(sampler_uniform(seed, static_cast<uint32_t>(site), static_cast<uint64_t>(r) * 4 + c));
        big_l[c] = l[c] - log(-log(u));
        z = big_l[c] > z ? big_l[c] : z;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        child_unperturbed[c * n + r] = l[c];
        child_perturbed[c * n + r] = -log(exp(-parent) - exp(-z) + exp(-big_l[c]));
    }
}

template <typename T>
__global__ void __launch_bounds__(kSamplerThreads)
    select_kernel(int64_t n_sel, const int64_t* __restrict__ index, int64_t n_parent, int site, const uint64_t* __restrict__ prefix,
                  int words, const T* __restrict__ child_unperturbed, const T* __restrict__ child_perturbed, uint64_t* __restrict__ out_prefix,
                  T* __restrict__ out_unperturbed, T* __restrict__ out_perturbed) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (k >= n_sel) {
        return;
    }
    const int64_t child = index[k];
    const int64_t c = child / n_parent;
    const int64_t r = child - c * n_parent;
    for (int w = 0; w < words; ++w) {
        out_prefix[k * words + w] = prefix[r * words + w];
    }
    const int bit = 2 * site;
    out_prefix[k * words + (bit >> 6)] |= static_cast<uint64_t>(c >> 1) << (bit & 63);
    out_prefix[k * words + ((bit + 1) >> 6)] |= static_cast<uint64_t>(c & 1) << ((bit + 1) & 63);
    out_unperturbed[k] = child_unperturbed[child];
    out_perturbed[k] = child_perturbed[child];
}

// Network input of a sampling step: the prefix bits as values, [2 site][n] (bit-major).
template <typename T>
__global__ void __launch_bounds__(kSamplerThreads)
    prefix_inputs_kernel(int64_t n, int site, const uint64_t* __restrict__ prefix, int words, T* __restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (r >= n) {
        return;
    }
    for (int b = 0; b < 2 * site; ++b) {
        out[b * n + r] = static_cast<T>(prefix_bit(prefix + r * words, b));
    }
}

// Sampled prefixes (all sites) -> configs: physical site j takes sampled site ordering[j].
__global__ void __launch_bounds__(kSamplerThreads)
    prefix_configs_kernel(int64_t n, int n_sites, const uint64_t* __restrict__ prefix, int words, const int64_t* __restrict__ ordering,
                          int n_qubytes, uint8_t* __restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (r >= n) {
        return;
    }
    const uint64_t* row = prefix + r * words;
    uint8_t* dst = out + r * n_qubytes;
    for (int byte = 0; byte < n_qubytes; ++byte) {
        uint32_t v = 0;
        for (int b = 0; b < 8; ++b) {
            const int bit = 8 * byte + b;
            const int j = bit >> 1;
            const int64_t o = j < n_sites ? ordering[j] : -1;
            if (o >= 0 && o < n_sites) {  // a permutation of the sites; anything else reads as empty
                v |= static_cast<uint32_t>(prefix_bit(row, 2 * static_cast<int>(o) + (bit & 1))) << b;
            }
        }
        dst[byte] = static_cast<uint8_t>(v);
    }
}

// Network input of the forward: config bits in sampling order, [2 S][B] (bit-major).
template <typename T>
__global__ void __launch_bounds__(kSamplerThreads)
    config_inputs_kernel(int64_t n, int n_sites, const uint8_t* __restrict__ configs, int n_qubytes,
                         const int64_t* __restrict__ ordering_reversed, T* __restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (r >= n) {
        return;
    }
    const uint8_t* row = configs + r * n_qubytes;
    for (int k = 0; k < n_sites; ++k) {
        const int64_t phys = ordering_reversed[k];
        const bool in = phys >= 0 && phys < n_sites;  // a permutation of the sites; anything else reads as empty
        out[(2 * k) * n + r] = static_cast<T>(in ? config_bit(row, 2 * static_cast<int>(phys)) : 0);
        out[(2 * k + 1) * n + r] = static_cast<T>(in ? config_bit(row, 2 * static_cast<int>(phys) + 1) : 0);
    }
}

// Forward amplitude: sum over sites (sampling order) of the selected normalised logit, in double.
// With grad != nullptr it instead writes d amplitude / d logits scaled by grad[r]:
// (delta(c, selected) - exp(2 x[c])) grad[r] for allowed c, 0 otherwise.
template <typename T>
__global__ void __launch_bounds__(kSamplerThreads)
    amplitude_kernel(int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* __restrict__ configs, int n_qubytes,
                     const int64_t* __restrict__ ordering_reversed, const T* __restrict__ logits, double* __restrict__ out,
                     const double* __restrict__ grad, T* __restrict__ grad_logits) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kSamplerThreads + threadIdx.x;
    if (r >= n) {
        return;
    }
    const uint8_t* row = configs + r * n_qubytes;
    int up = 0, down = 0;
    double total = 0.0;
    const double g = grad ? grad[r] : 0.0;
    for (int k = 0; k < n_sites; ++k) {
        bool ok[4];
        allowed_children(k, up, down, n_sites, spin_up, spin_down, ok);
        T x[4];
        const T* lg = logits + static_cast<int64_t>(k) * 4 * n;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = lg[c * n + r];
        }
        normalise(x, ok);
        const int64_t phys = ordering_reversed[k];
        const bool in = phys >= 0 && phys < n_sites;  // a permutation of the sites; anything else reads as empty
        const int a = in ? config_bit(row, 2 * static_cast<int>(phys)) : 0;
        const int b = in ? config_bit(row, 2 * static_cast<int>(phys) + 1) : 0;
        const int sel = 2 * a + b;
        if (grad) {
            T* gl = grad_logits + static_cast<int64_t>(k) * 4 * n;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T p = ok[c] ? exp(2 * x[c]) : static_cast<T>(0);
                gl[c * n + r] = static_cast<T>(((c == sel ? 1.0 : 0.0) - static_cast<double>(p)) * g);
            }
        } else {
            total += static_cast<double>(x[sel]);
        }
        up += a;
        down += b;
    }
    if (!grad) {
        out[r] = total;
    }
}

// ---------------------------------------------------------------------------------------------
// Fused one-hidden-layer site networks (float32).  A site network maps the 0/1 prefix bits x to
// W2 silu(W1 x + b1) + b2; with binary inputs W1 x is the sum of the W1 columns of the set bits,
// so the first layer is a handful of vector adds per row instead of a dense product, and the
// hidden activations never leave registers.  A work-group of 16 waves owns 1024 rows; per site it
// stages the site's weights in LDS (W1 as [k][h], so a bit's column is one contiguous row) and
// every wave walks its 64 rows one at a time with all 64 lanes across the hidden units (lane l
// holds units l, l + 64, ...: consecutive lanes read consecutive LDS words), then reduces the four
// outputs across the wave.  Each lane keeps its own
// row's logits, so the masks, normalisation and the rest run row-parallel afterwards.
// ---------------------------------------------------------------------------------------------
constexpr int kFusedWaves = 16;
constexpr int kFusedThreads = 64 * kFusedWaves;

__device__ __forceinline__ uint64_t wave_read64(uint64_t v, int lane) {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), lane);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// DPP lane exchange within rows of 16 lanes (quad_perm, row_half_mirror, row_mirror controls).
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the wave, the same value in every lane: a butterfly inside each row of 16 lanes on the
// DPP network, then the four row sums read as scalars and added in row order.
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp<0x141>(v);  // row_half_mirror
    v += dpp<0x140>(v);  // row_mirror
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return ((r0 + r1) + r2) + r3;
}

// LDS image of one site network: w1 [k_rows][H], b1 [H], w2 [n_out][H], b2 [n_out].
struct SiteLds {
    float* w1;
    float* b1;
    float* w2;
    float* b2;
};

__device__ __forceinline__ SiteLds site_lds(float* base, int k_max, int hidden) {
    SiteLds s;
    s.w1 = base;
    s.b1 = s.w1 + static_cast<size_t>(k_max) * hidden;
    s.w2 = s.b1 + hidden;
    s.b2 = s.w2 + 4 * hidden;
    return s;
}

__device__ __forceinline__ void stage_site(const SiteLds& s, const float* w1, int k_rows, const float* b1, const float* w2, const float* b2,
                                           int n_out, int hidden) {
    for (int t = threadIdx.x; t < k_rows * hidden; t += blockDim.x) {
        s.w1[t] = w1[t];
    }
    for (int t = threadIdx.x; t < hidden; t += blockDim.x) {
        s.b1[t] = b1[t];
    }
    for (int t = threadIdx.x; t < n_out * hidden; t += blockDim.x) {
        s.w2[t] = w2[t];
    }
    if (threadIdx.x < n_out) {
        s.b2[threadIdx.x] = b2[threadIdx.x];
    }
}

// The lane's slice of a staged site network kept in registers: b1 and w2 for its hidden units.
template <int HPL>
struct SiteRegs {
    float b1[HPL];
    float w2[4][HPL];
    float b2[4];
};

template <int HPL>
__device__ __forceinline__ SiteRegs<HPL> site_regs(const SiteLds& s, int hidden, int n_out, int lane) {
    SiteRegs<HPL> r;
#pragma unroll
    for (int j = 0; j < HPL; ++j) {
        r.b1[j] = s.b1[lane + 64 * j];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            r.w2[c][j] = c < n_out ? s.w2[c * hidden + lane + 64 * j] : 0.f;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        r.b2[c] = c < n_out ? s.b2[c] : 0.f;
    }
    return r;
}

// The n_out outputs of a site network on one row (bits s0/s1, wave-uniform, already cut to the
// network's inputs), summed across the wave; every lane receives them.
template <int HPL>
__device__ __forceinline__ void site_row(const SiteLds& s, const SiteRegs<HPL>& g, int hidden, int n_out, uint64_t s0, uint64_t s1, int lane,
                                         float out[4]) {
    float acc[HPL];
#pragma unroll
    for (int j = 0; j < HPL; ++j) {
        acc[j] = g.b1[j];
    }
    for (int half = 0; half < 2; ++half) {
        uint64_t m = half ? s1 : s0;
        while (m) {
            const int k = 64 * half + __builtin_ctzll(m);
            m &= m - 1;
            const float* col = s.w1 + static_cast<size_t>(k) * hidden + lane;
#pragma unroll
            for (int j = 0; j < HPL; ++j) {
                acc[j] += col[64 * j];
            }
        }
    }
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < HPL; ++j) {
        const float h = acc[j] * __builtin_amdgcn_rcpf(1.0f + __expf(-acc[j]));  // SiLU
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            part[c] += h * g.w2[c][j];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        out[c] = c < n_out ? wave_sum(part[c]) + g.b2[c] : 0.f;
    }
}

__device__ __forceinline__ void cut_bits(uint64_t& s0, uint64_t& s1, int n_bits) {
    if (n_bits < 64) {
        s0 &= n_bits == 0 ? 0ull : (~0ull >> (64 - n_bits));
        s1 = 0;
    } else if (n_bits < 128) {
        s1 &= n_bits == 64 ? 0ull : (~0ull >> (128 - n_bits));
    }
}

// Forward of the whole network: log-amplitude (double) and phase per config.
template <int HPL>
__global__ void __launch_bounds__(kFusedThreads)
    network_fused_kernel(int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* __restrict__ configs, int n_qubytes,
                         const int64_t* __restrict__ ordering_reversed, const float* __restrict__ w1, const float* __restrict__ b1,
                         const float* __restrict__ w2, const float* __restrict__ b2, const float* __restrict__ wp1,
                         const float* __restrict__ bp1, const float* __restrict__ wp2, const float* __restrict__ bp2,
                         double* __restrict__ amplitude, float* __restrict__ phase) {
    extern __shared__ float lds[];
    constexpr int H = 64 * HPL;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kFusedThreads + threadIdx.x;
    const bool valid = row < n;
    // the own row's bits in sampling order
    uint64_t b0 = 0, b1w = 0;
    if (valid) {
        const uint8_t* cr = configs + row * n_qubytes;
        for (int k = 0; k < n_sites; ++k) {
            const int64_t phys = ordering_reversed[k];
            if (phys < 0 || phys >= n_sites) {
                continue;  // a permutation of the sites; anything else reads as empty
            }
#pragma unroll
            for (int sb = 0; sb < 2; ++sb) {
                const uint64_t bit = static_cast<uint64_t>(config_bit(cr, 2 * static_cast<int>(phys) + sb));
                const int pos = 2 * k + sb;
                if (pos < 64) {
                    b0 |= bit << pos;
                } else {
                    b1w |= bit << (pos - 64);
                }
            }
        }
    }
    const int64_t wave_row0 = static_cast<int64_t>(blockIdx.x) * kFusedThreads + 64 * wave;
    const int rows_here = static_cast<int>(n - wave_row0 < 0 ? 0 : (n - wave_row0 > 64 ? 64 : n - wave_row0));
    const SiteLds s = site_lds(lds, 2 * n_sites, H);
    double total = 0.0;
    int up = 0, down = 0;
    for (int site = 0; site <= n_sites; ++site) {
        const bool is_phase = site == n_sites;
        const int k_rows = is_phase ? 2 * n_sites : 2 * site;
        const int n_out = is_phase ? 1 : 4;
        __syncthreads();
        if (is_phase) {
            stage_site(s, wp1, k_rows, bp1, wp2, bp2, 1, H);
        } else {
            stage_site(s, w1 + static_cast<size_t>(site) * (site - 1) * H, k_rows, b1 + static_cast<size_t>(site) * H,
                       w2 + static_cast<size_t>(site) * 4 * H, b2 + 4 * site, 4, H);
        }
        __syncthreads();
        const SiteRegs<HPL> g = site_regs<HPL>(s, H, n_out, lane);
        float mine[4] = {0.f, 0.f, 0.f, 0.f};
        for (int r = 0; r < rows_here; ++r) {
            uint64_t s0 = wave_read64(b0, r), s1 = wave_read64(b1w, r);
            cut_bits(s0, s1, k_rows);
            float out[4];
            site_row<HPL>(s, g, H, n_out, s0, s1, lane, out);
            if (lane == r) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    mine[c] = out[c];
                }
            }
        }
        if (!valid) {
            continue;
        }
        if (is_phase) {
            phase[row] = mine[0];
        } else {
            bool ok[4];
            allowed_children(site, up, down, n_sites, spin_up, spin_down, ok);
            normalise(mine, ok);
            const int pos = 2 * site;
            const uint64_t w = pos < 64 ? b0 : b1w;
            const int a = static_cast<int>((w >> (pos & 63)) & 1ull);
            const int b = static_cast<int>((w >> ((pos + 1) & 63)) & 1ull);
            total += static_cast<double>(mine[2 * a + b]);
            up += a;
            down += b;
        }
    }
    if (valid) {
        amplitude[row] = total;
    }
}

// One sampling step at `site` with the site network fused in: children log-probabilities and
// Gumbel-perturbed values, as children_kernel computes them from given logits.
template <int HPL>
__global__ void __launch_bounds__(kFusedThreads)
    sampler_fused_kernel(int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* __restrict__ prefix, int words,
                         const float* __restrict__ w1, const float* __restrict__ b1, const float* __restrict__ w2,
                         const float* __restrict__ b2, const float* __restrict__ unperturbed, const float* __restrict__ perturbed,
                         const float* __restrict__ uniform, uint64_t seed, float* __restrict__ child_unperturbed,
                         float* __restrict__ child_perturbed) {
    extern __shared__ float lds[];
    constexpr int H = 64 * HPL;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kFusedThreads + threadIdx.x;
    const bool valid = row < n;
    uint64_t b0 = 0, b1w = 0;
    if (valid) {
        b0 = prefix[row * words];
        b1w = words > 1 ? prefix[row * words + 1] : 0ull;
    }
    const int64_t wave_row0 = static_cast<int64_t>(blockIdx.x) * kFusedThreads + 64 * wave;
    const int rows_here = static_cast<int>(n - wave_row0 < 0 ? 0 : (n - wave_row0 > 64 ? 64 : n - wave_row0));
    const SiteLds s = site_lds(lds, 2 * site, H);
    stage_site(s, w1 + static_cast<size_t>(site) * (site - 1) * H, 2 * site, b1 + static_cast<size_t>(site) * H,
               w2 + static_cast<size_t>(site) * 4 * H, b2 + 4 * site, 4, H);
    __syncthreads();
    const SiteRegs<HPL> g = site_regs<HPL>(s, H, 4, lane);
    float mine[4] = {0.f, 0.f, 0.f, 0.f};
    for (int r = 0; r < rows_here; ++r) {
        uint64_t s0 = wave_read64(b0, r), s1 = wave_read64(b1w, r);
        cut_bits(s0, s1, 2 * site);
        float out[4];
        site_row<HPL>(s, g, H, 4, s0, s1, lane, out);
        if (lane == r) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                mine[c] = out[c];
            }
        }
    }
    if (!valid) {
        return;
    }
    const int up = __popcll(b0 & 0x5555555555555555ull) + __popcll(b1w & 0x5555555555555555ull);
    const int down = __popcll(b0 & 0xAAAAAAAAAAAAAAAAull) + __popcll(b1w & 0xAAAAAAAAAAAAAAAAull);
    bool ok[4];
    allowed_children(site, up, down, n_sites, spin_up, spin_down, ok);
    normalise(mine, ok);
    const float base = unperturbed[row];
    const float parent = perturbed[row];
    float l[4], big_l[4];
    float z = -INFINITY;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        l[c] = base + 2 * mine[c];
        const float u = uniform ? uniform[c * n + row] : static_cast<float>(sampler_uniform(seed, static_cast<uint32_t>(site), static_cast<uint64_t>(row) * 4 + c));
        big_l[c] = l[c] - log(-log(u));
        z = big_l[c] > z ? big_l[c] : z;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        child_unperturbed[c * n + row] = l[c];
        child_perturbed[c * n + row] = -log(exp(-parent) - exp(-z) + exp(-big_l[c]));
    }
}

template <typename F>
int by_hidden(int hidden, F&& f) {
    switch (hidden) {
        case 64:
            return f(std::integral_constant<int, 1>{});
        case 128:
            return f(std::integral_constant<int, 2>{});
        case 256:
            return f(std::integral_constant<int, 4>{});
        case 512:
            return f(std::integral_constant<int, 8>{});
        default:
            return set_error(QMB_ERR_UNSUPPORTED, "the fused network kernels take hidden widths 64, 128, 256 or 512, got %d", hidden);
    }
}

size_t fused_lds_bytes(int k_max, int hidden) { return (static_cast<size_t>(k_max) * hidden + 5 * static_cast<size_t>(hidden) + 4) * sizeof(float); }

constexpr size_t kFusedMaxLds = 160 * 1024;

template <typename F>
int by_dtype(int dtype, F&& f) {
    switch (dtype) {
        case 0:
            return f(float{});
        case 1:
            return f(double{});
        default:
            return set_error(QMB_ERR_ARGUMENT, "dtype must be 0 (float32) or 1 (float64), got %d", dtype);
    }
}

int check_sites(int n_sites, int spin_up, int spin_down) {
    if (n_sites < 1 || n_sites > 128) {
        return set_error(QMB_ERR_UNSUPPORTED, "n_sites must be in [1, 128], got %d", n_sites);
    }
    if (spin_up < 0 || spin_down < 0 || spin_up > n_sites || spin_down > n_sites) {
        return set_error(QMB_ERR_ARGUMENT, "electron counts (%d, %d) do not fit %d sites", spin_up, spin_down, n_sites);
    }
    return QMB_OK;
}

}  // namespace

int network_fused_supported(int n_sites, int hidden) {
    if (n_sites < 1 || n_sites > 64 || (hidden != 64 && hidden != 128 && hidden != 256 && hidden != 512)) {
        return 0;
    }
    return fused_lds_bytes(2 * n_sites, hidden) <= kFusedMaxLds ? 1 : 0;
}

int network_forward_fused(int64_t n, int n_sites, int hidden, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                          const int64_t* ordering_reversed, const float* w1, const float* b1, const float* w2, const float* b2,
                          const float* wp1, const float* bp1, const float* wp2, const float* bp2, double* amplitude, float* phase,
                          hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, spin_up, spin_down));
    if (!network_fused_supported(n_sites, hidden)) {
        return set_error(QMB_ERR_UNSUPPORTED, "no fused network kernel for %d sites and hidden width %d", n_sites, hidden);
    }
    if (n < 0 || n_qubytes < (2 * n_sites + 7) / 8) {
        return set_error(QMB_ERR_ARGUMENT, "configs of %d bytes cannot hold %d sites", n_qubytes, n_sites);
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!configs || !ordering_reversed || !w1 || !b1 || !w2 || !b2 || !wp1 || !bp1 || !wp2 || !bp2 || !amplitude || !phase) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    const size_t lds = fused_lds_bytes(2 * n_sites, hidden);
    const unsigned grid = static_cast<unsigned>((n + kFusedThreads - 1) / kFusedThreads);
    return by_hidden(hidden, [&](auto hc) -> int {
        constexpr int HPL = decltype(hc)::value;
        QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&network_fused_kernel<HPL>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(lds)));
        network_fused_kernel<HPL><<<grid, kFusedThreads, lds, stream>>>(n, n_sites, spin_up, spin_down, configs, n_qubytes, ordering_reversed,
                                                                       w1, b1, w2, b2, wp1, bp1, wp2, bp2, amplitude, phase);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int sampler_step_fused(int64_t n, int site, int n_sites, int hidden, int spin_up, int spin_down, const uint64_t* prefix, int words,
                       const float* w1, const float* b1, const float* w2, const float* b2, const float* unperturbed, const float* perturbed,
                       const float* uniform, uint64_t seed, float* child_unperturbed, float* child_perturbed, hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, spin_up, spin_down));
    if (!network_fused_supported(n_sites, hidden)) {
        return set_error(QMB_ERR_UNSUPPORTED, "no fused network kernel for %d sites and hidden width %d", n_sites, hidden);
    }
    if (site < 0 || site >= n_sites || words < (2 * n_sites + 63) / 64 || words > 2 || n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "bad sampling step: site %d of %d, %d prefix words, %lld rows", site, n_sites, words,
                         static_cast<long long>(n));
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!prefix || !w1 || !b1 || !w2 || !b2 || !unperturbed || !perturbed || !child_unperturbed || !child_perturbed) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    const size_t lds = fused_lds_bytes(2 * site, hidden);
    const unsigned grid = static_cast<unsigned>((n + kFusedThreads - 1) / kFusedThreads);
    return by_hidden(hidden, [&](auto hc) -> int {
        constexpr int HPL = decltype(hc)::value;
        QMB_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&sampler_fused_kernel<HPL>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(fused_lds_bytes(2 * n_sites, hidden))));
        sampler_fused_kernel<HPL><<<grid, kFusedThreads, lds, stream>>>(n, site, n_sites, spin_up, spin_down, prefix, words, w1, b1, w2, b2,
                                                                       unperturbed, perturbed, uniform, seed, child_unperturbed,
                                                                       child_perturbed);
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int sampler_children(int dtype, int64_t n, int site, int n_sites, int spin_up, int spin_down, const uint64_t* prefix, int words,
                     const void* logits, const void* unperturbed, const void* perturbed, const void* uniform, uint64_t seed,
                     void* child_unperturbed, void* child_perturbed, hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, spin_up, spin_down));
    if (site < 0 || site >= n_sites || words < (2 * n_sites + 63) / 64 || n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "bad sampling step: site %d of %d, %d prefix words, %lld rows", site, n_sites, words,
                         static_cast<long long>(n));
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!prefix || !logits || !unperturbed || !perturbed || !child_unperturbed || !child_perturbed) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    return by_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        children_kernel<T><<<sampler_blocks(n), kSamplerThreads, 0, stream>>>(
            n, site, n_sites, spin_up, spin_down, prefix, words, static_cast<const T*>(logits), static_cast<const T*>(unperturbed),
            static_cast<const T*>(perturbed), static_cast<const T*>(uniform), seed, static_cast<T*>(child_unperturbed),
            static_cast<T*>(child_perturbed));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int sampler_select(int dtype, int64_t n_sel, const int64_t* index, int64_t n_parent, int site, const uint64_t* prefix, int words,
                   const void* child_unperturbed, const void* child_perturbed, uint64_t* out_prefix, void* out_unperturbed,
                   void* out_perturbed, hipStream_t stream) {
    if (n_sel < 0 || n_sel > 4 * n_parent || site < 0 || 2 * site + 1 >= 64 * words) {
        return set_error(QMB_ERR_ARGUMENT, "bad selection: %lld of %lld children, site %d, %d words", static_cast<long long>(n_sel),
                         static_cast<long long>(4 * n_parent), site, words);
    }
    if (n_sel == 0) {
        return QMB_OK;
    }
    if (!index || !prefix || !child_unperturbed || !child_perturbed || !out_prefix || !out_unperturbed || !out_perturbed) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    return by_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        select_kernel<T><<<sampler_blocks(n_sel), kSamplerThreads, 0, stream>>>(
            n_sel, index, n_parent, site, prefix, words, static_cast<const T*>(child_unperturbed), static_cast<const T*>(child_perturbed),
            out_prefix, static_cast<T*>(out_unperturbed), static_cast<T*>(out_perturbed));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int sampler_prefix_inputs(int dtype, int64_t n, int site, const uint64_t* prefix, int words, void* out, hipStream_t stream) {
    if (n < 0 || site < 0 || 2 * site > 64 * words) {
        return set_error(QMB_ERR_ARGUMENT, "bad prefix inputs: site %d, %d words", site, words);
    }
    if (n == 0 || site == 0) {
        return QMB_OK;
    }
    if (!prefix || !out) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    return by_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        prefix_inputs_kernel<T><<<sampler_blocks(n), kSamplerThreads, 0, stream>>>(n, site, prefix, words, static_cast<T*>(out));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int sampler_prefix_configs(int64_t n, int n_sites, const uint64_t* prefix, int words, const int64_t* ordering, int n_qubytes,
                           uint8_t* out, hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, 0, 0));
    if (n < 0 || words < (2 * n_sites + 63) / 64 || n_qubytes < (2 * n_sites + 7) / 8) {
        return set_error(QMB_ERR_ARGUMENT, "bad prefix conversion: %d sites, %d words, %d bytes", n_sites, words, n_qubytes);
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!prefix || !ordering || !out) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    prefix_configs_kernel<<<sampler_blocks(n), kSamplerThreads, 0, stream>>>(n, n_sites, prefix, words, ordering, n_qubytes, out);
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

int network_inputs(int dtype, int64_t n, int n_sites, const uint8_t* configs, int n_qubytes, const int64_t* ordering_reversed, void* out,
                   hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, 0, 0));
    if (n < 0 || n_qubytes < (2 * n_sites + 7) / 8) {
        return set_error(QMB_ERR_ARGUMENT, "configs of %d bytes cannot hold %d sites", n_qubytes, n_sites);
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!configs || !ordering_reversed || !out) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    return by_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        config_inputs_kernel<T><<<sampler_blocks(n), kSamplerThreads, 0, stream>>>(n, n_sites, configs, n_qubytes, ordering_reversed,
                                                                                  static_cast<T*>(out));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

int network_amplitude(int dtype, int64_t n, int n_sites, int spin_up, int spin_down, const uint8_t* configs, int n_qubytes,
                      const int64_t* ordering_reversed, const void* logits, double* out, const double* grad, void* grad_logits,
                      hipStream_t stream) {
    QMB_TRY(check_sites(n_sites, spin_up, spin_down));
    if (n < 0 || n_qubytes < (2 * n_sites + 7) / 8) {
        return set_error(QMB_ERR_ARGUMENT, "configs of %d bytes cannot hold %d sites", n_qubytes, n_sites);
    }
    if ((grad == nullptr) != (grad_logits == nullptr)) {
        return set_error(QMB_ERR_ARGUMENT, "grad and grad_logits go together");
    }
    if (n == 0) {
        return QMB_OK;
    }
    if (!configs || !ordering_reversed || !logits || (!grad && !out)) {
        return set_error(QMB_ERR_ARGUMENT, "null pointer argument");
    }
    return by_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        amplitude_kernel<T><<<sampler_blocks(n), kSamplerThreads, 0, stream>>>(n, n_sites, spin_up, spin_down, configs, n_qubytes,
                                                                              ordering_reversed, static_cast<const T*>(logits), out, grad,
                                                                              static_cast<T*>(grad_logits));
        QMB_CHECK_LAUNCH();
        return QMB_OK;
    });
}

}  // namespace qmb
