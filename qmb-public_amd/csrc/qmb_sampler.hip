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
