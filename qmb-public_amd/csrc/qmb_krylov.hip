// qmb_krylov.hip -- the vector work of one Lanczos step, fused into three HBM passes.
//
// The reference's loop (qmb/haar.py:123-142, qmb/rldiag.py:23-44) runs, after each application
// w = H v_k:
//     alpha_k = Re(w^H v_k);  w -= alpha_k v_k + beta_{k-1} v_{k-1};  beta_k = |w|;  v_{k+1} = w / beta_k
// as five separate tensor operations (a dot, two axpys with temporaries, a norm, a division).  Here
// the same arithmetic takes three kernels over the complex128 vectors (float64 pairs):
//   1. dot:    per-block partial sums of Re(conj(w) v);
//   2. update: every block reduces the dot partials to alpha (same order in every block, so the
//              value is identical everywhere and from run to run), updates w in place and writes
//              per-block partial sums of |w|^2;
//   3. scale:  every block reduces the norm partials, v_next = w / |w|.
// Block 0 stores alpha and |w| in `scalars` for the host, which reads both with one copy.
// All reductions run in a fixed order, so a step is deterministic.  The work is HBM-bound:
// 32 + 64 (48 without v_prev) + 32 bytes per amplitude.

#include "qmb_internal.hpp"

namespace qmb {
namespace {

constexpr int kKrylovThreads = 256;
constexpr int kKrylovMaxBlocks = 1024;

int krylov_blocks(int64_t n) {
    const int64_t per_block = static_cast<int64_t>(kKrylovThreads) * 8;
    const int64_t g = (n + per_block - 1) / per_block;
    return static_cast<int>(g < 1 ? 1 : (g > kKrylovMaxBlocks ? kKrylovMaxBlocks : g));
}

// Sum over the block, result valid in every thread.  Butterfly within each wave, then the four
// wave sums in index order: a fixed association, independent of timing.
__device__ double block_sum(double x, double* lds) {
    for (int m = 32; m >= 1; m >>= 1) {
        x += __shfl_xor(x, m);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[wave] = x;
    }
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < kKrylovThreads / 64; ++k) {
        s += lds[k];
    }
    __syncthreads();
    return s;
}

// Sum of partials[0..g) in a fixed order, result valid in every thread.
__device__ double reduce_partials(const double* partials, int g, double* lds) {
    double s = 0.0;
    for (int b = threadIdx.x; b < g; b += kKrylovThreads) {
        s += partials[b];
    }
    return block_sum(s, lds);
}

__global__ void __launch_bounds__(kKrylovThreads) lanczos_dot_kernel(const double2* __restrict__ w, const double2* __restrict__ v,
                                                                     int64_t n, double* __restrict__ partial) {
    __shared__ double lds[kKrylovThreads / 64];
    double s = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kKrylovThreads;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kKrylovThreads + threadIdx.x; i < n; i += stride) {
        const double2 a = w[i];
        const double2 b = v[i];
        s += a.x * b.x + a.y * b.y;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kKrylovThreads) lanczos_update_kernel(double2* __restrict__ w, const double2* __restrict__ v,
                                                                        const double2* __restrict__ v_prev, double beta, int64_t n,
                                                                        const double* __restrict__ dot_partial,
                                                                        double* __restrict__ norm_partial, double* __restrict__ scalars) {
    __shared__ double lds[kKrylovThreads / 64];
    const double alpha = reduce_partials(dot_partial, gridDim.x, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[0] = alpha;
    }
    double s = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kKrylovThreads;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kKrylovThreads + threadIdx.x; i < n; i += stride) {
        double2 x = w[i];
        const double2 y = v[i];
        x.x -= alpha * y.x;
        x.y -= alpha * y.y;
        if (v_prev) {
            const double2 z = v_prev[i];
            x.x -= beta * z.x;
            x.y -= beta * z.y;
        }
        w[i] = x;
        s += x.x * x.x + x.y * x.y;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        norm_partial[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kKrylovThreads) lanczos_scale_kernel(const double2* __restrict__ w, int64_t n,
                                                                       const double* __restrict__ norm_partial,
                                                                       double2* __restrict__ v_next, double* __restrict__ scalars) {
    __shared__ double lds[kKrylovThreads / 64];
    const double norm = sqrt(reduce_partials(norm_partial, gridDim.x, lds));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[1] = norm;
    }
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kKrylovThreads;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kKrylovThreads + threadIdx.x; i < n; i += stride) {
        const double2 x = w[i];
        // a zero residual ends the recurrence on the host; its v_next is never used, keep it finite
        v_next[i] = norm > 0.0 ? make_double2(x.x / norm, x.y / norm) : make_double2(0.0, 0.0);
    }
}

}  // namespace

size_t lanczos_bytes() {
    Carver cv;
    cv.take<double>(kKrylovMaxBlocks);
    cv.take<double>(kKrylovMaxBlocks);
    return cv.off;
}

int lanczos_step(int64_t n, const double* v, const double* v_prev, double beta, double* w, double* v_next, double* scalars,
                 void* workspace, size_t bytes, hipStream_t stream) {
    if (n < 0) {
        return set_error(QMB_ERR_ARGUMENT, "n must be non-negative, got %lld", static_cast<long long>(n));
    }
    if (!scalars || (n > 0 && (!v || !w || !v_next))) {
        return set_error(QMB_ERR_ARGUMENT, "v, w, v_next and scalars must be non-null");
    }
    if (!workspace || bytes < lanczos_bytes()) {
        return set_error(QMB_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", bytes, lanczos_bytes());
    }
    if (n > 0 && (v_next == v || v_next == w || (v_prev && (v_next == v_prev || w == v_prev)) || w == v)) {
        return set_error(QMB_ERR_ARGUMENT, "v, v_prev, w and v_next must be distinct vectors");
    }
    Carver cv{static_cast<char*>(workspace)};
    double* dot_partial = cv.take<double>(kKrylovMaxBlocks);
    double* norm_partial = cv.take<double>(kKrylovMaxBlocks);
    const int g = krylov_blocks(n);
    auto* w2 = reinterpret_cast<double2*>(w);
    const auto* v2 = reinterpret_cast<const double2*>(v);
    const auto* vp2 = reinterpret_cast<const double2*>(v_prev);
    hipLaunchKernelGGL(lanczos_dot_kernel, dim3(g), dim3(kKrylovThreads), 0, stream, w2, v2, n, dot_partial);
    QMB_CHECK_LAUNCH();
    hipLaunchKernelGGL(lanczos_update_kernel, dim3(g), dim3(kKrylovThreads), 0, stream, w2, v2, vp2, beta, n, dot_partial, norm_partial,
                       scalars);
    QMB_CHECK_LAUNCH();
    hipLaunchKernelGGL(lanczos_scale_kernel, dim3(g), dim3(kKrylovThreads), 0, stream, w2, n, norm_partial,
                       reinterpret_cast<double2*>(v_next), scalars);
    QMB_CHECK_LAUNCH();
    return QMB_OK;
}

}  // namespace qmb
