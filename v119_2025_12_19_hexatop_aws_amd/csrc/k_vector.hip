// Vector kernels and deterministic reductions.
#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

// Sum over a block of RED_THREADS; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < RED_THREADS / 64; ++i) t += lds[i];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double block_max(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double t = -INFINITY;
    if (threadIdx.x == 0)
        for (int i = 0; i < RED_THREADS / 64; ++i) t = fmax(t, lds[i]);
    __syncthreads();
    return t;
}

template <typename T>
__global__ void k_fill(int64_t n, T v, T* __restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A* __restrict__ x, B* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (B)x[i];
}

template <typename DT>
__device__ __forceinline__ float dinv_f(DT v) { return (float)v; }
template <>
__device__ __forceinline__ float dinv_f<uint16_t>(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

template <typename T, typename DT>
__global__ void k_scale_dinv(int64_t n, const T* __restrict__ b, const DT* __restrict__ dinv, T omega,
                             T* __restrict__ x, int64_t vs) {
    b += blockIdx.y * vs;   // right-hand side blockIdx.y of a batch (dinv shared)
    x += blockIdx.y * vs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (sizeof(DT) == sizeof(double)) x[i] = omega * (T)dinv[i] * b[i];
        else x[i] = omega * (T)dinv_f<DT>(dinv[i]) * b[i];
    }
}
template <>
__global__ void k_scale_dinv<double, double>(int64_t n, const double* __restrict__ b,
                                             const double* __restrict__ dinv, double omega,
                                             double* __restrict__ x, int64_t vs) {
    b += blockIdx.y * vs;
    x += blockIdx.y * vs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = omega * dinv[i] * b[i];
}

// bf16 with round-to-nearest-even.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename DT>
__device__ __forceinline__ DT store_dinv(double v) { return (DT)v; }
template <>
__device__ __forceinline__ uint16_t store_dinv<uint16_t>(double v) { return to_bf16((float)v); }

// dinv = 1/(diag + shift) (clamp_mode 0: the GPU path adds the shift,
// GPUSolver.jl:353) or 1/max(diag, shift) (clamp_mode 1: the CPU path
// clamps, CPUSolver.jl:212); 0 on held DOFs.
template <typename DT>
__global__ void k_make_dinv(int64_t n, const double* __restrict__ diag, const uint8_t* __restrict__ mask,
                            double shift, int clamp_mode, DT* __restrict__ dinv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const bool held = (mask[i / 3] >> (i % 3)) & 1;
        const double d = clamp_mode ? (diag[i] <= shift ? shift : diag[i]) : diag[i] + shift;
        dinv[i] = held ? store_dinv<DT>(0.0) : store_dinv<DT>(1.0 / d);
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_dot2(int64_t n, const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
        s += a[i] * b[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// max over free DOFs of diag; also flags NaN: a NaN anywhere yields NaN.
__global__ __launch_bounds__(RED_THREADS) void k_max_free(int64_t nd, const double* __restrict__ d,
                                                          const uint8_t* __restrict__ mask,
                                                          double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double m = -INFINITY;
    int nan = 0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < nd; i += (int64_t)RED_BLOCKS * RED_THREADS) {
        if ((mask[i / 3] >> (i % 3)) & 1) continue;
        const double v = d[i];
        if (isnan(v)) nan = 1;
        else m = fmax(m, v);
    }
    m = block_max(nan ? INFINITY : m, lds);
    // NaN is reported as +inf in the partial and resolved below
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
    double nn = block_sum((double)nan, lds);
    if (threadIdx.x == 0) partials[RED_BLOCKS + blockIdx.x] = nn;
}

__global__ __launch_bounds__(RED_THREADS) void k_final_max(const double* __restrict__ partials,
                                                           double* __restrict__ out) {
    __shared__ double lds[RED_THREADS / 64];
    double m = -INFINITY, nn = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) {
        m = fmax(m, partials[i]);
        nn += partials[RED_BLOCKS + i];
    }
    m = block_max(m, lds);
    nn = block_sum(nn, lds);
    if (threadIdx.x == 0) out[0] = nn > 0.0 ? NAN : m;
}

__global__ __launch_bounds__(RED_THREADS) void k_final_sum(int nq, const double* __restrict__ partials,
                                                           double* __restrict__ out) {
    __shared__ double lds[RED_THREADS / 64];
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) s += partials[q * RED_BLOCKS + i];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[q] = s;
    }
}

template <typename T>
__global__ void k_axpy(int64_t n, T a, const T* __restrict__ x, T* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

}  // namespace

template <typename T>
void launch_fill(hipStream_t s, int64_t n, T v, T* x) {
    k_fill<T><<<grid_for(n), 256, 0, s>>>(n, v, x);
}
template <typename A, typename B>
void launch_convert(hipStream_t s, int64_t n, const A* x, B* y) {
    k_convert<A, B><<<grid_for(n), 256, 0, s>>>(n, x, y);
}
template <typename T, typename DT>
void launch_scale_dinv(hipStream_t s, int64_t n, const T* b, const DT* dinv, T omega, T* x, int nb) {
    k_scale_dinv<T, DT><<<dim3(grid_for(n), nb), 256, 0, s>>>(n, b, dinv, omega, x, batch_stride(n));
}
template <typename DT>
void launch_make_dinv(hipStream_t s, int64_t n, const double* diag, const uint8_t* mask, double shift,
                      int clamp_mode, DT* dinv) {
    k_make_dinv<DT><<<grid_for(n), 256, 0, s>>>(n, diag, mask, shift, clamp_mode, dinv);
}
void launch_dot2(hipStream_t s, int64_t n, const double* a, const double* b, double* partials) {
    k_dot2<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, a, b, partials);
}
void launch_max_free(hipStream_t s, const Grid3& g, const double* diag, const uint8_t* mask,
                     double* partials, double* out) {
    k_max_free<<<RED_BLOCKS, RED_THREADS, 0, s>>>(3 * g.nn, diag, mask, partials);
    k_final_max<<<1, RED_THREADS, 0, s>>>(partials, out);
}
void launch_final_sum(hipStream_t s, int nq, const double* partials, double* out) {
    k_final_sum<<<1, RED_THREADS, 0, s>>>(nq, partials, out);
}
template <typename T>
void launch_axpy(hipStream_t s, int64_t n, T a, const T* x, T* y) {
    k_axpy<T><<<grid_for(n), 256, 0, s>>>(n, a, x, y);
}

template void launch_fill<double>(hipStream_t, int64_t, double, double*);
template void launch_fill<float>(hipStream_t, int64_t, float, float*);
template void launch_fill<uint8_t>(hipStream_t, int64_t, uint8_t, uint8_t*);
template void launch_convert<double, float>(hipStream_t, int64_t, const double*, float*);
template void launch_convert<float, double>(hipStream_t, int64_t, const float*, double*);
template void launch_scale_dinv<double, double>(hipStream_t, int64_t, const double*, const double*,
                                                double, double*, int);
template void launch_scale_dinv<float, float>(hipStream_t, int64_t, const float*, const float*, float,
                                              float*, int);
template void launch_scale_dinv<float, uint16_t>(hipStream_t, int64_t, const float*, const uint16_t*,
                                                 float, float*, int);
template void launch_make_dinv<double>(hipStream_t, int64_t, const double*, const uint8_t*, double, int,
                                       double*);
template void launch_make_dinv<float>(hipStream_t, int64_t, const double*, const uint8_t*, double, int,
                                      float*);
template void launch_make_dinv<uint16_t>(hipStream_t, int64_t, const double*, const uint8_t*, double, int,
                                         uint16_t*);
template void launch_axpy<double>(hipStream_t, int64_t, double, const double*, double*);
template void launch_axpy<float>(hipStream_t, int64_t, float, const float*, float*);

}  // namespace hexa
