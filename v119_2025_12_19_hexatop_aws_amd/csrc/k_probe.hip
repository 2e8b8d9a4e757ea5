// Device ceilings measured on the GPU the library runs on: a STREAM-style
// copy (b = a) and triad (a = b + s c) over three arrays of 'bytes' each, and
// an FP64 fused-multiply-add loop (independent chains per lane, no memory
// traffic).  The solve loop's kernels are rated against these beside the
// vendor figures (8 TB/s HBM3E, 78.6 TFLOP/s FP64).
#include "handle.h"
#include "hexa_internal.h"

namespace {

// One 16-byte element per lane over a grid that covers the arrays: measured
// the fastest form on MI355X (6.1 TB/s copy at 2 GiB per array, against 4.6
// for a grid-stride loop over 32 blocks per CU and 5.4-5.8 for 4-8 elements
// per lane).
__global__ void k_copy(int64_t n2, const double2* __restrict__ a, double2* __restrict__ b) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n2) b[i] = a[i];
}

__global__ void k_triad(int64_t n2, double s, const double2* __restrict__ b, const double2* __restrict__ c,
                        double2* __restrict__ a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n2) {
        const double2 x = b[i], y = c[i];
        a[i] = double2{__builtin_fma(s, y.x, x.x), __builtin_fma(s, y.y, x.y)};
    }
}

constexpr int FMA_CHAINS = 16;
constexpr int FMA_STEPS = 1024;

// FMA_CHAINS independent chains of FMA_STEPS fused multiply-adds per lane;
// the sum is stored only when it equals the sentinel, which it never does,
// so the chains cannot be removed.
__global__ __launch_bounds__(256) void k_fma(double a, double b, double sentinel, double* out) {
    double v[FMA_CHAINS];
#pragma unroll
    for (int c = 0; c < FMA_CHAINS; ++c) v[c] = (double)(threadIdx.x + c);
    for (int s = 0; s < FMA_STEPS; ++s) {
#pragma unroll
        for (int c = 0; c < FMA_CHAINS; ++c) v[c] = __builtin_fma(v[c], a, b);
    }
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < FMA_CHAINS; ++c) t += v[c];
    if (t == sentinel) out[0] = t;
}

}  // namespace

extern "C" int32_t hexa_device_ceilings(int64_t bytes, double* copy_gbps, double* triad_gbps, double* fma_tflops) {
    using namespace hexa;
    if (bytes < (1 << 20) || bytes > ((int64_t)1 << 36)) {
        set_error("probe array size must be in [1 MiB, 64 GiB]");
        return HEXA_E_INVALID;
    }
    const int64_t n2 = bytes / 16;
    double2 *a = nullptr, *b = nullptr, *c = nullptr;
    double* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int32_t rc = HEXA_OK;
    auto fail = [&](hipError_t e, const char* what) {
        set_error("%s failed: %s", what, hipGetErrorString(e));
        rc = e == hipErrorOutOfMemory ? HEXA_E_OOM : HEXA_E_DEVICE;
    };
    {
        DeviceOpGuard lock;
        hipError_t e;
        if ((e = hipMalloc(&a, n2 * 16)) != hipSuccess || (e = hipMalloc(&b, n2 * 16)) != hipSuccess ||
            (e = hipMalloc(&c, n2 * 16)) != hipSuccess || (e = hipMalloc(&out, 16)) != hipSuccess ||
            (e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
            fail(e, "probe set-up");
    }
    int dev = 0, ncu = 256;
    if (rc == HEXA_OK && hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned mblocks = (unsigned)((n2 + 255) / 256);   // n2 <= 2^32: at most 2^24 blocks
    auto time_ms = [&](auto&& launch, int reps) -> double {
        launch();
        launch();
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            fail(e, "probe timing");
            return 0.0;
        }
        return (double)ms / reps;
    };
    if (rc == HEXA_OK) {
        k_copy<<<mblocks, 256, 0, st>>>(n2, b, a);   // first touch
        k_copy<<<mblocks, 256, 0, st>>>(n2, a, c);
        const double tc = time_ms([&] { k_copy<<<mblocks, 256, 0, st>>>(n2, a, b); }, 20);
        const double tt = time_ms([&] { k_triad<<<mblocks, 256, 0, st>>>(n2, 3.0, b, c, a); }, 20);
        const unsigned fblocks = (unsigned)ncu * 16;
        const double tf = time_ms([&] { k_fma<<<fblocks, 256, 0, st>>>(1.0000001, 1e-9, -1.0, out); }, 5);
        if (rc == HEXA_OK && hipGetLastError() != hipSuccess) fail(hipErrorLaunchFailure, "probe launch");
        if (rc == HEXA_OK) {
            if (copy_gbps) *copy_gbps = 2.0 * n2 * 16 / (tc * 1e-3) / 1e9;
            if (triad_gbps) *triad_gbps = 3.0 * n2 * 16 / (tt * 1e-3) / 1e9;
            if (fma_tflops)
                *fma_tflops = 2.0 * FMA_CHAINS * FMA_STEPS * (double)fblocks * 256 / (tf * 1e-3) / 1e12;
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    DeviceOpGuard lock;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    for (void* p : {(void*)a, (void*)b, (void*)c, (void*)out})
        if (p) (void)hipFree(p);
    return rc;
}
