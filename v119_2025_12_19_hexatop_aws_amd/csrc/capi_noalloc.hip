// hexa_oc_update_noalloc: the OC multisection update of hexa_oc_update on
// device arrays, for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its device scratch -- the multisection's OC_NL x RED_BLOCKS partial sums,
// the lambda bracket and the four reported values -- is a file-scope
// __device__ array of fixed size, present on every device the library is
// loaded on for the life of the process.  The per-element (q, lo, hi) values
// that hexa_oc_update keeps in work arrays are formed from x, dc and dv
// inside each pass instead, so the scratch does not grow with the design.
// All calls on one device share that device's array: calls that overlap on
// one device (different streams or threads without ordering) race.
#include <mutex>

#include "handle.h"
#include "hexa_internal.h"

namespace {

constexpr int OC_PART = hexa::OC_NL * hexa::RED_BLOCKS;   // partial sums
constexpr int OC_BRACKET = 16;                            // bracket (doubles)
constexpr int OC_OUT = 4;                                 // max change, volume, lambda, passes
constexpr int SCRATCH_DOUBLES = OC_PART + OC_BRACKET + OC_OUT;
// the volume target volfrac * n and the pass sums count elements in FP64
constexpr int64_t MAX_ELEMS = (int64_t)1 << 53;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

// The device address of this device's scratch array (looked up once per
// device, with that device current); the array's name is only a host-side
// stand-in and must not be passed to a kernel.
int32_t scratch_ptr(double** out) {
    int dev = 0;
    HX_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        hexa::set_error("device %d out of range for the scratch table", dev);
        return HEXA_E_DEVICE;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        HX_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            hexa::set_error("scratch symbol has no device address");
            return HEXA_E_DEVICE;
        }
        g_ptr[dev] = (double*)p;
    }
    *out = g_ptr[dev];
    return HEXA_OK;
}

}  // namespace

extern "C" int32_t hexa_oc_update_noalloc(int64_t n, double* d_x, const double* d_dc, const double* d_dv,
                                          const double* d_w, double volfrac, double move, double bisect_tol,
                                          void* stream, double* lambda_out) {
    using namespace hexa;
    if (!d_x || !d_dc || !d_dv || !d_w) { set_error("null device array"); return HEXA_E_INVALID; }
    if (n < 1 || n > MAX_ELEMS) {
        set_error("element count %lld outside [1, 2^53] (the volume target is counted in FP64)", (long long)n);
        return HEXA_E_INVALID;
    }
    if (!(volfrac > 0 && volfrac <= 1) || !(move > 0) || !(bisect_tol > 0)) {
        set_error("invalid OC parameters");
        return HEXA_E_INVALID;
    }
    static_assert(OC_BRACKET * (int)sizeof(double) >= 64, "bracket slot too small");
    if (oc_bracket_bytes() > OC_BRACKET * (int)sizeof(double)) {
        set_error("OC bracket does not fit its scratch slot");
        return HEXA_E_DEVICE;
    }
    double* base = nullptr;
    HX_TRY(scratch_ptr(&base));
    hipStream_t s = (hipStream_t)stream;
    double* partials = base;
    void* bracket = base + OC_PART;
    double* out4 = base + OC_PART + OC_BRACKET;
    launch_oc_direct(s, n, d_dc, d_dv, d_w, volfrac, move, bisect_tol, 48, d_x, bracket, partials, out4);
    HX_TRY(check_launch());
    if (lambda_out) {
        double out[OC_OUT];
        HX_CHECK(hipMemcpyAsync(out, out4, sizeof out, hipMemcpyDeviceToHost, s));
        HX_CHECK(hipStreamSynchronize(s));
        *lambda_out = out[2];
    }
    return HEXA_OK;
}
