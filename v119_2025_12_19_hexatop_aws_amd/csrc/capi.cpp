// libhexa C ABI: handle management and the linear-solve driver.
//
// The solve follows the control flow of the reference GPU driver
// gpu_matrix_free_cg_solve (R/src/Solvers/GPUSolver.jl:182-465) or, with
// HEXA_SEMANTICS_CPU, of the CPU matrix-free solver matrix_free_cg_solve
// (R/src/Solvers/CPUSolver.jl:192-255), with every vector and scalar kept on
// the device.  The multigrid preconditioner is a geometric V-cycle over the
// structured hierarchy (damped Jacobi smoothing, trilinear transfers,
// rediscretised coarse operators, dense direct solve on the coarsest grid),
// in FP64 or in FP32 with BF16-stored smoother diagonals.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "handle.h"
#include "hexa_internal.h"
#include "ke.h"

namespace hexa {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

std::shared_mutex& capture_mutex() {
    static std::shared_mutex m;
    return m;
}

}  // namespace hexa

using namespace hexa;

namespace hexa {

int32_t dalloc_bytes(hexa_handle* h, void** p, size_t bytes) {
    if (*p) return HEXA_OK;
    hipError_t e;
    {
        DeviceOpGuard lock;
        e = hipMalloc(p, bytes < 16 ? 16 : bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        return HEXA_E_OOM;
    }
    h->allocs.push_back(*p);
    h->bytes_allocated += (int64_t)(bytes < 16 ? 16 : bytes);
    return HEXA_OK;
}

int32_t check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("kernel launch failed: %s", hipGetErrorString(e));
        return HEXA_E_DEVICE;
    }
    return HEXA_OK;
}

int32_t need_ready(hexa_handle* h) {
    if (!h) { set_error("null handle"); return HEXA_E_INVALID; }
    if (!h->have_problem) { set_error("hexa_set_problem has not been called"); return HEXA_E_INVALID; }
    if (!h->have_density) { set_error("hexa_set_density has not been called"); return HEXA_E_INVALID; }
    return HEXA_OK;
}

int32_t poll_state(hexa_handle* h, int k, bool* halted) {
    const int slot = k % RING;
    HX_CHECK(hipMemcpyAsync(&h->h_state[slot], h->d_state, sizeof(PcgState), hipMemcpyDeviceToHost, h->stream));
    HX_CHECK(hipEventRecord(h->ring_ev[slot], h->stream));
    *halted = false;
    if (k - LOOKAHEAD >= 1) {
        const int kk = k - LOOKAHEAD;
        HX_CHECK(hipEventSynchronize(h->ring_ev[kk % RING]));
        *halted = h->h_state[kk % RING].halted != PCG_RUNNING;
    }
    return HEXA_OK;
}

}  // namespace hexa

namespace {

// Elasticity operator on one grid: factored row-wave kernel for the exact
// brick stiffness, dense-Ke kernel otherwise (e.g. the reference's
// f32-widened Ke).
// With dot_part, the factored kernel also writes per-block partial sums of
// (u, y); the return value is their count (0: not provided).
// nb > 1: a batch of load cases, vectors batch_stride(3 nn) apart.
template <typename T, typename DT>
int elastic_k(hexa_handle* h, bool factored, const Grid3& g, const FactCoef<T>& coef, const T* Ke, OpMode mode,
              const T* fac, const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y,
              double* dot_part, int nb) {
    if (factored)
        return launch_elastic_rw<T, DT>(h->stream, g, mode, coef, fac, mask, u, b, dinv, omega, shift, y, dot_part,
                                        nb, h->vc_halt, h->vc_halt_stride);
    const int64_t vs = batch_stride(3 * g.nn);
    for (int rb = 0; rb < nb; ++rb)
        launch_elastic_op<T, DT>(h->stream, g, mode, Ke, fac, mask, u ? u + rb * vs : u, b ? b + rb * vs : b, dinv,
                                 omega, shift, y + rb * vs);
    return 0;
}

// The handle's operator (h->factored decides the kernel).
template <typename T, typename DT>
int elastic(hexa_handle* h, const Grid3& g, const FactCoef<T>& coef, const T* Ke, OpMode mode, const T* fac,
            const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y,
            double* dot_part = nullptr, int nb = 1) {
    return elastic_k<T, DT>(h, h->factored, g, coef, Ke, mode, fac, mask, u, b, dinv, omega, shift, y, dot_part, nb);
}

// The V-cycle's level operator (h->vc_factored decides the kernel).
template <typename T, typename DT>
int vc_elastic(hexa_handle* h, const Grid3& g, const FactCoef<T>& coef, const T* Ke, OpMode mode, const T* fac,
               const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y, int nb = 1,
               double* dot_part = nullptr) {
    return elastic_k<T, DT>(h, h->vc_factored, g, coef, Ke, mode, fac, mask, u, b, dinv, omega, shift, y, dot_part,
                            nb);
}

int32_t ensure_fac32(hexa_handle* h) {
    if (h->fac32_valid) return HEXA_OK;
    HX_TRY(dalloc(h, &h->d_fac32, (size_t)h->g.ne));
    HX_TRY(dalloc(h, &h->d_Ke32, 576));
    float k32[576];
    for (int i = 0; i < 576; ++i) k32[i] = (float)h->Ke[i];
    HX_CHECK(hipMemcpyAsync(h->d_Ke32, k32, sizeof k32, hipMemcpyHostToDevice, h->stream));
    launch_convert<double, float>(h->stream, h->g.ne, h->d_fac, h->d_fac32);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(h->stream));
    h->fac32_valid = true;
    return HEXA_OK;
}

// ------------------------------------------------------------------ MG
}  // namespace

// Number of multigrid levels for an nx x ny x nz grid: coarsen (ceil(n/2) per
// axis) at least once, and until the coarsest grid has at most
// opt.mg_coarse_dofs (default HEXA_COARSE_DOFS) DOFs, or is a single element;
// or to opt.mg_levels.  The coarsest level's DOF count goes to *coarse_dofs.
namespace hexa {
int mg_level_count(int nx, int ny, int nz, const hexa_options& opt, int64_t* coarse_dofs) {
    const int want = opt.mg_levels;
    int64_t cap = opt.mg_coarse_dofs > 0 ? opt.mg_coarse_dofs : HEXA_COARSE_DOFS;
    if (opt.mg_coarse_dofs == 0) {
        const char* e = getenv("HEXA_COARSE_DOFS");   // tuning override of the default
        if (e && atoll(e) > 0) cap = std::min<int64_t>(atoll(e), 20000);
    }
    int L = 1;
    auto dofs = [&] { return 3LL * (nx + 1) * (ny + 1) * (nz + 1); };
    while ((want > 0 ? L < want : (L < 2 || dofs() > cap)) && (nx > 1 || ny > 1 || nz > 1)) {
        nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
        ++L;
    }
    if (coarse_dofs) *coarse_dofs = dofs();
    return L;
}
}  // namespace hexa

namespace {

int choose_levels(const hexa_handle* h) { return mg_level_count(h->g.nx, h->g.ny, h->g.nz, h->opt, nullptr); }

bool mixed(const hexa_handle* h) { return h->opt.precision == HEXA_PREC_MIXED; }

// Builds the hierarchy's shift-independent part: grids, Ke, factors, masks,
// diagonals, work vectors.
int32_t mg_build(hexa_handle* h) {
    ++h->graph_gen;
    // the grid transfers index nodes with 32 bits
    if (h->g.nn > (int64_t)UINT32_MAX) {
        set_error("multigrid needs fewer than 2^32 nodes");
        return HEXA_E_INVALID;
    }
    const int L = choose_levels(h);
    const bool mx = mixed(h);
    // a rebuild (after every density change) reuses the levels' buffers: the
    // level grids depend only on the handle's grid and options
    if ((int)h->levels.size() != L) h->levels.assign(L, Level{});
    for (int l = 0; l < L; ++l) {
        Level& lv = h->levels[l];
        lv.g = l == 0 ? h->g
                      : make_grid((h->levels[l - 1].g.nx + 1) / 2, (h->levels[l - 1].g.ny + 1) / 2,
                                  (h->levels[l - 1].g.nz + 1) / 2);
        lv.kscale = std::ldexp(1.0, l);
        const int64_t nd = 3 * lv.g.nn;
        double ke[576];
        for (int i = 0; i < 576; ++i) ke[i] = lv.kscale * h->Ke_vc[i];
        HX_TRY(dalloc(h, &lv.Ke64, 576));
        HX_CHECK(hipMemcpyAsync(lv.Ke64, ke, sizeof ke, hipMemcpyHostToDevice, h->stream));
        lv.c64 = make_fact_coef<double>(lv.kscale * h->grid.dx, lv.kscale * h->grid.dy, lv.kscale * h->grid.dz, h->opt.nu);
        lv.c32 = make_fact_coef<float>(lv.kscale * h->grid.dx, lv.kscale * h->grid.dy, lv.kscale * h->grid.dz, h->opt.nu);
        if (mx) {
            float k32[576];
            for (int i = 0; i < 576; ++i) k32[i] = (float)ke[i];
            HX_TRY(dalloc(h, &lv.Ke32, 576));
            HX_CHECK(hipMemcpyAsync(lv.Ke32, k32, sizeof k32, hipMemcpyHostToDevice, h->stream));
        }
        if (l == 0) {
            lv.fac64 = h->d_fac;
            lv.mask = h->d_mask;
        } else {
            Level& up = h->levels[l - 1];
            HX_TRY(dalloc(h, &lv.fac64, (size_t)lv.g.ne));
            HX_TRY(dalloc(h, &lv.mask, (size_t)lv.g.nn));
            launch_coarsen_factor<double>(h->stream, up.g, lv.g, up.fac64, lv.fac64);
            launch_coarsen_mask(h->stream, up.g, lv.g, up.mask, lv.mask);
        }
        if (mx) {
            if (l == 0) {
                HX_TRY(ensure_fac32(h));
                lv.fac32 = h->d_fac32;
            } else {
                HX_TRY(dalloc(h, &lv.fac32, (size_t)lv.g.ne));
                launch_convert<double, float>(h->stream, lv.g.ne, lv.fac64, lv.fac32);
            }
        }
        HX_TRY(dalloc(h, &lv.diag, (size_t)nd));
        launch_diag<double>(h->stream, lv.g, lv.Ke64, lv.fac64, lv.mask, 0, lv.diag);
        const size_t es = mx ? sizeof(float) : sizeof(double);
        const bool coarsest = (l == L - 1);
        // the level's work vectors hold a batch of load cases when nrhs > 1
        const size_t nv = h->opt.nrhs > 1 ? (size_t)batch_stride(nd) * h->opt.nrhs : (size_t)nd;
        if (!coarsest) {
            HX_TRY(dalloc(h, &lv.dinv, (size_t)nd * (mx ? sizeof(uint16_t) : sizeof(double))));
            HX_TRY(dalloc(h, &lv.xt, nv * es));
            HX_TRY(dalloc(h, &lv.r, nv * es));
        }
        if (l > 0 || mx) HX_TRY(dalloc(h, &lv.b, nv * es));
        if (l > 0 || mx) HX_TRY(dalloc(h, &lv.x, nv * es));
        HX_TRY(check_launch());
    }
    Level& c = h->levels[L - 1];
    h->n_coarse = 3 * c.g.nn;
    if (h->n_coarse > 20000) {
        set_error("coarsest multigrid level has %lld DOFs (> 20000); use more levels (mg_levels)",
                  (long long)h->n_coarse);
        return HEXA_E_INVALID;
    }
    HX_TRY(dalloc(h, &h->d_Ainv, (size_t)(h->n_coarse * h->n_coarse)));
    HX_TRY(dalloc(h, &h->d_gj_row, (size_t)gj_row_doubles(h->n_coarse)));
    HX_TRY(dalloc(h, &h->d_gj_col, (size_t)gj_col_doubles(h->n_coarse)));
    return HEXA_OK;
}

// Shift-dependent part: smoother diagonals and the coarsest inverse.  The
// shift s of the fine operator is carried to level l as 8^l s (row sum of
// P^T P for the trilinear transfer).
int32_t mg_set_shift(hexa_handle* h, double s) {
    if (h->mg_shift == s) return HEXA_OK;
    const bool mx = mixed(h);
    const int L = (int)h->levels.size();
    for (int l = 0; l < L; ++l) {
        Level& lv = h->levels[l];
        lv.shift = s * std::pow(8.0, l);
        if (l < L - 1) {
            if (mx) launch_make_dinv<uint16_t>(h->stream, 3 * lv.g.nn, lv.diag, lv.mask, lv.shift, 0,
                                               (uint16_t*)lv.dinv);
            else launch_make_dinv<double>(h->stream, 3 * lv.g.nn, lv.diag, lv.mask, lv.shift, 0,
                                          (double*)lv.dinv);
        }
    }
    Level& c = h->levels[L - 1];
    launch_fill<double>(h->stream, h->n_coarse * h->n_coarse, 0.0, h->d_Ainv);
    launch_coarse_assemble(h->stream, c.g, c.Ke64, c.fac64, c.mask, c.shift, h->d_Ainv);
    // the inversion's launches replay from a graph captured once per handle
    // and coarse size
    const char* ge = getenv("HEXA_GRAPHS");
    if (ge && ge[0] == '0') {
        launch_gauss_jordan(h->stream, h->n_coarse, h->d_Ainv, h->d_gj_row, h->d_gj_col);
    } else {
        HX_TRY(check_launch());
        if (!(h->gj_exec && h->gj_n == h->n_coarse && h->gj_A == h->d_Ainv)) {
            CaptureGuard lock;
            if (h->gj_exec) {
                HX_CHECK(hipGraphExecDestroy(h->gj_exec));
                h->gj_exec = nullptr;
            }
            HX_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            launch_gauss_jordan(h->stream, h->n_coarse, h->d_Ainv, h->d_gj_row, h->d_gj_col);
            const hipError_t el = hipGetLastError();
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &graph);
            if (el != hipSuccess || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                set_error("coarse inverse capture failed: %s", hipGetErrorString(el != hipSuccess ? el : e));
                return HEXA_E_DEVICE;
            }
            const hipError_t ei = hipGraphInstantiate(&h->gj_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                h->gj_exec = nullptr;
                set_error("coarse inverse graph instantiation failed: %s", hipGetErrorString(ei));
                return HEXA_E_DEVICE;
            }
            h->gj_n = h->n_coarse;
            h->gj_A = h->d_Ainv;
        }
        HX_CHECK(hipGraphLaunch(h->gj_exec, h->stream));
    }
    HX_TRY(check_launch());
    h->mg_shift = s;
    return HEXA_OK;
}

// nb > 1: a batch of load cases (b, out and the level vectors hold nb
// vectors, batch_stride(3 nn) apart on each level).
// With dots, the last smoothing sweep also writes partial sums of
// (b, out) there; the return value is their count (0: not written, e.g.
// with no post-smoothing sweep).
template <typename T, typename DT>
int vcycle_level(hexa_handle* h, int l, const T* b, T* out, int nb = 1, double* dots = nullptr) {
    Level& lv = h->levels[l];
    const int L = (int)h->levels.size();
    hipStream_t s = h->stream;
    const int64_t nd = 3 * lv.g.nn;
    if (l == L - 1) {
        launch_dense_gemv<T>(s, h->n_coarse, h->d_Ainv, b, out, nb, h->vc_halt, h->vc_halt_stride);
        return 0;
    }
    const T* Ke = std::is_same<T, float>::value ? (const T*)lv.Ke32 : (const T*)lv.Ke64;
    const T* fac = std::is_same<T, float>::value ? (const T*)lv.fac32 : (const T*)lv.fac64;
    const FactCoef<T>& coef = *reinterpret_cast<const FactCoef<T>*>(
        std::is_same<T, float>::value ? (const void*)&lv.c32 : (const void*)&lv.c64);
    const DT* dinv = (const DT*)lv.dinv;
    const T omega = (T)h->opt.mg_omega;
    const T shift = (T)lv.shift;
    const int pre = std::max(1, h->opt.mg_pre), post = std::max(0, h->opt.mg_post);
    // the first sweep from x = 0 is omega D^-1 b; with a second pre-sweep the
    // factored kernel does both in one pass (OP_SMOOTH0) without forming it
    const bool fuse0 = pre >= 2 && h->vc_factored && elastic_rw_supported(lv.g, std::max(sizeof(T), sizeof(DT)));
    const int first = fuse0 ? 2 : 1;
    const int swaps = (pre - first) + post;
    T* xa = (swaps % 2 == 0) ? out : (T*)lv.xt;
    T* xb = (swaps % 2 == 0) ? (T*)lv.xt : out;
    if (fuse0) vc_elastic<T, DT>(h, lv.g, coef, Ke, OP_SMOOTH0, fac, lv.mask, nullptr, b, dinv, omega, shift, xa, nb);
    else launch_scale_dinv<T, DT>(s, nd, b, dinv, omega, xa, nb);
    for (int it = first; it < pre; ++it) {
        vc_elastic<T, DT>(h, lv.g, coef, Ke, OP_SMOOTH, fac, lv.mask, xa, b, dinv, omega, shift, xb, nb);
        std::swap(xa, xb);
    }
    // coarse corrections: one (V-cycle), or two on every level but the
    // finest (W-cycle), the second from the residual the first left
    Level& nx = h->levels[l + 1];
    const int corrections = (h->opt.mg_cycle == HEXA_CYCLE_W && l > 0) ? 2 : 1;
    for (int cc = 0; cc < corrections; ++cc) {
        vc_elastic<T, DT>(h, lv.g, coef, Ke, OP_RESID, fac, lv.mask, xa, b, dinv, omega, shift, (T*)lv.r, nb);
        launch_restrict<T>(s, lv.g, nx.g, (const T*)lv.r, nx.mask, (T*)nx.b, nb, h->vc_halt, h->vc_halt_stride);
        vcycle_level<T, DT>(h, l + 1, (const T*)nx.b, (T*)nx.x, nb);
        launch_prolong_add<T>(s, lv.g, nx.g, (const T*)nx.x, lv.mask, xa, nb, h->vc_halt,
                              h->vc_halt_stride);
    }
    int nq = 0;
    for (int it = 0; it < post; ++it) {
        nq = vc_elastic<T, DT>(h, lv.g, coef, Ke, OP_SMOOTH, fac, lv.mask, xa, b, dinv, omega, shift, xb, nb,
                               it == post - 1 ? dots : nullptr);
        std::swap(xa, xb);
    }
    return nq;
}

// z = M^-1 r (FP64 in and out); nb load cases batch_stride(3 nn) apart.
// dots: as vcycle_level (FP64 V-cycle only; the return value is the count).
int apply_mg(hexa_handle* h, const double* r, double* z, int nb = 1, double* dots = nullptr) {
    if (h->levels.size() == 1) {
        launch_dense_gemv<double>(h->stream, h->n_coarse, h->d_Ainv, r, z, nb);
        return 0;
    }
    if (mixed(h)) {
        Level& l0 = h->levels[0];
        const int64_t nd = 3 * l0.g.nn, vs = batch_stride(nd);
        for (int rb = 0; rb < nb; ++rb)
            launch_convert<double, float>(h->stream, nd, r + rb * vs, (float*)l0.b + rb * vs);
        vcycle_level<float, uint16_t>(h, 0, (const float*)l0.b, (float*)l0.x, nb);
        for (int rb = 0; rb < nb; ++rb)
            launch_convert<float, double>(h->stream, nd, (const float*)l0.x + rb * vs, z + rb * vs);
        return 0;
    }
    return vcycle_level<double, double>(h, 0, r, z, nb, dots);
}


// Reference two-level preconditioner set-up (GPUGeometricMultigrid.jl:253-315):
// coarse grid of floor(n/2) (at least 1) bricks of twice the size, the
// inverse coarse diagonal from the x-entry diagonal of the coarse brick's
// Float32 Ke with E = 1 and nu = 0.3 (as :299-302), an all-free coarse mask.
int32_t refmg_build(hexa_handle* h) {
    ++h->graph_gen;
    if (h->g.nn > (int64_t)UINT32_MAX) {
        set_error("multigrid needs fewer than 2^32 nodes");
        return HEXA_E_INVALID;
    }
    if (h->refmg_ready) return HEXA_OK;
    const Grid3& f = h->g;
    h->rg = make_grid(std::max(1, f.nx / 2), std::max(1, f.ny / 2), std::max(1, f.nz / 2));
    const size_t ndc = (size_t)(3 * h->rg.nn);
    HX_TRY(dalloc(h, &h->d_rc, ndc));
    HX_TRY(dalloc(h, &h->d_xc, ndc));
    HX_TRY(dalloc(h, &h->d_dcinv, ndc));
    HX_TRY(dalloc_bytes(h, (void**)&h->d_rcmask, (size_t)h->rg.nn));
    double Kc[576], kd[8];
    element_stiffness(2.0 * h->grid.dx, 2.0 * h->grid.dy, 2.0 * h->grid.dz, 0.3, HEXA_KE_F32_WIDENED, Kc);
    for (int a = 0; a < 8; ++a) kd[a] = Kc[(3 * a) * 24 + 3 * a];
    HX_CHECK(hipMemsetAsync(h->d_rcmask, 0, (size_t)h->rg.nn, h->stream));
    launch_refmg_dcinv(h->stream, f, h->rg, h->d_mask, kd, h->d_dcinv);
    HX_TRY(check_launch());
    h->refmg_ready = true;
    return HEXA_OK;
}

// z = minv r + P (D_c + 1e-9)^-1 P^T r on the free DOFs
// (GPUGeometricMultigrid.jl:317-346).
void apply_refmg(hexa_handle* h, const double* r, double* z) {
    hipStream_t s = h->stream;
    const int64_t nd = 3 * h->g.nn;
    launch_scale_dinv<double, double>(s, nd, r, h->d_minv, 1.0, z);
    launch_restrict<double>(s, h->g, h->rg, r, h->d_rcmask, h->d_rc);
    launch_scale_dinv<double, double>(s, 3 * h->rg.nn, h->d_rc, h->d_dcinv, 1.0, h->d_xc);
    launch_prolong_add<double>(s, h->g, h->rg, h->d_xc, h->d_mask, z);
}

// b = f masked to the free DOFs.
__global__ void k_mask_copy(int64_t nd, const double* f, const uint8_t* __restrict__ mask, double* b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = ((mask[i / 3] >> (i % 3)) & 1) ? 0.0 : f[i];
}

unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

struct AttemptResult {
    int halted = PCG_RUNNING;
    int iterations = 0;
    double final_rel = 0.0, best_rel = INFINITY;
};

// While a PCG iteration is launched or captured: the operator and V-cycle
// kernels read the halted flags of the nb load cases' states, so that work
// queued after a case's solve ended returns at once (every other kernel of
// the iteration checks the state itself).
struct HaltScope {
    hexa_handle* h;
    HaltScope(hexa_handle* hh, int nb) : h(hh) {
        h->vc_halt = &h->d_state[0].halted;
        h->vc_halt_stride = nb > 1 ? (int)(sizeof(PcgState) / sizeof(int32_t)) : 0;
    }
    ~HaltScope() {
        h->vc_halt = nullptr;
        h->vc_halt_stride = 0;
    }
};

// One PCG attempt with shift s from the current d_x.  Returns the final
// device state (synchronised).
// pk: 0 Jacobi, 1 multigrid V-cycle, 2 reference two-level preconditioner.
int32_t run_attempt(hexa_handle* h, const hexa_solve_params* prm, int pk, double s,
                    AttemptResult* res, int64_t* matvecs) {
    const bool use_mg = pk == 1;
    hipStream_t st = h->stream;
    const int64_t nd = 3 * h->g.nn;
    const bool cpu = prm->semantics == HEXA_SEMANTICS_CPU;
    const int check_every = prm->check_every > 0 ? prm->check_every : (cpu ? 1 : 10);

    // minv and the MG shift
    launch_make_dinv<double>(st, nd, h->d_diag, h->d_mask, s, cpu ? 1 : 0, h->d_minv);
    if (use_mg) HX_TRY(mg_set_shift(h, s));

    // state, r = b - A x, z = M^-1 r, p = z, rz
    launch_dot2(st, nd, h->d_b, h->d_b, h->d_partials);
    launch_pcg_init_state(st, h->d_state, h->d_partials, prm->tol);
    elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_RESID, h->d_fac, h->d_mask, h->d_x, h->d_b,
                            nullptr, 0.0, s, h->d_r);
    ++*matvecs;
    if (pk == 1) apply_mg(h, h->d_r, h->d_z);
    else if (pk == 2) apply_refmg(h, h->d_r, h->d_z);
    else launch_scale_dinv<double, double>(st, nd, h->d_r, h->d_minv, 1.0, h->d_z);
    HX_CHECK(hipMemcpyAsync(h->d_p, h->d_z, nd * sizeof(double), hipMemcpyDeviceToDevice, st));
    launch_pcg_rz(st, nd, h->d_r, h->d_z, h->d_partials);
    launch_pcg_set_rz(st, h->d_state, h->d_partials);
    HX_TRY(check_launch());

    int k_done = 0;
    bool stop = false;
    // mixed V-cycle: the update writes its FP32 input and the dots and the
    // p-update read its FP32 output, so no FP64 copy of z is formed
    const bool mx = use_mg && mixed(h) && h->levels.size() > 1;
    float* r32 = mx ? (float*)h->levels[0].b : nullptr;
    const float* z32 = mx ? (const float*)h->levels[0].x : nullptr;
    // there, with CPU semantics (no best-iterate copy of x after the check),
    // x += alpha p moves from the update to the p update, which reads p anyway
    const bool defer_x = mx && cpu;
    double* const vdots = h->d_dotpart;
    // One iteration.  Everything that changes between iterations (the
    // iteration count, the check schedule, the deferred x update) is decided
    // on the device, so the launches are the same every iteration.
    auto iteration = [&]() -> int32_t {
        // the operator and the V-cycle of an iteration queued after the
        // solve ended do nothing (every other kernel of the iteration checks
        // the state itself)
        HaltScope halt_scope(h, 1);
        // Ap = (K + sI) p, with the per-block partials of (p, Ap)
        const int nq = elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, h->d_mask,
                                               h->d_p, nullptr, nullptr, 0.0, s, h->d_Ap, h->d_dotpart);
        if (nq > 0) {
            launch_pcg_alpha_n(st, h->d_state, h->d_dotpart, nq, cpu ? 0 : 1,
                               h->d_partials + (MAX_RED_Q - 1) * RED_BLOCKS);
        } else {
            launch_pcg_pAp(st, nd, h->d_state, h->d_p, h->d_Ap, h->d_partials);
            launch_pcg_alpha(st, h->d_state, h->d_partials, cpu ? 0 : 1);
        }
        launch_pcg_update(st, nd, h->d_state, defer_x ? nullptr : h->d_x, h->d_p, h->d_r, h->d_Ap,
                          pk == 0 ? h->d_minv : nullptr, h->d_partials, h->d_upart, r32);
        if (pk == 2) {
            // z = M r and (r, z) for the standard beta (GPUSolver.jl:409-417)
            apply_refmg(h, h->d_r, h->d_z);
            launch_pcg_rz(st, nd, h->d_r, h->d_z, h->d_partials + RED_BLOCKS);
        }
        launch_pcg_check(st, h->d_state, h->d_partials, cpu ? 1 : check_every, cpu ? 0 : 1, cpu ? 1 : 0,
                         use_mg ? 0 : 1);
        // (copies only when the check improved the best residual)
        if (!cpu) launch_pcg_best_copy(st, nd, h->d_state, h->d_x, h->d_best);
        // beta: the V-cycle (symmetric: as many sweeps after the coarse
        // correction as before) ends with a smoothing sweep that forms the
        // (r, z) partials (with the FP32 residual it smooths, in the mixed
        // V-cycle), and beta is the standard (r, z)_new / (r, z); with the
        // dense-Ke V-cycle kernels, the flexible beta from a separate pass
        // over r, z and Ap
        double* const bscr = h->d_partials + (MAX_RED_Q - 1) * RED_BLOCKS;
        if (mx) {
            const int nz = vcycle_level<float, uint16_t>(h, 0, r32, (float*)z32, 1, vdots);
            if (nz > 0) {
                launch_pcg_beta_n(st, h->d_state, h->d_dotpart, nz, bscr);
            } else {
                launch_pcg_flex_dots_z32(st, nd, h->d_state, h->d_r, z32, h->d_Ap, h->d_partials);
                launch_pcg_flex_beta(st, h->d_state, h->d_partials);
            }
            launch_pcg_p_update_z32(st, nd, h->d_state, h->d_p, z32, defer_x ? h->d_x : nullptr);
        } else if (use_mg) {
            const int nz = apply_mg(h, h->d_r, h->d_z, 1, vdots);
            if (nz > 0) {
                launch_pcg_beta_n(st, h->d_state, h->d_dotpart, nz, bscr);
            } else {
                launch_pcg_flex_dots(st, nd, h->d_state, h->d_r, h->d_z, h->d_Ap, h->d_partials);
                launch_pcg_flex_beta(st, h->d_state, h->d_partials);
            }
            launch_pcg_p_update(st, nd, h->d_state, h->d_p, h->d_z, nullptr, nullptr);
        } else if (pk == 2) {
            launch_pcg_p_update(st, nd, h->d_state, h->d_p, h->d_z, nullptr, nullptr);
        } else {
            launch_pcg_p_update(st, nd, h->d_state, h->d_p, nullptr, h->d_r, h->d_minv);
        }
        return check_launch();
    };
    // Replayed as a graph (one launch per iteration instead of tens; what
    // small grids are bound by), captured once per configuration.
    // HEXA_GRAPHS=0 launches eagerly.
    const char* ge = getenv("HEXA_GRAPHS");
    const bool graphs = !(ge && ge[0] == '0');
    if (graphs && !(h->pcg_exec && h->pcg_key.pk == pk && h->pcg_key.cpu == (int)cpu &&
                    h->pcg_key.every == check_every && h->pcg_key.s == s && h->pcg_key.gen == h->graph_gen)) {
        CaptureGuard lock;
        if (h->pcg_exec) {
            HX_CHECK(hipGraphExecDestroy(h->pcg_exec));
            h->pcg_exec = nullptr;
        }
        HX_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int32_t rc = iteration();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != HEXA_OK || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (rc != HEXA_OK) return rc;
            set_error("PCG iteration capture failed: %s", hipGetErrorString(e));
            return HEXA_E_DEVICE;
        }
        const hipError_t ei = hipGraphInstantiate(&h->pcg_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            h->pcg_exec = nullptr;
            set_error("PCG iteration graph instantiation failed: %s", hipGetErrorString(ei));
            return HEXA_E_DEVICE;
        }
        h->pcg_key.pk = pk;
        h->pcg_key.cpu = (int)cpu;
        h->pcg_key.every = check_every;
        h->pcg_key.s = s;
        h->pcg_key.gen = h->graph_gen;
    }
    for (int k = 1; k <= prm->max_iter && !stop; ++k) {
        if (graphs) HX_CHECK(hipGraphLaunch(h->pcg_exec, st));
        else HX_TRY(iteration());
        ++*matvecs;
        k_done = k;
        HX_TRY(poll_state(h, k, &stop));
    }
    HX_CHECK(hipMemcpyAsync(&h->h_state[0], h->d_state, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    const PcgState& fs = h->h_state[0];
    res->halted = fs.halted;
    res->iterations = fs.halted != PCG_RUNNING ? fs.last_k : k_done;
    res->final_rel = fs.final_rel;
    res->best_rel = fs.best_rel;
    return HEXA_OK;
}


int32_t solve_one(hexa_handle* h, int rhs, double* u_inout, const hexa_solve_params* prm,
                  hexa_solve_report* rep) {
    hipStream_t st = h->stream;
    const int64_t nd = 3 * h->g.nn;
    const bool cpu = prm->semantics == HEXA_SEMANTICS_CPU;
    hexa_solve_report R{};
    R.method = cpu ? HEXA_METHOD_JACOBI_CPU : HEXA_METHOD_JACOBI_GPU;
    const double t0 = now_ms();

    // diagonal and its maximum over free DOFs (GPUSolver.jl:297-315)
    const int xentry = h->opt.diag_mode == HEXA_DIAG_X_ENTRY && !cpu;
    launch_diag<double>(st, h->g, h->d_Ke64, h->d_fac, h->d_mask, xentry, h->d_diag);
    launch_max_free(st, h->g, h->d_diag, h->d_mask, h->d_partials, h->d_scal);
    k_mask_copy<<<grid_for(nd), 256, 0, st>>>(nd, h->d_f + rhs * nd, h->d_mask, h->d_b);
    HX_TRY(check_launch());
    double maxd = 0.0;
    HX_CHECK(hipMemcpyAsync(&maxd, h->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (std::isnan(maxd) || std::isinf(maxd)) {
        launch_fill<double>(st, nd, 0.0, h->d_x);
        HX_CHECK(hipStreamSynchronize(st));
        if (u_inout) std::memset(u_inout, 0, sizeof(double) * nd);
        R.rel_res = NAN;
        R.method = HEXA_METHOD_FAILED_CORRUPT_INPUT;
        if (rep) *rep = R;
        set_error("stiffness diagonal contains NaN/Inf");
        return HEXA_E_NAN_INPUT;
    }

    // multigrid kind: 1 V-cycle, 2 reference two-level preconditioner
    const int mg_kind = prm->precond == HEXA_PRECOND_MG ? 1 : (prm->precond == HEXA_PRECOND_REF_MG ? 2 : 0);
    const bool want_mg = mg_kind != 0;
    bool use_mg = want_mg;
    if (mg_kind == 1 && (h->precond_ready != HEXA_PRECOND_MG)) {
        int32_t s = mg_build(h);
        if (s != HEXA_OK) {
            if (s == HEXA_E_OOM) return s;
            use_mg = false;  // MG set-up failure falls back to Jacobi (GPUSolver.jl:259-267)
        } else {
            h->precond_ready = HEXA_PRECOND_MG;
            h->mg_shift = -1.0;
        }
    } else if (mg_kind == 2) {
        int32_t s = refmg_build(h);
        if (s != HEXA_OK) {
            if (s == HEXA_E_OOM) return s;
            use_mg = false;
        }
    }
    // the first attempt's smoother and coarse inverse belong to the set-up
    if (use_mg && mg_kind == 1) HX_TRY(mg_set_shift(h, prm->shift_factor * maxd));
    // z-chunk lengths of the operator kernels, timed once per grid and kernel
    // (d_x, d_best and the level buffers are scratch at this point)
    if (h->factored) tune_elastic_rw<double>(st, h->g, h->coef64, h->d_fac, h->d_x, h->d_best);
    if (h->vc_factored) {
        if (use_mg && mg_kind == 1)
            for (size_t l = 0; l + 1 < h->levels.size(); ++l) {
                Level& lv = h->levels[l];
                if (mixed(h))
                    tune_elastic_rw<float>(st, lv.g, lv.c32, lv.fac32, (const float*)lv.r, (float*)lv.xt);
                else
                    tune_elastic_rw<double>(st, lv.g, lv.c64, lv.fac64, (const double*)lv.r, (double*)lv.xt);
            }
        HX_TRY(check_launch());
    }
    // warm start: from the host buffer, or (u_inout == NULL) from the
    // device-resident solution of the previous solve on this handle
    const bool warm = prm->warm_start && (u_inout || h->have_solution);
    if (warm) {
        HX_TRY(dalloc(h, &h->d_guess, (size_t)nd));
        if (u_inout)
            HX_CHECK(hipMemcpyAsync(h->d_guess, u_inout, nd * sizeof(double), hipMemcpyHostToDevice, st));
        else
            HX_CHECK(hipMemcpyAsync(h->d_guess, h->d_x, nd * sizeof(double), hipMemcpyDeviceToDevice, st));
        k_mask_copy<<<grid_for(nd), 256, 0, st>>>(nd, h->d_guess, h->d_mask, h->d_guess);
    }
    launch_fill<double>(st, nd, 0.0, h->d_x);
    launch_fill<double>(st, nd, 0.0, h->d_best);
    HX_CHECK(hipStreamSynchronize(st));
    R.setup_ms = now_ms() - t0;
    const double t1 = now_ms();

    int64_t matvecs = 0;
    bool success = false;
    double final_rel = 0.0;
    int iters = 0, attempt_used = 1;
    double shift_used = 0.0;
    if (cpu) {
        // CPUSolver.jl:192-255: one attempt, x0 = 0, clamp, every-iteration check;
        // a zero load on the free DOFs returns u = 0 at once (:222-224)
        const double s = prm->shift_factor * maxd;
        launch_dot2(st, nd, h->d_b, h->d_b, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal + 3);
        HX_TRY(check_launch());
        double bb = 0.0;
        HX_CHECK(hipMemcpyAsync(&bb, h->d_scal + 3, sizeof(double), hipMemcpyDeviceToHost, st));
        HX_CHECK(hipStreamSynchronize(st));
        shift_used = s;
        if (bb == 0.0) {
            success = true;
            R.method = HEXA_METHOD_NONE;
        } else {
            AttemptResult ar;
            HX_TRY(run_attempt(h, prm, use_mg ? mg_kind : 0, s, &ar, &matvecs));
            success = ar.halted == PCG_CONVERGED;
            final_rel = ar.final_rel;
            iters = ar.iterations;
        }
    } else {
        // GPUSolver.jl:332-444
        while (true) {
            double cur = prm->shift_factor;
            bool ok = false, fallback = false;
            const int attempts = std::max(1, prm->max_shift_attempts);
            double best_rel = INFINITY;
            for (int attempt = 1; attempt <= attempts; ++attempt) {
                const double s = cur * maxd;
                attempt_used = attempt;
                shift_used = s;
                if (warm && attempt == 1 && use_mg == want_mg) {
                    HX_CHECK(hipMemcpyAsync(h->d_x, h->d_guess, nd * sizeof(double), hipMemcpyDeviceToDevice, st));
                } else if (attempt > 1 || use_mg != want_mg) {
                    launch_fill<double>(st, nd, 0.0, h->d_x);
                }
                AttemptResult ar;
                HX_TRY(run_attempt(h, prm, use_mg ? mg_kind : 0, s, &ar, &matvecs));
                iters = ar.iterations;
                final_rel = ar.final_rel;
                best_rel = ar.best_rel;
                const bool diverged = ar.halted == PCG_NAN || ar.halted == PCG_DIVERGED;
                if (diverged && use_mg) { fallback = true; break; }
                if (ar.halted == PCG_CONVERGED) { ok = true; break; }
                if (diverged && attempt != attempts) { cur *= prm->shift_multiplier; continue; }
                if (prm->stagnation_tol > 0.0 && best_rel < prm->stagnation_tol) {
                    ok = true;
                    HX_CHECK(hipMemcpyAsync(h->d_x, h->d_best, nd * sizeof(double), hipMemcpyDeviceToDevice, st));
                    break;
                }
                cur *= prm->shift_multiplier;
            }
            if (fallback) { use_mg = false; continue; }
            success = ok;
            if (!ok) HX_CHECK(hipMemcpyAsync(h->d_x, h->d_best, nd * sizeof(double), hipMemcpyDeviceToDevice, st));
            break;
        }
        R.method = use_mg ? HEXA_METHOD_GMG_GPU : HEXA_METHOD_JACOBI_GPU;
    }
    HX_CHECK(hipStreamSynchronize(st));
    R.solve_ms = now_ms() - t1;

    // compliance f.u and the true relative residual of the shifted system
    launch_dot2(st, nd, h->d_b, h->d_x, h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal);
    elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_RESID, h->d_fac, h->d_mask, h->d_x, h->d_b,
                            nullptr, 0.0, shift_used, h->d_r);
    launch_dot2(st, nd, h->d_r, h->d_r, h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal + 1);
    launch_dot2(st, nd, h->d_b, h->d_b, h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal + 2);
    HX_TRY(check_launch());
    double sc[3];
    HX_CHECK(hipMemcpyAsync(sc, h->d_scal, sizeof sc, hipMemcpyDeviceToHost, st));
    if (u_inout) HX_CHECK(hipMemcpyAsync(u_inout, h->d_x, nd * sizeof(double), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    h->have_solution = true;
    R.compliance = sc[0];
    R.true_rel_res = sc[2] > 0 ? std::sqrt(sc[1] / sc[2]) : 0.0;
    R.rel_res = final_rel;
    R.iterations = iters;
    R.attempt = attempt_used;
    R.shift_used = shift_used;
    R.converged = success ? 1 : 0;
    R.matvecs = matvecs;
    R.mg_levels = use_mg ? (mg_kind == 1 ? (int32_t)h->levels.size() : 2) : 0;
    h->last_mg = use_mg ? mg_kind : 0;
    if (rep) *rep = R;
    return success ? HEXA_OK : HEXA_E_NOT_CONVERGED;
}

// ------------------------------------------------------------ load cases
// Several load cases solved together: the PCG of CPUSolver.jl:192-255 with
// the V-cycle preconditioner, one recurrence per load case, advanced in lock
// step with every kernel of an iteration working on all load cases in one
// launch.  Each load case keeps its own state, scalars and partial sums in
// the same order as a single solve, so its iterates are bitwise those of
// solve_one.  An iteration launches as many kernels as a single solve's, each
// over all load cases: the coarse levels, where a single solve's kernels
// are too small to fill the GPU, gain the most.
constexpr int32_t BATCH_UNSUPPORTED = 1;
const int32_t k_converged_flag = PCG_CONVERGED;

// Device states of the nb load cases of iteration k - LOOKAHEAD: all halted?
int32_t poll_batch(hexa_handle* h, int k, int nb, bool* all_halted) {
    const int slot = k % RING;
    HX_CHECK(hipMemcpyAsync(&h->h_state[slot * nb], h->d_state, sizeof(PcgState) * nb, hipMemcpyDeviceToHost,
                            h->stream));
    HX_CHECK(hipEventRecord(h->ring_ev[slot], h->stream));
    *all_halted = false;
    if (k - LOOKAHEAD >= 1) {
        const int kk = (k - LOOKAHEAD) % RING;
        HX_CHECK(hipEventSynchronize(h->ring_ev[kk]));
        bool all = true;
        for (int r = 0; r < nb; ++r) all = all && h->h_state[kk * nb + r].halted != PCG_RUNNING;
        *all_halted = all;
    }
    return HEXA_OK;
}

int32_t run_attempt_batched(hexa_handle* h, const hexa_solve_params* prm, double s, int nb, const bool* active,
                            AttemptResult* res) {
    hipStream_t st = h->stream;
    const int64_t nd = 3 * h->g.nn, vs = batch_stride(nd);
    const int PS = BPART_STRIDE;
    launch_make_dinv<double>(st, nd, h->d_diag, h->d_mask, s, 1, h->d_minv);
    HX_TRY(mg_set_shift(h, s));
    for (int rb = 0; rb < nb; ++rb) {
        launch_dot2(st, nd, h->d_b + rb * vs, h->d_b + rb * vs, h->d_partials);
        launch_pcg_init_state(st, h->d_state + rb, h->d_partials, prm->tol);
    }
    elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_RESID, h->d_fac, h->d_mask, h->d_x, h->d_b, nullptr,
                            0.0, s, h->d_r, nullptr, nb);
    apply_mg(h, h->d_r, h->d_z, nb);
    HX_CHECK(hipMemcpyAsync(h->d_p, h->d_z, (size_t)(vs * nb) * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int rb = 0; rb < nb; ++rb) {
        launch_pcg_rz(st, nd, h->d_r + rb * vs, h->d_z + rb * vs, h->d_partials);
        launch_pcg_set_rz(st, h->d_state + rb, h->d_partials);
        // a zero load on the free DOFs is solved by u = 0 (CPUSolver.jl:222-224)
        if (!active[rb])
            HX_CHECK(hipMemcpyAsync(&h->d_state[rb].halted, &k_converged_flag, sizeof(int32_t),
                                    hipMemcpyHostToDevice, st));
    }
    HX_TRY(check_launch());
    const bool mx = mixed(h) && h->levels.size() > 1;
    float* r32 = mx ? (float*)h->levels[0].b : nullptr;
    float* z32 = mx ? (float*)h->levels[0].x : nullptr;
    const int64_t pstride = elastic_rw_partials(h->g);
    // the V-cycle's (r, z) partials (from its FP32 or FP64 level-0 kernels)
    const int64_t vstride = elastic_rw_partials(h->g, mx ? sizeof(float) : sizeof(double));
    double* const vdots = h->d_dotpart;
    auto iteration = [&]() -> int32_t {
        // a load case's operator and V-cycle work stops once its solve ended
        HaltScope halt_scope(h, nb);
        const int nq = elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, h->d_mask, h->d_p,
                                               nullptr, nullptr, 0.0, s, h->d_Ap, h->d_dotpart, nb);
        if (nq <= 0) {
            set_error("batched load cases need the row-wave operator");
            return HEXA_E_INVALID;
        }
        launch_pcg_alpha_n(st, h->d_state, h->d_dotpart, nq, 0, h->d_bscr, nb, pstride);
        // (mixed: x += alpha p is deferred to the p update, which reads p anyway)
        launch_pcg_update(st, nd, h->d_state, mx ? nullptr : h->d_x, h->d_p, h->d_r, h->d_Ap, nullptr, h->d_bpart,
                          h->d_upart, r32, nb, vs, PS);
        launch_pcg_check(st, h->d_state, h->d_bpart, 1, 0, 1, 0, nb, PS);
        // beta as in the single solve: from the V-cycle's last smoothing
        // sweep's (r, z) partials where it forms them
        if (mx) {
            const int nz = vcycle_level<float, uint16_t>(h, 0, r32, z32, nb, vdots);
            if (nz > 0) {
                launch_pcg_beta_n(st, h->d_state, h->d_dotpart, nz, h->d_bscr, nb, vstride);
            } else {
                launch_pcg_flex_dots_z32(st, nd, h->d_state, h->d_r, z32, h->d_Ap, h->d_bpart, nb, vs, PS);
                launch_pcg_flex_beta(st, h->d_state, h->d_bpart, nb, PS);
            }
            launch_pcg_p_update_z32(st, nd, h->d_state, h->d_p, z32, h->d_x, nb, vs);
        } else {
            const int nz = apply_mg(h, h->d_r, h->d_z, nb, vdots);
            if (nz > 0) {
                launch_pcg_beta_n(st, h->d_state, h->d_dotpart, nz, h->d_bscr, nb, vstride);
            } else {
                launch_pcg_flex_dots(st, nd, h->d_state, h->d_r, h->d_z, h->d_Ap, h->d_bpart, nb, vs, PS);
                launch_pcg_flex_beta(st, h->d_state, h->d_bpart, nb, PS);
            }
            launch_pcg_p_update(st, nd, h->d_state, h->d_p, h->d_z, nullptr, nullptr, nb, vs);
        }
        return check_launch();
    };
    const char* ge = getenv("HEXA_GRAPHS");
    const bool graphs = !(ge && ge[0] == '0');
    if (graphs && !(h->bpcg_exec && h->bpcg_key.nb == nb && h->bpcg_key.s == s && h->bpcg_key.gen == h->graph_gen)) {
        CaptureGuard lock;
        if (h->bpcg_exec) {
            HX_CHECK(hipGraphExecDestroy(h->bpcg_exec));
            h->bpcg_exec = nullptr;
        }
        HX_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int32_t rc = iteration();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != HEXA_OK || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (rc != HEXA_OK) return rc;
            set_error("batched PCG iteration capture failed: %s", hipGetErrorString(e));
            return HEXA_E_DEVICE;
        }
        const hipError_t ei = hipGraphInstantiate(&h->bpcg_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            h->bpcg_exec = nullptr;
            set_error("batched PCG iteration graph instantiation failed: %s", hipGetErrorString(ei));
            return HEXA_E_DEVICE;
        }
        h->bpcg_key.nb = nb;
        h->bpcg_key.s = s;
        h->bpcg_key.gen = h->graph_gen;
    }
    int k_done = 0;
    bool stop = false;
    for (int k = 1; k <= prm->max_iter && !stop; ++k) {
        if (graphs) HX_CHECK(hipGraphLaunch(h->bpcg_exec, st));
        else HX_TRY(iteration());
        k_done = k;
        HX_TRY(poll_batch(h, k, nb, &stop));
    }
    HX_CHECK(hipMemcpyAsync(h->h_state, h->d_state, sizeof(PcgState) * nb, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    for (int rb = 0; rb < nb; ++rb) {
        const PcgState& fs = h->h_state[rb];
        res[rb].halted = fs.halted;
        res[rb].iterations = fs.halted != PCG_RUNNING ? fs.last_k : k_done;
        res[rb].final_rel = fs.final_rel;
        res[rb].best_rel = fs.best_rel;
    }
    return HEXA_OK;
}

// All load cases of the handle with CPU semantics and the V-cycle, batched.
// Returns BATCH_UNSUPPORTED (nothing solved) when the hierarchy cannot be
// built, so that the caller solves them one by one (with the Jacobi fallback
// of solve_one).
int32_t solve_batched(hexa_handle* h, double* u_inout, const hexa_solve_params* prm, hexa_solve_report* rep) {
    hipStream_t st = h->stream;
    const int nb = h->opt.nrhs;
    const int64_t nd = 3 * h->g.nn, vs = batch_stride(nd);
    const double t0 = now_ms();
    launch_diag<double>(st, h->g, h->d_Ke64, h->d_fac, h->d_mask, 0, h->d_diag);
    launch_max_free(st, h->g, h->d_diag, h->d_mask, h->d_partials, h->d_scal);
    for (int rb = 0; rb < nb; ++rb) {
        k_mask_copy<<<grid_for(nd), 256, 0, st>>>(nd, h->d_f + rb * nd, h->d_mask, h->d_b + rb * vs);
        launch_dot2(st, nd, h->d_b + rb * vs, h->d_b + rb * vs, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal + 16 + rb);
    }
    HX_TRY(check_launch());
    std::vector<double> sc(16 + nb);
    HX_CHECK(hipMemcpyAsync(sc.data(), h->d_scal, sizeof(double) * (16 + nb), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    const double maxd = sc[0];
    if (std::isnan(maxd) || std::isinf(maxd)) return BATCH_UNSUPPORTED;   // solve_one reports it per case
    if (h->precond_ready != HEXA_PRECOND_MG) {
        const int32_t sb = mg_build(h);
        if (sb == HEXA_E_OOM) return sb;
        if (sb != HEXA_OK) return BATCH_UNSUPPORTED;
        h->precond_ready = HEXA_PRECOND_MG;
        h->mg_shift = -1.0;
    }
    const double s = prm->shift_factor * maxd;
    HX_TRY(mg_set_shift(h, s));
    tune_elastic_rw<double>(st, h->g, h->coef64, h->d_fac, h->d_x, h->d_best);
    for (size_t l = 0; l + 1 < h->levels.size(); ++l) {
        Level& lv = h->levels[l];
        if (mixed(h)) tune_elastic_rw<float>(st, lv.g, lv.c32, lv.fac32, (const float*)lv.r, (float*)lv.xt);
        else tune_elastic_rw<double>(st, lv.g, lv.c64, lv.fac64, (const double*)lv.r, (double*)lv.xt);
    }
    launch_fill<double>(st, vs * nb, 0.0, h->d_x);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(st));
    const double setup_ms = now_ms() - t0;
    const double t1 = now_ms();
    std::vector<AttemptResult> ar(nb);
    std::unique_ptr<bool[]> active(new bool[nb]);
    for (int rb = 0; rb < nb; ++rb) active[rb] = sc[16 + rb] != 0.0;
    HX_TRY(run_attempt_batched(h, prm, s, nb, active.get(), ar.data()));
    const double solve_ms = now_ms() - t1;
    bool all_ok = true;
    for (int rb = 0; rb < nb; ++rb) {
        double* x = h->d_x + rb * vs;
        double* b = h->d_b + rb * vs;
        double* r = h->d_r + rb * vs;
        launch_dot2(st, nd, b, x, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal);
        elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_RESID, h->d_fac, h->d_mask, x, b, nullptr, 0.0, s,
                                r);
        launch_dot2(st, nd, r, r, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal + 1);
        launch_dot2(st, nd, b, b, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal + 2);
        HX_TRY(check_launch());
        double v[3];
        HX_CHECK(hipMemcpyAsync(v, h->d_scal, sizeof v, hipMemcpyDeviceToHost, st));
        if (u_inout) HX_CHECK(hipMemcpyAsync(u_inout + rb * nd, x, nd * sizeof(double), hipMemcpyDeviceToHost, st));
        HX_CHECK(hipStreamSynchronize(st));
        hexa_solve_report R{};
        const bool none = !active[rb];
        const bool ok = none || ar[rb].halted == PCG_CONVERGED;
        all_ok = all_ok && ok;
        R.method = none ? HEXA_METHOD_NONE : HEXA_METHOD_JACOBI_CPU;
        R.compliance = v[0];
        R.true_rel_res = v[2] > 0 ? std::sqrt(v[1] / v[2]) : 0.0;
        R.rel_res = none ? 0.0 : ar[rb].final_rel;
        R.iterations = none ? 0 : ar[rb].iterations;
        R.attempt = 1;
        R.shift_used = s;
        R.converged = ok ? 1 : 0;
        R.matvecs = none ? 0 : ar[rb].iterations + 1;
        R.mg_levels = (int32_t)h->levels.size();
        // the batch's set-up and solve time, shared evenly by its load cases
        R.setup_ms = setup_ms / nb;
        R.solve_ms = solve_ms / nb;
        if (rep) rep[rb] = R;
    }
    h->have_solution = true;
    h->last_mg = 1;
    return all_ok ? HEXA_OK : HEXA_E_NOT_CONVERGED;
}

}  // namespace

// =================================================================== C ABI
namespace hexa {
// The reduction, scalar and PCG-state buffers that the solver, the filter
// and the element kernels share, at the sizes of the largest user (the
// batched solve of nrhs load cases), whichever call allocates them first.
int32_t ensure_shared_buffers(hexa_handle* h) {
    const int nb = std::max(1, h->opt.nrhs);
    HX_TRY(dalloc(h, &h->d_partials, (size_t)(MAX_RED_Q * RED_BLOCKS)));
    // the PCG update's per-block sums (two per block and load case)
    HX_TRY(dalloc(h, &h->d_upart, (size_t)(2 * pcg_update_blocks(3 * h->g.nn)) * nb));
    HX_TRY(dalloc(h, &h->d_scal, 16 + (size_t)nb));
    HX_TRY(dalloc(h, &h->d_state, (size_t)nb));
    return HEXA_OK;
}

int32_t ensure_solver_vectors(hexa_handle* h) {
    const size_t nd = (size_t)(3 * h->g.nn);
    // with several load cases the iterate vectors hold a batch of them,
    // batch_stride(nd) entries apart (the batched PCG)
    const int nb = h->opt.nrhs;
    const size_t nv = nb > 1 ? (size_t)batch_stride((int64_t)nd) * nb : nd;
    HX_TRY(dalloc(h, &h->d_x, nv));
    HX_TRY(dalloc(h, &h->d_r, nv));
    HX_TRY(dalloc(h, &h->d_p, nv));
    HX_TRY(dalloc(h, &h->d_Ap, nv));
    HX_TRY(dalloc(h, &h->d_z, nv));
    HX_TRY(dalloc(h, &h->d_minv, nd));
    HX_TRY(dalloc(h, &h->d_best, nd));
    HX_TRY(dalloc(h, &h->d_b, nv));
    HX_TRY(dalloc(h, &h->d_diag, nd));
    HX_TRY(ensure_shared_buffers(h));
    // (p, Ap) partials of the FP64 operator and (r, z) partials of the
    // V-cycle's FP32 or FP64 last sweep, which may tile the grid differently
    HX_TRY(dalloc(h, &h->d_dotpart,
                  (size_t)std::max<int64_t>({1, elastic_rw_partials(h->g, sizeof(double)),
                                             elastic_rw_partials(h->g, sizeof(float))}) * nb));
    if (nb > 1) {
        HX_TRY(dalloc(h, &h->d_bpart, (size_t)BPART_STRIDE * nb));
        HX_TRY(dalloc(h, &h->d_bscr, (size_t)RED_BLOCKS * nb));
    }
    return HEXA_OK;
}

int32_t update_factor(hexa_handle* h) {
    hipStream_t st = h->stream;
    HX_TRY(dalloc(h, &h->d_fac, (size_t)h->g.ne));
    const uint8_t* act = (h->have_problem && h->d_active) ? h->d_active : nullptr;
    launch_stiffness_factor<double>(st, h->g.ne, h->d_rho, act, h->opt.E0, h->opt.Emin, h->opt.penal,
                                    h->d_fac);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(st));
    h->have_density = true;
    h->fac32_valid = false;
    h->precond_ready = -1;
    return HEXA_OK;
}
}  // namespace hexa

extern "C" {

const char* hexa_last_error(void) { return g_err.c_str(); }

int32_t hexa_version(void) { return 100; }

void hexa_default_options(hexa_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->E0 = 1.0;
    o->Emin = 0.0;
    o->penal = 1.0;
    o->nu = 0.3;
    o->ke_mode = HEXA_KE_FP64;
    o->helm_element = HEXA_HELM_BUG_COMPAT;
    o->precision = HEXA_PREC_FP64;
    o->matvec_kernel = HEXA_MATVEC_AUTO;
    o->mg_levels = 0;
    o->mg_pre = 2;
    o->mg_post = 2;
    o->mg_omega = 0.6;
    o->nrhs = 1;
    o->diag_mode = HEXA_DIAG_EXACT;
    o->mg_cycle = HEXA_CYCLE_V;
    o->mg_coarse_dofs = 0;
}

void hexa_default_solve_params(hexa_solve_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->tol = 1e-8;
    p->max_iter = 30000;
    p->shift_factor = 1e-10;
    p->max_shift_attempts = 3;
    p->shift_multiplier = 10.0;
    p->stagnation_tol = 0.0;
    p->precond = HEXA_PRECOND_JACOBI;
    p->semantics = HEXA_SEMANTICS_GPU;
    p->warm_start = 0;
    p->check_every = 0;
}

int32_t hexa_device_info(char* name, int32_t len, int64_t* total_bytes, int32_t* n_cu) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device: %s", hipGetErrorString(e));
        return HEXA_E_DEVICE;
    }
    int dev = 0;
    HX_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HX_CHECK(hipGetDeviceProperties(&p, dev));
    if (name && len > 0) {
        std::snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (total_bytes) *total_bytes = (int64_t)p.totalGlobalMem;
    if (n_cu) *n_cu = p.multiProcessorCount;
    return HEXA_OK;
}

int32_t hexa_create(const hexa_grid* grid, const hexa_options* opt, hexa_handle** out) {
    if (!grid || !out) { set_error("null argument"); return HEXA_E_INVALID; }
    *out = nullptr;
    // node rows are launch-grid rows (y and z launch dimensions hold at most 65535)
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || grid->ny > 65534 || grid->nz > 65534 ||
        !(grid->dx > 0) || !(grid->dy > 0) || !(grid->dz > 0)) {
        set_error("invalid grid %d x %d x %d (h = %g, %g, %g)", grid->nx, grid->ny, grid->nz, grid->dx,
                  grid->dy, grid->dz);
        return HEXA_E_INVALID;
    }
    const int64_t nn = (int64_t)(grid->nx + 1) * (grid->ny + 1) * (grid->nz + 1);
    if (3 * nn > ((int64_t)1 << 40)) { set_error("grid too large"); return HEXA_E_INVALID; }
    hexa_options o;
    if (opt) o = *opt; else hexa_default_options(&o);
    if (o.nrhs < 1 || o.nrhs > 64) { set_error("nrhs must be in [1, 64]"); return HEXA_E_INVALID; }
    if (!(o.mg_omega > 0.0 && o.mg_omega < 2.0)) { set_error("mg_omega must be in (0, 2)"); return HEXA_E_INVALID; }
    // The V-cycle preconditions CG, which needs it symmetric positive: as
    // many Jacobi sweeps after the coarse correction as before.  (On an
    // 8x4x4 cantilever with the golden test's density, V(2,0) and V(2,1)
    // ended in NaN with the flexible beta too.)
    if (o.mg_pre < 1 || o.mg_pre > 16 || o.mg_post != o.mg_pre) {
        set_error("mg_pre must be in [1,16] and mg_post equal to it (a symmetric V-cycle)");
        return HEXA_E_INVALID;
    }
    if (!(o.penal > 0) || o.E0 < o.Emin) { set_error("invalid material law"); return HEXA_E_INVALID; }
    if (o.mg_cycle != HEXA_CYCLE_V && o.mg_cycle != HEXA_CYCLE_W) { set_error("invalid mg_cycle"); return HEXA_E_INVALID; }
    if (o.mg_coarse_dofs < 0 || o.mg_coarse_dofs > 20000) {
        set_error("mg_coarse_dofs must be in [0, 20000]");
        return HEXA_E_INVALID;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) { set_error("no HIP device available"); return HEXA_E_DEVICE; }
    hexa_handle* h = new hexa_handle();
    h->grid = *grid;
    h->opt = o;
    h->g = make_grid(grid->nx, grid->ny, grid->nz);
    if (!element_stiffness(grid->dx, grid->dy, grid->dz, o.nu, o.ke_mode, h->Ke)) {
        delete h;
        set_error("invalid element (nu must be in (-1, 0.5))");
        return HEXA_E_INVALID;
    }
    if (o.matvec_kernel == HEXA_MATVEC_FACTORED && o.ke_mode != HEXA_KE_FP64) {
        delete h;
        set_error("the factored element kernel needs ke_mode = HEXA_KE_FP64 (exact brick stiffness)");
        return HEXA_E_INVALID;
    }
    h->factored = o.matvec_kernel == HEXA_MATVEC_FACTORED ||
                  (o.matvec_kernel == HEXA_MATVEC_AUTO && o.ke_mode == HEXA_KE_FP64);
    h->vc_factored = o.matvec_kernel != HEXA_MATVEC_DENSE;
    if (h->vc_factored) (void)element_stiffness(grid->dx, grid->dy, grid->dz, o.nu, HEXA_KE_FP64, h->Ke_vc);
    else std::memcpy(h->Ke_vc, h->Ke, sizeof h->Ke_vc);
    h->coef64 = make_fact_coef<double>(grid->dx, grid->dy, grid->dz, o.nu);
    h->coef32 = make_fact_coef<float>(grid->dx, grid->dy, grid->dz, o.nu);
    auto fail = [&](int32_t s) { hexa_destroy(h); return s; };
    if (hipGetDevice(&h->device) != hipSuccess) return fail(HEXA_E_DEVICE);
    bool ok;
    {
        DeviceOpGuard lock;
        ok = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreate(&h->ev_a) == hipSuccess && hipEventCreate(&h->ev_b) == hipSuccess &&
             hipHostMalloc((void**)&h->h_state, sizeof(PcgState) * RING * o.nrhs, hipHostMallocDefault) == hipSuccess;
        for (int i = 0; ok && i < RING; ++i)
            ok = hipEventCreateWithFlags(&h->ring_ev[i], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        set_error("stream, event or pinned buffer creation failed");
        return fail(HEXA_E_DEVICE);
    }
    int32_t s = dalloc(h, &h->d_Ke64, 576);
    if (s != HEXA_OK) return fail(s);
    if (hipMemcpyAsync(h->d_Ke64, h->Ke, sizeof h->Ke, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("element matrix upload failed");
        return fail(HEXA_E_DEVICE);
    }
    *out = h;
    return HEXA_OK;
}

int32_t hexa_destroy(hexa_handle* h) {
    if (!h) return HEXA_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    DeviceOpGuard lock;
    if (h->pcg_exec) hipGraphExecDestroy(h->pcg_exec);
    if (h->bpcg_exec) hipGraphExecDestroy(h->bpcg_exec);
    if (h->filt_exec) hipGraphExecDestroy(h->filt_exec);
    if (h->gj_exec) hipGraphExecDestroy(h->gj_exec);
    for (void* p : h->allocs) hipFree(p);
    if (h->h_state) hipHostFree(h->h_state);
    for (int i = 0; i < RING; ++i)
        if (h->ring_ev[i]) hipEventDestroy(h->ring_ev[i]);
    if (h->ev_a) hipEventDestroy(h->ev_a);
    if (h->ev_b) hipEventDestroy(h->ev_b);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return HEXA_OK;
}

int32_t hexa_set_problem(hexa_handle* h, const uint8_t* fixed_dof, const double* f,
                         const uint8_t* active_elem) {
    if (!h || !f) { set_error("null argument"); return HEXA_E_INVALID; }
    return set_problem_impl(h, fixed_dof, f, 0, nullptr, nullptr, active_elem);
}

}  // extern "C"

namespace hexa {

// hexa_set_problem with the load given either densely (f, nd x nrhs) or as
// nsp entries f[sidx[i]] = sval[i] of the first right-hand side, the rest 0
// (uploads a few values instead of the whole vector).
int32_t set_problem_impl(hexa_handle* h, const uint8_t* fixed_dof, const double* f, int64_t nsp,
                         const int64_t* sidx, const double* sval, const uint8_t* active_elem) {
    const int64_t nd = 3 * h->g.nn;
    for (int64_t i = 0; i < nsp; ++i)
        if (sidx[i] < 0 || sidx[i] >= nd) { set_error("load index out of range"); return HEXA_E_INVALID; }
    HX_TRY(dalloc(h, &h->d_mask, (size_t)h->g.nn));
    HX_TRY(dalloc(h, &h->d_f, (size_t)(nd * h->opt.nrhs)));
    HX_TRY(dalloc(h, &h->d_fixed, (size_t)nd));
    hipStream_t st = h->stream;
    if (fixed_dof) HX_CHECK(hipMemcpyAsync(h->d_fixed, fixed_dof, (size_t)nd, hipMemcpyHostToDevice, st));
    else launch_fill<uint8_t>(st, nd, 0, h->d_fixed);
    if (active_elem) {
        HX_TRY(dalloc(h, &h->d_active, (size_t)h->g.ne));
        HX_CHECK(hipMemcpyAsync(h->d_active, active_elem, (size_t)h->g.ne, hipMemcpyHostToDevice, st));
    } else if (h->d_active) {
        launch_fill<uint8_t>(st, h->g.ne, 1, h->d_active);   // a previous active set no longer applies
    }
    const uint8_t* act = active_elem ? h->d_active : nullptr;
    launch_node_mask(st, h->g, h->d_fixed, act, h->d_mask);
    if (f) {
        HX_CHECK(hipMemcpyAsync(h->d_f, f, sizeof(double) * nd * h->opt.nrhs, hipMemcpyHostToDevice, st));
    } else {
        launch_fill<double>(st, nd * h->opt.nrhs, 0.0, h->d_f);
        for (int64_t i = 0; i < nsp; ++i)
            HX_CHECK(hipMemcpyAsync(h->d_f + sidx[i], sval + i, sizeof(double), hipMemcpyHostToDevice, st));
    }
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(st));
    h->have_problem = true;
    h->precond_ready = -1;
    h->refmg_ready = false;   // the coarse diagonal depends on the excluded nodes
    h->have_density = false;  // the factor depends on the active set
    h->have_solution = false;
    return HEXA_OK;
}

}  // namespace hexa

extern "C" {

int32_t hexa_prune_voids(hexa_handle* h, double min_stiffness_threshold, int64_t* n_active) {
    if (!h) { set_error("null handle"); return HEXA_E_INVALID; }
    if (!h->have_problem || !h->d_rho) { set_error("hexa_set_problem and hexa_set_density come first"); return HEXA_E_INVALID; }
    hipStream_t st = h->stream;
    const int64_t ne = h->g.ne;
    HX_TRY(dalloc(h, &h->d_active, (size_t)ne));
    int64_t na = ne;
    if (min_stiffness_threshold >= 0.0) {
        // Solver.jl:86-89: Float32 threshold * 1.01f, strict comparison
        const float thr = (float)min_stiffness_threshold * 1.01f;
        HX_TRY(ensure_solver_vectors(h));
        launch_active_from_rho(st, ne, h->d_rho, thr, h->d_active, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal + 5);
        HX_TRY(check_launch());
        double cnt = 0.0;
        HX_CHECK(hipMemcpyAsync(&cnt, h->d_scal + 5, sizeof(double), hipMemcpyDeviceToHost, st));
        HX_CHECK(hipStreamSynchronize(st));
        if (cnt == 0.0) { set_error("MeshPruner: no active elements"); return HEXA_E_INVALID; }
        na = (int64_t)cnt;
        // prune only when fewer than 99% of the elements are active (:90)
        if ((double)na >= (double)ne * 0.99) na = ne;
    }
    if (na == ne) launch_fill<uint8_t>(st, ne, 1, h->d_active);
    launch_node_mask(st, h->g, h->d_fixed, h->d_active, h->d_mask);
    HX_TRY(check_launch());
    h->precond_ready = -1;
    h->refmg_ready = false;
    if (n_active) *n_active = na;
    return update_factor(h);
}

int32_t hexa_set_density(hexa_handle* h, const double* rho) {
    if (!h || !rho) { set_error("null argument"); return HEXA_E_INVALID; }
    HX_TRY(dalloc(h, &h->d_rho, (size_t)h->g.ne));
    HX_CHECK(hipMemcpyAsync(h->d_rho, rho, sizeof(double) * h->g.ne, hipMemcpyHostToDevice, h->stream));
    return update_factor(h);
}

int32_t hexa_set_density_dev(hexa_handle* h, const double* d_rho) {
    if (!h || !d_rho) { set_error("null argument"); return HEXA_E_INVALID; }
    HX_TRY(dalloc(h, &h->d_rho, (size_t)h->g.ne));
    HX_CHECK(hipMemcpyAsync(h->d_rho, d_rho, sizeof(double) * h->g.ne, hipMemcpyDeviceToDevice, h->stream));
    return update_factor(h);
}

int32_t hexa_solve(hexa_handle* h, double* u_inout, const hexa_solve_params* prm, hexa_solve_report* rep) {
    HX_TRY(need_ready(h));
    hexa_solve_params p;
    if (prm) p = *prm; else hexa_default_solve_params(&p);
    if (!(p.tol > 0) || p.max_iter < 1 || p.shift_factor < 0 || !(p.shift_multiplier > 0)) {
        set_error("invalid solve parameters");
        return HEXA_E_INVALID;
    }
    HX_TRY(ensure_solver_vectors(h));
    const int64_t nd = 3 * h->g.nn;
    if (h->opt.nrhs > 1 && p.rhs_mode != HEXA_RHS_SEQUENTIAL && p.semantics == HEXA_SEMANTICS_CPU &&
        p.precond == HEXA_PRECOND_MG && h->factored && elastic_rw_supported(h->g, sizeof(double))) {
        const int32_t sb = solve_batched(h, u_inout, &p, rep);
        if (sb != BATCH_UNSUPPORTED) return sb;
    }
    int32_t worst = HEXA_OK;
    for (int r = 0; r < h->opt.nrhs; ++r) {
        int32_t s = solve_one(h, r, u_inout ? u_inout + r * nd : nullptr, &p, rep ? rep + r : nullptr);
        if (s == HEXA_E_NAN_INPUT || s == HEXA_E_OOM || s == HEXA_E_DEVICE || s == HEXA_E_INVALID) return s;
        if (s != HEXA_OK) worst = s;
    }
    return worst;
}

int32_t hexa_get_displacement(hexa_handle* h, double* u) {
    if (!h || !u || !h->d_x) { set_error("no solution available"); return HEXA_E_INVALID; }
    HX_CHECK(hipMemcpyAsync(u, h->d_x, sizeof(double) * 3 * h->g.nn, hipMemcpyDeviceToHost, h->stream));
    HX_CHECK(hipStreamSynchronize(h->stream));
    return HEXA_OK;
}

int32_t hexa_displacement_dev(hexa_handle* h, const double** d_u) {
    if (!h || !d_u || !h->d_x) { set_error("no solution available"); return HEXA_E_INVALID; }
    *d_u = h->d_x;
    return HEXA_OK;
}

int32_t hexa_matvec(hexa_handle* h, const double* x, double* y) {
    if (!h || !x || !y) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!h->have_density) { set_error("hexa_set_density has not been called"); return HEXA_E_INVALID; }
    const int64_t nd = 3 * h->g.nn;
    HX_TRY(dalloc(h, &h->d_p, (size_t)nd));
    HX_TRY(dalloc(h, &h->d_Ap, (size_t)nd));
    hipStream_t st = h->stream;
    HX_CHECK(hipMemcpyAsync(h->d_p, x, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, nullptr, h->d_p, nullptr,
                            nullptr, 0.0, 0.0, h->d_Ap);
    HX_TRY(check_launch());
    HX_CHECK(hipMemcpyAsync(y, h->d_Ap, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    return HEXA_OK;
}

int32_t hexa_matvec_dev(hexa_handle* h, const void* d_x, void* d_y, int32_t precision, int32_t reps) {
    if (!h || !d_x || !d_y || reps < 1) { set_error("invalid argument"); return HEXA_E_INVALID; }
    if (!h->have_density) { set_error("hexa_set_density has not been called"); return HEXA_E_INVALID; }
    hipStream_t st = h->stream;
    if (precision == HEXA_PREC_MIXED) HX_TRY(ensure_fac32(h));
    // (the chunk tuning reads x and writes y, as the product does)
    if (h->factored) {
        if (precision == HEXA_PREC_MIXED)
            tune_elastic_rw<float>(st, h->g, h->coef32, h->d_fac32, (const float*)d_x, (float*)d_y);
        else
            tune_elastic_rw<double>(st, h->g, h->coef64, h->d_fac, (const double*)d_x, (double*)d_y);
        HX_TRY(check_launch());
    }
    HX_CHECK(hipEventRecord(h->ev_a, st));
    for (int r = 0; r < reps; ++r) {
        if (precision == HEXA_PREC_MIXED)
            elastic<float, float>(h, h->g, h->coef32, h->d_Ke32, OP_APPLY, h->d_fac32, nullptr,
                                  (const float*)d_x, nullptr, nullptr, 0.0f, 0.0f, (float*)d_y);
        else
            elastic<double, double>(h, h->g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, nullptr,
                                    (const double*)d_x, nullptr, nullptr, 0.0, 0.0, (double*)d_y);
    }
    HX_CHECK(hipEventRecord(h->ev_b, st));
    HX_TRY(check_launch());
    HX_CHECK(hipEventSynchronize(h->ev_b));
    float ms = 0.f;
    HX_CHECK(hipEventElapsedTime(&ms, h->ev_a, h->ev_b));
    h->last_kernel_ms = (double)ms / reps;
    return HEXA_OK;
}

int32_t hexa_apply_precond(hexa_handle* h, const double* r, double* z) {
    if (!h || !r || !z) { set_error("null argument"); return HEXA_E_INVALID; }
    const bool ref = h->last_mg == 2 && h->refmg_ready;
    if (!ref && h->precond_ready != HEXA_PRECOND_MG) {
        set_error("no multigrid set-up: solve with HEXA_PRECOND_MG or HEXA_PRECOND_REF_MG first");
        return HEXA_E_INVALID;
    }
    const int64_t nd = 3 * h->g.nn;
    hipStream_t st = h->stream;
    HX_CHECK(hipMemcpyAsync(h->d_r, r, sizeof(double) * nd, hipMemcpyHostToDevice, st));
    if (ref) apply_refmg(h, h->d_r, h->d_z);
    else apply_mg(h, h->d_r, h->d_z);
    HX_TRY(check_launch());
    HX_CHECK(hipMemcpyAsync(z, h->d_z, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    return HEXA_OK;
}

int32_t hexa_last_kernel_ms(hexa_handle* h, double* ms) {
    if (!h || !ms) { set_error("null argument"); return HEXA_E_INVALID; }
    *ms = h->last_kernel_ms;
    return HEXA_OK;
}

int32_t hexa_stream_sync(hexa_handle* h) {
    if (!h) { set_error("null handle"); return HEXA_E_INVALID; }
    HX_CHECK(hipStreamSynchronize(h->stream));
    return HEXA_OK;
}

int32_t hexa_kernel_report(hexa_handle* h, int32_t reps, hexa_kernel_stat* out, int32_t cap, int32_t* count) {
    if (!h || !out || !count || cap < 1 || reps < 1) { set_error("invalid argument"); return HEXA_E_INVALID; }
    *count = 0;
    if (h->precond_ready != HEXA_PRECOND_MG || h->levels.size() < 2) {
        set_error("hexa_kernel_report needs a multigrid solve (HEXA_PRECOND_MG, two levels or more) first");
        return HEXA_E_INVALID;
    }
    hipStream_t st = h->stream;
    const Grid3& g = h->g;
    const int64_t nn = g.nn, ne = g.ne, nd = 3 * nn;
    const bool mx = mixed(h);
    // scratch vectors only: the solution (d_x), load (d_b) and diagonals stay
    for (double* v : {h->d_p, h->d_Ap, h->d_z, h->d_best}) launch_fill<double>(st, nd, 0.5, v);
    HX_TRY(check_launch());
    int n = 0;
    auto timed = [&](const char* name, int level, double bytes, auto&& launch) -> int32_t {
        if (n >= cap) return HEXA_OK;
        launch();   // warm
        HX_CHECK(hipEventRecord(h->ev_a, st));
        for (int r = 0; r < reps; ++r) launch();
        HX_CHECK(hipEventRecord(h->ev_b, st));
        HX_TRY(check_launch());
        HX_CHECK(hipEventSynchronize(h->ev_b));
        float ms = 0.f;
        HX_CHECK(hipEventElapsedTime(&ms, h->ev_a, h->ev_b));
        hexa_kernel_stat& o = out[n++];
        std::memset(&o, 0, sizeof o);
        std::snprintf(o.name, sizeof o.name, "%s", name);
        o.level = level;
        o.launches = reps;
        o.ms = (double)ms / reps;
        o.bytes = bytes;
        o.gbps = bytes / (o.ms * 1e-3) / 1e9;
        return HEXA_OK;
    };
    const double b64 = 8.0 * ne + 48.0 * nn;
    HX_TRY(timed("K(rho)u fp64", 0, b64, [&] {
        elastic<double, double>(h, g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, nullptr, h->d_p, nullptr, nullptr,
                                0.0, 0.0, h->d_Ap);
    }));
    HX_TRY(timed("K(rho)u fp64 + (p,Ap) partials", 0, b64 + nn, [&] {
        elastic<double, double>(h, g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, h->d_mask, h->d_p, nullptr, nullptr,
                                0.0, 0.0, h->d_Ap, h->d_dotpart);
    }));
    const int nb = h->opt.nrhs;
    if (nb > 1 && h->factored && elastic_rw_supported(g, sizeof(double))) {
        // the load cases of a batch in one launch: element factors and flags
        // are algorithmic bytes once
        char nm[40];
        std::snprintf(nm, sizeof nm, "K(rho)u fp64, %d load cases at once", nb);
        HX_TRY(timed(nm, 0, 8.0 * ne + 48.0 * nn * nb, [&] {
            elastic<double, double>(h, g, h->coef64, h->d_Ke64, OP_APPLY, h->d_fac, nullptr, h->d_p, nullptr, nullptr,
                                    0.0, 0.0, h->d_Ap, nullptr, nb);
        }));
    }
    // level 0 of the V-cycle and the transfers to and from level 1
    Level& l0 = h->levels[0];
    Level& l1 = h->levels[1];
    const int L = (int)h->levels.size();
    const double es = mx ? 4.0 : 8.0, ds = mx ? 2.0 : 8.0;   // vector and inverse-diagonal entry sizes
    const double vec = 3.0 * es * nn, dinv = 3.0 * ds * nn, fac = es * ne, msk = (double)nn;
    const double omega = h->opt.mg_omega;
    if (L > 1) {
        // (the FP64 V-cycle's level 0 reads and writes the caller's vectors:
        // it is timed on the solver's scratch vectors)
        auto sweep = [&](OpMode mode) {
            if (mx)
                vc_elastic<float, uint16_t>(h, g, l0.c32, l0.Ke32, mode, l0.fac32, l0.mask, (const float*)l0.x,
                                            (const float*)l0.b, (const uint16_t*)l0.dinv, (float)omega,
                                            (float)l0.shift, (float*)l0.xt);
            else
                vc_elastic<double, double>(h, g, l0.c64, l0.Ke64, mode, l0.fac64, l0.mask, h->d_best, h->d_z,
                                           (const double*)l0.dinv, omega, l0.shift, (double*)l0.xt);
        };
        if (h->vc_factored && elastic_rw_supported(g, mx ? 4 : 8))
            HX_TRY(timed(mx ? "smooth from 0, fp32/bf16" : "smooth from 0, fp64", 0, dinv + vec * 2 + fac + msk,
                         [&] { sweep(OP_SMOOTH0); }));
        HX_TRY(timed(mx ? "smooth, fp32/bf16" : "smooth, fp64", 0, dinv + vec * 3 + fac + msk,
                     [&] { sweep(OP_SMOOTH); }));
        if (h->vc_factored && elastic_rw_supported(g, mx ? 4 : 8)) {
            // the V-cycle's last sweep, which also forms the (r, z) partials of beta
            const double parts = 8.0 * (double)elastic_rw_partials(g, mx ? sizeof(float) : sizeof(double));
            HX_TRY(timed(mx ? "last smooth + (r,z) partials, fp32/bf16" : "last smooth + (r,z) partials, fp64", 0,
                         dinv + vec * 3 + fac + msk + parts, [&] {
                             if (mx)
                                 vc_elastic<float, uint16_t>(h, g, l0.c32, l0.Ke32, OP_SMOOTH, l0.fac32, l0.mask,
                                                             (const float*)l0.x, (const float*)l0.b,
                                                             (const uint16_t*)l0.dinv, (float)omega, (float)l0.shift,
                                                             (float*)l0.xt, 1, h->d_dotpart);
                             else
                                 vc_elastic<double, double>(h, g, l0.c64, l0.Ke64, OP_SMOOTH, l0.fac64, l0.mask,
                                                            h->d_best, h->d_z, (const double*)l0.dinv, omega, l0.shift,
                                                            (double*)l0.xt, 1, h->d_dotpart);
                         }));
        }
        if (nb > 1 && mx) {
            char nm[40];
            std::snprintf(nm, sizeof nm, "smooth fp32/bf16, %d load cases", nb);
            HX_TRY(timed(nm, 0, dinv + vec * 3 * nb + fac + msk, [&] {
                vc_elastic<float, uint16_t>(h, g, l0.c32, l0.Ke32, OP_SMOOTH, l0.fac32, l0.mask, (const float*)l0.x,
                                            (const float*)l0.b, (const uint16_t*)l0.dinv, (float)omega,
                                            (float)l0.shift, (float*)l0.xt, nb);
            }));
        }
        HX_TRY(timed(mx ? "residual, fp32" : "residual, fp64", 0, vec * 3 + fac + msk, [&] { sweep(OP_RESID); }));
        const double cvec = 3.0 * es * l1.g.nn;
        HX_TRY(timed(mx ? "restrict, fp32" : "restrict, fp64", 0, vec + cvec + l1.g.nn, [&] {
            if (mx) launch_restrict<float>(st, g, l1.g, (const float*)l0.xt, l1.mask, (float*)l1.b);
            else launch_restrict<double>(st, g, l1.g, (const double*)l0.xt, l1.mask, (double*)l1.b);
        }));
        HX_TRY(timed(mx ? "prolong + add, fp32" : "prolong + add, fp64", 0, 2 * vec + cvec + msk, [&] {
            if (mx) launch_prolong_add<float>(st, g, l1.g, (const float*)l1.x, l0.mask, (float*)l0.xt);
            else launch_prolong_add<double>(st, g, l1.g, (const double*)l1.x, l0.mask, (double*)l0.xt);
        }));
    }
    // the PCG recurrence as the mixed (or FP64) multigrid solve runs it
    launch_pcg_state_for_timing(st, h->d_state);
    HX_TRY(check_launch());
    float* r32 = mx ? (float*)l0.b : nullptr;
    const float* z32 = mx ? (const float*)l0.x : nullptr;
    if (mx) {
        HX_TRY(timed("pcg update (r, r32)", 0, 28.0 * nd, [&] {
            launch_pcg_update(st, nd, h->d_state, nullptr, h->d_p, h->d_z, h->d_Ap, nullptr, h->d_partials,
                              h->d_upart, r32);
        }));
        HX_TRY(timed("pcg flex dots z32 (no last sweep)", 0, 20.0 * nd, [&] {
            launch_pcg_flex_dots_z32(st, nd, h->d_state, h->d_z, z32, h->d_Ap, h->d_partials);
        }));
        HX_TRY(timed("pcg p update + deferred x (z32)", 0, 36.0 * nd, [&] {
            launch_pcg_p_update_z32(st, nd, h->d_state, h->d_p, z32, h->d_best);
        }));
    } else {
        HX_TRY(timed("pcg update (x, r)", 0, 48.0 * nd, [&] {
            launch_pcg_update(st, nd, h->d_state, h->d_best, h->d_p, h->d_z, h->d_Ap, nullptr, h->d_partials,
                              h->d_upart);
        }));
        HX_TRY(timed("pcg flexible dots (no last sweep)", 0, 24.0 * nd, [&] {
            launch_pcg_flex_dots(st, nd, h->d_state, h->d_z, h->d_best, h->d_Ap, h->d_partials);
        }));
        HX_TRY(timed("pcg p update", 0, 24.0 * nd, [&] {
            launch_pcg_p_update(st, nd, h->d_state, h->d_p, h->d_best, nullptr, nullptr);
        }));
    }
    if (h->d_Ae && h->d_hp && h->d_hAp)
        HX_TRY(timed("Helmholtz stencil, fp64", 0, 16.0 * nn, [&] { launch_helm_apply(st, g, h->d_Ae, h->d_hp, h->d_hAp); }));
    if (h->d_e0 && h->d_bracket && h->d_rho) {
        double* e[7] = {h->d_e0, h->d_e1, h->d_e2, h->d_e3, h->d_e4, h->d_e5, h->d_e6};
        launch_fill<double>(st, ne, 0.5, e[0]);
        launch_fill<double>(st, ne, -1.0, e[1]);
        launch_fill<double>(st, ne, 1.0, e[2]);
        launch_fill<double>(st, ne, 1.0, e[3]);
        HX_TRY(check_launch());
        // one multisection pass reads q, lo, hi and w (OC_NL volumes at once)
        HX_TRY(timed("OC multisection (16 passes)", 0, ne * (48.0 + 16 * 32.0 + 48.0), [&] {
            launch_oc(st, ne, e[1], e[2], e[3], 0.3, 0.2, 1e-300, 16, e[0], e[4], e[5], e[6], h->d_bracket,
                      h->d_partials, h->d_scal + 8);
        }));
        HX_TRY(timed("SIMP sensitivity", 0, 24.0 * nn + 16.0 * ne, [&] {
            if (h->factored)
                launch_simp_sens_fact(st, g, h->coef64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, e[1],
                                      h->d_partials);
            else
                launch_simp_sens(st, g, h->d_Ke64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, e[1],
                                 h->d_partials);
        }));
    }
    HX_CHECK(hipStreamSynchronize(st));
    *count = n;
    return HEXA_OK;
}

}  // extern "C"
