// Internal declarations shared by the libhexa translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <mutex>
#include <shared_mutex>

#include "../../include/hexa.h"

namespace hexa {

void set_error(const char* fmt, ...);

// A stream capture fails, and so does the other thread's call, when another
// thread of the process meanwhile makes a HIP call that is not ordered on a
// stream of its own (allocation and freeing, stream, event and graph
// creation and destruction, synchronous copies, launches on the legacy
// stream).  Every such call in the library holds this lock shared; a capture
// holds it exclusively from the destruction of the graph it replaces to the
// instantiation of the new one.  Neither is taken while the other is held
// by the same thread.
std::shared_mutex& capture_mutex();
struct CaptureGuard {
    std::unique_lock<std::shared_mutex> lk{capture_mutex()};
};
struct DeviceOpGuard {
    std::shared_lock<std::shared_mutex> lk{capture_mutex()};
};

#define HX_CHECK(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::hexa::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                              __FILE__, __LINE__);                                      \
            return e_ == hipErrorOutOfMemory ? HEXA_E_OOM : HEXA_E_DEVICE;              \
        }                                                                               \
    } while (0)

#define HX_TRY(expr)                      \
    do {                                  \
        int32_t s_ = (expr);              \
        if (s_ != HEXA_OK) return s_;     \
    } while (0)

// Reductions: every reducing kernel writes one partial per block into a
// [quantity][RED_BLOCKS] array; a single-block kernel then sums them in a
// fixed order, so every reduction is deterministic run to run and
// independent of what else shares the GPU.
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr int OC_NL = 32;      // lambda points per multisection pass of the OC update
constexpr int MAX_RED_Q = 32;  // partial-sum rows per reduction buffer

// Node mask bits.  Bits 0..2: component held.  Bit 3: node excluded (no
// active element touches it; all three components are then held).
constexpr uint8_t MASK_ALL = 0x7;

// Operator application modes of the fused elasticity kernel.
// OP_SMOOTH0: one smoothing sweep started from x0 = omega D^-1 b, without
// forming x0 (factored row-wave kernel only; u is not read).
enum OpMode { OP_APPLY = 0, OP_RESID = 1, OP_SMOOTH = 2, OP_SMOOTH0 = 3 };

// HEXA_FALLBACK_KERNELS=1 routes every elasticity product to the kernels that
// otherwise run only on grids too large for the row-wave and matrix-core
// kernels (node planes over 2 GB, 2^31 nodes or more): the factored tile
// kernel and the node-centric dense kernel.  Used by the tests to cover them.
inline bool fallback_kernels_forced() {
    const char* e = getenv("HEXA_FALLBACK_KERNELS");
    return e && e[0] == '1';
}

// Entries between the right-hand sides of a batch of vectors of n entries
// (rounded up so that every right-hand side starts 512-byte aligned).
inline int64_t batch_stride(int64_t n) { return (n + 63) & ~(int64_t)63; }

// Grid description of one level.
struct Grid3 {
    int nx, ny, nz;     // elements
    int64_t nn;         // nodes
    int64_t ne;         // elements
};

inline Grid3 make_grid(int nx, int ny, int nz) {
    Grid3 g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    g.ne = (int64_t)nx * ny * nz;
    return g;
}

// Factored element operator (exact brick stiffness).  In the per-axis
// sum/difference ("modal") basis of each displacement component the hex8
// brick stiffness is a diagonal per component and mode plus, for each
// ordered component pair (i,j) with third axis d, two couplings per mode
// parity q of d.  Coefficients carry the 1/8 of the inverse butterfly.
template <typename T>
struct FactCoef {
    T D[3][8];    // modal diagonal, per output component, mode m = mx + 2my + 4mz
    T cl[3][2];   // lambda-term coupling, per third axis d and its mode q
    T cm[3][2];   // mu-term coupling
};

// Device-side state of one PCG recurrence.  The iteration's scalars live on
// the device, so the host never waits inside an iteration; it reads this
// struct back only to decide when to stop.
enum PcgHalt { PCG_RUNNING = 0, PCG_CONVERGED = 1, PCG_BREAKDOWN = 2, PCG_NAN = 3, PCG_DIVERGED = 4 };
struct PcgState {
    double bnorm;      // ||b||
    double rz;         // (r, z) of the current direction
    double pAp;
    double alpha, beta;
    double rr;         // ||r||^2 after the last update
    double rel;        // last checked ||r||/||b||
    double best_rel;
    double final_rel;
    double zAp;
    double tol;
    int32_t halted;    // PcgHalt
    int32_t improved;  // last check improved best_rel
    int32_t last_k;    // last iteration that updated x
    int32_t x_iter;    // iteration whose x update k_update deferred
    int32_t k;         // iterations started (advanced by the alpha kernels)
    int32_t x_done;    // last iteration whose deferred x update was applied
    double rzb[2];     // (r, z) by iteration parity (the fused filter PCG)
};

// ---- kernel launchers (implemented in the .hip files) -----------------
// Elasticity (k_elastic.hip)
// (runs on the matrix cores, k_elastic_mfma.hip; OP_SMOOTH0 is not provided
// by the dense kernels)
template <typename T, typename DT>
void launch_elastic_op(hipStream_t s, const Grid3& g, OpMode mode, const T* Ke, const T* fac,
                       const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega,
                       T shift, T* y);
bool elastic_mfma_supported(const Grid3& g);
template <typename T, typename DT>
void launch_elastic_mfma(hipStream_t s, const Grid3& g, OpMode mode, const T* Ke, const T* fac, const uint8_t* mask,
                         const T* u, const T* b, const DT* dinv, T omega, T shift, T* y);
template <typename T, typename DT>
void launch_elastic_fact(hipStream_t s, const Grid3& g, OpMode mode, const FactCoef<T>& coef, const T* fac,
                         const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y);
// Returns the number of partial sums written to dot_part (0 when dot_part
// is null or the kernel cannot provide them): of (u, y) for OP_APPLY, of
// (b, y) for OP_SMOOTH; their values, and so their fixed-order sum, do not
// depend on the kernel's z-chunk length.
// nb > 1: nb right-hand sides in one launch, vectors batch_stride(3 nn)
// apart and partial sums elastic_rw_partials(g, sizeof(T)) apart; fac, mask
// and dinv are shared.
template <typename T, typename DT>
int launch_elastic_rw(hipStream_t s, const Grid3& g, OpMode mode, const FactCoef<T>& coef, const T* fac,
                      const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y,
                      double* dot_part = nullptr, int nb = 1, const int32_t* halt = nullptr, int hs = 0);
// Whether the row-wave kernel takes grid g (its buffer descriptors address one
// node plane of elem_bytes-sized values with 31-bit byte counts); otherwise
// launch_elastic_rw falls back to the tile kernel, which has no OP_SMOOTH0.
bool elastic_rw_supported(const Grid3& g, size_t elem_bytes);
// Number of partial sums launch_elastic_rw writes for grid g with
// elem_bytes-sized values (the FP64 and FP32 kernels may tile a grid
// differently).
int64_t elastic_rw_partials(const Grid3& g, size_t elem_bytes = 8);
// Picks the z-chunk length of the row-wave kernels for grid g by timing
// (process-wide, once per grid and kernel); u is read, y overwritten.
template <typename T>
void tune_elastic_rw(hipStream_t s, const Grid3& g, const FactCoef<T>& coef, const T* fac, const T* u, T* y);
template <typename T>
FactCoef<T> make_fact_coef(double hx, double hy, double hz, double nu);
template <typename T>
void launch_stiffness_factor(hipStream_t s, int64_t ne, const double* rho, const uint8_t* active,
                             double E0, double Emin, double penal, T* fac);
template <typename T>
void launch_diag(hipStream_t s, const Grid3& g, const double* Ke, const T* fac, const uint8_t* mask,
                 int x_entry, T* diag);
void launch_node_mask(hipStream_t s, const Grid3& g, const uint8_t* fixed_dof, const uint8_t* active,
                      uint8_t* mask);
template <typename T>
void launch_coarsen_factor(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* ffac, T* cfac);
void launch_coarsen_mask(hipStream_t s, const Grid3& fine, const Grid3& coarse, const uint8_t* fmask,
                         uint8_t* cmask);

// Multigrid transfers (k_mg.hip)
// (nb > 1: a batch of right-hand sides, vectors batch_stride(3 nn) apart on
// each level).  halt, when set, points to a solve's halted flag (of load
// case rb at halt[rb * hs]): a launch queued after the solve ended returns at
// once (the V-cycle kernels and the iteration's operator take it).
template <typename T>
void launch_restrict(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* rf,
                     const uint8_t* cmask, T* rc, int nb = 1, const int32_t* halt = nullptr, int hs = 0);
template <typename T>
void launch_prolong_add(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* xc,
                        const uint8_t* fmask, T* xf, int nb = 1, const int32_t* halt = nullptr, int hs = 0);
// active[e] = (float)rho[e] > thr; RED_BLOCKS per-block counts in partials
void launch_active_from_rho(hipStream_t s, int64_t ne, const double* rho, float thr, uint8_t* active,
                            double* partials);
void launch_refmg_dcinv(hipStream_t s, const Grid3& fine, const Grid3& coarse, const uint8_t* fmask,
                        const double* kd8, double* dcinv);
void launch_coarse_assemble(hipStream_t s, const Grid3& g, const double* Ke, const double* fac,
                            const uint8_t* mask, double shift, double* A);
// Multigrid level count of an nx x ny x nz grid under opt, and the coarsest
// level's DOF count (capi.cpp).
int mg_level_count(int nx, int ny, int nz, const hexa_options& opt, int64_t* coarse_dofs);
// In-place inverse of the SPD n x n matrix A; rowbuf holds gj_row_doubles(n)
// and colbuf gj_col_doubles(n) values.
constexpr int GJ_PANEL = 32;
inline int64_t gj_row_doubles(int64_t n) { return GJ_PANEL * n + GJ_PANEL * GJ_PANEL; }
inline int64_t gj_col_doubles(int64_t n) { return GJ_PANEL * n; }
void launch_gauss_jordan(hipStream_t s, int64_t n, double* A, double* rowbuf, double* colbuf);
template <typename T>
void launch_dense_gemv(hipStream_t s, int64_t n, const double* Ainv, const T* x, T* y, int nb = 1,
                       const int32_t* halt = nullptr, int hs = 0);

// Vectors and reductions (k_vector.hip)
template <typename T>
void launch_fill(hipStream_t s, int64_t n, T v, T* x);
template <typename A, typename B>
void launch_convert(hipStream_t s, int64_t n, const A* x, B* y);
template <typename T, typename DT>
void launch_scale_dinv(hipStream_t s, int64_t n, const T* b, const DT* dinv, T omega, T* x, int nb = 1);
template <typename DT>
void launch_make_dinv(hipStream_t s, int64_t n, const double* diag, const uint8_t* mask, double shift,
                      int clamp_mode, DT* dinv);
void launch_dot2(hipStream_t s, int64_t n, const double* a, const double* b, double* partials);
void launch_max_free(hipStream_t s, const Grid3& g, const double* diag, const uint8_t* mask,
                     double* partials, double* out);
void launch_final_sum(hipStream_t s, int nq, const double* partials, double* out);
template <typename T>
void launch_axpy(hipStream_t s, int64_t n, T a, const T* x, T* y);

// Helmholtz filter (k_filter.hip)
void launch_helm_apply(hipStream_t s, const Grid3& g, const double* Ae, const double* x, double* y);
// One iteration of the fused filter PCG (three launches; state and vectors
// as the filter's PCG, partials >= FUSED_PAP_MAX + 512 values); returns -1
// when the grid needs more stencil blocks than FUSED_PAP_MAX.
constexpr int FUSED_PAP_MAX = 16384;
static_assert(FUSED_PAP_MAX + 512 <= MAX_RED_Q * RED_BLOCKS, "fused filter partials exceed the reduction buffer");
int launch_filter_fused_iteration(hipStream_t s, const Grid3& g, const double* Ae, PcgState* st, double* x,
                                  double* p, double* r, double* Ap, const double* dinv, double* partials);
// the 27-point stencil of an interior node from the 8x8 element matrix
// (written after Ae's 64 values for launch_helm_apply)
void helm_stencil(const double* Ae, double* S);
void launch_helm_rhs_diag(hipStream_t s, const Grid3& g, const double* Ae, const double* rho, double w, double* b,
                          double* dinv);
void launch_helm_extract(hipStream_t s, const Grid3& g, const double* x, double scale, double* out);
// element copies between a grid and a sub-box of it (the blocked filter)
void launch_box_gather(hipStream_t s, const Grid3& g, const Grid3& sub, int ox, int oy, int oz, const double* src,
                       double* dst);
void launch_box_scatter(hipStream_t s, const Grid3& g, const Grid3& sub, int ox, int oy, int oz, int cx, int cy,
                        int cz, int sx, int sy, int sz, const double* src, double* dst);

// Optimisation step (k_topopt.hip)
void launch_stress(hipStream_t s, const Grid3& g, double dx, double dy, double dz, double E, double nu,
                   const double* U, const double* rho, float* principal, float* vm, float* l1, float* dmax,
                   float* dmin, float* voigt);
void launch_simp_sens(hipStream_t s, const Grid3& g, const double* Ke, const double* rho, double E0, double Emin,
                      double penal, const double* U, double* dc, double* partials);
// The same with the exact brick stiffness in factored form (coef at E = 1).
void launch_simp_sens_fact(hipStream_t s, const Grid3& g, const FactCoef<double>& coef, const double* rho, double E0,
                           double Emin, double penal, const double* U, double* dc, double* partials);
int oc_bracket_bytes();
void launch_oc(hipStream_t s, int64_t n, const double* dc, const double* dv, const double* w, double volfrac,
               double move, double tol, int max_passes, double* x, double* q, double* lo, double* hi, void* bracket,
               double* partials, double* out4);
// The same update with q, lo and hi formed from x, dc and dv inside the
// passes (no per-element work arrays; bitwise the same result as launch_oc):
// scratch is the bracket (oc_bracket_bytes) and OC_NL * RED_BLOCKS partials.
void launch_oc_direct(hipStream_t s, int64_t n, const double* dc, const double* dv, const double* w, double volfrac,
                      double move, double tol, int max_passes, double* x, void* bracket, double* partials,
                      double* out4);
void launch_upd_proposed(hipStream_t s, int64_t n, const float* l1, const uint8_t* prot, const float* orig, float allow,
                         float E, float mind, double* prop);
void launch_regrid(hipStream_t s, int onx, int ony, int onz, int nnx, int nny, int nnz, float ndx, float ndy,
                   float ndz, float lx, float ly, float lz, int hard, const float* rho, const float* alpha, float* out,
                   float* aout);
void launch_nan_count(hipStream_t s, int64_t n, const double* x, double* partials);
void launch_nan_count(hipStream_t s, int64_t n, const float* x, double* partials);
void launch_upd_apply(hipStream_t s, int64_t n, const double* filt, const double* prop, int use_prop,
                      const uint8_t* prot, float mind, float maxd, float* dens, double* partials);
void launch_upd_candidates(hipStream_t s, int64_t n, const uint8_t* prot, const float* orig, float mind, float maxd,
                           float thr, float* dens, uint8_t* cand, double* partials);
void launch_radix_hist(hipStream_t s, int64_t n, const uint8_t* cand, const float* dens, uint32_t prefix, int shift,
                       unsigned long long* hist);
void launch_tie_count(hipStream_t s, int64_t n, const uint8_t* cand, const float* dens, uint32_t key, int64_t* counts);
void launch_cull(hipStream_t s, int64_t n, const uint8_t* cand, uint32_t key_t, int64_t ties, const int64_t* tie_base,
                 float mind, float* dens);
void launch_upd_finish(hipStream_t s, int64_t n, const uint8_t* prot, const float* orig, float mind, int hard,
                       float* dens);

// PCG (k_pcg.hip).  All take the device state and do nothing once it halted.
// Those with an nb argument run nb right-hand sides in one launch: states
// st[0..nb), vectors vs entries apart, partial sums ps apart (per right-hand
// side); nb = 1 is the single solve.
void launch_pcg_init_state(hipStream_t s, PcgState* st, const double* partials, double tol);
void launch_pcg_state_for_timing(hipStream_t s, PcgState* st);
void launch_pcg_rz(hipStream_t s, int64_t n, const double* r, const double* z, double* partials);
void launch_pcg_set_rz(hipStream_t s, PcgState* st, const double* partials);
void launch_pcg_pAp(hipStream_t s, int64_t n, PcgState* st, const double* p, const double* Ap,
                    double* partials);
void launch_pcg_alpha(hipStream_t s, PcgState* st, const double* partials, int breakdown_check);
// Per-block partial sums of the update: upart holds 2 pcg_update_blocks(n)
// values per right-hand side; they are summed to RED_BLOCKS range sums in
// rows 0 (r, r) and 1 (r, minv r) of partials.
int64_t pcg_update_blocks(int64_t n);
void launch_pcg_update(hipStream_t s, int64_t n, PcgState* st, double* x, const double* p, double* r,
                       const double* Ap, const double* minv, double* partials, double* upart, float* r32 = nullptr,
                       int nb = 1, int64_t vs = 0, int ps = 0);
// The residual test runs when k % check_every == 0 (check_every > 0), or
// k == 1 with check_first; k is the state's iteration count.
void launch_pcg_check(hipStream_t s, PcgState* st, const double* partials, int check_every, int check_first,
                      int cpu_semantics, int jacobi, int nb = 1, int ps = 0);
void launch_pcg_flex_dots(hipStream_t s, int64_t n, const PcgState* st, const double* r, const double* z,
                          const double* Ap, double* partials, int nb = 1, int64_t vs = 0, int ps = 0);
void launch_pcg_flex_beta(hipStream_t s, PcgState* st, const double* partials, int nb = 1, int ps = 0);
void launch_pcg_p_update(hipStream_t s, int64_t n, const PcgState* st, double* p, const double* z,
                         const double* r, const double* minv, int nb = 1, int64_t vs = 0);
// alpha from the n (p, Ap) partials of the operator kernel, summed in a fixed
// order (through RED_BLOCKS range sums in scratch when n > RED_BLOCKS).
void launch_pcg_alpha_n(hipStream_t s, PcgState* st, const double* partials, int64_t n, int breakdown_check,
                        double* scratch, int nb = 1, int64_t pstride = 0);
// beta = (r, z)_new / (r, z) from the n (r, z) partials of the V-cycle's last
// smoothing sweep (fixed order, as launch_pcg_alpha_n), and rz = (r, z)_new.
void launch_pcg_beta_n(hipStream_t s, PcgState* st, const double* partials, int64_t n, double* scratch,
                       int nb = 1, int64_t pstride = 0);
void launch_pcg_flex_dots_z32(hipStream_t s, int64_t n, const PcgState* st, const double* r, const float* z,
                              const double* Ap, double* partials, int nb = 1, int64_t vs = 0, int ps = 0);
void launch_pcg_p_update_z32(hipStream_t s, int64_t n, const PcgState* st, double* p, const float* z,
                             double* x = nullptr, int nb = 1, int64_t vs = 0);
void launch_pcg_best_copy(hipStream_t s, int64_t n, const PcgState* st, const double* x, double* best);

// Library-internal entry points shared by the C API files and the job server
// (capi.cpp, capi_topopt.cpp).
int32_t ensure_shared_buffers(hexa_handle* h);
int32_t set_problem_impl(hexa_handle* h, const uint8_t* fixed_dof, const double* f, int64_t nsp,
                         const int64_t* sidx, const double* sval, const uint8_t* active_elem);
int32_t oc_check(hexa_handle* h, const hexa_oc_params* p);
int32_t oc_step_core(hexa_handle* h, const hexa_oc_params* p, hexa_oc_report* rep, double t0);

}  // namespace hexa
