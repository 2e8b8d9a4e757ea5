// The libhexa handle and host helpers shared by the C-ABI translation units.
#pragma once

#include <chrono>
#include <vector>

#include "hexa_internal.h"

namespace hexa {

constexpr int RING = 4;       // pinned state slots for the look-ahead poll
constexpr int LOOKAHEAD = 2;  // iterations queued beyond the last one inspected
constexpr int BPART_STRIDE = 4 * RED_BLOCKS;   // partial sums per load case of a batch

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Level {
    Grid3 g{};
    double kscale = 1.0;            // Ke(level) = kscale * Ke(fine)
    double shift = 0.0;
    double* Ke64 = nullptr;
    float* Ke32 = nullptr;
    FactCoef<double> c64{};
    FactCoef<float> c32{};
    double* fac64 = nullptr;        // owned for l >= 1
    float* fac32 = nullptr;
    uint8_t* mask = nullptr;        // owned for l >= 1
    double* diag = nullptr;
    void* dinv = nullptr;
    void* b = nullptr;              // rhs (owned for l >= 1, or l == 0 in mixed mode)
    void* x = nullptr;
    void* xt = nullptr;
    void* r = nullptr;
};

}  // namespace hexa

struct hexa_handle {
    hexa_grid grid{};
    hexa_options opt{};
    hexa::Grid3 g{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    hipEvent_t ring_ev[hexa::RING] = {};
    double last_kernel_ms = 0.0;
    // one PCG iteration captured as a graph, for the configuration in
    // pcg_key; graph_gen advances whenever a preconditioner is rebuilt
    hipGraphExec_t pcg_exec = nullptr;
    struct {
        int pk = -1, cpu = 0, every = 0;
        double s = 0.0;
        int64_t gen = -1;
    } pcg_key;
    int64_t graph_gen = 0;
    // the coarsest level's Gauss-Jordan sweep (two launches per pivot)
    // captured as a graph, for n_coarse = gj_n on d_Ainv = gj_A
    hipGraphExec_t gj_exec = nullptr;
    int64_t gj_n = -1;
    const double* gj_A = nullptr;
    hipGraphExec_t filt_exec = nullptr;   // one Helmholtz-filter PCG iteration
    // one iteration of the batched load-case PCG, for the configuration in
    // bpcg_key (nb right-hand sides, shift, hierarchy generation)
    hipGraphExec_t bpcg_exec = nullptr;
    struct {
        int nb = -1;
        double s = 0.0;
        int64_t gen = -1;
    } bpcg_key;
    int filt_every = 0;
    bool filt_fused = false;

    double Ke[576];
    bool factored = false;          // exact brick Ke -> factored element kernel
    // The V-cycle's level operators: the exact brick Ke on the factored
    // kernel unless the dense kernel was asked for (HEXA_MATVEC_DENSE), also
    // when the outer operator uses the dense f32-widened Ke: a preconditioner
    // only has to approximate the operator, and the two element matrices
    // differ by Float32 rounding (2e-6 relative).
    bool vc_factored = false;
    double Ke_vc[576];              // the V-cycle's element matrix at level 0
    hexa::FactCoef<double> coef64{};
    hexa::FactCoef<float> coef32{};
    double* d_Ke64 = nullptr;
    float* d_Ke32 = nullptr;

    uint8_t* d_fixed = nullptr;    // 3*nn
    uint8_t* d_active = nullptr;   // ne or null
    uint8_t* d_mask = nullptr;     // nn
    double* d_f = nullptr;         // nrhs * 3nn
    double* d_rho = nullptr;
    double* d_fac = nullptr;
    float* d_fac32 = nullptr;
    bool have_problem = false, have_density = false, fac32_valid = false, have_solution = false;

    double *d_x = nullptr, *d_r = nullptr, *d_p = nullptr, *d_Ap = nullptr, *d_z = nullptr;
    double *d_minv = nullptr, *d_best = nullptr, *d_b = nullptr, *d_diag = nullptr, *d_guess = nullptr;
    double* d_partials = nullptr;
    double* d_upart = nullptr;     // the PCG update's per-block sums (2 pcg_update_blocks(3 nn) per load case)
    double* d_dotpart = nullptr;   // (p, Ap) partials of the fused operator, then (r, z) partials of the
                                   // V-cycle's last smoothing sweep (per load case)
    double* d_bpart = nullptr;     // batched PCG: BPART_STRIDE partials per load case
    double* d_bscr = nullptr;      // batched PCG: RED_BLOCKS alpha range sums per load case
    double* d_scal = nullptr;
    hexa::PcgState* d_state = nullptr;
    // while a PCG iteration is launched or captured: its state's halted flag,
    // passed to the operator and V-cycle kernels so that iterations queued
    // after the solve ended skip their work (null otherwise)
    const int32_t* vc_halt = nullptr;
    int vc_halt_stride = 0;         // int32 entries between the flags of a batch's load cases
    hexa::PcgState* h_state = nullptr;   // pinned ring

    // preconditioner set-up of the last solve
    int precond_ready = -1;        // HEXA_PRECOND_* or -1
    std::vector<hexa::Level> levels;
    double* d_Ainv = nullptr;
    double *d_gj_row = nullptr, *d_gj_col = nullptr;
    int64_t n_coarse = 0;
    double mg_shift = -1.0;         // shift the smoother/coarse inverse were built for
    int last_mg = -1;               // multigrid kind the last solve applied
    // reference two-level preconditioner: floor(n/2) coarse grid, coarse
    // residual/correction and inverse coarse diagonal, an all-free coarse mask
    bool refmg_ready = false;
    hexa::Grid3 rg{};
    double* d_rc = nullptr;
    double* d_xc = nullptr;
    double* d_dcinv = nullptr;
    uint8_t* d_rcmask = nullptr;

    // Helmholtz filter and optimisation-step work arrays
    double* d_Ae = nullptr;         // 64
    double *d_hb = nullptr, *d_hx = nullptr, *d_hr = nullptr, *d_hp = nullptr, *d_hAp = nullptr, *d_hdinv = nullptr;
    double *d_e0 = nullptr, *d_e1 = nullptr, *d_e2 = nullptr, *d_e3 = nullptr, *d_e4 = nullptr, *d_e5 = nullptr;
    double *d_e6 = nullptr, *d_ones = nullptr;
    float* d_stress = nullptr;      // 17 nE floats of stress outputs
    double w_radius = -1.0;         // radius the filter weights d_w were computed for
    double* d_w = nullptr;
    float *d_f0 = nullptr, *d_f1 = nullptr, *d_f2 = nullptr;
    // protected mask and original density of the reference update, resident
    // for the device loop (hexa_set_update_fields)
    uint8_t* d_prot = nullptr;
    float* d_orig = nullptr;
    bool have_update_fields = false;
    uint8_t *d_u0 = nullptr, *d_u1 = nullptr;
    void* d_bracket = nullptr;
    unsigned long long* d_hist = nullptr;
    int64_t* d_ties = nullptr;

    std::vector<void*> allocs;
    int64_t bytes_allocated = 0;
};

namespace hexa {

int32_t dalloc_bytes(hexa_handle* h, void** p, size_t bytes);
inline int32_t dalloc(hexa_handle* h, void** p, size_t bytes) { return dalloc_bytes(h, p, bytes); }
template <typename T>
inline int32_t dalloc(hexa_handle* h, T** p, size_t count) {
    return dalloc_bytes(h, reinterpret_cast<void**>(p), count * sizeof(T));
}
int32_t check_launch();
int32_t need_ready(hexa_handle* h);
// Recomputes the stiffness multiplier from d_rho (after the density changed).
int32_t update_factor(hexa_handle* h);
int32_t ensure_solver_vectors(hexa_handle* h);
// Waits for the device state of iteration k - LOOKAHEAD (ring slot) and
// reports whether it halted; records iteration k's state copy.
int32_t poll_state(hexa_handle* h, int k, bool* halted);
// Helmholtz filter of an element field on the handle's grid (device arrays).
// fused: the three-launch iteration of launch_filter_fused_iteration
// (residual checked every iteration; check_every must be 1).
// reg: a diagonal shift reg I added to the filter operator (0 for the
// reference's GPU filter, 1e-9 for its CPU one).
int32_t filter_dev(hexa_handle* h, const double* d_in, double* d_out, double radius, double tol, int max_iter,
                   int check_every, int* iters, double* rel, bool fused = false, double reg = 0.0);

}  // namespace hexa
