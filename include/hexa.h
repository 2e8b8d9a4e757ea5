/*
 * hexa.h -- C ABI of libhexa, the MI355X (gfx950) implementation of HEXA
 * TopOpt's inner solve loop: matrix-free hex8 K(rho)*u, the PCG driver with
 * a geometric-multigrid V-cycle, the Helmholtz density filter, the stress
 * "sensitivity" with the reference density update, and SIMP + OC.
 *
 * Reference interfaces replaced (R/ = v119_2025_12_19_HEXATOP_AWS/):
 *   GPUSolver.solve_system_gpu            R/src/Solvers/GPUSolver.jl:467-477
 *     (and gpu_matrix_free_cg_solve       R/src/Solvers/GPUSolver.jl:182-465)
 *   GPUGeometricMultigrid.setup_multigrid / apply_mg_vcycle!
 *                                         R/src/Solvers/GPUGeometricMultigrid.jl:253-346
 *   GPUHelmholtz.apply_gpu_filter!        R/src/Optimization/GPUHelmholtz.jl:422-443
 *   Stress.compute_stress_field           R/src/Core/Stress.jl:287-374
 *   TopologyOptimization.update_density!  R/src/Optimization/TopOpt.jl:211-482
 *   compliance dot(F, U)                  R/src/Main.jl:453
 * plus SIMP sensitivity / OC (no reference counterpart), batched load cases
 * and a device-resident job server.
 *
 * Conventions
 *   - Structured grid of nx*ny*nz bricks of size dx*dy*dz (the reference's
 *     generate_mesh, R/src/Mesh/Mesh.jl:23-60), 0-based:
 *       node (i,j,k)    = i + (nx+1)*(j + (ny+1)*k)
 *       element (i,j,k) = i + nx*(j + ny*k)
 *       DOF             = 3*node + component   (interleaved, as R's U and F)
 *   - A pruned reference mesh (R/src/Mesh/MeshPruner.jl:10-87) is passed as
 *     the full box plus an active-element mask; nodes that no active element
 *     touches are dropped from the system, which is the same linear system
 *     as the renumbered one.
 *   - Every function returns an int32 status (HEXA_OK or a negative code);
 *     hexa_last_error() returns a thread-local message.  No function calls
 *     exit/abort.  Host buffers are borrowed for the duration of a call.
 *   - A handle owns its device state and one HIP stream; use one handle from
 *     one thread at a time.  Distinct handles are independent.
 *   - Functions with the _dev suffix take device pointers that must stay
 *     valid until the call returns; the call is ordered on the handle's
 *     stream and synchronises before returning unless stated otherwise.
 */
#ifndef HEXA_H
#define HEXA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define HEXA_OK               0
#define HEXA_E_INVALID       (-1)  /* bad argument / shape / state            */
#define HEXA_E_OOM           (-2)  /* device memory (or budget) exhausted     */
#define HEXA_E_NAN_INPUT     (-3)  /* NaN/Inf in the stiffness diagonal       */
#define HEXA_E_NOT_CONVERGED (-4)  /* solve stopped above tolerance           */
#define HEXA_E_DEVICE        (-5)  /* HIP runtime error / no device           */

/* ---- enumerations ---------------------------------------------------- */
enum { HEXA_PREC_FP64 = 0, HEXA_PREC_MIXED = 1 };            /* V-cycle precision */
enum { HEXA_KE_FP64 = 0, HEXA_KE_F32_WIDENED = 1 };           /* element matrix    */
enum { HEXA_HELM_BUG_COMPAT = 0, HEXA_HELM_CORRECTED = 1 };    /* filter element    */
/* HEXA_PRECOND_MG: geometric multigrid V-cycle (this library's);
 * HEXA_PRECOND_REF_MG: the reference's additive two-level preconditioner
 * z = D^-1 r + P (D_c + 1e-9)^-1 P^T r (GPUGeometricMultigrid.jl:253-346). */
enum { HEXA_PRECOND_JACOBI = 0, HEXA_PRECOND_MG = 1, HEXA_PRECOND_REF_MG = 2 };
enum { HEXA_MATVEC_AUTO = 0, HEXA_MATVEC_DENSE = 1, HEXA_MATVEC_FACTORED = 2 };
enum { HEXA_SEMANTICS_GPU = 0, HEXA_SEMANTICS_CPU = 1 };
enum { HEXA_DIAG_EXACT = 0, HEXA_DIAG_X_ENTRY = 1 };
/* Multigrid cycle: HEXA_CYCLE_V visits every level once per application;
 * HEXA_CYCLE_W makes two coarse corrections on every level but the finest,
 * each from the residual of the iterate the previous one left (levels l >= 2
 * are visited 2^(l-1) times). */
enum { HEXA_CYCLE_V = 0, HEXA_CYCLE_W = 1 };
/* Default bound on the coarsest level's DOFs (its inverse is dense).  On
 * solid/void designs the V-cycle's rediscretised coarse levels lose the
 * structure's members; a coarsest grid of up to 2200 DOFs keeps the 60x20x20
 * cantilever's MG-PCG at 13-24 iterations over 20 SIMP/OC steps, where 400
 * let it grow from 17 to 49. */
#define HEXA_COARSE_DOFS 2200
/* Several load cases (nrhs > 1): HEXA_RHS_AUTO solves them together, each
 * kernel of a PCG iteration working on all of them in one launch, when the
 * solve is a V-cycle-preconditioned one with CPU semantics on the factored
 * operator, and one after another otherwise; HEXA_RHS_SEQUENTIAL always one
 * after another.  Both give each load case bitwise the same result. */
enum { HEXA_RHS_AUTO = 0, HEXA_RHS_SEQUENTIAL = 1 };
/* method reported by a solve; names follow GPUSolver.jl:314,463 */
enum { HEXA_METHOD_JACOBI_GPU = 0, HEXA_METHOD_GMG_GPU = 1,
       HEXA_METHOD_FAILED_CORRUPT_INPUT = 2, HEXA_METHOD_JACOBI_CPU = 3,
       HEXA_METHOD_NONE = 4 /* zero load on the free DOFs (CPUSolver.jl:222-224) */ };

typedef struct hexa_handle hexa_handle;

typedef struct {
    int32_t nx, ny, nz;        /* elements per axis (>= 1; ny, nz <= 65534) */
    double dx, dy, dz;         /* element edge lengths (> 0)               */
} hexa_grid;

typedef struct {
    /* element stiffness multiplier  Emin + rho^penal * (E0 - Emin);
     * the reference's linear E*rho is E0=E, Emin=0, penal=1              */
    double E0, Emin, penal, nu;
    int32_t ke_mode;           /* HEXA_KE_*                                */
    int32_t helm_element;      /* HEXA_HELM_*                              */
    int32_t precision;         /* HEXA_PREC_*  (MG V-cycle arithmetic)      */
    int32_t matvec_kernel;     /* HEXA_MATVEC_*                            */
    int32_t mg_levels;         /* 0 = automatic                            */
    int32_t mg_pre, mg_post;   /* damped-Jacobi sweeps, equal, 1..16        */
    double mg_omega;           /* Jacobi damping                            */
    int32_t nrhs;              /* load cases solved together (>= 1)        */
    int32_t diag_mode;         /* HEXA_DIAG_*                               */
    int32_t mg_cycle;          /* HEXA_CYCLE_*                              */
    int32_t mg_coarse_dofs;    /* coarsen until the coarsest level has at most
                                  this many DOFs (0 = HEXA_COARSE_DOFS);
                                  ignored when mg_levels > 0               */
    int64_t reserved[7];
} hexa_options;

typedef struct {
    double tol;                /* relative residual ||r||/||f_free||        */
    int32_t max_iter;
    double shift_factor;       /* s = shift_factor * max diag(K_ff)         */
    int32_t max_shift_attempts;
    double shift_multiplier;
    double stagnation_tol;     /* accept best x below this (0 = off)       */
    int32_t precond;           /* HEXA_PRECOND_*                           */
    int32_t semantics;         /* HEXA_SEMANTICS_*                          */
    int32_t warm_start;        /* use u_inout as the attempt-1 guess       */
    int32_t check_every;       /* 0 = semantics default (GPU 10, CPU 1)    */
    int32_t rhs_mode;          /* HEXA_RHS_*: how nrhs > 1 load cases run  */
    int32_t pad;
    int64_t reserved[3];
} hexa_solve_params;

typedef struct {
    double rel_res;            /* last checked ||r||/||f_free||             */
    double true_rel_res;       /* ||f - A u||/||f|| recomputed in FP64      */
    double shift_used;
    double compliance;         /* f . u                                     */
    double setup_ms, solve_ms;
    int32_t iterations;
    int32_t attempt;
    int32_t method;            /* HEXA_METHOD_*                             */
    int32_t converged;
    int32_t mg_levels;
    int32_t pad;
    int64_t matvecs;
} hexa_solve_report;

/* ---- library / handle ------------------------------------------------- */
const char* hexa_last_error(void);
int32_t hexa_version(void);
void hexa_default_options(hexa_options* opt);
void hexa_default_solve_params(hexa_solve_params* prm);
/* Device name (NUL-terminated, truncated to len), total bytes, CU count. */
int32_t hexa_device_info(char* name, int32_t len, int64_t* total_bytes, int32_t* n_cu);

/* Replaces the per-call workspace set-up of GPUSolver.jl:246-256 and the
 * three module-global caches (GPUSolver.jl:63, GPUGeometricMultigrid.jl:37,
 * GPUHelmholtz.jl:41) by one handle with device-resident state. */
int32_t hexa_create(const hexa_grid* grid, const hexa_options* opt, hexa_handle** out);
int32_t hexa_destroy(hexa_handle* h);

/* Supports and loads.  fixed_dof: 3*nN bytes, nonzero = held (the
 * reference's bc_indicator > 0, GPUSolver.jl:170-180).  f: nrhs * 3*nN,
 * load case r at f + r*3*nN.  active_elem: nE bytes or NULL (all active). */
int32_t hexa_set_problem(hexa_handle* h, const uint8_t* fixed_dof, const double* f,
                         const uint8_t* active_elem);

/* Element densities (nE), FP64.  The stiffness multiplier is formed on the
 * device (GPUSolver.jl:254 for the reference's linear law). */
int32_t hexa_set_density(hexa_handle* h, const double* rho);
int32_t hexa_set_density_dev(hexa_handle* h, const double* d_rho);

/* Void pruning on the device, replacing the host-side
 * MeshPruner.prune_system step of Solver.solve_system (Solver.jl:85-116,
 * MeshPruner.jl:10-87): with the handle's current density, elements with
 * Float32(rho) <= Float32(min_stiffness_threshold) * 1.01f become inactive
 * when fewer than 99% of the elements would stay active (otherwise none
 * is); nodes no active element touches are excluded.  The node numbering
 * is unchanged, so a warm start needs no remap (Solver.jl:97-108).
 * min_stiffness_threshold < 0 makes every element active again.
 * n_active (may be NULL) receives the number of active elements. */
int32_t hexa_prune_voids(hexa_handle* h, double min_stiffness_threshold, int64_t* n_active);

/* Linear solve (K_ff(rho) + s I) u = f_free.  Replaces
 * GPUSolver.solve_system_gpu (GPUSolver.jl:467-477).  u_inout: nrhs*3*nN
 * host buffer, read as the warm start when prm->warm_start, written with
 * the solution (held DOFs = 0).  May be NULL: the solution then stays on
 * the device (hexa_get_displacement).  Returns HEXA_OK when every load
 * case reached tol, HEXA_E_NOT_CONVERGED when the best iterate was
 * returned instead (GPUSolver.jl:439-443), HEXA_E_NAN_INPUT for a NaN/Inf
 * diagonal (GPUSolver.jl:310-315; u is then zero).  rep may be NULL; for
 * nrhs > 1 rep points to nrhs reports. */
int32_t hexa_solve(hexa_handle* h, double* u_inout, const hexa_solve_params* prm,
                   hexa_solve_report* rep);
int32_t hexa_get_displacement(hexa_handle* h, double* u);
/* Device pointer of the solution (FP64), valid until the next solve or
 * destroy.  Load cases solved together (HEXA_RHS_AUTO) are all kept, load
 * case r at offset r*S with S = 3*nN rounded up to a multiple of 64; solved
 * one after another, the last one is kept at offset 0. */
int32_t hexa_displacement_dev(hexa_handle* h, const double** d_u);

/* y = K(rho) x on all DOFs (no supports, no shift): the bare operator of
 * matvec_unstructured_atomic_kernel! (GPUSolver.jl:91-128) on the full
 * vector.  Host arrays of 3*nN. */
int32_t hexa_matvec(hexa_handle* h, const double* x, double* y);
/* Same on device arrays; precision HEXA_PREC_FP64 (double*) or
 * HEXA_PREC_MIXED (float*); 'reps' back-to-back launches (for timing);
 * no synchronisation. */
int32_t hexa_matvec_dev(hexa_handle* h, const void* d_x, void* d_y, int32_t precision,
                        int32_t reps);
/* Apply the preconditioner of the last solve set-up to r (3*nN host
 * arrays, FP64): z = M^-1 r. */
int32_t hexa_apply_precond(hexa_handle* h, const double* r, double* z);

/* Ceilings of this GPU, measured: STREAM-style copy and triad over arrays of
 * 'bytes' each (GB/s counting every array read and written), and an FP64
 * FMA loop (TFLOP/s, 2 flops per FMA).  Allocates and frees its own arrays.
 * Any output may be NULL.  No reference counterpart. */
int32_t hexa_device_ceilings(int64_t bytes, double* copy_gbps, double* triad_gbps, double* fma_tflops);

/* Per-kernel time and achieved HBM rate of the solve loop's kernels on the
 * handle's grid (no reference counterpart: the reference reports no kernel
 * times).  Each kernel runs 'reps' times on the handle's stream between HIP
 * events, on the handle's own buffers, after a solve with HEXA_PRECOND_MG
 * (the V-cycle kernels of level 0 and the transfers to level 1 are timed at
 * the handle's V-cycle precision).  The Helmholtz stencil is included when a
 * filter has run on the handle, the OC multisection and SIMP sensitivity
 * when an OC step or update has.  'bytes' is the kernel's algorithmic HBM
 * traffic per launch: every array it must read or write, once.  The
 * solution, density and load stay as they were; the solver's scratch
 * vectors and the OC work arrays are overwritten.  Fills out[0..cap) and sets
 * *count to the number of entries written. */
typedef struct {
    char name[40];
    int32_t level;             /* multigrid level (0 = the handle's grid)     */
    int32_t launches;          /* launches timed                              */
    double ms;                 /* GPU time per launch                         */
    double bytes;              /* algorithmic HBM bytes per launch            */
    double gbps;               /* bytes / time, GB/s                          */
} hexa_kernel_stat;
int32_t hexa_kernel_report(hexa_handle* h, int32_t reps, hexa_kernel_stat* out, int32_t cap, int32_t* count);

/* Milliseconds of GPU time of the last hexa_matvec_dev batch (HIP events on
 * the handle's stream), averaged per launch. */
int32_t hexa_last_kernel_ms(hexa_handle* h, double* ms);
int32_t hexa_stream_sync(hexa_handle* h);

/* ---- Helmholtz density filter ------------------------------------------
 * Replaces GPUHelmholtz.apply_gpu_filter! (GPUHelmholtz.jl:422-443) with the
 * solve semantics of solve_helmholtz_on_gpu (:190-266): (R^2 K + M) x = b,
 * b_n = sum rho_e V/8, Jacobi-PCG from zero, ||r||/||b|| checked every
 * check_every iterations (reference: 10) up to max_iter (reference: 200),
 * output rho~_e = mean of the element's 8 nodal values.  The element matrix
 * follows opt->helm_element.  rho_in/rho_out: nE host arrays (may alias).
 * A zero right-hand side returns rho_in unchanged with 0 iterations. */
typedef struct {
    int32_t iterations;
    int32_t pad;
    double rel_res;            /* last checked residual; 0 if none checked */
    double ms;
} hexa_filter_report;
int32_t hexa_filter(hexa_handle* h, const double* rho_in, double* rho_out, double radius, double tol,
                    int32_t max_iter, int32_t check_every, hexa_filter_report* rep);

/* The reference's CPU filter fallback, run when its GPU filter is unavailable
 * (TopOpt.jl:319-351): (R^2 K + M + 1e-9 I) x = b solved exactly (here: PCG
 * to a relative residual of 1e-13) on the whole box when it has fewer than
 * min_blocked elements (<= 0: the reference's 100000), else the domain
 * decomposition of apply_domain_decomposed_cpu_filter! (TopOpt.jl:129-209):
 * blocks of about target_block elements (<= 0: 50000; ceil(n / cbrt(target))
 * blocks per axis), each solved alone on its box widened by
 * max(2, ceil(3 radius / min h)) cells and clipped to the grid, the block
 * cores written back.  That result approximates the global filter near the
 * block faces.  rho_in/rho_out: nE host arrays; rep: largest block's
 * iterations and residual. */
int32_t hexa_filter_fallback(hexa_handle* h, const double* rho_in, double* rho_out, double radius,
                             int64_t min_blocked, int32_t target_block, hexa_filter_report* rep);

/* ---- element stress ------------------------------------------------------
 * Replaces Stress.compute_stress_field (Stress.jl:287-374): centroid stress
 * of every element with E_e = E * rho_e, von Mises, principal values
 * (descending), l1 = sum |sigma_i|, directions of the largest/smallest;
 * rho < 1e-6 -> zeros.  u: 3nN host array or NULL (the handle's last
 * solution); rho: nE host array or NULL (the handle's density).  Outputs
 * are Float32 like the reference (3nE, nE, nE, 3nE, 3nE, 6nE); any may be NULL. */
int32_t hexa_stress(hexa_handle* h, const double* u, const double* rho, double E, float* principal,
                    float* vonmises, float* l1, float* dir_max, float* dir_min, float* voigt);

/* ---- SIMP sensitivity and OC update (no reference counterpart) ---------
 * hexa_sensitivity: dc/drho~_e = -p rho~^(p-1) (E0-Emin) u_e' K0 u_e for the
 * handle's density (the physical field) and last solution; also returns
 * the compliance.  dc: nE host array. */
int32_t hexa_sensitivity(hexa_handle* h, double* dc, double* compliance);
/* Optimality-criteria update of x (nE, host, in/out):
 * x_new = clamp(x sqrt(-dc/(lambda dv)), max(0,x-move), min(1,x+move)),
 * lambda by multisection on [0, 1e9] until (l2-l1)/(l1+l2) <= bisect_tol
 * such that sum w x_new = volfrac * nE.  Returns lambda in *lambda_out. */
int32_t hexa_oc_update(hexa_handle* h, double* x, const double* dc, const double* dv, const double* w,
                       double volfrac, double move, double bisect_tol, double* lambda_out);

/* hexa_oc_update_noalloc: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the OC update of hexa_oc_update on
 * device arrays (d_x: nE doubles, in/out; d_dc, d_dv, d_w: nE doubles), with
 * no handle and no device allocation.  Its scratch (the multisection's
 * 32 x 1024 partial sums and the lambda bracket, 262,304 bytes) is a fixed
 * array in device memory that every device the library is loaded on holds
 * until the process ends; the per-element move-limit values are formed
 * inside each pass, so the scratch does not grow with n (n in [1, 2^53]:
 * the volume target is counted in FP64).  Same x and lambda as
 * hexa_oc_update, bitwise.  Work is queued on stream (a hipStream_t, or NULL
 * for the null stream); lambda_out (host) may be NULL, in which case the call
 * does not synchronise (capture-safe); otherwise it waits for the stream.
 * All calls on one device share its scratch: calls that overlap on one
 * device race. */
int32_t hexa_oc_update_noalloc(int64_t n, double* d_x, const double* d_dc, const double* d_dv, const double* d_w,
                               double volfrac, double move, double bisect_tol, void* stream, double* lambda_out);

typedef struct {
    double volfrac, move, rmin;   /* rmin = Helmholtz radius of the density filter */
    double filter_tol, bisect_tol;
    int32_t filter_max_iter;
    int32_t pad;
    hexa_solve_params solve;      /* warm_start uses the previous step's u */
} hexa_oc_params;
typedef struct {
    double compliance, volume, change, lambda;
    double solve_ms, step_ms;
    int32_t solve_iterations, filter_iterations;
    int32_t status;               /* status of the linear solve */
    int32_t pad;
} hexa_oc_report;
/* One SIMP/OC iteration on the device: x~ = F(x) (PDE filter), solve
 * K(x~) u = f, dc~, dc = F^T dc~, dv = F^T 1, OC update of the design x
 * (nE, host, in/out).  The filter is linear and symmetric, so F^T = F and
 * the volume of the filtered design is sum (F 1)_e x_e. */
int32_t hexa_oc_step(hexa_handle* h, double* x, const hexa_oc_params* p, hexa_oc_report* rep);
/* The same step with the design in device memory (d_x, nE doubles, in/out):
 * no host transfer of x.  No counterpart in the reference; for callers that
 * keep the design on the GPU across iterations (the job server). */
int32_t hexa_oc_step_dev(hexa_handle* h, double* d_x, const hexa_oc_params* p, hexa_oc_report* rep);

/* ---- reference stress-ratio density update -------------------------------
 * Replaces TopologyOptimization.update_density! (TopOpt.jl:211-482) on the
 * handle's grid: proposed density from l1, radius schedule, Helmholtz
 * filter with R/2.5, clamp, mean change, threshold schedule, culling of the
 * lowest candidates in stable (density, index) order, optional "hard"
 * projection, protected elements restored.  density (nE, Float32, in/out),
 * l1 (nE), protected_mask (nE bytes), original_density (nE). */
typedef struct {
    double E, l1_allow, min_density, max_density, max_domain_dim, filter_tol;
    double R_init_perc, R_interm_perc, R_final_perc, R_interm_iter_perc; /* percents */
    double min_radius_factor, final_threshold, max_culling_ratio;
    int32_t iter, n_iters, hard, pad;
} hexa_update_params;
typedef struct {
    double mean_change, R_length, threshold, filter_res;
    int32_t filter_iters, skipped;  /* skipped: NaN in l1 (TopOpt.jl:230-236) */
    int64_t culled;
} hexa_update_report;
int32_t hexa_update_density(hexa_handle* h, float* density, const float* l1, const uint8_t* protected_mask,
                            const float* original_density, const hexa_update_params* p,
                            hexa_update_report* rep);

/* The reference's outer iteration kept on the device (Main.jl:439-493):
 * solve (K(rho) + s I) u = f with the handle's density (after void pruning
 * when prune_threshold >= 0, as Solver.jl:85-116), the stress l1 of that u
 * (Stress.jl:159-285, material E), and update_density! (TopOpt.jl:211-482)
 * of the handle's density, which then becomes the handle's density.  No
 * field crosses PCIe: the protected flags and original densities are set
 * once with hexa_set_update_fields; density_out (nE Float32, may be NULL)
 * receives a copy of the new density.  Returns the solve's status. */
int32_t hexa_set_update_fields(hexa_handle* h, const uint8_t* protected_mask, const float* original_density);
int32_t hexa_ref_step(hexa_handle* h, const hexa_solve_params* solve, double E, double prune_threshold,
                      const hexa_update_params* update, hexa_solve_report* solve_report,
                      hexa_update_report* update_report, float* density_out);

/* ---- files and regridding from device fields ----------------------------
 * hexa_write_vti: VTK XML image data (.vti appended if missing) with the
 * layout of ExportVTK.export_solution_vti (ExportVTK.jl:16-111): Density,
 * L1_Stress, VonMises and, when write_vectors, PrincipalValues and the
 * max/min principal directions, computed from the handle's device solution
 * (stress with material E).  density: nE Float32 host array, or NULL for
 * the handle's density.  Origin and the handle's spacing go to the header.
 * hexa_write_bin: the .bin checkpoint of Postprocessing.export_binary_for_web
 * (Postprocessing.jl:370-471): elements with density >= threshold, strided
 * down to max_export_cells when > 0; full_field = 1 writes every element
 * instead, which is what Configuration.load_checkpoint (Configuration.jl:
 * 32-78) restarts from.  json: the config JSON (NULL = "{}"). */
int32_t hexa_write_vti(hexa_handle* h, const char* path, const float* density, double E, double origin_x,
                       double origin_y, double origin_z, int32_t write_vectors);
int32_t hexa_write_bin(hexa_handle* h, const char* path, const float* density, double E, int32_t iter,
                       float radius, float threshold, int64_t max_export_cells, int32_t full_field,
                       const char* json);
/* Density (and optional alpha) remap onto a refined grid, device arrays
 * (MeshRefiner.jl:185-237): new element centroids mapped into the old cell
 * index space of a domain len_x * len_y * len_z, trilinear interpolation of
 * the clamped neighbours, hard != 0 projects (> 0.4 -> 1, < 0.2 -> 1e-4);
 * alpha takes the nearest-lower neighbour.  stream: a hipStream_t or NULL. */
int32_t hexa_regrid_density(const hexa_grid* old_grid, const float* d_old_density, const float* d_old_alpha,
                            const hexa_grid* new_grid, double len_x, double len_y, double len_z, int32_t hard,
                            float* d_new_density, float* d_new_alpha, void* stream);
/* The same with host arrays (device copies made and freed inside). */
int32_t hexa_regrid_density_host(const hexa_grid* old_grid, const float* old_density, const float* old_alpha,
                                 const hexa_grid* new_grid, double len_x, double len_y, double len_z, int32_t hard,
                                 float* new_density, float* new_alpha);

/* ---- job server (no reference counterpart) ------------------------------
 * Topology-optimisation jobs of mixed mesh sizes served on one GPU.  Each
 * job: a box of nx*ny*nz unit bricks, all DOFs of one face held
 * (support_face 0..5 = x0, x1, y0, y1, z0, z1), n_loads (1-4) unit point
 * loads at seeded nodes of the opposite face with seeded directions, SIMP
 * p=3, Emin=1e-9, uniform start rho = volfrac, oc_iters iterations of
 * hexa_oc_step (MG-preconditioned solves to tol).  Workers take queued jobs
 * first-fit when a worker is free and the job's estimated device bytes fit
 * in the remaining budget (the admission rule generalises the VRAM guard of
 * R/src/Utils/Helpers.jl:177-244); jobs run concurrently on their own
 * streams.  Results do not depend on what else runs. */
typedef struct hexa_server hexa_server;
typedef struct {
    int32_t nx, ny, nz;
    int32_t support_face;
    int32_t n_loads;
    int32_t oc_iters;
    int32_t precision;         /* HEXA_PREC_* */
    int32_t pad;
    double volfrac, move, rmin, tol;
    uint64_t seed;
} hexa_job;
typedef struct {
    int64_t id;
    int32_t status, oc_iters_done;
    int32_t solve_iterations, pad;
    double compliance, volume;
    double t_submit_ms, t_start_ms, t_end_ms;   /* server clock */
    int64_t est_bytes, device_bytes;
} hexa_job_result;
/* Estimated device bytes of a handle on this grid (used for admission). */
int64_t hexa_estimate_bytes(const hexa_grid* grid, const hexa_options* opt);
int32_t hexa_server_create(int32_t workers, int64_t mem_budget_bytes, hexa_server** out);
int32_t hexa_server_submit(hexa_server* s, const hexa_job* job, int64_t* id);
/* Blocks until the job ends; x_out (nE) receives the final design or NULL.
 * One wait per job: the wait consumes the job's result, and a second wait
 * on an id already waited for (or being waited for) returns HEXA_E_INVALID. */
int32_t hexa_server_wait(hexa_server* s, int64_t id, hexa_job_result* res, double* x_out);
double hexa_server_clock_ms(hexa_server* s);
int32_t hexa_server_destroy(hexa_server* s);

#ifdef __cplusplus
}
#endif
#endif /* HEXA_H */
