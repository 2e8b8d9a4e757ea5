// libhexa C ABI, optimisation step: Helmholtz filter, element stress, SIMP
// sensitivity, optimality-criteria update, one full SIMP/OC iteration, and
// the reference's stress-ratio density update.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handle.h"
#include "hexa_internal.h"
#include "ke.h"

using namespace hexa;

namespace hexa {

// (R^2 K + M) x = b, b_n = sum rho_e V/8, Jacobi-PCG from zero with the
// residual checked every check_every iterations; d_out = element means of x.
// R/src/Optimization/GPUHelmholtz.jl:190-266.
int32_t filter_dev(hexa_handle* h, const double* d_in, double* d_out, double radius, double tol, int max_iter,
                   int check_every, int* iters, double* rel, bool fused, double reg) {
    hipStream_t st = h->stream;
    const Grid3& g = h->g;
    const int64_t nn = g.nn;
    HX_TRY(dalloc(h, &h->d_Ae, 64 + 27 + 1));
    HX_TRY(dalloc(h, &h->d_hb, (size_t)nn));
    HX_TRY(dalloc(h, &h->d_hx, (size_t)nn));
    HX_TRY(dalloc(h, &h->d_hr, (size_t)nn));
    HX_TRY(dalloc(h, &h->d_hp, (size_t)nn));
    HX_TRY(dalloc(h, &h->d_hAp, (size_t)nn));
    HX_TRY(dalloc(h, &h->d_hdinv, (size_t)nn));
    HX_TRY(ensure_shared_buffers(h));
    double K8[64], M8[64], Ae[64 + 27 + 1];
    helmholtz_element(h->grid.dx, h->grid.dy, h->grid.dz, h->opt.helm_element == HEXA_HELM_BUG_COMPAT, K8, M8);
    for (int i = 0; i < 64; ++i) Ae[i] = radius * radius * K8[i] + M8[i];
    helm_stencil(Ae, Ae + 64);
    Ae[64 + 27] = reg;
    HX_CHECK(hipMemcpyAsync(h->d_Ae, Ae, sizeof Ae, hipMemcpyHostToDevice, st));
    const double w = h->grid.dx * h->grid.dy * h->grid.dz / 8.0;
    launch_helm_rhs_diag(st, g, h->d_Ae, d_in, w, h->d_hb, h->d_hdinv);
    launch_dot2(st, nn, h->d_hb, h->d_hb, h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal);
    launch_pcg_init_state(st, h->d_state, h->d_partials, tol);
    HX_TRY(check_launch());
    double bb = 0.0;
    HX_CHECK(hipMemcpyAsync(&bb, h->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (bb == 0.0) {
        if (d_out != d_in)
            HX_CHECK(hipMemcpyAsync(d_out, d_in, sizeof(double) * g.ne, hipMemcpyDeviceToDevice, st));
        *iters = 0;
        *rel = 0.0;
        return HEXA_OK;
    }
    launch_fill<double>(st, nn, 0.0, h->d_hx);
    HX_CHECK(hipMemcpyAsync(h->d_hr, h->d_hb, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    launch_scale_dinv<double, double>(st, nn, h->d_hr, h->d_hdinv, 1.0, h->d_hp);
    launch_pcg_rz(st, nn, h->d_hr, h->d_hp, h->d_partials);
    launch_pcg_set_rz(st, h->d_state, h->d_partials);
    int k_done = 0;
    bool stop = false;
    fused = fused && check_every == 1;
    auto iteration = [&]() -> int32_t {
        if (fused && launch_filter_fused_iteration(st, g, h->d_Ae, h->d_state, h->d_hx, h->d_hp, h->d_hr, h->d_hAp,
                                                   h->d_hdinv, h->d_partials) == 0)
            return check_launch();
        launch_helm_apply(st, g, h->d_Ae, h->d_hp, h->d_hAp);
        launch_pcg_pAp(st, nn, h->d_state, h->d_hp, h->d_hAp, h->d_partials);
        launch_pcg_alpha(st, h->d_state, h->d_partials, 0);
        launch_pcg_update(st, nn, h->d_state, h->d_hx, h->d_hp, h->d_hr, h->d_hAp, h->d_hdinv, h->d_partials,
                          h->d_upart);
        launch_pcg_check(st, h->d_state, h->d_partials, check_every, 0, 1, 1);
        launch_pcg_p_update(st, nn, h->d_state, h->d_hp, nullptr, h->d_hr, h->d_hdinv);
        return check_launch();
    };
    // replayed from a captured graph, as the elasticity PCG (the buffers are
    // the handle's own, so only the check period keys it)
    const char* ge = getenv("HEXA_GRAPHS");
    const bool graphs = !(ge && ge[0] == '0');
    if (graphs && !(h->filt_exec && h->filt_every == check_every && h->filt_fused == fused)) {
        CaptureGuard lock;
        if (h->filt_exec) {
            HX_CHECK(hipGraphExecDestroy(h->filt_exec));
            h->filt_exec = nullptr;
        }
        HX_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int32_t rc = iteration();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != HEXA_OK || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (rc != HEXA_OK) return rc;
            set_error("filter iteration capture failed: %s", hipGetErrorString(e));
            return HEXA_E_DEVICE;
        }
        const hipError_t ei = hipGraphInstantiate(&h->filt_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            h->filt_exec = nullptr;
            set_error("filter iteration graph instantiation failed: %s", hipGetErrorString(ei));
            return HEXA_E_DEVICE;
        }
        h->filt_every = check_every;
        h->filt_fused = fused;
    }
    for (int k = 1; k <= max_iter && !stop; ++k) {
        if (graphs) HX_CHECK(hipGraphLaunch(h->filt_exec, st));
        else HX_TRY(iteration());
        k_done = k;
        HX_TRY(poll_state(h, k, &stop));
    }
    HX_CHECK(hipMemcpyAsync(&h->h_state[0], h->d_state, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    launch_helm_extract(st, g, h->d_hx, 0.125, d_out);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(st));
    const PcgState& fs = h->h_state[0];
    *iters = fs.halted != PCG_RUNNING ? fs.last_k : k_done;
    *rel = fs.final_rel;
    return HEXA_OK;
}

}  // namespace hexa

namespace {

int32_t ensure_elem(hexa_handle* h) {
    const size_t ne = (size_t)h->g.ne;
    HX_TRY(dalloc(h, &h->d_e0, ne));
    HX_TRY(dalloc(h, &h->d_e1, ne));
    HX_TRY(dalloc(h, &h->d_e2, ne));
    HX_TRY(dalloc(h, &h->d_e3, ne));
    HX_TRY(dalloc(h, &h->d_e4, ne));
    HX_TRY(dalloc(h, &h->d_e5, ne));
    HX_TRY(dalloc(h, &h->d_e6, ne));
    HX_TRY(ensure_shared_buffers(h));
    HX_TRY(dalloc_bytes(h, &h->d_bracket, (size_t)oc_bracket_bytes()));
    return HEXA_OK;
}

// w = F(1): the filtered volume weights, cached per radius.
int32_t filter_weights(hexa_handle* h, double radius, double tol, int max_iter) {
    if (h->w_radius == radius && h->d_w) return HEXA_OK;
    HX_TRY(dalloc(h, &h->d_w, (size_t)h->g.ne));
    HX_TRY(dalloc(h, &h->d_ones, (size_t)h->g.ne));
    launch_fill<double>(h->stream, h->g.ne, 1.0, h->d_ones);
    int it;
    double rel;
    HX_TRY(filter_dev(h, h->d_ones, h->d_w, radius, tol, max_iter, 1, &it, &rel, true));
    h->w_radius = radius;
    return HEXA_OK;
}

// Radius schedule of update_density! (TopOpt.jl:265-300), in float.
float radius_schedule(const hexa_update_params* P, double dx, double dy, double dz) {
    const float mdd = (float)P->max_domain_dim;
    const float Ri = (float)P->R_init_perc / 100.0f * mdd, Rm = (float)P->R_interm_perc / 100.0f * mdd,
                Rf = (float)P->R_final_perc / 100.0f * mdd;
    const float ip = (float)P->R_interm_iter_perc / 100.0f;
    int iter_interm = (int)std::nearbyint(ip * (float)P->n_iters);  // Julia round: ties to even
    if (iter_interm < 1) iter_interm = 1;
    const int calc = P->iter < P->n_iters ? P->iter : P->n_iters;
    float R;
    if (calc <= iter_interm) {
        const float t = iter_interm > 1 ? (float)(calc - 1) / (float)(iter_interm - 1) : 0.0f;
        R = Ri * (1.0f - t) + Rm * t;
    } else {
        const float t = P->n_iters > iter_interm ? (float)(calc - iter_interm) / (float)(P->n_iters - iter_interm) : 0.0f;
        R = Rm * (1.0f - t) + Rf * t;
    }
    const float avg = ((float)dx + (float)dy + (float)dz) / 3.0f;
    const float Rsafe = (float)P->min_radius_factor * avg;
    return R < Rsafe ? Rsafe : R;
}

}  // namespace

extern "C" {

int32_t hexa_filter(hexa_handle* h, const double* rho_in, double* rho_out, double radius, double tol,
                    int32_t max_iter, int32_t check_every, hexa_filter_report* rep) {
    if (!h || !rho_in || !rho_out) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!(radius >= 0) || !(tol > 0) || max_iter < 0) { set_error("invalid filter parameters"); return HEXA_E_INVALID; }
    HX_TRY(ensure_elem(h));
    const double t0 = now_ms();
    HX_CHECK(hipMemcpyAsync(h->d_e0, rho_in, sizeof(double) * h->g.ne, hipMemcpyHostToDevice, h->stream));
    int it = 0;
    double rel = 0.0;
    HX_TRY(filter_dev(h, h->d_e0, h->d_e1, radius, tol, max_iter, check_every, &it, &rel));
    HX_CHECK(hipMemcpyAsync(rho_out, h->d_e1, sizeof(double) * h->g.ne, hipMemcpyDeviceToHost, h->stream));
    HX_CHECK(hipStreamSynchronize(h->stream));
    if (rep) {
        rep->iterations = it;
        rep->pad = 0;
        rep->rel_res = rel;
        rep->ms = now_ms() - t0;
    }
    return HEXA_OK;
}

// The reference's CPU filter (TopOpt.jl:319-351): below min_blocked
// elements the whole box is solved, (R^2 K + M + 1e-9 I) x = b with a direct
// factorisation (:32-127); from min_blocked elements on, the box is split
// into blocks of about target_block elements, each solved alone on its box
// widened by max(2, ceil(3R / min h)) halo cells (clipped to the box), and
// the block cores are copied back (:129-209).  The direct solves are PCG
// solves here, to a relative residual of 1e-13.
int32_t hexa_filter_fallback(hexa_handle* h, const double* rho_in, double* rho_out, double radius,
                             int64_t min_blocked, int32_t target_block, hexa_filter_report* rep) {
    if (!h || !rho_in || !rho_out) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!(radius >= 0)) { set_error("invalid filter radius"); return HEXA_E_INVALID; }
    if (min_blocked <= 0) min_blocked = 100000;
    if (target_block <= 0) target_block = 50000;
    constexpr double REG = 1e-9, TOL = 1e-13;
    constexpr int MAX_IT = 20000;
    HX_TRY(ensure_elem(h));
    const double t0 = now_ms();
    hipStream_t st = h->stream;
    const Grid3& g = h->g;
    HX_CHECK(hipMemcpyAsync(h->d_e0, rho_in, sizeof(double) * g.ne, hipMemcpyHostToDevice, st));
    int it_max = 0;
    double rel_max = 0.0;
    if (g.ne < min_blocked) {
        HX_TRY(filter_dev(h, h->d_e0, h->d_e1, radius, TOL, MAX_IT, 1, &it_max, &rel_max, true, REG));
    } else {
        HX_CHECK(hipStreamSynchronize(st));
        const double cube = std::cbrt((double)target_block);
        const int n[3] = {g.nx, g.ny, g.nz};
        int chunks[3], blk[3];
        for (int a = 0; a < 3; ++a) {
            chunks[a] = std::max(1, (int)std::ceil(n[a] / cube));
            blk[a] = (n[a] + chunks[a] - 1) / chunks[a];
        }
        const double hmin = std::min(h->grid.dx, std::min(h->grid.dy, h->grid.dz));
        const int halo = std::max(2, (int)std::ceil(3.0 * radius / hmin));
        // one handle per distinct block box size, made on first use
        std::vector<std::pair<std::array<int, 3>, hexa_handle*>> subs;
        int32_t rc = HEXA_OK;
        for (int cz = 0; cz < chunks[2] && rc == HEXA_OK; ++cz)
            for (int cy = 0; cy < chunks[1] && rc == HEXA_OK; ++cy)
                for (int cx = 0; cx < chunks[0] && rc == HEXA_OK; ++cx) {
                    const int c[3] = {cx, cy, cz};
                    int s0[3], s1[3], h0[3], h1[3];
                    bool empty = false;
                    for (int a = 0; a < 3; ++a) {
                        s0[a] = c[a] * blk[a];                       // core [s0, s1), 0-based
                        s1[a] = std::min((c[a] + 1) * blk[a], n[a]);
                        empty = empty || s0[a] >= s1[a];
                        h0[a] = std::max(0, s0[a] - halo);
                        h1[a] = std::min(n[a], s1[a] + halo);
                    }
                    if (empty) continue;
                    const std::array<int, 3> dims = {h1[0] - h0[0], h1[1] - h0[1], h1[2] - h0[2]};
                    hexa_handle* sh = nullptr;
                    for (auto& kv : subs)
                        if (kv.first == dims) sh = kv.second;
                    if (!sh) {
                        hexa_grid sg{dims[0], dims[1], dims[2], h->grid.dx, h->grid.dy, h->grid.dz};
                        rc = hexa_create(&sg, &h->opt, &sh);
                        if (rc != HEXA_OK) break;
                        subs.emplace_back(dims, sh);
                        rc = ensure_elem(sh);
                        if (rc != HEXA_OK) break;
                    }
                    launch_box_gather(sh->stream, g, sh->g, h0[0], h0[1], h0[2], h->d_e0, sh->d_e0);
                    int it = 0;
                    double rel = 0.0;
                    rc = filter_dev(sh, sh->d_e0, sh->d_e1, radius, TOL, MAX_IT, 1, &it, &rel, true, REG);
                    if (rc != HEXA_OK) break;
                    launch_box_scatter(sh->stream, g, sh->g, h0[0], h0[1], h0[2], s0[0] - h0[0], s0[1] - h0[1],
                                       s0[2] - h0[2], s1[0] - s0[0], s1[1] - s0[1], s1[2] - s0[2], sh->d_e1, h->d_e1);
                    rc = check_launch();
                    it_max = std::max(it_max, it);
                    rel_max = std::max(rel_max, rel);
                }
        for (auto& kv : subs) {
            if (hipStreamSynchronize(kv.second->stream) != hipSuccess && rc == HEXA_OK) {
                set_error("blocked filter: stream synchronisation failed");
                rc = HEXA_E_DEVICE;
            }
            hexa_destroy(kv.second);
        }
        HX_TRY(rc);
    }
    HX_CHECK(hipMemcpyAsync(rho_out, h->d_e1, sizeof(double) * g.ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (rep) {
        rep->iterations = it_max;
        rep->pad = 0;
        rep->rel_res = rel_max;
        rep->ms = now_ms() - t0;
    }
    return HEXA_OK;
}

int32_t hexa_stress(hexa_handle* h, const double* u, const double* rho, double E, float* principal,
                    float* vonmises, float* l1, float* dir_max, float* dir_min, float* voigt) {
    if (!h) { set_error("null handle"); return HEXA_E_INVALID; }
    if (!u && !h->have_solution) { set_error("no displacement: pass u or solve first"); return HEXA_E_INVALID; }
    if (!rho && !h->have_density) { set_error("no density: pass rho or set a density first"); return HEXA_E_INVALID; }
    hipStream_t st = h->stream;
    const int64_t ne = h->g.ne, nd = 3 * h->g.nn;
    HX_TRY(ensure_elem(h));
    HX_TRY(dalloc(h, &h->d_stress, (size_t)(17 * ne)));
    const double* du = h->d_x;
    if (u) {
        HX_TRY(dalloc(h, &h->d_z, (size_t)nd));
        HX_CHECK(hipMemcpyAsync(h->d_z, u, sizeof(double) * nd, hipMemcpyHostToDevice, st));
        du = h->d_z;
    }
    const double* dr = h->d_rho;
    if (rho) {
        HX_CHECK(hipMemcpyAsync(h->d_e0, rho, sizeof(double) * ne, hipMemcpyHostToDevice, st));
        dr = h->d_e0;
    }
    float* base = h->d_stress;
    float* dp = principal ? base : nullptr;
    float* dv = vonmises ? base + 3 * ne : nullptr;
    float* dl = l1 ? base + 4 * ne : nullptr;
    float* dmx = dir_max ? base + 5 * ne : nullptr;
    float* dmn = dir_min ? base + 8 * ne : nullptr;
    float* dvo = voigt ? base + 11 * ne : nullptr;
    launch_stress(st, h->g, h->grid.dx, h->grid.dy, h->grid.dz, E, h->opt.nu, du, dr, dp, dv, dl, dmx, dmn, dvo);
    HX_TRY(check_launch());
    if (principal) HX_CHECK(hipMemcpyAsync(principal, dp, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
    if (vonmises) HX_CHECK(hipMemcpyAsync(vonmises, dv, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    if (l1) HX_CHECK(hipMemcpyAsync(l1, dl, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    if (dir_max) HX_CHECK(hipMemcpyAsync(dir_max, dmx, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
    if (dir_min) HX_CHECK(hipMemcpyAsync(dir_min, dmn, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
    if (voigt) HX_CHECK(hipMemcpyAsync(voigt, dvo, sizeof(float) * 6 * ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    return HEXA_OK;
}

int32_t hexa_sensitivity(hexa_handle* h, double* dc, double* compliance) {
    if (!h || !dc) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!h->have_solution || !h->have_density) { set_error("solve first"); return HEXA_E_INVALID; }
    HX_TRY(ensure_elem(h));
    hipStream_t st = h->stream;
    if (h->factored)
        launch_simp_sens_fact(st, h->g, h->coef64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, h->d_e2,
                              h->d_partials);
    else
        launch_simp_sens(st, h->g, h->d_Ke64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, h->d_e2,
                         h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal);
    HX_TRY(check_launch());
    double c = 0.0;
    HX_CHECK(hipMemcpyAsync(&c, h->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipMemcpyAsync(dc, h->d_e2, sizeof(double) * h->g.ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (compliance) *compliance = c;
    return HEXA_OK;
}

int32_t hexa_oc_update(hexa_handle* h, double* x, const double* dc, const double* dv, const double* w,
                       double volfrac, double move, double bisect_tol, double* lambda_out) {
    if (!h || !x || !dc || !dv || !w) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!(volfrac > 0 && volfrac <= 1) || !(move > 0) || !(bisect_tol > 0)) {
        set_error("invalid OC parameters");
        return HEXA_E_INVALID;
    }
    HX_TRY(ensure_elem(h));
    hipStream_t st = h->stream;
    const int64_t ne = h->g.ne;
    HX_CHECK(hipMemcpyAsync(h->d_e0, x, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_e1, dc, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_e2, dv, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_e3, w, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    launch_oc(st, ne, h->d_e1, h->d_e2, h->d_e3, volfrac, move, bisect_tol, 48, h->d_e0, h->d_e4, h->d_e5, h->d_e6,
              h->d_bracket, h->d_partials, h->d_scal);
    HX_TRY(check_launch());
    double out[4];
    HX_CHECK(hipMemcpyAsync(out, h->d_scal, sizeof out, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipMemcpyAsync(x, h->d_e0, sizeof(double) * ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (lambda_out) *lambda_out = out[2];
    return HEXA_OK;
}

}  // extern "C"

namespace hexa {

int32_t oc_check(hexa_handle* h, const hexa_oc_params* p) {
    if (!h->have_problem) { set_error("hexa_set_problem has not been called"); return HEXA_E_INVALID; }
    if (!(p->volfrac > 0 && p->volfrac <= 1) || !(p->move > 0) || !(p->rmin >= 0) || !(p->filter_tol > 0) ||
        !(p->bisect_tol > 0) || p->filter_max_iter < 1) {
        set_error("invalid OC step parameters");
        return HEXA_E_INVALID;
    }
    HX_TRY(ensure_elem(h));
    HX_TRY(dalloc(h, &h->d_rho, (size_t)h->g.ne));
    return HEXA_OK;
}

// One SIMP/OC iteration on the design held in h->d_e0 (in/out); returns with
// the stream synchronized and rep filled.
int32_t oc_step_core(hexa_handle* h, const hexa_oc_params* p, hexa_oc_report* rep, double t0) {
    hipStream_t st = h->stream;
    const int64_t ne = h->g.ne;
    hexa_oc_report R{};
    // physical density x~ = F(x)
    int fit = 0;
    double frel = 0.0;
    HX_TRY(filter_dev(h, h->d_e0, h->d_rho, p->rmin, p->filter_tol, p->filter_max_iter, 1, &fit, &frel, true));
    R.filter_iterations = fit;
    HX_TRY(update_factor(h));
    // K(x~) u = f
    hexa_solve_report sr{};
    int32_t ss = hexa_solve(h, nullptr, &p->solve, &sr);
    R.status = ss;
    R.solve_iterations = sr.iterations;
    R.solve_ms = sr.solve_ms;
    if (ss == HEXA_E_NAN_INPUT || ss == HEXA_E_OOM || ss == HEXA_E_DEVICE || ss == HEXA_E_INVALID) {
        if (rep) *rep = R;
        return ss;
    }
    // dc~ and compliance, dc = F^T dc~ = F dc~, dv = F^T 1 = w
    if (h->factored)
        launch_simp_sens_fact(st, h->g, h->coef64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, h->d_e2,
                              h->d_partials);
    else
        launch_simp_sens(st, h->g, h->d_Ke64, h->d_rho, h->opt.E0, h->opt.Emin, h->opt.penal, h->d_x, h->d_e2,
                         h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal + 4);
    HX_TRY(check_launch());
    HX_TRY(filter_dev(h, h->d_e2, h->d_e3, p->rmin, p->filter_tol, p->filter_max_iter, 1, &fit, &frel, true));
    HX_TRY(filter_weights(h, p->rmin, p->filter_tol, p->filter_max_iter));
    launch_oc(st, ne, h->d_e3, h->d_w, h->d_w, p->volfrac, p->move, p->bisect_tol, 48, h->d_e0, h->d_e4, h->d_e5,
              h->d_e6, h->d_bracket, h->d_partials, h->d_scal);
    HX_TRY(check_launch());
    double out[5];
    HX_CHECK(hipMemcpyAsync(out, h->d_scal, sizeof out, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    R.change = out[0];
    R.volume = out[1] / (double)ne;
    R.lambda = out[2];
    R.compliance = out[4];
    R.step_ms = now_ms() - t0;
    if (rep) *rep = R;
    return ss;
}

}  // namespace hexa

extern "C" {

int32_t hexa_oc_step(hexa_handle* h, double* x, const hexa_oc_params* p, hexa_oc_report* rep) {
    if (!h || !x || !p) { set_error("null argument"); return HEXA_E_INVALID; }
    const double t0 = now_ms();
    HX_TRY(oc_check(h, p));
    const size_t bytes = sizeof(double) * h->g.ne;
    HX_CHECK(hipMemcpyAsync(h->d_e0, x, bytes, hipMemcpyHostToDevice, h->stream));
    const int32_t ss = oc_step_core(h, p, rep, t0);
    if (ss != HEXA_OK && ss != HEXA_E_NOT_CONVERGED) return ss;
    HX_CHECK(hipMemcpyAsync(x, h->d_e0, bytes, hipMemcpyDeviceToHost, h->stream));
    HX_CHECK(hipStreamSynchronize(h->stream));
    return ss;
}

int32_t hexa_oc_step_dev(hexa_handle* h, double* d_x, const hexa_oc_params* p, hexa_oc_report* rep) {
    if (!h || !d_x || !p) { set_error("null argument"); return HEXA_E_INVALID; }
    const double t0 = now_ms();
    HX_TRY(oc_check(h, p));
    const size_t bytes = sizeof(double) * h->g.ne;
    HX_CHECK(hipMemcpyAsync(h->d_e0, d_x, bytes, hipMemcpyDeviceToDevice, h->stream));
    const int32_t ss = oc_step_core(h, p, rep, t0);
    if (ss != HEXA_OK && ss != HEXA_E_NOT_CONVERGED) return ss;
    HX_CHECK(hipMemcpyAsync(d_x, h->d_e0, bytes, hipMemcpyDeviceToDevice, h->stream));
    HX_CHECK(hipStreamSynchronize(h->stream));
    return ss;
}

}  // extern "C"

namespace {

// update_density! (TopOpt.jl:211-482) on device fields: density d_f0 (in/out),
// l1 d_f1, original density d_f2, protected flags d_u0; the element work
// arrays must exist (ensure_elem).
int32_t update_core(hexa_handle* h, const hexa_update_params* P, hexa_update_report* rep) {
    const int64_t ne = h->g.ne;
    hexa_update_report R{};
    hipStream_t st = h->stream;
    // TopOpt.jl:230-236: NaN in l1 skips the update
    launch_nan_count(st, ne, (const float*)h->d_f1, h->d_partials);
    launch_final_sum(st, 1, h->d_partials, h->d_scal);
    HX_TRY(check_launch());
    double nanc = 0.0;
    HX_CHECK(hipMemcpyAsync(&nanc, h->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    if (nanc > 0.0) {
        R.skipped = 1;
        if (rep) *rep = R;
        return HEXA_OK;
    }
    HX_TRY(dalloc(h, &h->d_u1, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_hist, 256));
    HX_TRY(dalloc(h, &h->d_ties, (size_t)(2 * RED_BLOCKS)));
    const float mind = (float)P->min_density, maxd = (float)P->max_density;
    launch_upd_proposed(st, ne, h->d_f1, h->d_u0, h->d_f2, (float)P->l1_allow, (float)P->E, mind, h->d_e0);
    const float R_len = radius_schedule(P, h->grid.dx, h->grid.dy, h->grid.dz);
    const float Reff = R_len / 2.5f;
    int use_prop = 1;
    if (Reff > 1e-4f) {
        int it = 0;
        double rel = 0.0;
        HX_TRY(filter_dev(h, h->d_e0, h->d_e1, (double)Reff, P->filter_tol, 200, 10, &it, &rel));
        R.filter_iters = it;
        R.filter_res = rel;
        launch_nan_count(st, ne, h->d_e1, h->d_partials);
        launch_final_sum(st, 1, h->d_partials, h->d_scal);
        double nn = 0.0;
        HX_CHECK(hipMemcpyAsync(&nn, h->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
        HX_CHECK(hipStreamSynchronize(st));
        use_prop = nn > 0.0;  // TopOpt.jl:361-364
    }
    launch_upd_apply(st, ne, h->d_e1, h->d_e0, use_prop, h->d_u0, mind, maxd, h->d_f0, h->d_partials);
    launch_final_sum(st, 2, h->d_partials, h->d_scal);
    float thr;
    if (P->iter > P->n_iters) thr = (float)P->final_threshold;
    else {
        float pr = (float)P->iter / (float)P->n_iters;
        pr = pr < 0.f ? 0.f : (pr > 1.f ? 1.f : pr);
        thr = (float)P->final_threshold * pr;
    }
    launch_upd_candidates(st, ne, h->d_u0, h->d_f2, mind, maxd, thr, h->d_f0, h->d_u1, h->d_partials);
    launch_final_sum(st, 2, h->d_partials, h->d_scal + 2);
    HX_TRY(check_launch());
    double sc[4];
    HX_CHECK(hipMemcpyAsync(sc, h->d_scal, sizeof sc, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipStreamSynchronize(st));
    R.mean_change = sc[1] > 0 ? sc[0] / sc[1] : 0.0;
    const int64_t active = (int64_t)sc[2], ncand = (int64_t)sc[3];
    const int64_t maxcull = (int64_t)std::floor((double)((float)active * (float)P->max_culling_ratio));
    std::vector<int64_t> base(RED_BLOCKS, 0);
    uint32_t key_t = 0xFFFFFFFFu;
    int64_t ties = 0;
    if (ncand > maxcull) {
        // radix select of the maxcull-th smallest candidate key
        uint32_t prefix = 0;
        int64_t need = maxcull;
        if (need > 0) {
            for (int shift = 24; shift >= 0; shift -= 8) {
                unsigned long long hist[256];
                launch_fill<double>(st, 256, 0.0, reinterpret_cast<double*>(h->d_hist));
                launch_radix_hist(st, ne, h->d_u1, h->d_f0, prefix, shift, h->d_hist);
                HX_TRY(check_launch());
                HX_CHECK(hipMemcpyAsync(hist, h->d_hist, sizeof hist, hipMemcpyDeviceToHost, st));
                HX_CHECK(hipStreamSynchronize(st));
                int64_t cum = 0;
                int d = 0;
                for (; d < 256; ++d) {
                    if (cum + (int64_t)hist[d] >= need) break;
                    cum += (int64_t)hist[d];
                }
                need -= cum;
                prefix |= (uint32_t)d << shift;
            }
            key_t = prefix;
            ties = need;
            launch_tie_count(st, ne, h->d_u1, h->d_f0, key_t, h->d_ties);
            std::vector<int64_t> cnt(RED_BLOCKS);
            HX_CHECK(hipMemcpyAsync(cnt.data(), h->d_ties, sizeof(int64_t) * RED_BLOCKS, hipMemcpyDeviceToHost, st));
            HX_CHECK(hipStreamSynchronize(st));
            int64_t run = 0;
            for (int b = 0; b < RED_BLOCKS; ++b) { base[b] = run; run += cnt[b]; }
        } else {
            key_t = 0;  // nothing below key 0: cull none
            ties = 0;
        }
        R.culled = maxcull;
    } else {
        R.culled = ncand;
    }
    HX_CHECK(hipMemcpyAsync(h->d_ties + RED_BLOCKS, base.data(), sizeof(int64_t) * RED_BLOCKS,
                            hipMemcpyHostToDevice, st));
    launch_cull(st, ne, h->d_u1, key_t, ties, h->d_ties + RED_BLOCKS, mind, h->d_f0);
    launch_upd_finish(st, ne, h->d_u0, h->d_f2, mind, P->hard, h->d_f0);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(st));
    R.R_length = R_len;
    R.threshold = thr;
    if (rep) *rep = R;
    return HEXA_OK;
}

int32_t upload_update_inputs(hexa_handle* h, const float* l1, const uint8_t* prot, const float* orig) {
    const int64_t ne = h->g.ne;
    hipStream_t st = h->stream;
    HX_TRY(dalloc(h, &h->d_f1, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_f2, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_u0, (size_t)ne));
    if (l1) HX_CHECK(hipMemcpyAsync(h->d_f1, l1, sizeof(float) * ne, hipMemcpyHostToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_f2, orig, sizeof(float) * ne, hipMemcpyHostToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_u0, prot, (size_t)ne, hipMemcpyHostToDevice, st));
    return HEXA_OK;
}

}  // namespace

extern "C" {

int32_t hexa_update_density(hexa_handle* h, float* density, const float* l1, const uint8_t* prot,
                            const float* orig, const hexa_update_params* P, hexa_update_report* rep) {
    if (!h || !density || !l1 || !prot || !orig || !P) { set_error("null argument"); return HEXA_E_INVALID; }
    if (P->n_iters < 1 || !(P->l1_allow != 0) || !(P->E != 0)) { set_error("invalid update parameters"); return HEXA_E_INVALID; }
    const int64_t ne = h->g.ne;
    HX_TRY(ensure_elem(h));
    HX_TRY(ensure_solver_vectors(h));
    HX_TRY(dalloc(h, &h->d_f0, (size_t)ne));
    HX_TRY(upload_update_inputs(h, l1, prot, orig));
    hipStream_t st = h->stream;
    HX_CHECK(hipMemcpyAsync(h->d_f0, density, sizeof(float) * ne, hipMemcpyHostToDevice, st));
    hexa_update_report R{};
    HX_TRY(update_core(h, P, &R));
    if (!R.skipped) {
        HX_CHECK(hipMemcpyAsync(density, h->d_f0, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
        HX_CHECK(hipStreamSynchronize(st));
    }
    if (rep) *rep = R;
    return HEXA_OK;
}

int32_t hexa_set_update_fields(hexa_handle* h, const uint8_t* prot, const float* orig) {
    if (!h || !prot || !orig) { set_error("null argument"); return HEXA_E_INVALID; }
    const int64_t ne = h->g.ne;
    HX_TRY(dalloc(h, &h->d_prot, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_orig, (size_t)ne));
    HX_CHECK(hipMemcpyAsync(h->d_prot, prot, (size_t)ne, hipMemcpyHostToDevice, h->stream));
    HX_CHECK(hipMemcpyAsync(h->d_orig, orig, sizeof(float) * ne, hipMemcpyHostToDevice, h->stream));
    HX_CHECK(hipStreamSynchronize(h->stream));
    h->have_update_fields = true;
    return HEXA_OK;
}

int32_t hexa_ref_step(hexa_handle* h, const hexa_solve_params* sp, double E, double prune_threshold,
                      const hexa_update_params* P, hexa_solve_report* srep, hexa_update_report* urep,
                      float* density_out) {
    if (!h || !P) { set_error("null argument"); return HEXA_E_INVALID; }
    if (!h->have_update_fields) { set_error("hexa_set_update_fields comes first"); return HEXA_E_INVALID; }
    if (P->n_iters < 1 || !(P->l1_allow != 0) || !(P->E != 0)) { set_error("invalid update parameters"); return HEXA_E_INVALID; }
    const int64_t ne = h->g.ne;
    hipStream_t st = h->stream;
    // 1. solve with the handle's density (pruned first when asked)
    if (prune_threshold >= 0.0) HX_TRY(hexa_prune_voids(h, prune_threshold, nullptr));
    int32_t ss = hexa_solve(h, nullptr, sp, srep);
    if (ss != HEXA_OK && ss != HEXA_E_NOT_CONVERGED) return ss;
    // 2. stress l1 of the device solution (Stress.jl:159-285)
    HX_TRY(ensure_elem(h));
    HX_TRY(dalloc(h, &h->d_stress, (size_t)(17 * ne)));
    HX_TRY(dalloc(h, &h->d_f0, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_f1, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_f2, (size_t)ne));
    HX_TRY(dalloc(h, &h->d_u0, (size_t)ne));
    launch_stress(st, h->g, h->grid.dx, h->grid.dy, h->grid.dz, E, h->opt.nu, h->d_x, h->d_rho, nullptr, nullptr,
                  h->d_f1, nullptr, nullptr, nullptr);
    // 3. update_density! on device copies of the density, flags and originals
    launch_convert<double, float>(st, ne, h->d_rho, h->d_f0);
    HX_CHECK(hipMemcpyAsync(h->d_f2, h->d_orig, sizeof(float) * ne, hipMemcpyDeviceToDevice, st));
    HX_CHECK(hipMemcpyAsync(h->d_u0, h->d_prot, (size_t)ne, hipMemcpyDeviceToDevice, st));
    HX_TRY(check_launch());
    hexa_update_report R{};
    HX_TRY(update_core(h, P, &R));
    if (urep) *urep = R;
    if (R.skipped) return ss;
    // 4. the new density becomes the handle's density
    launch_convert<float, double>(st, ne, h->d_f0, h->d_rho);
    HX_TRY(check_launch());
    if (density_out) HX_CHECK(hipMemcpyAsync(density_out, h->d_f0, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    HX_TRY(update_factor(h));
    return ss;
}

}  // extern "C"
