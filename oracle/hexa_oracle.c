/*
 * hexa_oracle.c -- CPU restatement of HEXA TopOpt's inner solve loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * is compared against.  It is loaded by the tests, by the package's smoke
 * entry and by the bench's cpu_baseline leg, and by nothing else: the
 * product path never links or calls it.
 *
 * Every function restates the algorithm of the reference file:line given in
 * its comment (R/ = v119_2025_12_19_HEXATOP_AWS/).  The reference is Julia
 * (CUDA.jl); no line of it is reused here, only its arithmetic.
 *
 * Conventions (0-based, structured grid of nx*ny*nz unit bricks):
 *   node (i,j,k)    -> i + (nx+1)*(j + (ny+1)*k)           R/src/Mesh/Mesh.jl:13-15
 *   element (i,j,k) -> i + nx*(j + ny*k)                   R/src/Mesh/Mesh.jl:44-55
 *   element corners n1..n8 = (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)
 *   DOF of node n, component c -> 3n+c                     R/src/Solvers/GPUSolver.jl:101-108
 *   fixed_dof[3n+c] != 0 marks a constrained DOF           R/src/Solvers/GPUSolver.jl:170-180
 *
 * Parity pins: see tests/test_oracle_*.py.  Every function here is pinned
 * there by closed forms, invariants, brute force (scipy direct solves) or
 * committed golden vectors.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_OK 0
#define OR_E_INVALID (-1)

static const int CORNER[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},
                                 {0,0,1},{1,0,1},{1,1,1},{0,1,1}};

static inline int64_t node_id(int64_t i, int64_t j, int64_t k, int nx, int ny) {
    return i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k);
}

static void elem_nodes(int64_t e, int nx, int ny, int64_t out[8]) {
    int64_t i = e % nx, j = (e / nx) % ny, k = e / ((int64_t)nx * ny);
    for (int a = 0; a < 8; ++a)
        out[a] = node_id(i + CORNER[a][0], j + CORNER[a][1], k + CORNER[a][2], nx, ny);
}

/* ------------------------------------------------------------------ RNG */
/* SplitMix64(seed, index): the seeded input generator shared by tests and
 * bench (same constants as the package's problems.py). */
uint64_t or_splitmix64(uint64_t seed, uint64_t index) {
    uint64_t z = seed + (index + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double or_uniform(uint64_t seed, uint64_t index) {
    return (double)(or_splitmix64(seed, index) >> 11) * (1.0 / 9007199254740992.0);
}

/* ----------------------------------------------------------- element math */
/* Trilinear shape functions on [-1,1]^3 with NAT_COORDS ordering.
 * R/src/Core/Element.jl:8-17 (NAT_COORDS), :23-43 (shape_functions). */
static const double NAT[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},
                                 {-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};

static void shape_d(double xi, double eta, double zeta, double N[8], double dN[8][3]) {
    for (int i = 0; i < 8; ++i) {
        double tx = 1.0 + xi * NAT[i][0], ty = 1.0 + eta * NAT[i][1], tz = 1.0 + zeta * NAT[i][2];
        N[i] = 0.125 * tx * ty * tz;
        dN[i][0] = 0.125 * NAT[i][0] * ty * tz;
        dN[i][1] = 0.125 * tx * NAT[i][1] * tz;
        dN[i][2] = 0.125 * tx * ty * NAT[i][2];
    }
}
static void shape_f(float xi, float eta, float zeta, float N[8], float dN[8][3]) {
    for (int i = 0; i < 8; ++i) {
        float tx = 1.0f + xi * (float)NAT[i][0], ty = 1.0f + eta * (float)NAT[i][1],
              tz = 1.0f + zeta * (float)NAT[i][2];
        N[i] = 0.125f * tx * ty * tz;
        dN[i][0] = 0.125f * (float)NAT[i][0] * ty * tz;
        dN[i][1] = 0.125f * tx * (float)NAT[i][1] * tz;
        dN[i][2] = 0.125f * tx * ty * (float)NAT[i][2];
    }
}

static double det3d(const double J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
         - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
         + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
static void inv3d(const double J[3][3], double d, double I[3][3]) {
    I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
    I[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / d;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
    I[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / d;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
    I[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / d;
    I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
    I[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / d;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
}
static float det3f(const float J[3][3]) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
         - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
         + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
static void inv3f(const float J[3][3], float d, float I[3][3]) {
    I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
    I[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / d;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
    I[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / d;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
    I[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / d;
    I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
    I[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / d;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
}

/* 24x24 hex8 stiffness of the dx*dy*dz brick centred at the origin, E=1.
 * R/src/Core/Element.jl:45-71 (material_matrix), :73-102
 * (hex_element_stiffness, 2x2x2 Gauss at +-1/sqrt(3), xi outermost),
 * :104-111 (get_canonical_stiffness).  f32mode=1 evaluates everything in
 * float and widens the result, as R/src/Solvers/GPUSolver.jl:248-249 does;
 * the reference's BLAS accumulation order is not reproduced, so that mode
 * agrees with the reference to float rounding only. */
int or_ke_hex8(double dx, double dy, double dz, double nu, int f32mode, double* Ke) {
    if (!(dx > 0 && dy > 0 && dz > 0) || !(nu > -1.0 && nu < 0.5)) return OR_E_INVALID;
    memset(Ke, 0, sizeof(double) * 576);
    if (!f32mode) {
        double nodes[8][3];
        for (int i = 0; i < 8; ++i) {
            nodes[i][0] = CORNER[i][0] * dx - dx / 2;
            nodes[i][1] = CORNER[i][1] * dy - dy / 2;
            nodes[i][2] = CORNER[i][2] * dz - dz / 2;
        }
        double f = 1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double c1 = (1.0 - nu) * f, c2 = nu * f, c3 = ((1.0 - 2.0 * nu) / 2.0) * f;
        double D[6][6] = {{0}};
        D[0][0] = D[1][1] = D[2][2] = c1;
        D[0][1] = D[0][2] = D[1][0] = D[1][2] = D[2][0] = D[2][1] = c2;
        D[3][3] = D[4][4] = D[5][5] = c3;
        double a = 1.0 / sqrt(3.0), g[2] = {-a, a};
        for (int p = 0; p < 2; ++p) for (int q = 0; q < 2; ++q) for (int s = 0; s < 2; ++s) {
            double N[8], dN[8][3], J[3][3], I[3][3], dNx[8][3], B[6][24] = {{0}}, DB[6][24];
            shape_d(g[p], g[q], g[s], N, dN);
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) {
                double acc = 0; for (int i = 0; i < 8; ++i) acc += dN[i][r] * nodes[i][c];
                J[r][c] = acc;
            }
            double d = det3d(J);
            if (d <= 0) return OR_E_INVALID;
            inv3d(J, d, I);
            for (int i = 0; i < 8; ++i) for (int r = 0; r < 3; ++r) {
                double acc = 0; for (int c = 0; c < 3; ++c) acc += dN[i][c] * I[r][c];
                dNx[i][r] = acc;
            }
            for (int i = 0; i < 8; ++i) {
                int o = 3 * i;
                B[0][o] = dNx[i][0]; B[1][o + 1] = dNx[i][1]; B[2][o + 2] = dNx[i][2];
                B[3][o] = dNx[i][1]; B[3][o + 1] = dNx[i][0];
                B[4][o + 1] = dNx[i][2]; B[4][o + 2] = dNx[i][1];
                B[5][o] = dNx[i][2]; B[5][o + 2] = dNx[i][0];
            }
            for (int r = 0; r < 6; ++r) for (int c = 0; c < 24; ++c) {
                double acc = 0; for (int t = 0; t < 6; ++t) acc += D[r][t] * B[t][c];
                DB[r][c] = acc;
            }
            for (int r = 0; r < 24; ++r) for (int c = 0; c < 24; ++c) {
                double acc = 0; for (int t = 0; t < 6; ++t) acc += B[t][r] * DB[t][c];
                Ke[r * 24 + c] += acc * d;
            }
        }
    } else {
        float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz, fnu = (float)nu;
        float nodes[8][3];
        for (int i = 0; i < 8; ++i) {
            nodes[i][0] = (CORNER[i][0] ? fdx : 0.0f) - fdx / 2;
            nodes[i][1] = (CORNER[i][1] ? fdy : 0.0f) - fdy / 2;
            nodes[i][2] = (CORNER[i][2] ? fdz : 0.0f) - fdz / 2;
        }
        float inv_den = 1.0f / ((1.0f + fnu) * (1.0f - 2.0f * fnu));
        float fac = 1.0f * inv_den;
        float c1 = (1.0f - fnu) * fac, c2 = fnu * fac, c3 = ((1.0f - 2.0f * fnu) / 2.0f) * fac;
        float D[6][6] = {{0}};
        D[0][0] = D[1][1] = D[2][2] = c1;
        D[0][1] = D[0][2] = D[1][0] = D[1][2] = D[2][0] = D[2][1] = c2;
        D[3][3] = D[4][4] = D[5][5] = c3;
        float a = 1.0f / sqrtf(3.0f), g[2] = {-a, a};
        float K32[24][24] = {{0}};
        for (int p = 0; p < 2; ++p) for (int q = 0; q < 2; ++q) for (int s = 0; s < 2; ++s) {
            float N[8], dN[8][3], J[3][3], I[3][3], dNx[8][3], B[6][24] = {{0}}, BD[24][6];
            shape_f(g[p], g[q], g[s], N, dN);
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) {
                float acc = 0; for (int i = 0; i < 8; ++i) acc += dN[i][r] * nodes[i][c];
                J[r][c] = acc;
            }
            float d = det3f(J);
            if (d <= 0) return OR_E_INVALID;
            inv3f(J, d, I);
            for (int i = 0; i < 8; ++i) for (int r = 0; r < 3; ++r) {
                float acc = 0; for (int c = 0; c < 3; ++c) acc += dN[i][c] * I[r][c];
                dNx[i][r] = acc;
            }
            for (int i = 0; i < 8; ++i) {
                int o = 3 * i;
                B[0][o] = dNx[i][0]; B[1][o + 1] = dNx[i][1]; B[2][o + 2] = dNx[i][2];
                B[3][o] = dNx[i][1]; B[3][o + 1] = dNx[i][0];
                B[4][o + 1] = dNx[i][2]; B[4][o + 2] = dNx[i][1];
                B[5][o] = dNx[i][2]; B[5][o + 2] = dNx[i][0];
            }
            /* ((B' * D) * B) * detJ, left to right as Julia parses it */
            for (int r = 0; r < 24; ++r) for (int c = 0; c < 6; ++c) {
                float acc = 0; for (int t = 0; t < 6; ++t) acc += B[t][r] * D[t][c];
                BD[r][c] = acc;
            }
            for (int r = 0; r < 24; ++r) for (int c = 0; c < 24; ++c) {
                float acc = 0; for (int t = 0; t < 6; ++t) acc += BD[r][t] * B[t][c];
                K32[r][c] += acc * d;
            }
        }
        for (int r = 0; r < 24; ++r) for (int c = 0; c < 24; ++c) Ke[r * 24 + c] = (double)K32[r][c];
    }
    return OR_OK;
}

/* Scalar hex8 Laplacian Ke8 and mass Me8 for the Helmholtz filter.
 * R/src/Core/Element.jl:113-134.  bug_compat=1 keeps the reference's node
 * table, whose top face sits at z=+dz instead of +dz/2 (Element.jl:116), so
 * the element spans 1.5*dz; bug_compat=0 uses the true brick. */
int or_helm_mats(double dx, double dy, double dz, int bug_compat, double* Ke8, double* Me8) {
    if (!(dx > 0 && dy > 0 && dz > 0)) return OR_E_INVALID;
    double nodes[8][3];
    for (int i = 0; i < 8; ++i) {
        nodes[i][0] = CORNER[i][0] ? dx / 2 : -dx / 2;
        nodes[i][1] = CORNER[i][1] ? dy / 2 : -dy / 2;
        nodes[i][2] = CORNER[i][2] ? (bug_compat ? dz : dz / 2) : -dz / 2;
    }
    memset(Ke8, 0, sizeof(double) * 64);
    memset(Me8, 0, sizeof(double) * 64);
    double a = 1.0 / sqrt(3.0), g[2] = {-a, a};
    for (int p = 0; p < 2; ++p) for (int q = 0; q < 2; ++q) for (int s = 0; s < 2; ++s) {
        double N[8], dN[8][3], J[3][3], I[3][3], dNx[8][3];
        shape_d(g[p], g[q], g[s], N, dN);
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) {
            double acc = 0; for (int i = 0; i < 8; ++i) acc += dN[i][r] * nodes[i][c];
            J[r][c] = acc;
        }
        double d = det3d(J);
        inv3d(J, d, I);
        for (int i = 0; i < 8; ++i) for (int r = 0; r < 3; ++r) {
            double acc = 0; for (int c = 0; c < 3; ++c) acc += dN[i][c] * I[r][c];
            dNx[i][r] = acc;
        }
        for (int r = 0; r < 8; ++r) for (int c = 0; c < 8; ++c) {
            Ke8[r * 8 + c] += (dNx[r][0] * dNx[c][0] + dNx[r][1] * dNx[c][1] + dNx[r][2] * dNx[c][2]) * d;
            Me8[r * 8 + c] += N[r] * N[c] * d;
        }
    }
    return OR_OK;
}

/* ------------------------------------------------------- stiffness factor */
/* Per-element stiffness multiplier.  Reference: E*rho (GPUSolver.jl:254,
 * CPUSolver.jl:84), i.e. Emin=0, penal=1; elements with rho < threshold are
 * skipped by the CPU path (CPUSolver.jl:82) -> factor 0.  SIMP (new, no
 * reference): Emin + rho^p (E0-Emin). */
void or_stiffness_factor(int64_t n, const double* rho, double E0, double Emin, double penal,
                         double threshold, const uint8_t* active, double* out) {
    for (int64_t e = 0; e < n; ++e) {
        double r = rho[e];
        int on = (r >= threshold) && (!active || active[e]);
        out[e] = on ? (Emin + (penal == 1.0 ? r : pow(r, penal)) * (E0 - Emin)) : 0.0;
    }
}

/* ------------------------------------------------------------ K(rho) * x */
/* y = sum_e factor_e * Ke * x_e, element by element, with one full-length
 * buffer per task and a serial reduction in task order.
 * R/src/Solvers/CPUSolver.jl:56-122 (apply_stiffness).  Elements with
 * factor 0 are skipped (CPUSolver.jl:82).  nchunks <= 0 -> 16 tasks (a
 * fixed count, so the rounding does not depend on the machine's thread
 * count; the reference uses one task per Julia thread). */
void or_apply_stiffness(int nx, int ny, int nz, const double* Ke, const double* factor,
                        const double* x, double* y, int nchunks) {
    int64_t nE = (int64_t)nx * ny * nz;
    int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    if (nchunks <= 0) nchunks = 16;
    int64_t chunk = (nE + nchunks - 1) / nchunks;
    double** parts = (double**)calloc((size_t)nchunks, sizeof(double*));
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nchunks; ++t) {
        int64_t e0 = (int64_t)t * chunk, e1 = e0 + chunk < nE ? e0 + chunk : nE;
        if (e0 >= e1) continue;
        double* loc = (double*)calloc((size_t)ndof, sizeof(double));
        parts[t] = loc;
        double ue[24], fe[24];
        int64_t nd[8];
        for (int64_t e = e0; e < e1; ++e) {
            double f = factor[e];
            if (f == 0.0) continue;
            elem_nodes(e, nx, ny, nd);
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) ue[3 * a + c] = x[3 * nd[a] + c];
            for (int r = 0; r < 24; ++r) {
                double acc = 0;
                for (int c = 0; c < 24; ++c) acc += Ke[r * 24 + c] * ue[c];
                fe[r] = acc * f;
            }
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) loc[3 * nd[a] + c] += fe[3 * a + c];
        }
    }
    memset(y, 0, sizeof(double) * (size_t)ndof);
    for (int t = 0; t < nchunks; ++t) {
        if (!parts[t]) continue;
        double* p = parts[t];
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < ndof; ++d) y[d] += p[d];
        free(p);
    }
    free(parts);
}

/* diag(K(rho)).  mode 0: exact per-component local diagonal, skipping
 * factor-0 elements (CPUSolver.jl:142-190).  mode 1: the reference GPU
 * kernel, which adds the x-entry Ke[3a,3a] to all three components
 * (GPUSolver.jl:130-144). */
void or_diag(int nx, int ny, int nz, const double* Ke, const double* factor, int mode, double* diag) {
    int64_t nE = (int64_t)nx * ny * nz;
    int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    memset(diag, 0, sizeof(double) * (size_t)ndof);
    int64_t nd[8];
    for (int64_t e = 0; e < nE; ++e) {
        double f = factor[e];
        if (f == 0.0 && mode == 0) continue;
        elem_nodes(e, nx, ny, nd);
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) {
                int l = mode == 1 ? 3 * a : 3 * a + c;
                diag[3 * nd[a] + c] += Ke[l * 24 + l] * f;
            }
    }
}

/* Sums are formed over fixed blocks of SUM_BLOCK terms, block sums then
 * added in block order: the result depends neither on the thread count nor
 * on the order in which threads finish (an OpenMP reduction clause combines
 * thread partials in completion order). */
#define SUM_BLOCK 8192
static double dotm(int64_t n, const double* a, const double* b) {
    const int64_t nb = (n + SUM_BLOCK - 1) / SUM_BLOCK;
    double* part = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nb; ++k) {
        const int64_t i1 = (k + 1) * SUM_BLOCK < n ? (k + 1) * SUM_BLOCK : n;
        double t = 0;
        for (int64_t i = k * SUM_BLOCK; i < i1; ++i) t += a[i] * b[i];
        part[k] = t;
    }
    double s = 0;
    for (int64_t k = 0; k < nb; ++k) s += part[k];
    free(part);
    return s;
}

/* A x with A = P_f (K + shift I) P_f, P_f = projection on free DOFs. */
static void apply_A(int nx, int ny, int nz, const double* Ke, const double* factor,
                    const uint8_t* fixed, double shift, const double* x, double* y, int nchunks) {
    int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    or_apply_stiffness(nx, ny, nz, Ke, factor, x, y, nchunks);
#pragma omp parallel for schedule(static)
    for (int64_t d = 0; d < ndof; ++d) y[d] = fixed[d] ? 0.0 : y[d] + shift * x[d];
}

/* Jacobi-PCG with the CPU matrix-free semantics, in FP64.
 * R/src/Solvers/CPUSolver.jl:192-255 (matrix_free_cg_solve):
 *   shift = shift_factor * max(diag_free)                       (:204-205)
 *   diag_free clamped from below to shift (not added)           (:212,242)
 *   x0 = 0, r = f_free                                          (:197,211)
 *   Ap = K p + shift p                                          (:228)
 *   ||r||/||f_free|| < tol checked after every update           (:233-239)
 * Vectors are full length with the constrained DOFs held at zero, which is
 * the same arithmetic as the reference's free-DOF compaction.
 * out: x (full, constrained = 0); returns iterations; *rel_out, *shift_out. */
int or_pcg_cpu(int nx, int ny, int nz, const double* Ke, const double* factor,
               const uint8_t* fixed, const double* f, double tol, int max_iter,
               double shift_factor, double* x, double* rel_out, double* shift_out, int nchunks) {
    int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    double* diag = (double*)malloc(sizeof(double) * ndof);
    double* r = (double*)malloc(sizeof(double) * ndof);
    double* z = (double*)malloc(sizeof(double) * ndof);
    double* p = (double*)malloc(sizeof(double) * ndof);
    double* Ap = (double*)malloc(sizeof(double) * ndof);
    or_diag(nx, ny, nz, Ke, factor, 0, diag);
    double maxd = -INFINITY;
    for (int64_t d = 0; d < ndof; ++d) if (!fixed[d] && diag[d] > maxd) maxd = diag[d];
    double shift = shift_factor * maxd;
    for (int64_t d = 0; d < ndof; ++d) {
        x[d] = 0.0;
        r[d] = fixed[d] ? 0.0 : f[d];
        if (!fixed[d] && diag[d] <= shift) diag[d] = shift;
        z[d] = fixed[d] ? 0.0 : r[d] / diag[d];
        p[d] = z[d];
    }
    double rz = dotm(ndof, r, z);
    double nf = sqrt(dotm(ndof, r, r));
    double rel = 0.0;
    int it = 0;
    if (nf == 0.0) { *rel_out = 0.0; *shift_out = shift; free(diag); free(r); free(z); free(p); free(Ap); return 0; }
    for (it = 1; it <= max_iter; ++it) {
        apply_A(nx, ny, nz, Ke, factor, fixed, shift, p, Ap, nchunks);
        double alpha = rz / dotm(ndof, p, Ap);
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < ndof; ++d) { x[d] += alpha * p[d]; r[d] -= alpha * Ap[d]; }
        rel = sqrt(dotm(ndof, r, r)) / nf;
        if (rel < tol) break;
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < ndof; ++d) z[d] = fixed[d] ? 0.0 : r[d] / diag[d];
        double rz_new = dotm(ndof, r, z);
        double beta = rz_new / rz;
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < ndof; ++d) p[d] = z[d] + beta * p[d];
        rz = rz_new;
    }
    if (it > max_iter) it = max_iter;
    *rel_out = rel;
    *shift_out = shift;
    free(diag); free(r); free(z); free(p); free(Ap);
    return it;
}

static void refmg_setup(int nx, int ny, int nz, const double* kd8, const uint8_t* present,
                        int* ncx, int* ncy, int* ncz, double** dc_out);
static void refmg_apply(int nx, int ny, int nz, int cx, int cy, int cz, const double* dc,
                        const uint8_t* fixed, const double* minv, const double* r, double* z);

/* Jacobi-PCG with the GPU driver's control flow, in FP64.
 * R/src/Solvers/GPUSolver.jl:182-465 (gpu_matrix_free_cg_solve), Jacobi
 * branch: shift = cur_shift*max(diag_free) ADDED to the diagonal (:343,353);
 * warm start on attempt 1 only (:347-351); residual checked at k==1 and
 * k%10==0 (:379); best-x snapshot (:384); NaN/divergence -> next attempt
 * with shift*multiplier (:387-401,422-425); breakdown |pAp|<1e-20 (:374);
 * stagnation accept (:426-430); best-x fallback (:439-443); NaN/Inf in the
 * diagonal -> "Failed_Corrupt_Input" (:310-315).  diag_mode as or_diag.
 * Returns status: 0 converged, 1 not converged (best x returned),
 * 2 corrupt input.  info[0]=rel_res, info[1]=last iteration,
 * info[2]=attempt used, info[3]=shift used, info[4]=1 if the reference
 * 2-level preconditioner fell back to Jacobi. */
int or_pcg_gpu_semantics(int nx, int ny, int nz, const double* Ke, const double* factor,
                         const uint8_t* fixed, const double* f, const double* u_guess,
                         double tol, int max_iter, double shift_factor, int max_attempts,
                         double shift_multiplier, double stagnation_tol, int diag_mode,
                         const double* refmg_kd8, const uint8_t* present,
                         double* x, double* info, int nchunks) {
    int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    double* diag = (double*)malloc(sizeof(double) * ndof);
    double* minv = (double*)malloc(sizeof(double) * ndof);
    double* r = (double*)malloc(sizeof(double) * ndof);
    double* z = (double*)malloc(sizeof(double) * ndof);
    double* p = (double*)malloc(sizeof(double) * ndof);
    double* Ap = (double*)malloc(sizeof(double) * ndof);
    double* best = (double*)calloc((size_t)ndof, sizeof(double));
    double* b = (double*)malloc(sizeof(double) * ndof);
    or_diag(nx, ny, nz, Ke, factor, diag_mode, diag);
    double maxd = -INFINITY;
    int bad = 0;
    for (int64_t d = 0; d < ndof; ++d) if (!fixed[d]) {
        if (isnan(diag[d])) bad = 1;
        else if (diag[d] > maxd) maxd = diag[d];
    }
    int status = 1;
    info[0] = info[1] = info[2] = info[3] = 0.0;
    if (bad || isnan(maxd) || isinf(maxd)) {
        memset(x, 0, sizeof(double) * ndof);
        info[0] = NAN;
        status = 2;
        goto done;
    }
    for (int64_t d = 0; d < ndof; ++d) { b[d] = fixed[d] ? 0.0 : f[d]; x[d] = 0.0; }
    double nb = sqrt(dotm(ndof, b, b));
    double cur = shift_factor, final_rel = 0.0;
    int success = 0;
    /* the reference 2-level preconditioner when refmg_kd8 is given; it falls
     * back to Jacobi on NaN/divergence (:396-399,434-437) */
    const int want_mg = refmg_kd8 != NULL;
    int use_mg = want_mg, cx = 0, cy = 0, cz = 0;
    double* dc = NULL;
    if (want_mg) refmg_setup(nx, ny, nz, refmg_kd8, present, &cx, &cy, &cz, &dc);
    info[4] = 0.0;
restart:
    cur = shift_factor;
    for (int attempt = 1; attempt <= max_attempts; ++attempt) {
        double s = cur * maxd;
        int diverged = 0, ok = 0;
        info[2] = attempt; info[3] = s;
        if (u_guess && attempt == 1 && use_mg == want_mg) {
            for (int64_t d = 0; d < ndof; ++d) x[d] = fixed[d] ? 0.0 : u_guess[d];
        } else if (attempt > 1 || use_mg != want_mg) {
            memset(x, 0, sizeof(double) * ndof);
        }
        for (int64_t d = 0; d < ndof; ++d) minv[d] = fixed[d] ? 0.0 : 1.0 / (diag[d] + s);
        apply_A(nx, ny, nz, Ke, factor, fixed, s, x, Ap, nchunks);
        for (int64_t d = 0; d < ndof; ++d) { r[d] = b[d] - Ap[d]; z[d] = r[d] * minv[d]; }
        if (use_mg) refmg_apply(nx, ny, nz, cx, cy, cz, dc, fixed, minv, r, z);
        for (int64_t d = 0; d < ndof; ++d) p[d] = z[d];
        double rz = dotm(ndof, r, z), best_rel = INFINITY;
        for (int k = 1; k <= max_iter; ++k) {
            apply_A(nx, ny, nz, Ke, factor, fixed, s, p, Ap, nchunks);
            double den = dotm(ndof, p, Ap);
            info[1] = k;
            if (fabs(den) < 1e-20) break;
            double alpha = rz / den;
#pragma omp parallel for schedule(static)
            for (int64_t d = 0; d < ndof; ++d) { x[d] += alpha * p[d]; r[d] -= alpha * Ap[d]; }
            if (k % 10 == 0 || k == 1) {
                double rel = sqrt(dotm(ndof, r, r)) / nb;
                final_rel = rel;
                if (rel < best_rel) { best_rel = rel; memcpy(best, x, sizeof(double) * ndof); }
                if (rel < tol) { ok = 1; break; }
                if (isnan(rel) || isinf(rel)) { diverged = 1; break; }
                if (rel > 10.0 || (k > 500 && rel > best_rel * 100.0)) { diverged = 1; break; }
            }
            if (use_mg) {
                refmg_apply(nx, ny, nz, cx, cy, cz, dc, fixed, minv, r, z);
            } else {
#pragma omp parallel for schedule(static)
                for (int64_t d = 0; d < ndof; ++d) z[d] = r[d] * minv[d];
            }
            double rz_new = dotm(ndof, r, z);
            double beta = rz_new / rz;
#pragma omp parallel for schedule(static)
            for (int64_t d = 0; d < ndof; ++d) p[d] = z[d] + beta * p[d];
            rz = rz_new;
        }
        if (ok) { success = 1; break; }
        if (diverged && use_mg) { use_mg = 0; info[4] = 1.0; goto restart; }
        if (diverged && attempt != max_attempts) { cur *= shift_multiplier; continue; }
        if (stagnation_tol > 0.0 && best_rel < stagnation_tol) {
            success = 1;
            memcpy(x, best, sizeof(double) * ndof);
            break;
        }
        cur *= shift_multiplier;
    }
    if (!success) memcpy(x, best, sizeof(double) * ndof);
    info[0] = final_rel;
    status = success ? 0 : 1;
    free(dc);
done:
    free(diag); free(minv); free(r); free(z); free(p); free(Ap); free(best); free(b);
    return status;
}

/* ---------------------------------------- reference 2-level preconditioner */
/* R/src/Solvers/GPUGeometricMultigrid.jl:253-315 (set-up) and :317-346
 * (apply): coarse grid of floor(n/2) (at least 1) bricks of twice the size;
 * fine node coordinates divided by the coarse size (here exactly i/2, j/2,
 * k/2); coarse element density 1 where a present fine node falls in it
 * (floor of the normalised coordinates, :209-219), 0.001 elsewhere; coarse
 * diagonal D_c = 1 + sum over coarse elements of rho_c * kd (kd: the x-entry
 * diagonal of the coarse brick's Ke, used for all three components, :41-54);
 * apply: z = minv r, r_c = P^T r by trilinear weights from the normalised
 * coordinates (:57-126), x_c = r_c / (D_c + 1e-9), z += P x_c on the free
 * DOFs (:128-190, 339-343).  kd8: the 8 corner x-diagonal entries (corner
 * order as the element's node order); present: per fine node, or NULL for
 * all present. */
static void refmg_setup(int nx, int ny, int nz, const double* kd8, const uint8_t* present,
                        int* ncx, int* ncy, int* ncz, double** dc_out) {
    const int cx = nx / 2 > 1 ? nx / 2 : 1, cy = ny / 2 > 1 ? ny / 2 : 1, cz = nz / 2 > 1 ? nz / 2 : 1;
    const int64_t nNc = (int64_t)(cx + 1) * (cy + 1) * (cz + 1), nEc = (int64_t)cx * cy * cz;
    double* rho_c = (double*)malloc(sizeof(double) * nEc);
    for (int64_t e = 0; e < nEc; ++e) rho_c[e] = 0.001;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const int64_t n = i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k);
                if (present && !present[n]) continue;
                const int ix = (int)floor(0.5 * i), iy = (int)floor(0.5 * j), iz = (int)floor(0.5 * k);
                if (ix < cx && iy < cy && iz < cz) rho_c[ix + (int64_t)cx * (iy + (int64_t)cy * iz)] = 1.0;
            }
    double* dc = (double*)malloc(sizeof(double) * 3 * nNc);
    for (int64_t d = 0; d < 3 * nNc; ++d) dc[d] = 1.0;
    int64_t nd[8];
    for (int64_t e = 0; e < nEc; ++e) {
        elem_nodes(e, cx, cy, nd);
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) dc[3 * nd[a] + c] += kd8[a] * rho_c[e];
    }
    free(rho_c);
    *ncx = cx; *ncy = cy; *ncz = cz;
    *dc_out = dc;
}

static void refmg_apply(int nx, int ny, int nz, int cx, int cy, int cz, const double* dc,
                        const uint8_t* fixed, const double* minv, const double* r, double* z) {
    const int64_t nN = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t nNc = (int64_t)(cx + 1) * (cy + 1) * (cz + 1);
    double* rc = (double*)calloc((size_t)(3 * nNc), sizeof(double));
    for (int64_t d = 0; d < 3 * nN; ++d) z[d] = fixed[d] ? 0.0 : r[d] * minv[d];
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1)
            for (int64_t d = 0; d < 3 * nNc; ++d) rc[d] = rc[d] / (dc[d] + 1e-9);
        for (int64_t n = 0; n < nN; ++n) {
            const int i = (int)(n % (nx + 1)), j = (int)((n / (nx + 1)) % (ny + 1)), k = (int)(n / ((int64_t)(nx + 1) * (ny + 1)));
            const double x = 0.5 * i, y = 0.5 * j, zz = 0.5 * k;
            const int i0 = (int)floor(x), j0 = (int)floor(y), k0 = (int)floor(zz);
            const double tx = x - i0, ty = y - j0, tz = zz - k0;
            double acc[3] = {0, 0, 0};
            for (int q = 0; q < 8; ++q) {
                const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
                const int I = i0 + a, J = j0 + b, K = k0 + c;
                if (I > cx || J > cy || K > cz) continue;
                const double w = (a ? tx : 1 - tx) * (b ? ty : 1 - ty) * (c ? tz : 1 - tz);
                const int64_t m = I + (int64_t)(cx + 1) * (J + (int64_t)(cy + 1) * K);
                for (int comp = 0; comp < 3; ++comp) {
                    if (pass == 0) rc[3 * m + comp] += w * (fixed[3 * n + comp] ? 0.0 : r[3 * n + comp]);
                    else acc[comp] += w * rc[3 * m + comp];
                }
            }
            if (pass == 1)
                for (int comp = 0; comp < 3; ++comp)
                    if (!fixed[3 * n + comp]) z[3 * n + comp] += acc[comp];
        }
    }
    free(rc);
}

void or_refmg_apply(int nx, int ny, int nz, const double* kd8, const uint8_t* present,
                    const uint8_t* fixed, const double* minv, const double* r, double* z) {
    int cx, cy, cz;
    double* dc;
    refmg_setup(nx, ny, nz, kd8, present, &cx, &cy, &cz, &dc);
    refmg_apply(nx, ny, nz, cx, cy, cz, dc, fixed, minv, r, z);
    free(dc);
}

/* ------------------------------------------------------------ regridding */
/* Density remap onto a refined grid, R/src/Mesh/MeshRefiner.jl:185-237 with
 * interpolate_trilinear_projected (:44-75), in the reference's Float32
 * operation order (this file is built without FMA contraction).  Grids are
 * element counts; lengths are the domain extents; alpha may be NULL. */
void or_regrid(int onx, int ony, int onz, int nnx, int nny, int nnz, float ndx, float ndy, float ndz,
               float lx, float ly, float lz, int hard, const float* rho, const float* alpha, float* out,
               float* aout) {
    const int64_t n_new = (int64_t)nnx * nny * nnz, sl = (int64_t)onx * ony;
    for (int64_t e = 0; e < n_new; ++e) {
        const int64_t ix = e % nnx + 1, iy = (e / nnx) % nny + 1, iz = e / ((int64_t)nnx * nny) + 1;
        const float cx = ((float)ix - 0.5f) * ndx, cy = ((float)iy - 0.5f) * ndy, cz = ((float)iz - 0.5f) * ndz;
        float u = (cx / lx) * (float)onx, v = (cy / ly) * (float)ony, w = (cz / lz) * (float)onz;
        u = u < 0.5f ? 0.5f : (u > (float)onx - 0.5f ? (float)onx - 0.5f : u);
        v = v < 0.5f ? 0.5f : (v > (float)ony - 0.5f ? (float)ony - 0.5f : v);
        w = w < 0.5f ? 0.5f : (w > (float)onz - 0.5f ? (float)onz - 0.5f : w);
        const int64_t u0 = (int64_t)floorf(u - 0.5f), v0 = (int64_t)floorf(v - 0.5f), w0 = (int64_t)floorf(w - 0.5f);
        const float tx = (u - 0.5f) - (float)u0, ty = (v - 0.5f) - (float)v0, tz = (w - 0.5f) - (float)w0;
#define CL(a, n) ((a) < 1 ? (int64_t)1 : ((a) > (n) ? (int64_t)(n) : (a)))
        const int64_t i0 = CL(u0 + 1, onx), i1 = CL(u0 + 2, onx), j0 = CL(v0 + 1, ony), j1 = CL(v0 + 2, ony),
                      k0 = CL(w0 + 1, onz), k1 = CL(w0 + 2, onz);
#undef CL
#define AT(i, j, k) rho[((i) - 1) + ((j) - 1) * onx + ((k) - 1) * sl]
        const float ctx = tx < 0.f ? 0.f : (tx > 1.f ? 1.f : tx), cty = ty < 0.f ? 0.f : (ty > 1.f ? 1.f : ty),
                    ctz = tz < 0.f ? 0.f : (tz > 1.f ? 1.f : tz);
        const float c00 = AT(i0, j0, k0) * (1 - ctx) + AT(i1, j0, k0) * ctx;
        const float c10 = AT(i0, j1, k0) * (1 - ctx) + AT(i1, j1, k0) * ctx;
        const float c01 = AT(i0, j0, k1) * (1 - ctx) + AT(i1, j0, k1) * ctx;
        const float c11 = AT(i0, j1, k1) * (1 - ctx) + AT(i1, j1, k1) * ctx;
#undef AT
        const float c0 = c00 * (1 - cty) + c10 * cty;
        const float c1 = c01 * (1 - cty) + c11 * cty;
        float val = c0 * (1 - ctz) + c1 * ctz;
        if (hard) val = val > 0.4f ? 1.0f : (val < 0.2f ? 0.0001f : val);
        out[e] = val;
        if (aout) aout[e] = alpha ? alpha[(i0 - 1) + (j0 - 1) * onx + (k0 - 1) * sl] : 0.0f;
    }
}

/* ------------------------------------------------------- Helmholtz filter */
static void helm_apply(int nx, int ny, int nz, const double* Ae, const double* x, double* y) {
    int64_t nN = (int64_t)(nx + 1) * (ny + 1) * (nz + 1), nE = (int64_t)nx * ny * nz;
    memset(y, 0, sizeof(double) * nN);
    int64_t nd[8];
    for (int64_t e = 0; e < nE; ++e) {
        elem_nodes(e, nx, ny, nd);
        double xe[8];
        for (int a = 0; a < 8; ++a) xe[a] = x[nd[a]];
        for (int r = 0; r < 8; ++r) {
            double acc = 0;
            for (int c = 0; c < 8; ++c) acc += Ae[r * 8 + c] * xe[c];
            y[nd[r]] += acc;
        }
    }
}

/* PDE (Helmholtz) density filter: (R^2 K + M) x = b, b_n = sum_{e ∋ n}
 * rho_e V/8, Jacobi-PCG from x=0 checking ||r||/||b|| every check_every
 * iterations (reference: 10) up to max_iter (reference: 200), output
 * rho~_e = mean of the 8 nodal x.  R/src/Optimization/GPUHelmholtz.jl:59-70
 * (rhs), :73-97 (matvec), :99-110 (extract), :112-123,153 (diagonal),
 * :190-266 (solve).  Returns iterations; *rel_out = last checked residual
 * (0 if none was checked, as the reference initialises it). */
int or_filter_reg(int nx, int ny, int nz, double dx, double dy, double dz, double radius,
                  int bug_compat, double reg, const double* rho, double* out, double tol, int max_iter,
                  int check_every, double* rel_out);
int or_filter(int nx, int ny, int nz, double dx, double dy, double dz, double radius,
              int bug_compat, const double* rho, double* out, double tol, int max_iter,
              int check_every, double* rel_out) {
    return or_filter_reg(nx, ny, nz, dx, dy, dz, radius, bug_compat, 0.0, rho, out, tol, max_iter, check_every,
                         rel_out);
}

/* The same with reg I added to the operator: the reference's CPU filter
 * factorises R^2 K + M + 1e-9 I (R/src/Optimization/TopOpt.jl:72-77). */
int or_filter_reg(int nx, int ny, int nz, double dx, double dy, double dz, double radius,
                  int bug_compat, double reg, const double* rho, double* out, double tol, int max_iter,
                  int check_every, double* rel_out) {
    int64_t nN = (int64_t)(nx + 1) * (ny + 1) * (nz + 1), nE = (int64_t)nx * ny * nz;
    double Ke8[64], Me8[64], Ae[64];
    or_helm_mats(dx, dy, dz, bug_compat, Ke8, Me8);
    for (int i = 0; i < 64; ++i) Ae[i] = radius * radius * Ke8[i] + Me8[i];
    double* b = (double*)calloc((size_t)nN, sizeof(double));
    double* dg = (double*)calloc((size_t)nN, sizeof(double));
    double* x = (double*)calloc((size_t)nN, sizeof(double));
    double* r = (double*)malloc(sizeof(double) * nN);
    double* z = (double*)malloc(sizeof(double) * nN);
    double* p = (double*)malloc(sizeof(double) * nN);
    double* Ap = (double*)malloc(sizeof(double) * nN);
    double w = dx * dy * dz / 8.0;
    int64_t nd[8];
    for (int64_t e = 0; e < nE; ++e) {
        elem_nodes(e, nx, ny, nd);
        for (int a = 0; a < 8; ++a) { b[nd[a]] += rho[e] * w; dg[nd[a]] += Ae[a * 8 + a]; }
    }
    for (int64_t n = 0; n < nN; ++n) dg[n] += reg;
    double nb = sqrt(dotm(nN, b, b));
    int it = 0;
    double rel = 0.0;
    if (nb == 0.0) {
        memcpy(out, rho, sizeof(double) * nE);
        *rel_out = 0.0;
        goto done;
    }
    for (int64_t n = 0; n < nN; ++n) { r[n] = b[n]; z[n] = r[n] / dg[n]; p[n] = z[n]; }
    double rho_old = dotm(nN, r, z);
    for (it = 1; it <= max_iter; ++it) {
        helm_apply(nx, ny, nz, Ae, p, Ap);
        for (int64_t n = 0; n < nN; ++n) Ap[n] += reg * p[n];
        double alpha = rho_old / dotm(nN, p, Ap);
        for (int64_t n = 0; n < nN; ++n) { x[n] += alpha * p[n]; r[n] -= alpha * Ap[n]; }
        if (check_every > 0 && it % check_every == 0) {
            rel = sqrt(dotm(nN, r, r)) / nb;
            if (rel < tol) break;
        }
        for (int64_t n = 0; n < nN; ++n) z[n] = r[n] / dg[n];
        double rho_new = dotm(nN, r, z);
        double beta = rho_new / rho_old;
        for (int64_t n = 0; n < nN; ++n) p[n] = z[n] + beta * p[n];
        rho_old = rho_new;
    }
    if (it > max_iter) it = max_iter;
    for (int64_t e = 0; e < nE; ++e) {
        elem_nodes(e, nx, ny, nd);
        double s = 0;
        for (int a = 0; a < 8; ++a) s += x[nd[a]];
        out[e] = s / 8.0;
    }
    *rel_out = rel;
done:
    free(b); free(dg); free(x); free(r); free(z); free(p); free(Ap);
    return it;
}

/* ---------------------------------------------------------------- stress */
/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations in FP64
 * (textbook; converges to machine precision).  Returns values sorted
 * descending and the eigenvectors of the largest and smallest. */
static void eig3(double a[3][3], double val[3], double vmax[3], double vmin[3]) {
    double v[3][3] = {{1,0,0},{0,1,0},{0,0,1}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        double sc = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]) + off;
        if (off <= 1e-300 || off <= 1e-17 * sc) break;
        for (int p = 0; p < 2; ++p) for (int q = p + 1; q < 3; ++q) {
            if (a[p][q] == 0.0) continue;
            double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
            double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) {
                double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq; a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk; a[q][k] = s * apk + c * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq; v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    int idx[3] = {0, 1, 2};
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 2 - i; ++j)
        if (a[idx[j]][idx[j]] < a[idx[j + 1]][idx[j + 1]]) { int t = idx[j]; idx[j] = idx[j + 1]; idx[j + 1] = t; }
    for (int i = 0; i < 3; ++i) val[i] = a[idx[i]][idx[i]];
    for (int k = 0; k < 3; ++k) { vmax[k] = v[k][idx[0]]; vmin[k] = v[k][idx[2]]; }
}

/* Centroid stress of every element of the structured grid.
 * R/src/Core/Stress.jl:159-285 (stress_kernel!): J from the 8 node
 * coordinates with dN/dxi = +-1/8 at the centroid, strain from U,
 * sigma = D(E*rho) eps (Voigt xx,yy,zz,xy,yz,xz), von Mises, principal
 * values (descending) and l1 = sum |sigma_i|; rho < 1e-6 -> all zero
 * (Stress.jl:166-181).  Outputs are nE-major triples like the reference's
 * 3 x nE arrays.  Any output pointer may be NULL. */
void or_stress(int nx, int ny, int nz, double dx, double dy, double dz, double E, double nu,
               const double* U, const double* rho, double* principal, double* vm,
               double* l1, double* dmax, double* dmin, double* voigt) {
    int64_t nE = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nE; ++e) {
        double dens = rho[e];
        if (dens < 1e-6) {
            if (principal) principal[3 * e] = principal[3 * e + 1] = principal[3 * e + 2] = 0;
            if (vm) vm[e] = 0;
            if (l1) l1[e] = 0;
            if (dmax) dmax[3 * e] = dmax[3 * e + 1] = dmax[3 * e + 2] = 0;
            if (dmin) dmin[3 * e] = dmin[3 * e + 1] = dmin[3 * e + 2] = 0;
            if (voigt) for (int c = 0; c < 6; ++c) voigt[6 * e + c] = 0;
            continue;
        }
        int64_t nd[8];
        elem_nodes(e, nx, ny, nd);
        int64_t i = e % nx, j = (e / nx) % ny, k = e / ((int64_t)nx * ny);
        double J[3][3] = {{0}}, I[3][3];
        double dxi[8][3];
        for (int a = 0; a < 8; ++a) {
            dxi[a][0] = CORNER[a][0] ? 0.125 : -0.125;
            dxi[a][1] = CORNER[a][1] ? 0.125 : -0.125;
            dxi[a][2] = CORNER[a][2] ? 0.125 : -0.125;
            double X[3] = {(i + CORNER[a][0]) * dx, (j + CORNER[a][1]) * dy, (k + CORNER[a][2]) * dz};
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) J[r][c] += dxi[a][r] * X[c];
        }
        double d = det3d(J);
        if (d <= 1e-12) d = 1.0;
        inv3d(J, d, I);
        double ex = 0, ey = 0, ez = 0, gxy = 0, gyz = 0, gxz = 0;
        for (int a = 0; a < 8; ++a) {
            double Nx = dxi[a][0] * I[0][0] + dxi[a][1] * I[1][0] + dxi[a][2] * I[2][0];
            double Ny = dxi[a][0] * I[0][1] + dxi[a][1] * I[1][1] + dxi[a][2] * I[2][1];
            double Nz = dxi[a][0] * I[0][2] + dxi[a][1] * I[1][2] + dxi[a][2] * I[2][2];
            double ux = U[3 * nd[a]], uy = U[3 * nd[a] + 1], uz = U[3 * nd[a] + 2];
            ex += Nx * ux; ey += Ny * uy; ez += Nz * uz;
            gxy += Ny * ux + Nx * uy; gyz += Nz * uy + Ny * uz; gxz += Nz * ux + Nx * uz;
        }
        double f = E * dens / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double c1 = (1.0 - nu) * f, c2 = nu * f, c3 = (0.5 - nu) * f;
        double sxx = c1 * ex + c2 * ey + c2 * ez, syy = c2 * ex + c1 * ey + c2 * ez,
               szz = c2 * ex + c2 * ey + c1 * ez;
        double sxy = c3 * gxy, syz = c3 * gyz, sxz = c3 * gxz;
        if (voigt) { voigt[6*e]=sxx; voigt[6*e+1]=syy; voigt[6*e+2]=szz; voigt[6*e+3]=sxy; voigt[6*e+4]=syz; voigt[6*e+5]=sxz; }
        double q = 0.5 * ((sxx - syy) * (sxx - syy) + (syy - szz) * (syy - szz) + (szz - sxx) * (szz - sxx))
                 + 3.0 * (sxy * sxy + syz * syz + sxz * sxz);
        if (vm) vm[e] = sqrt(q > 0 ? q : 0);
        double A[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}}, val[3], v1[3], v3[3];
        eig3(A, val, v1, v3);
        if (principal) for (int c = 0; c < 3; ++c) principal[3 * e + c] = val[c];
        if (dmax) for (int c = 0; c < 3; ++c) dmax[3 * e + c] = v1[c];
        if (dmin) for (int c = 0; c < 3; ++c) dmin[3 * e + c] = v3[c];
        if (l1) l1[e] = fabs(val[0]) + fabs(val[1]) + fabs(val[2]);
    }
}

/* ------------------------------------------------ reference density update */
typedef struct {
    double E, l1_allow, min_density, max_density, max_domain_dim, filter_tol;
    double R_init_perc, R_interm_perc, R_final_perc, R_interm_iter_perc;   /* iter perc in % */
    double min_radius_factor, final_threshold, max_culling_ratio;
    int iter, n_iters, hard, bug_compat;
    double dx, dy, dz;
    int nx, ny, nz;
} or_update_params;

typedef struct { float d; int64_t idx; } cand_t;
static int cmp_cand2(const void* a, const void* b) {
    const cand_t* x = (const cand_t*)a; const cand_t* y = (const cand_t*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

/* Stress-ratio density update.  R/src/Optimization/TopOpt.jl:211-482
 * (update_density!): proposed = max((l1/allow)/E, min) off the protected
 * set (:255-263); radius schedule init->interm->final with the
 * minimum-radius guard and R_eff = R/2.5 (:265-300); Helmholtz filter with
 * R_eff (:310-356, here the FP64 PCG of or_filter to filter_tol, 200
 * iterations, check every 10, i.e. the GPU filter's semantics); NaN guard
 * (:361-364); rho = max(rho~, min), mean change (:374-411); threshold
 * schedule, clamp to max, culling of at most floor(active*ratio) lowest
 * candidates in a stable (density, index) order (:416-461); "hard" update
 * (:463-473); protected restore (:475-479).  density is float like the
 * reference's Vector{Float32}.  out[0..5] = mean_change, R_length,
 * current_threshold, filter_iters, filter_res, culled.  Returns 1 if the
 * NaN guard skipped the update (:230-236), else 0. */
int or_update_density(float* density, const float* l1, const uint8_t* prot, const float* orig,
                      const or_update_params* P, double* out) {
    int64_t nE = (int64_t)P->nx * P->ny * P->nz;
    for (int64_t e = 0; e < nE; ++e) if (isnan(l1[e])) { for (int i = 0; i < 6; ++i) out[i] = 0; return 1; }
    float E = (float)P->E, allow = (float)P->l1_allow, mind = (float)P->min_density, maxd = (float)P->max_density;
    double* prop = (double*)malloc(sizeof(double) * nE);
    double* filt = (double*)malloc(sizeof(double) * nE);
    float* propf = (float*)malloc(sizeof(float) * nE);
    for (int64_t e = 0; e < nE; ++e) {
        if (!prot[e]) { float v = (l1[e] / allow) / E; propf[e] = v > mind ? v : mind; }
        else propf[e] = orig[e];
        prop[e] = propf[e];
    }
    float mdd = (float)P->max_domain_dim;
    float Ri = (float)P->R_init_perc / 100.0f * mdd, Rm = (float)P->R_interm_perc / 100.0f * mdd,
          Rf = (float)P->R_final_perc / 100.0f * mdd;
    float ip = (float)P->R_interm_iter_perc / 100.0f;
    int iter_interm = (int)nearbyintf(ip * (float)P->n_iters);
    if (iter_interm < 1) iter_interm = 1;
    int calc = P->iter < P->n_iters ? P->iter : P->n_iters;
    float R;
    if (calc <= iter_interm) {
        float t = iter_interm > 1 ? (float)(calc - 1) / (float)(iter_interm - 1) : 0.0f;
        R = Ri * (1.0f - t) + Rm * t;
    } else {
        float t = P->n_iters > iter_interm ? (float)(calc - iter_interm) / (float)(P->n_iters - iter_interm) : 0.0f;
        R = Rm * (1.0f - t) + Rf * t;
    }
    float avg = ((float)P->dx + (float)P->dy + (float)P->dz) / 3.0f;
    float Rsafe = (float)P->min_radius_factor * avg;
    if (R < Rsafe) R = Rsafe;
    float Reff = R / 2.5f;
    int fit = 0; double fres = 0.0;
    if (Reff > 1e-4f) {
        fit = or_filter(P->nx, P->ny, P->nz, P->dx, P->dy, P->dz, (double)Reff, P->bug_compat,
                        prop, filt, P->filter_tol, 200, 10, &fres);
        int bad = 0;
        for (int64_t e = 0; e < nE; ++e) if (isnan(filt[e])) { bad = 1; break; }
        if (bad) memcpy(filt, prop, sizeof(double) * nE);
    } else {
        memcpy(filt, prop, sizeof(double) * nE);
    }
    double change = 0.0; int64_t nact = 0;
    for (int64_t e = 0; e < nE; ++e) {
        if (prot[e]) continue;
        float old = density[e];
        float fv = (float)filt[e];
        float nv = fv > mind ? fv : mind;
        density[e] = nv;
        float eff = nv < maxd ? nv : maxd;
        change += fabsf(eff - old);
        if (nv > mind) nact++;
    }
    double mean_change = nact > 0 ? change / (double)nact : 0.0;
    float thr;
    if (P->iter > P->n_iters) thr = (float)P->final_threshold;
    else {
        float pr = (float)P->iter / (float)P->n_iters;
        if (pr < 0) pr = 0;
        if (pr > 1) pr = 1;
        thr = (float)P->final_threshold * pr;
    }
    cand_t* cand = (cand_t*)malloc(sizeof(cand_t) * (nE > 0 ? nE : 1));
    int64_t nc = 0, active = 0;
    for (int64_t e = 0; e < nE; ++e) {
        if (!prot[e]) {
            if (density[e] > maxd) density[e] = maxd;
            if (density[e] > mind) {
                active++;
                if (density[e] < thr) { cand[nc].d = density[e]; cand[nc].idx = e; nc++; }
            }
        } else if (orig[e] > mind) active++;
    }
    int64_t maxcull = (int64_t)floor((double)((float)active * (float)P->max_culling_ratio));
    int64_t culled;
    if (nc > maxcull) {
        qsort(cand, (size_t)nc, sizeof(cand_t), cmp_cand2);
        for (int64_t i = 0; i < maxcull; ++i) density[cand[i].idx] = mind;
        culled = maxcull;
    } else {
        for (int64_t i = 0; i < nc; ++i) density[cand[i].idx] = mind;
        culled = nc;
    }
    if (P->hard)
        for (int64_t e = 0; e < nE; ++e) if (!prot[e] && density[e] > mind) density[e] = 1.0f;
    for (int64_t e = 0; e < nE; ++e) if (prot[e]) density[e] = orig[e];
    out[0] = mean_change; out[1] = R; out[2] = thr; out[3] = fit; out[4] = fres; out[5] = (double)culled;
    free(prop); free(filt); free(propf); free(cand);
    return 0;
}

/* ------------------------------------------------------- SIMP + OC (new) */
/* Compliance sensitivity of modified SIMP with respect to the physical
 * density rho~ (no reference counterpart; Andreassen et al. 2011, top88;
 * Liu & Tovar 2014, top3d):  E_e = Emin + rho~^p (E0-Emin),
 * dc/drho~_e = -p rho~^(p-1) (E0-Emin) u_e' K0 u_e.  Returns the
 * compliance sum_e E_e u_e' K0 u_e.  Elements are not skipped. */
double or_simp_sensitivity(int nx, int ny, int nz, const double* Ke, const double* rho_phys,
                           double E0, double Emin, double penal, const double* U, double* dc) {
    int64_t nE = (int64_t)nx * ny * nz;
    double* ce_all = (double*)malloc(sizeof(double) * (size_t)(nE > 0 ? nE : 1));
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nE; ++e) {
        int64_t nd[8];
        elem_nodes(e, nx, ny, nd);
        double ue[24];
        for (int a = 0; a < 8; ++a) for (int k = 0; k < 3; ++k) ue[3 * a + k] = U[3 * nd[a] + k];
        double ce = 0;
        for (int r = 0; r < 24; ++r) {
            double acc = 0;
            for (int k = 0; k < 24; ++k) acc += Ke[r * 24 + k] * ue[k];
            ce += ue[r] * acc;
        }
        double rp = rho_phys[e];
        ce_all[e] = (Emin + pow(rp, penal) * (E0 - Emin)) * ce;
        dc[e] = -penal * pow(rp, penal - 1.0) * (E0 - Emin) * ce;
    }
    double* ones = (double*)malloc(sizeof(double) * (size_t)(nE > 0 ? nE : 1));
    for (int64_t e = 0; e < nE; ++e) ones[e] = 1.0;
    const double c = dotm(nE, ce_all, ones);
    free(ones);
    free(ce_all);
    return c;
}

/* Optimality-criteria update with move limit and bisection on the
 * Lagrange multiplier (top88/top3d):
 *   x_new = clamp(x * sqrt(-dc / (lambda * dv)), max(0, x-move), min(1, x+move))
 * with the volume of the FILTERED new design, sum_e w_e x_new_e where
 * w = F^T 1 (exact because the filter F is linear), driven to
 * volfrac * nE.  Bisection on [0, 1e9] until (l2-l1)/(l1+l2) < bisect_tol.
 * Returns the number of bisection steps; *lambda_out = final midpoint. */
int or_oc_update(int64_t nE, const double* x, const double* dc, const double* dv, const double* w,
                 double volfrac, double move, double bisect_tol, double* x_new, double* lambda_out) {
    double l1 = 0.0, l2 = 1e9, lmid = 0.0;
    double target = volfrac * (double)nE;
    int steps = 0;
    while ((l2 - l1) / (l1 + l2) > bisect_tol && steps < 400) {
        lmid = 0.5 * (l2 + l1);
        double vol = 0.0;
        for (int64_t e = 0; e < nE; ++e) {
            double be = -dc[e] / (dv[e] * lmid);
            double xn = x[e] * sqrt(be > 0 ? be : 0.0);
            double lo = x[e] - move, hi = x[e] + move;
            if (lo < 0.0) lo = 0.0;
            if (hi > 1.0) hi = 1.0;
            if (xn < lo) xn = lo;
            if (xn > hi) xn = hi;
            x_new[e] = xn;
            vol += w[e] * xn;
        }
        if (vol > target) l1 = lmid; else l2 = lmid;
        steps++;
    }
    lmid = 0.5 * (l2 + l1);
    for (int64_t e = 0; e < nE; ++e) {
        double be = -dc[e] / (dv[e] * lmid);
        double xn = x[e] * sqrt(be > 0 ? be : 0.0);
        double lo = x[e] - move, hi = x[e] + move;
        if (lo < 0.0) lo = 0.0;
        if (hi > 1.0) hi = 1.0;
        if (xn < lo) xn = lo;
        if (xn > hi) xn = hi;
        x_new[e] = xn;
    }
    *lambda_out = lmid;
    return steps;
}
