// Optimisation-step kernels: element stress and its l1 norm, SIMP compliance
// sensitivity, the optimality-criteria update, and the reference's
// stress-ratio density update with culling.
#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

__device__ constexpr int TCX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int TCY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int TCZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

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

__device__ __forceinline__ void elem_disp(const double* __restrict__ U, const Grid3& g, int64_t e, double (&ue)[24]) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    const int i = (int)(e % g.nx), j = (int)((e / g.nx) % g.ny), k = (int)(e / ((int64_t)g.nx * g.ny));
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t n = (i + TCX[a]) + (int64_t)NX * ((j + TCY[a]) + (int64_t)NY * (k + TCZ[a]));
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * a + c] = U[3 * n + c];
    }
}

// Symmetric 3x3 eigenvalues (descending) and the eigenvectors of the
// largest and smallest, by cyclic Jacobi rotations in FP64.
__device__ void eig3_sym(double a[3][3], double val[3], double vmax[3], double vmin[3]) {
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double sc = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]) + off;
        if (off <= 1e-300 || off <= 1e-17 * sc) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int o0 = 0, o1 = 1, o2 = 2, t;
    if (a[o0][o0] < a[o1][o1]) { t = o0; o0 = o1; o1 = t; }
    if (a[o1][o1] < a[o2][o2]) { t = o1; o1 = o2; o2 = t; }
    if (a[o0][o0] < a[o1][o1]) { t = o0; o0 = o1; o1 = t; }
    val[0] = a[o0][o0]; val[1] = a[o1][o1]; val[2] = a[o2][o2];
    for (int k = 0; k < 3; ++k) { vmax[k] = v[k][o0]; vmin[k] = v[k][o2]; }
}

// Centroid stress of each brick (R/src/Core/Stress.jl:159-285): strain from
// the 8 nodal displacements with dN/dx = +-1/(4 dx) at the centroid,
// sigma = D(E rho) eps, von Mises, principal values, l1 = sum |sigma_i|;
// rho < 1e-6 -> zeros.
__global__ void k_stress(Grid3 g, double dx, double dy, double dz, double E, double nu,
                         const double* __restrict__ U, const double* __restrict__ rho,
                         float* __restrict__ principal, float* __restrict__ vm, float* __restrict__ l1,
                         float* __restrict__ dmax, float* __restrict__ dmin, float* __restrict__ voigt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double dens = rho[e];
        if (dens < 1e-6) {
            if (principal) principal[3 * e] = principal[3 * e + 1] = principal[3 * e + 2] = 0.f;
            if (vm) vm[e] = 0.f;
            if (l1) l1[e] = 0.f;
            if (dmax) dmax[3 * e] = dmax[3 * e + 1] = dmax[3 * e + 2] = 0.f;
            if (dmin) dmin[3 * e] = dmin[3 * e + 1] = dmin[3 * e + 2] = 0.f;
            if (voigt) for (int c = 0; c < 6; ++c) voigt[6 * e + c] = 0.f;
            continue;
        }
        double ue[24];
        elem_disp(U, g, e, ue);
        double ex = 0, ey = 0, ez = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const double Nx = (TCX[a] ? 0.25 : -0.25) / dx, Ny = (TCY[a] ? 0.25 : -0.25) / dy,
                         Nz = (TCZ[a] ? 0.25 : -0.25) / dz;
            const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
            ex += Nx * ux; ey += Ny * uy; ez += Nz * uz;
            gxy += Ny * ux + Nx * uy; gyz += Nz * uy + Ny * uz; gxz += Nz * ux + Nx * uz;
        }
        const double f = E * dens / ((1.0 + nu) * (1.0 - 2.0 * nu));
        const double c1 = (1.0 - nu) * f, c2 = nu * f, c3 = (0.5 - nu) * f;
        const double sxx = c1 * ex + c2 * ey + c2 * ez, syy = c2 * ex + c1 * ey + c2 * ez,
                     szz = c2 * ex + c2 * ey + c1 * ez;
        const double sxy = c3 * gxy, syz = c3 * gyz, sxz = c3 * gxz;
        if (voigt) {
            voigt[6 * e] = (float)sxx; voigt[6 * e + 1] = (float)syy; voigt[6 * e + 2] = (float)szz;
            voigt[6 * e + 3] = (float)sxy; voigt[6 * e + 4] = (float)syz; voigt[6 * e + 5] = (float)sxz;
        }
        const double q = 0.5 * ((sxx - syy) * (sxx - syy) + (syy - szz) * (syy - szz) + (szz - sxx) * (szz - sxx)) +
                         3.0 * (sxy * sxy + syz * syz + sxz * sxz);
        if (vm) vm[e] = (float)sqrt(q > 0 ? q : 0.0);
        double A[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}}, val[3], v1[3], v3[3];
        eig3_sym(A, val, v1, v3);
        if (principal) for (int c = 0; c < 3; ++c) principal[3 * e + c] = (float)val[c];
        if (dmax) for (int c = 0; c < 3; ++c) dmax[3 * e + c] = (float)v1[c];
        if (dmin) for (int c = 0; c < 3; ++c) dmin[3 * e + c] = (float)v3[c];
        if (l1) l1[e] = (float)(fabs(val[0]) + fabs(val[1]) + fabs(val[2]));
    }
}

// SIMP sensitivity: ce = u_e' K0 u_e, dc_e = -p rho^(p-1) (E0-Emin) ce,
// compliance partials sum (Emin + rho^p (E0-Emin)) ce.  K0 is dense (576 in
// scalar registers via uniform loads).
__global__ __launch_bounds__(RED_THREADS) void k_simp_sens(Grid3 g, const double* __restrict__ Ke,
                                                           const double* __restrict__ rho, double E0, double Emin,
                                                           double penal, const double* __restrict__ U,
                                                           double* __restrict__ dc, double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double csum = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < g.ne; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        double ue[24];
        elem_disp(U, g, e, ue);
        double ce = 0.0;
#pragma unroll
        for (int r = 0; r < 24; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 24; ++c) acc += Ke[r * 24 + c] * ue[c];
            ce += ue[r] * acc;
        }
        const double rp = rho[e];
        const double pw = penal == 3.0 ? rp * rp * rp : pow(rp, penal);
        const double pm = penal == 3.0 ? rp * rp : pow(rp, penal - 1.0);
        csum += (Emin + pw * (E0 - Emin)) * ce;
        dc[e] = -penal * pm * (E0 - Emin) * ce;
    }
    csum = block_sum(csum, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = csum;
}

// The same for the exact brick stiffness in its factored form (the row-wave
// operator's): u_e' K0 u_e = v' M v with v the per-axis sum/difference modes
// of the element's corner values and M the modal diagonal plus the
// cross-component couplings, about 100 flops per element instead of 1152.
__global__ __launch_bounds__(RED_THREADS) void k_simp_sens_fact(Grid3 g, FactCoef<double> coef,
                                                                const double* __restrict__ rho, double E0, double Emin,
                                                                double penal, const double* __restrict__ U,
                                                                double* __restrict__ dc, double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double csum = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < g.ne; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        double ue[24];
        elem_disp(U, g, e, ue);
        // modes m = mx + 2 my + 4 mz (bit set = difference along that axis);
        // corners n1..n8 = (0,0,0) (1,0,0) (1,1,0) (0,1,0), then z = 1
        double v[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double P[2][4];
#pragma unroll
            for (int z = 0; z < 2; ++z) {
                const double a = ue[3 * (4 * z + 0) + c], bx = ue[3 * (4 * z + 1) + c];
                const double d = ue[3 * (4 * z + 2) + c], cy = ue[3 * (4 * z + 3) + c];
                const double s0 = a + bx, d0 = bx - a, s1 = cy + d, d1 = d - cy;
                P[z][0] = s0 + s1;
                P[z][1] = d0 + d1;
                P[z][2] = s1 - s0;
                P[z][3] = d1 - d0;
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                v[c][m] = P[0][m] + P[1][m];
                v[c][m + 4] = P[1][m] - P[0][m];
            }
        }
        double ce = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double yv[8];
#pragma unroll
            for (int m = 1; m < 8; ++m) yv[m] = coef.D[c][m] * v[c][m];
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
                if (bb == c) continue;
                const int d = 3 - c - bb;
#pragma unroll
                for (int qq = 0; qq < 2; ++qq) {
                    const int mq = qq << d;
                    yv[(1 << c) | mq] += coef.cl[d][qq] * v[bb][(1 << bb) | mq];
                    yv[(1 << bb) | mq] += coef.cm[d][qq] * v[bb][(1 << c) | mq];
                }
            }
#pragma unroll
            for (int m = 1; m < 8; ++m) ce += v[c][m] * yv[m];
        }
        const double rp = rho[e];
        const double pw = penal == 3.0 ? rp * rp * rp : pow(rp, penal);
        const double pm = penal == 3.0 ? rp * rp : pow(rp, penal - 1.0);
        csum += (Emin + pw * (E0 - Emin)) * ce;
        dc[e] = -penal * pm * (E0 - Emin) * ce;
    }
    csum = block_sum(csum, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = csum;
}

// OC_NL block sums at once, each in block_sum's order (wave butterfly, then
// the waves in order), with one barrier instead of one per sum: on return
// thread k < OC_NL holds sum k in v[0].
__device__ __forceinline__ void block_sums(double (&v)[OC_NL], double (&lw)[RED_THREADS / 64][OC_NL]) {
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < OC_NL; ++k) {
        double t = v[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if ((threadIdx.x & 63) == 0) lw[w][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < OC_NL) {
        double t = 0.0;
        for (int i = 0; i < RED_THREADS / 64; ++i) t += lw[i][threadIdx.x];
        v[0] = t;
    }
}

// ---------------------------------------------------------------- OC
// Multisection on the Lagrange multiplier: each pass evaluates the filtered
// volume at OC_NL points inside the bracket and keeps the sub-interval where
// it crosses the target; the bracket lives on the device.
struct OcBracket {
    double l1, l2, target, tol;
    int32_t done, passes;
};

// q_e = x_e sqrt(max(-dc_e/dv_e, 0)); lo/hi = move-limit box in [0,1]
__device__ __forceinline__ void oc_box(double x, double dc, double dv, double move, double& q, double& lo,
                                       double& hi) {
    const double be = -dc / dv;
    q = x * sqrt(be > 0 ? be : 0.0);
    lo = fmax(0.0, x - move);
    hi = fmin(1.0, x + move);
}

__global__ void k_oc_prep(int64_t n, const double* __restrict__ x, const double* __restrict__ dc,
                          const double* __restrict__ dv, double move, double* __restrict__ q,
                          double* __restrict__ lo, double* __restrict__ hi) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        oc_box(x[e], dc[e], dv[e], move, q[e], lo[e], hi[e]);
}

// Element e's (q, lo, hi): read from the arrays k_oc_prep wrote (DIRECT =
// false: a, b, c = q, lo, hi), or formed from the design, dc and dv on the
// fly (DIRECT = true: a, b, c = x, dc, dv; no per-element work arrays).  Both
// give the same values.
template <bool DIRECT>
__device__ __forceinline__ void oc_elem(const double* __restrict__ a, const double* __restrict__ b,
                                        const double* __restrict__ c, double move, int64_t e, double& q, double& lo,
                                        double& hi) {
    if (DIRECT) {
        oc_box(a[e], b[e], c[e], move, q, lo, hi);
    } else {
        q = a[e];
        lo = b[e];
        hi = c[e];
    }
}

template <bool DIRECT>
__global__ __launch_bounds__(RED_THREADS) void k_oc_eval(int64_t n, const OcBracket* br, const double* __restrict__ a,
                                                         const double* __restrict__ b, const double* __restrict__ c,
                                                         double move, const double* __restrict__ w,
                                                         double* __restrict__ partials) {
    if (br->done) return;
    __shared__ double lw[RED_THREADS / 64][OC_NL];
    const double l1 = br->l1, l2 = br->l2;
    double s[OC_NL];
#pragma unroll
    for (int k = 0; k < OC_NL; ++k) s[k] = 0.0;
    double inv[OC_NL];
#pragma unroll
    for (int k = 0; k < OC_NL; ++k) inv[k] = 1.0 / sqrt(l1 + (l2 - l1) * (double)(k + 1) / (double)(OC_NL + 1));
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < n; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        double qe, le, he;
        oc_elem<DIRECT>(a, b, c, move, e, qe, le, he);
        const double we = w[e];
#pragma unroll
        for (int k = 0; k < OC_NL; ++k) s[k] = fma(we, fmin(fmax(qe * inv[k], le), he), s[k]);
    }
    block_sums(s, lw);
    if (threadIdx.x < OC_NL) partials[threadIdx.x * RED_BLOCKS + blockIdx.x] = s[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_oc_pick(OcBracket* br, const double* __restrict__ partials) {
    if (br->done) return;
    __shared__ double lw[RED_THREADS / 64][OC_NL];
    __shared__ double vol[OC_NL];
    double s[OC_NL];
#pragma unroll
    for (int k = 0; k < OC_NL; ++k) {
        s[k] = 0.0;
        for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) s[k] += partials[k * RED_BLOCKS + i];
    }
    block_sums(s, lw);
    if (threadIdx.x < OC_NL) vol[threadIdx.x] = s[0];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double l1 = br->l1, l2 = br->l2;
    // volume decreases with lambda: new bracket around the first point whose
    // volume is not above the target (bisection keeps l1 when vol > target)
    int k = 0;
    while (k < OC_NL && vol[k] > br->target) ++k;
    const double a = k == 0 ? l1 : l1 + (l2 - l1) * (double)k / (double)(OC_NL + 1);
    const double b = k == OC_NL ? l2 : l1 + (l2 - l1) * (double)(k + 1) / (double)(OC_NL + 1);
    br->l1 = a;
    br->l2 = b;
    br->passes += 1;
    if ((b - a) / (a + b) <= br->tol || !(b > a)) br->done = 1;
}

// x_new at lambda = bracket midpoint; partial max |x_new - x|.  (With
// DIRECT, a is the design itself, read before it is overwritten.)
template <bool DIRECT>
__global__ __launch_bounds__(RED_THREADS) void k_oc_apply(int64_t n, const OcBracket* br, const double* a,
                                                          const double* __restrict__ b, const double* __restrict__ c,
                                                          double move, const double* __restrict__ w, double* x,
                                                          double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    const double lm = 0.5 * (br->l1 + br->l2);
    const double inv = 1.0 / sqrt(lm);
    double ch = 0.0, vol = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < n; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        double qe, le, he;
        oc_elem<DIRECT>(a, b, c, move, e, qe, le, he);
        const double xn = fmin(fmax(qe * inv, le), he);
        ch = fmax(ch, fabs(xn - x[e]));
        vol = fma(w[e], xn, vol);
        x[e] = xn;
    }
    for (int off = 32; off > 0; off >>= 1) ch = fmax(ch, __shfl_xor(ch, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = ch;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < RED_THREADS / 64; ++i) m = fmax(m, lds[i]);
        partials[blockIdx.x] = m;
    }
    __syncthreads();
    vol = block_sum(vol, lds);
    if (threadIdx.x == 0) partials[RED_BLOCKS + blockIdx.x] = vol;
}

__global__ __launch_bounds__(RED_THREADS) void k_oc_final(const OcBracket* br, const double* __restrict__ partials,
                                                          double* __restrict__ out) {
    __shared__ double lds[RED_THREADS / 64];
    double m = 0.0, v = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) {
        m = fmax(m, partials[i]);
        v += partials[RED_BLOCKS + i];
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    double mm = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < RED_THREADS / 64; ++i) mm = fmax(mm, lds[i]);
    __syncthreads();
    v = block_sum(v, lds);
    if (threadIdx.x == 0) {
        out[0] = mm;
        out[1] = v;
        out[2] = 0.5 * (br->l1 + br->l2);
        out[3] = (double)br->passes;
    }
}

__global__ void k_oc_init(OcBracket* br, double target, double tol) {
    br->l1 = 0.0;
    br->l2 = 1e9;
    br->target = target;
    br->tol = tol;
    br->done = 0;
    br->passes = 0;
}

// ------------------------------------------------------ reference update
// proposed = max((l1/allow)/E, min) off the protected set, else original
// (TopOpt.jl:255-263), in float as the reference.
__global__ void k_upd_proposed(int64_t n, const float* __restrict__ l1, const uint8_t* __restrict__ prot,
                               const float* __restrict__ orig, float allow, float E, float mind,
                               double* __restrict__ prop) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float v;
        if (!prot[e]) { v = (l1[e] / allow) / E; v = v > mind ? v : mind; }
        else v = orig[e];
        prop[e] = (double)v;
    }
}

// rho = max(rho~, min) with the mean-change and active-count partials
// (TopOpt.jl:374-411); a NaN in rho~ falls back to the unfiltered field
// (TopOpt.jl:361-364) through use_prop.
__global__ __launch_bounds__(RED_THREADS) void k_upd_apply(int64_t n, const double* __restrict__ filt,
                                                           const double* __restrict__ prop, int use_prop,
                                                           const uint8_t* __restrict__ prot, float mind, float maxd,
                                                           float* __restrict__ dens, double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double ch = 0.0, act = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < n; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        if (prot[e]) continue;
        const float old = dens[e];
        const float fv = (float)(use_prop ? prop[e] : filt[e]);
        const float nv = fv > mind ? fv : mind;
        dens[e] = nv;
        const float eff = nv < maxd ? nv : maxd;
        ch += (double)fabsf(eff - old);
        if (nv > mind) act += 1.0;
    }
    ch = block_sum(ch, lds);
    act = block_sum(act, lds);
    if (threadIdx.x == 0) { partials[blockIdx.x] = ch; partials[RED_BLOCKS + blockIdx.x] = act; }
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_nan_count(int64_t n, const T* __restrict__ x,
                                                           double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double c = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < n; e += (int64_t)RED_BLOCKS * RED_THREADS)
        if (isnan(x[e])) c += 1.0;
    c = block_sum(c, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = c;
}

// Clamp to max, count active elements and cull candidates
// (TopOpt.jl:427-447).  Each block owns a contiguous index range so that the
// tie ranking below can follow index order.
__global__ __launch_bounds__(RED_THREADS) void k_upd_candidates(int64_t n, int64_t chunk, const uint8_t* __restrict__ prot,
                                                                const float* __restrict__ orig, float mind, float maxd,
                                                                float thr, float* __restrict__ dens,
                                                                uint8_t* __restrict__ cand, double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double act = 0.0, nc = 0.0;
    const int64_t e0 = blockIdx.x * chunk, e1 = min(n, e0 + chunk);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += RED_THREADS) {
        uint8_t c = 0;
        if (!prot[e]) {
            float d = dens[e];
            if (d > maxd) { d = maxd; dens[e] = d; }
            if (d > mind) {
                act += 1.0;
                if (d < thr) { c = 1; nc += 1.0; }
            }
        } else if (orig[e] > mind) {
            act += 1.0;
        }
        cand[e] = c;
    }
    act = block_sum(act, lds);
    nc = block_sum(nc, lds);
    if (threadIdx.x == 0) { partials[blockIdx.x] = act; partials[RED_BLOCKS + blockIdx.x] = nc; }
}

// Radix-select histogram of the candidates' float bits (positive floats
// order like their bit patterns): counts of the next 8-bit digit among the
// candidates whose higher digits equal 'prefix'.
__global__ __launch_bounds__(RED_THREADS) void k_radix_hist(int64_t n, const uint8_t* __restrict__ cand,
                                                            const float* __restrict__ dens, uint32_t prefix,
                                                            int shift, unsigned long long* __restrict__ hist) {
    __shared__ unsigned int h[256];
    for (int t = threadIdx.x; t < 256; t += RED_THREADS) h[t] = 0;
    __syncthreads();
    const uint32_t himask = shift >= 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
    for (int64_t e = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; e < n; e += (int64_t)RED_BLOCKS * RED_THREADS) {
        if (!cand[e]) continue;
        const uint32_t key = __float_as_uint(dens[e]);
        if ((key & himask) != (prefix & himask)) continue;
        atomicAdd(&h[(key >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 256; t += RED_THREADS)
        if (h[t]) atomicAdd(&hist[t], (unsigned long long)h[t]);
}

// Per-block count of candidates whose key equals 'key' (block = contiguous chunk).
__global__ __launch_bounds__(RED_THREADS) void k_tie_count(int64_t n, int64_t chunk, const uint8_t* __restrict__ cand,
                                                           const float* __restrict__ dens, uint32_t key,
                                                           int64_t* __restrict__ counts) {
    __shared__ double lds[RED_THREADS / 64];
    double c = 0.0;
    const int64_t e0 = blockIdx.x * chunk, e1 = min(n, e0 + chunk);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += RED_THREADS)
        if (cand[e] && __float_as_uint(dens[e]) == key) c += 1.0;
    c = block_sum(c, lds);
    if (threadIdx.x == 0) counts[blockIdx.x] = (int64_t)c;
}

// Cull: every candidate with key < key_t, and the first 'ties' candidates
// with key == key_t in index order (stable (density, index) order,
// TopOpt.jl:451-461).  tie_base = exclusive prefix of the per-block tie counts.
__global__ __launch_bounds__(RED_THREADS) void k_cull(int64_t n, int64_t chunk, const uint8_t* __restrict__ cand,
                                                      uint32_t key_t, int64_t ties, const int64_t* __restrict__ tie_base,
                                                      float mind, float* __restrict__ dens) {
    __shared__ int64_t scan[RED_THREADS];
    const int64_t e0 = blockIdx.x * chunk, e1 = min(n, e0 + chunk);
    int64_t base = tie_base[blockIdx.x];
    for (int64_t t0 = e0; t0 < e1; t0 += RED_THREADS) {
        const int64_t e = t0 + threadIdx.x;
        const bool in = e < e1;
        const bool c = in && cand[e];
        const uint32_t key = c ? __float_as_uint(dens[e]) : 0xFFFFFFFFu;
        const int tie = (c && key == key_t) ? 1 : 0;
        scan[threadIdx.x] = tie;
        __syncthreads();
        for (int off = 1; off < RED_THREADS; off <<= 1) {
            const int64_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        const int64_t rank = base + scan[threadIdx.x] - tie;   // exclusive
        const int64_t tile_total = scan[RED_THREADS - 1];
        if (c && (key < key_t || (tie && rank < ties))) dens[e] = mind;
        base += tile_total;
        __syncthreads();
    }
}

__global__ void k_upd_finish(int64_t n, const uint8_t* __restrict__ prot, const float* __restrict__ orig, float mind,
                             int hard, float* __restrict__ dens) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        if (prot[e]) dens[e] = orig[e];
        else if (hard && dens[e] > mind) dens[e] = 1.0f;
    }
}

// Density remap onto a refined grid (R/src/Mesh/MeshRefiner.jl:185-237):
// the new element's centroid in the old grid's continuous cell-index space,
// clamped to [0.5, n - 0.5]; trilinear interpolation of the 8 surrounding
// old cells (clamped indices) with the clamped weights of
// interpolate_trilinear_projected (:44-75), optionally projected (> 0.4 -> 1,
// < 0.2 -> 1e-4); the alpha field takes the (0,0,0) neighbour.  Float32
// arithmetic in the reference's order, with contraction into FMAs off, so the
// result is the reference's bit for bit.
__global__ void k_regrid(int onx, int ony, int onz, int nnx, int nny, int nnz, float ndx, float ndy, float ndz,
                         float lx, float ly, float lz, int hard, const float* __restrict__ rho,
                         const float* __restrict__ alpha, float* __restrict__ out, float* __restrict__ aout) {
#pragma clang fp contract(off)
    const int64_t n_new = (int64_t)nnx * nny * nnz;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_new; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = e % nnx + 1, iy = (e / nnx) % nny + 1, iz = e / ((int64_t)nnx * nny) + 1;
        const float cx = ((float)ix - 0.5f) * ndx, cy = ((float)iy - 0.5f) * ndy, cz = ((float)iz - 0.5f) * ndz;
        const float u = fminf(fmaxf((cx / lx) * (float)onx, 0.5f), (float)onx - 0.5f);
        const float v = fminf(fmaxf((cy / ly) * (float)ony, 0.5f), (float)ony - 0.5f);
        const float w = fminf(fmaxf((cz / lz) * (float)onz, 0.5f), (float)onz - 0.5f);
        const int64_t u0 = (int64_t)floorf(u - 0.5f), v0 = (int64_t)floorf(v - 0.5f), w0 = (int64_t)floorf(w - 0.5f);
        const float tx = (u - 0.5f) - (float)u0, ty = (v - 0.5f) - (float)v0, tz = (w - 0.5f) - (float)w0;
        auto cl = [](int64_t a, int64_t n) { return a < 1 ? (int64_t)1 : (a > n ? n : a); };
        const int64_t i0 = cl(u0 + 1, onx), i1 = cl(u0 + 2, onx), j0 = cl(v0 + 1, ony), j1 = cl(v0 + 2, ony),
                      k0 = cl(w0 + 1, onz), k1 = cl(w0 + 2, onz);
        const int64_t sl = (int64_t)onx * ony;
        auto at = [&](int64_t i, int64_t j, int64_t k) { return rho[(i - 1) + (j - 1) * onx + (k - 1) * sl]; };
        const float ctx = fminf(fmaxf(tx, 0.0f), 1.0f), cty = fminf(fmaxf(ty, 0.0f), 1.0f),
                    ctz = fminf(fmaxf(tz, 0.0f), 1.0f);
        const float c00 = at(i0, j0, k0) * (1 - ctx) + at(i1, j0, k0) * ctx;
        const float c10 = at(i0, j1, k0) * (1 - ctx) + at(i1, j1, k0) * ctx;
        const float c01 = at(i0, j0, k1) * (1 - ctx) + at(i1, j0, k1) * ctx;
        const float c11 = at(i0, j1, k1) * (1 - ctx) + at(i1, j1, k1) * ctx;
        const float c0 = c00 * (1 - cty) + c10 * cty;
        const float c1 = c01 * (1 - cty) + c11 * cty;
        float val = c0 * (1 - ctz) + c1 * ctz;
        if (hard) val = val > 0.4f ? 1.0f : (val < 0.2f ? 0.0001f : val);
        out[e] = val;
        if (aout) aout[e] = alpha ? alpha[(i0 - 1) + (j0 - 1) * onx + (k0 - 1) * sl] : 0.0f;
    }
}

}  // namespace

void launch_stress(hipStream_t s, const Grid3& g, double dx, double dy, double dz, double E, double nu,
                   const double* U, const double* rho, float* principal, float* vm, float* l1, float* dmax,
                   float* dmin, float* voigt) {
    k_stress<<<grid_for(g.ne), 256, 0, s>>>(g, dx, dy, dz, E, nu, U, rho, principal, vm, l1, dmax, dmin, voigt);
}
void launch_simp_sens(hipStream_t s, const Grid3& g, const double* Ke, const double* rho, double E0, double Emin,
                      double penal, const double* U, double* dc, double* partials) {
    k_simp_sens<<<RED_BLOCKS, RED_THREADS, 0, s>>>(g, Ke, rho, E0, Emin, penal, U, dc, partials);
}
void launch_simp_sens_fact(hipStream_t s, const Grid3& g, const FactCoef<double>& coef, const double* rho, double E0,
                           double Emin, double penal, const double* U, double* dc, double* partials) {
    k_simp_sens_fact<<<RED_BLOCKS, RED_THREADS, 0, s>>>(g, coef, rho, E0, Emin, penal, U, dc, partials);
}
int oc_bracket_bytes() { return (int)sizeof(OcBracket); }
void launch_oc(hipStream_t s, int64_t n, const double* dc, const double* dv, const double* w, double volfrac,
               double move, double tol, int max_passes, double* x, double* q, double* lo, double* hi, void* bracket,
               double* partials, double* out4) {
    OcBracket* br = (OcBracket*)bracket;
    k_oc_init<<<1, 1, 0, s>>>(br, volfrac * (double)n, tol);
    k_oc_prep<<<grid_for(n), 256, 0, s>>>(n, x, dc, dv, move, q, lo, hi);
    for (int p = 0; p < max_passes; ++p) {
        k_oc_eval<false><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, br, q, lo, hi, move, w, partials);
        k_oc_pick<<<1, RED_THREADS, 0, s>>>(br, partials);
    }
    k_oc_apply<false><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, br, q, lo, hi, move, w, x, partials);
    k_oc_final<<<1, RED_THREADS, 0, s>>>(br, partials, out4);
}
void launch_oc_direct(hipStream_t s, int64_t n, const double* dc, const double* dv, const double* w, double volfrac,
                      double move, double tol, int max_passes, double* x, void* bracket, double* partials,
                      double* out4) {
    OcBracket* br = (OcBracket*)bracket;
    k_oc_init<<<1, 1, 0, s>>>(br, volfrac * (double)n, tol);
    for (int p = 0; p < max_passes; ++p) {
        k_oc_eval<true><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, br, x, dc, dv, move, w, partials);
        k_oc_pick<<<1, RED_THREADS, 0, s>>>(br, partials);
    }
    k_oc_apply<true><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, br, x, dc, dv, move, w, x, partials);
    k_oc_final<<<1, RED_THREADS, 0, s>>>(br, partials, out4);
}
void launch_upd_proposed(hipStream_t s, int64_t n, const float* l1, const uint8_t* prot, const float* orig, float allow,
                         float E, float mind, double* prop) {
    k_upd_proposed<<<grid_for(n), 256, 0, s>>>(n, l1, prot, orig, allow, E, mind, prop);
}
void launch_nan_count(hipStream_t s, int64_t n, const double* x, double* partials) {
    k_nan_count<double><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, x, partials);
}
void launch_nan_count(hipStream_t s, int64_t n, const float* x, double* partials) {
    k_nan_count<float><<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, x, partials);
}
void launch_upd_apply(hipStream_t s, int64_t n, const double* filt, const double* prop, int use_prop,
                      const uint8_t* prot, float mind, float maxd, float* dens, double* partials) {
    k_upd_apply<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, filt, prop, use_prop, prot, mind, maxd, dens, partials);
}
void launch_upd_candidates(hipStream_t s, int64_t n, const uint8_t* prot, const float* orig, float mind, float maxd,
                           float thr, float* dens, uint8_t* cand, double* partials) {
    const int64_t chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    k_upd_candidates<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, chunk, prot, orig, mind, maxd, thr, dens, cand, partials);
}
void launch_radix_hist(hipStream_t s, int64_t n, const uint8_t* cand, const float* dens, uint32_t prefix, int shift,
                       unsigned long long* hist) {
    k_radix_hist<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, cand, dens, prefix, shift, hist);
}
void launch_tie_count(hipStream_t s, int64_t n, const uint8_t* cand, const float* dens, uint32_t key, int64_t* counts) {
    const int64_t chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    k_tie_count<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, chunk, cand, dens, key, counts);
}
void launch_cull(hipStream_t s, int64_t n, const uint8_t* cand, uint32_t key_t, int64_t ties, const int64_t* tie_base,
                 float mind, float* dens) {
    const int64_t chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    k_cull<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, chunk, cand, key_t, ties, tie_base, mind, dens);
}
void launch_upd_finish(hipStream_t s, int64_t n, const uint8_t* prot, const float* orig, float mind, int hard,
                       float* dens) {
    k_upd_finish<<<grid_for(n), 256, 0, s>>>(n, prot, orig, mind, hard, dens);
}

void launch_regrid(hipStream_t s, int onx, int ony, int onz, int nnx, int nny, int nnz, float ndx, float ndy,
                   float ndz, float lx, float ly, float lz, int hard, const float* rho, const float* alpha, float* out,
                   float* aout) {
    const int64_t n = (int64_t)nnx * nny * nnz;
    int64_t b = (n + 255) / 256;
    if (b > 65535 * 8) b = 65535 * 8;
    k_regrid<<<(unsigned)(b < 1 ? 1 : b), 256, 0, s>>>(onx, ony, onz, nnx, nny, nnz, ndx, ndy, ndz, lx, ly, lz, hard,
                                                       rho, alpha, out, aout);
}

}  // namespace hexa
