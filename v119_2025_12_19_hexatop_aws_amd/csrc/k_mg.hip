// Geometric multigrid kernels: structured 2x trilinear transfers, the dense
// coarsest-level operator and its inverse.
//
// The reference's "multigrid" (R/src/Solvers/GPUGeometricMultigrid.jl:57-209)
// transfers residuals between a fine and one coarse grid through trilinear
// weights computed from node coordinates, with 24 atomics per fine node in
// the restriction.  On the structured grid the same trilinear operator P is
// the 2x tensor-product stencil (weights 1, 1/2, 1/4, 1/8), so both P and
// R = P^T are written here as gathers: no coordinates, no atomics, and R is
// the exact adjoint of P.
#include <algorithm>

#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

// Transfers run one thread per output node over the flattened node index
// (no idle lanes at row ends); the index is split into (i, j, k) with
// multiply-high divisions.
constexpr int TB = 256;

// n / d for 0 <= n < 2^32 as the high word of n * ceil(2^64 / d)
struct FastDiv {
    uint64_t m;
    uint32_t d;
};
inline FastDiv fast_div(uint32_t d) { return FastDiv{d > 1 ? ~0ull / d + 1 : 0, d}; }
__device__ __forceinline__ uint32_t operator/(uint32_t n, const FastDiv& f) {
    return f.d > 1 ? (uint32_t)__umul64hi((uint64_t)n, f.m) : n;
}

// A node's three components, loaded and stored as one vector access (12 or
// 24 contiguous bytes per lane, lanes contiguous along x).
template <typename T>
struct alignas(sizeof(T)) V3 {
    T x, y, z;
};
template <typename T>
__device__ __forceinline__ V3<T> ld(const T* p, int64_t node) { return *reinterpret_cast<const V3<T>*>(p + 3 * node); }
template <typename T>
__device__ __forceinline__ void st(T* p, int64_t node, const V3<T>& v) { *reinterpret_cast<V3<T>*>(p + 3 * node) = v; }

// Loads through buffer descriptors: an offset past the descriptor's size
// (OOB) or a descriptor of size 0 reads zeros, so the transfers' neighbour
// loads carry no per-lane condition.  (A load under a per-lane condition
// compiles to a branch and a wait of its own, which serialises the round
// trips of a thread's loads.)  A descriptor covers one node plane (at most a
// few MB, so 32-bit offsets suffice for any grid).
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t OOB = 0x80000000u;
__device__ __forceinline__ Rsrc rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
// the descriptor of node plane k of a 3-component node array (size 0 when
// k is outside [0, nz))
template <typename T>
__device__ __forceinline__ Rsrc plane3(const T* p, int k, int nz, int plane_nodes) {
    const bool in = (unsigned)k < (unsigned)nz;
    return rsrc(p + (in ? 3 * (int64_t)plane_nodes * k : 0), in ? plane_nodes * 3 * (int)sizeof(T) : 0);
}
__device__ __forceinline__ V3<float> ldb(Rsrc r, uint32_t off, float) {
    const u32x3 a = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
    return V3<float>{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)};
}
__device__ __forceinline__ V3<double> ldb(Rsrc r, uint32_t off, double) {
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16, 0, 0);
    return V3<double>{__builtin_bit_cast(double, u32x2{a.x, a.y}), __builtin_bit_cast(double, u32x2{a.z, a.w}),
                      __builtin_bit_cast(double, b)};
}

// r_c(I,J,K) = sum over the 3x3x3 fine nodes around (2I,2J,2K) of
// w(di) w(dj) w(dk) r_f, w(0) = 1, w(+-1) = 1/2; held coarse DOFs -> 0.
template <typename T>
__global__ __launch_bounds__(TB) void k_restrict(Grid3 f, Grid3 c, FastDiv dnx, FastDiv dny, uint32_t ncn,
                                                 const T* __restrict__ rf, const uint8_t* __restrict__ cmask,
                                                 T* __restrict__ rc, int64_t fs, int64_t cs,
                                                 const int32_t* __restrict__ halt, int hs) {
    if (halt && halt[blockIdx.y * hs]) return;   // the solve this V-cycle belongs to has ended
    rf += blockIdx.y * fs;   // right-hand side blockIdx.y of a batch
    rc += blockIdx.y * cs;
    const uint32_t n = blockIdx.x * TB + threadIdx.x;
    if (n >= ncn) return;
    const uint32_t row = n / dnx, K = row / dny;
    const int I = (int)(n - row * dnx.d), J = (int)(row - K * dny.d);
    const int FNX = f.nx + 1, FNY = f.ny + 1, FNZ = f.nz + 1;
    const uint8_t mb = cmask[n];
    V3<T> s{T(0), T(0), T(0)};
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
        const Rsrc rp = plane3(rf, 2 * (int)K + dk, FNZ, FNX * FNY);
        const T wk = dk == 0 ? T(1) : T(0.5);
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            const int j = 2 * J + dj;
            const T wjk = wk * (dj == 0 ? T(1) : T(0.5));
#pragma unroll
            for (int di = -1; di <= 1; ++di) {
                const int i = 2 * I + di;
                const bool in = (unsigned)j < (unsigned)FNY && (unsigned)i < (unsigned)FNX;
                const T w = wjk * (di == 0 ? T(1) : T(0.5));
                const V3<T> v = ldb(rp, in ? (uint32_t)(i + FNX * j) * (3 * sizeof(T)) : OOB, T(0));
                s.x += w * v.x;
                s.y += w * v.y;
                s.z += w * v.z;
            }
        }
    }
    st(rc, n, V3<T>{(mb & 1) ? T(0) : s.x, (mb & 2) ? T(0) : s.y, (mb & 4) ? T(0) : s.z});
}

// x_f += P x_c, masked by the fine supports.  One thread per pair of fine
// nodes (2m, 2m+1) of a row: the even node coincides with coarse node m
// (weight 1 along x), the odd one lies midway between m and m+1 (weights
// 1/2), so the pair reads each coarse column once.  Along y and z an even
// fine index takes coarse row/plane j/2, an odd one the two around it (1/2
// each).  Coarse nodes beyond the coarse grid (the last fine node of an odd
// axis under a floor(n/2) coarse grid) contribute nothing.
template <typename T>
__global__ __launch_bounds__(TB) void k_prolong_add(Grid3 f, Grid3 c, FastDiv dpr, FastDiv dny, uint32_t npairs,
                                                    const T* __restrict__ xc, const uint8_t* __restrict__ fmask,
                                                    T* __restrict__ xf, int64_t fs, int64_t cs,
                                                    const int32_t* __restrict__ halt, int hs) {
    if (halt && halt[blockIdx.y * hs]) return;
    xc += blockIdx.y * cs;   // right-hand side blockIdx.y of a batch
    xf += blockIdx.y * fs;
    const uint32_t q = blockIdx.x * TB + threadIdx.x;
    if (q >= npairs) return;
    const uint32_t row = q / dpr, k = row / dny;
    const int m = (int)(q - row * dpr.d), j = (int)(row - k * dny.d);
    const int FNX = f.nx + 1, CNX = c.nx + 1, CNY = c.ny + 1, CNZ = c.nz + 1;
    const int i = 2 * m;
    const bool odd_in = i + 1 < FNX;                      // the pair's odd node exists
    const int64_t n = (int64_t)row * FNX + i;
    const uint8_t mb0 = fmask[n], mb1 = odd_in ? fmask[n + 1] : (uint8_t)7;
    if ((mb0 & 7) == 7 && (mb1 & 7) == 7) return;
    const int j0 = j >> 1, k0 = (int)k >> 1;
    const bool jodd = j & 1, kodd = k & 1;
    // coarse columns m and m+1 of rows j0 (and j0+1 when j is odd) of planes
    // k0 (and k0+1 when k is odd): all 8 loads issued, those outside the
    // coarse grid or not needed read 0
    V3<T> s0{T(0), T(0), T(0)}, s1{T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const Rsrc rp = plane3(xc, (a == 0 || kodd) ? k0 + a : -1, CNZ, CNX * CNY);
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            const bool rin = (bb == 0 || jodd) && j0 + bb < CNY;
            const uint32_t roff = (uint32_t)(CNX * (j0 + bb) + m) * (3 * sizeof(T));
            const V3<T> v0 = ldb(rp, rin && m < CNX ? roff : OOB, T(0));
            const V3<T> v1 = ldb(rp, rin && m + 1 < CNX ? roff + 3 * sizeof(T) : OOB, T(0));
            s0.x += v0.x;
            s0.y += v0.y;
            s0.z += v0.z;
            s1.x += v1.x;
            s1.y += v1.y;
            s1.z += v1.z;
        }
    }
    const T wyz = (jodd ? T(0.5) : T(1)) * (kodd ? T(0.5) : T(1));
    if ((mb0 & 7) != 7) {
        V3<T> x = ld(xf, n);
        if (!(mb0 & 1)) x.x += wyz * s0.x;
        if (!(mb0 & 2)) x.y += wyz * s0.y;
        if (!(mb0 & 4)) x.z += wyz * s0.z;
        st(xf, n, x);
    }
    if ((mb1 & 7) != 7) {
        const T w = T(0.5) * wyz;
        V3<T> x = ld(xf, n + 1);
        if (!(mb1 & 1)) x.x += w * (s0.x + s1.x);
        if (!(mb1 & 2)) x.y += w * (s0.y + s1.y);
        if (!(mb1 & 4)) x.z += w * (s0.z + s1.z);
        st(xf, n + 1, x);
    }
}

// Reference 2-level preconditioner, coarse diagonal
// (R/src/Solvers/GPUGeometricMultigrid.jl:253-315): per coarse node,
// 1 + sum over its coarse elements of rho_c * kd[corner], with rho_c = 1 when
// a present fine node (mask bit 3 clear) has floor(i/2, j/2, k/2) equal to
// the element's index, else 0.001; stored as 1/(D + 1e-9) for all three
// components (:334).  Node-centric: no atomics.
struct Kd8 {
    double v[8];
};
__device__ constexpr int MCX_H[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int MCY_H[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int MCZ_H[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__global__ void k_refmg_dcinv(Grid3 f, Grid3 c, const uint8_t* __restrict__ fmask, Kd8 kd, double* __restrict__ dcinv) {
    const int CNX = c.nx + 1, CNY = c.ny + 1, FNX = f.nx + 1, FNY = f.ny + 1;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < c.nn;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(n % CNX), J = (int)((n / CNX) % CNY), K = (int)(n / ((int64_t)CNX * CNY));
        double d = 1.0;
        for (int q = 0; q < 8; ++q) {
            // element (I - cx, J - cy, K - cz) has this node as its corner q
            const int eI = I - MCX_H[q], eJ = J - MCY_H[q], eK = K - MCZ_H[q];
            if (eI < 0 || eJ < 0 || eK < 0 || eI >= c.nx || eJ >= c.ny || eK >= c.nz) continue;
            bool present = false;
            for (int t = 0; t < 8 && !present; ++t) {
                const int i = 2 * eI + (t & 1), j = 2 * eJ + ((t >> 1) & 1), k = 2 * eK + (t >> 2);
                if (i > f.nx || j > f.ny || k > f.nz) continue;
                present = !(fmask[i + (int64_t)FNX * (j + (int64_t)FNY * k)] & 8);
            }
            d += kd.v[q] * (present ? 1.0 : 0.001);
        }
        const double inv = 1.0 / (d + 1e-9);
        dcinv[3 * n] = inv;
        dcinv[3 * n + 1] = inv;
        dcinv[3 * n + 2] = inv;
    }
}

__device__ constexpr int MCX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int MCY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int MCZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ __forceinline__ int corner_id(int cx, int cy, int cz) {
    return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
}

// Dense coarsest-level matrix (A zero-filled beforehand), one row per
// thread (deterministic): row DOF
// 3n+c sums factor_e * Ke(l,:) over the elements around node n.  Held DOFs
// become identity rows/columns; the shift is added on the free diagonal.
__global__ void k_coarse_assemble(Grid3 g, const double* __restrict__ Ke, const double* __restrict__ fac,
                                  const uint8_t* __restrict__ mask, double shift, double* __restrict__ A) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    const int64_t nd = 3 * g.nn;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < nd;
         row += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = row / 3;
        const int c = (int)(row % 3);
        double* Ar = A + row * nd;  // zero-filled by the caller
        const bool held = (mask[n] >> c) & 1;
        if (held) { Ar[row] = 1.0; continue; }
        const int i = (int)(n % NX), j = (int)((n / NX) % NY), k = (int)(n / ((int64_t)NX * NY));
        for (int s = 0; s < 8; ++s) {
            const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
            const int ei = i + ex, ej = j + ey, ek = k + ez;
            if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny ||
                (unsigned)ek >= (unsigned)g.nz)
                continue;
            const double f = fac[ei + (int64_t)g.nx * (ej + (int64_t)g.ny * ek)];
            const int l = corner_id(-ex, -ey, -ez);
            for (int m = 0; m < 8; ++m) {
                const int64_t nm = (ei + MCX[m]) + (int64_t)NX * ((ej + MCY[m]) + (int64_t)NY * (ek + MCZ[m]));
                const uint8_t mm = mask[nm];
                for (int cc = 0; cc < 3; ++cc) {
                    if ((mm >> cc) & 1) continue;
                    Ar[3 * nm + cc] += f * Ke[(3 * l + c) * 24 + 3 * m + cc];
                }
            }
        }
        Ar[row] += shift;
    }
}

// In-place blocked Gauss-Jordan inversion of an SPD matrix (no pivoting is
// needed: every Schur complement of an SPD matrix is SPD), GJ_PANEL pivots
// per three launches.  With the pivot block p (rows and columns
// p0 .. p0+nb-1) and P = inv(A_pp):
//   A_ij <- A_ij - A_ip P A_pj   (i, j outside p)
//   A_pj <- P A_pj,  A_ip <- -A_ip P,  A_pp <- P.
//  * k_bgj_pivot (one workgroup): P by scalar Gauss-Jordan in LDS;
//  * k_bgj_panels (one workgroup per 32 rows / columns): R = P A_p. (the
//    block row, GJ_PANEL x n) and a copy C = A_.p (the block column,
//    n x GJ_PANEL), both read before the update overwrites them;
//  * k_bgj_update (one workgroup per 32 x 32 tile): outside the block row
//    and column, A_ij - (C R)_ij, a rank-32 update on the FP64 matrix cores
//    (each wave one 16 x 16 quarter, 8 steps of v_mfma_f64_16x16x4); the
//    block row and column tiles entry by entry.
// A 4131-DOF coarse grid takes 3 x 130 launches instead of 2 x 4131.
constexpr int GJ_NB = GJ_PANEL;
typedef double gj_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(GJ_NB* GJ_NB) void k_bgj_pivot(int64_t n, int64_t p0, int nb,
                                                          const double* __restrict__ A, double* __restrict__ P) {
    __shared__ double S[GJ_NB][GJ_NB + 1];
    __shared__ double rk[GJ_NB], ck[GJ_NB];
    const int t = threadIdx.x, r = t / GJ_NB, c = t % GJ_NB;
    const bool in = r < nb && c < nb;
    S[r][c] = in ? A[(p0 + r) * n + p0 + c] : 0.0;
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        if (t < GJ_NB) {
            rk[t] = S[k][t];
            ck[t] = S[t][k];
        }
        __syncthreads();
        if (in) {
            const double piv = 1.0 / rk[k];
            double v;
            if (r == k && c == k) v = piv;
            else if (r == k) v = rk[c] * piv;
            else if (c == k) v = -ck[r] * piv;
            else v = S[r][c] - ck[r] * rk[c] * piv;
            S[r][c] = v;
        }
        __syncthreads();
    }
    P[t] = in ? S[r][c] : 0.0;
}

__global__ __launch_bounds__(256) void k_bgj_panels(int64_t n, int64_t p0, int nb, const double* __restrict__ A,
                                                    const double* __restrict__ P, double* __restrict__ R,
                                                    double* __restrict__ C) {
    __shared__ double Ps[GJ_NB][GJ_NB + 1], As[GJ_NB][GJ_NB + 1];
    const int64_t t0 = (int64_t)blockIdx.x * GJ_NB;
    for (int e = threadIdx.x; e < GJ_NB * GJ_NB; e += 256) {
        const int a = e / GJ_NB, b = e % GJ_NB;
        Ps[a][b] = P[e];
        As[a][b] = (a < nb && t0 + b < n) ? A[(p0 + a) * n + t0 + b] : 0.0;
        // block column: rows t0.., columns p0..
        if (t0 + a < n) C[(t0 + a) * GJ_NB + b] = b < nb ? A[(t0 + a) * n + p0 + b] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < GJ_NB * GJ_NB; e += 256) {
        const int q = e / GJ_NB, b = e % GJ_NB;
        if (t0 + b >= n) continue;
        double acc = 0.0;
        for (int m = 0; m < nb; ++m) acc = __builtin_fma(Ps[q][m], As[m][b], acc);
        R[q * n + t0 + b] = acc;
    }
}

__global__ __launch_bounds__(256) void k_bgj_update(int64_t n, int64_t p0, int nb, double* __restrict__ A,
                                                    const double* __restrict__ R, const double* __restrict__ C,
                                                    const double* __restrict__ P) {
    __shared__ double Cs[GJ_NB][GJ_NB + 1], Rs[GJ_NB][GJ_NB + 1], Ps[GJ_NB][GJ_NB + 1];
    const int64_t i0 = (int64_t)blockIdx.y * GJ_NB, j0 = (int64_t)blockIdx.x * GJ_NB;
    const bool pivot_tile = i0 == p0 || j0 == p0;   // p0 is a multiple of GJ_NB
    for (int e = threadIdx.x; e < GJ_NB * GJ_NB; e += 256) {
        const int a = e / GJ_NB, b = e % GJ_NB;
        Cs[a][b] = i0 + a < n ? C[(i0 + a) * GJ_NB + b] : 0.0;
        Rs[a][b] = j0 + b < n ? R[a * n + j0 + b] : 0.0;
        if (pivot_tile) Ps[a][b] = P[e];
    }
    __syncthreads();
    if (!pivot_tile) {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int wr = 16 * (w >> 1), wc = 16 * (w & 1), lj = lane & 15, lk = lane >> 4;
        gj_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < GJ_NB / 4; ++st)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Cs[wr + lj][4 * st + lk], Rs[4 * st + lk][wc + lj], acc, 0, 0,
                                                       0);
        // accumulator r of lane group lk holds row lk + 4 r, column lj
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t i = i0 + wr + lk + 4 * r, j = j0 + wc + lj;
            if (i < n && j < n) A[i * n + j] -= acc[r];
        }
        return;
    }
    for (int e = threadIdx.x; e < GJ_NB * GJ_NB; e += 256) {
        const int a = e / GJ_NB, b = e % GJ_NB;
        const int64_t i = i0 + a, j = j0 + b;
        if (i >= n || j >= n) continue;
        const bool ip = i >= p0 && i < p0 + nb, jp = j >= p0 && j < p0 + nb;
        double v;
        if (ip && jp) {
            v = Ps[i - p0][j - p0];
        } else if (ip) {
            v = Rs[i - p0][b];
        } else if (jp) {
            double acc = 0.0;
            for (int m = 0; m < nb; ++m) acc = __builtin_fma(Cs[a][m], Ps[m][j - p0], acc);
            v = -acc;
        } else {
            double acc = 0.0;
            for (int m = 0; m < nb; ++m) acc = __builtin_fma(Cs[a][m], Rs[m][b], acc);
            v = A[i * n + j] - acc;
        }
        A[i * n + j] = v;
    }
}

// y = Ainv x with FP64 arithmetic, one wave per row (deterministic tree).
template <typename T>
__global__ __launch_bounds__(256) void k_dense_gemv(int64_t n, const double* __restrict__ Ainv,
                                                    const T* __restrict__ x, T* __restrict__ y, int64_t vs,
                                                    const int32_t* __restrict__ halt, int hs) {
    if (halt && halt[blockIdx.y * hs]) return;
    x += blockIdx.y * vs;   // right-hand side blockIdx.y of a batch
    y += blockIdx.y * vs;
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* Ar = Ainv + row * n;
    double s = 0.0;
    for (int64_t c = lane; c < n; c += 64) s += Ar[c] * (double)x[c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[row] = (T)s;
}

}  // namespace

template <typename T>
void launch_restrict(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* rf,
                     const uint8_t* cmask, T* rc, int nb, const int32_t* halt, int hs) {
    const uint32_t ncn = (uint32_t)coarse.nn;
    k_restrict<T><<<dim3((ncn + TB - 1) / TB, nb), TB, 0, s>>>(fine, coarse, fast_div(coarse.nx + 1),
                                                               fast_div(coarse.ny + 1), ncn, rf, cmask, rc,
                                                               batch_stride(3 * fine.nn), batch_stride(3 * coarse.nn),
                                                               halt, hs);
}

template <typename T>
void launch_prolong_add(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* xc,
                        const uint8_t* fmask, T* xf, int nb, const int32_t* halt, int hs) {
    const uint32_t ppr = (uint32_t)(fine.nx + 2) / 2;   // node pairs per fine row
    const uint32_t npairs = ppr * (uint32_t)((fine.ny + 1) * (fine.nz + 1));
    k_prolong_add<T><<<dim3((npairs + TB - 1) / TB, nb), TB, 0, s>>>(fine, coarse, fast_div(ppr), fast_div(fine.ny + 1),
                                                                     npairs, xc, fmask, xf, batch_stride(3 * fine.nn),
                                                                     batch_stride(3 * coarse.nn), halt, hs);
}

void launch_refmg_dcinv(hipStream_t s, const Grid3& fine, const Grid3& coarse, const uint8_t* fmask,
                        const double* kd8, double* dcinv) {
    Kd8 kd;
    for (int q = 0; q < 8; ++q) kd.v[q] = kd8[q];
    k_refmg_dcinv<<<grid_for(coarse.nn), 256, 0, s>>>(fine, coarse, fmask, kd, dcinv);
}

void launch_coarse_assemble(hipStream_t s, const Grid3& g, const double* Ke, const double* fac,
                            const uint8_t* mask, double shift, double* A) {
    k_coarse_assemble<<<grid_for(3 * g.nn, 64), 64, 0, s>>>(g, Ke, fac, mask, shift, A);
}

void launch_gauss_jordan(hipStream_t s, int64_t n, double* A, double* rowbuf, double* colbuf) {
    const unsigned tiles = (unsigned)((n + GJ_NB - 1) / GJ_NB);
    double* P = rowbuf + (int64_t)GJ_NB * n;
    for (int64_t p0 = 0; p0 < n; p0 += GJ_NB) {
        const int nb = (int)std::min<int64_t>(GJ_NB, n - p0);
        k_bgj_pivot<<<1, GJ_NB * GJ_NB, 0, s>>>(n, p0, nb, A, P);
        k_bgj_panels<<<tiles, 256, 0, s>>>(n, p0, nb, A, P, rowbuf, colbuf);
        k_bgj_update<<<dim3(tiles, tiles), 256, 0, s>>>(n, p0, nb, A, rowbuf, colbuf, P);
    }
}

template <typename T>
void launch_dense_gemv(hipStream_t s, int64_t n, const double* Ainv, const T* x, T* y, int nb, const int32_t* halt,
                       int hs) {
    k_dense_gemv<T><<<dim3((unsigned)((n + 3) / 4), nb), 256, 0, s>>>(n, Ainv, x, y, batch_stride(n), halt, hs);
}

template void launch_restrict<double>(hipStream_t, const Grid3&, const Grid3&, const double*,
                                      const uint8_t*, double*, int, const int32_t*, int);
template void launch_restrict<float>(hipStream_t, const Grid3&, const Grid3&, const float*,
                                     const uint8_t*, float*, int, const int32_t*, int);
template void launch_prolong_add<double>(hipStream_t, const Grid3&, const Grid3&, const double*,
                                         const uint8_t*, double*, int, const int32_t*, int);
template void launch_prolong_add<float>(hipStream_t, const Grid3&, const Grid3&, const float*,
                                        const uint8_t*, float*, int, const int32_t*, int);
template void launch_dense_gemv<double>(hipStream_t, int64_t, const double*, const double*, double*, int, const int32_t*, int);
template void launch_dense_gemv<float>(hipStream_t, int64_t, const double*, const float*, float*, int, const int32_t*, int);

}  // namespace hexa
