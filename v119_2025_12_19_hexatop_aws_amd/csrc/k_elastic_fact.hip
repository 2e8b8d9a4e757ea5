// Factored element operator for the exact brick stiffness (gfx950).
//
// For a brick, the hex8 element stiffness factorises per axis into 2x2
// one-dimensional integrals (mass, stiffness, mixed).  In the basis of
// per-axis sums and differences (u0+u1, u1-u0) those 2x2 factors are
// diagonal (mass, stiffness) or a single entry (mixed), so Ke u_e becomes:
// a 3-stage butterfly per component (24 add/sub), a diagonal scaling per
// component and mode (21 mul), 24 cross-component couplings, and the
// inverse butterfly -- about 210 flops instead of the 1152 of a dense 24x24
// product.  The result equals the Gauss-integrated Ke of the reference
// (R/src/Core/Element.jl:73-111) up to rounding.
//
// Work mapping: a 16x16 thread block owns a 14x14 column of nodes and marches
// along z over a chunk of node planes.  At each element layer every thread
// (except the last row/column) computes one element from the two node planes
// held in LDS, writes its 8 corner results to LDS, and the owning thread of
// each node sums the 4 in-plane elements' corner values; the layer below's
// upper-corner sums are carried in registers.  No atomics, no colours, and
// every node sums its 8 elements in the same order on every run.
#include "hexa_internal.h"

namespace hexa {
namespace {

#ifndef HEXA_FACT_PF
#define HEXA_FACT_PF 1
#endif
constexpr int BX = 16, BY = 16;
constexpr int PF = HEXA_FACT_PF;   // layers prefetched ahead

template <typename DT>
__device__ __forceinline__ float dinv_f(DT v) { return (float)v; }
template <>
__device__ __forceinline__ float dinv_f<uint16_t>(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
template <typename T, typename DT>
__device__ __forceinline__ T ld_dinv(const DT* p, int64_t i) {
    if constexpr (sizeof(DT) == sizeof(double)) return (T)p[i];
    else return (T)dinv_f<DT>(p[i]);
}

// Forward butterfly of one component: v[cx + 2cy + 4cz] -> modes m = mx + 2my + 4mz
template <typename T>
__device__ __forceinline__ void fwd8(T (&v)[8]) {
#pragma unroll
    for (int s = 1; s <= 4; s <<= 1) {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (a & s) continue;
            const T lo = v[a], hi = v[a | s];
            v[a] = lo + hi;
            v[a | s] = hi - lo;
        }
    }
}
// Inverse butterfly without the 1/8 (folded into the coefficients).
template <typename T>
__device__ __forceinline__ void inv8(T (&v)[8]) {
#pragma unroll
    for (int s = 1; s <= 4; s <<= 1) {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (a & s) continue;
            const T lo = v[a], hi = v[a | s];
            v[a] = lo - hi;
            v[a | s] = lo + hi;
        }
    }
}

// corner index of the reference's node order -> (cx + 2cy + 4cz)
__device__ constexpr int LEX[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// Element result in lexicographic corner order out[comp][cx+2cy+4cz] from
// inputs in the same order.
template <typename T>
__device__ __forceinline__ void elem_apply(T (&v)[3][8], const FactCoef<T>& c) {
#pragma unroll
    for (int i = 0; i < 3; ++i) fwd8(v[i]);
    T y[3][8];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        y[i][0] = T(0);
#pragma unroll
        for (int m = 1; m < 8; ++m) y[i][m] = c.D[i][m] * v[i][m];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            const int d = 3 - i - j;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int mq = q << d;
                // lambda term: in (m_i=0, m_j=1) -> out (m_i=1, m_j=0)
                y[i][(1 << i) | mq] += c.cl[d][q] * v[j][(1 << j) | mq];
                // mu term: in (m_i=1, m_j=0) -> out (m_i=0, m_j=1)
                y[i][(1 << j) | mq] += c.cm[d][q] * v[j][(1 << i) | mq];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        inv8(y[i]);
#pragma unroll
        for (int m = 0; m < 8; ++m) v[i][m] = y[i][m];
    }
}

template <typename T>
__device__ __forceinline__ void load_node(const T* __restrict__ u, int i, int j, int k, const Grid3& g,
                                          T (&o)[3]) {
    const bool in = (unsigned)i <= (unsigned)g.nx && (unsigned)j <= (unsigned)g.ny && (unsigned)k <= (unsigned)g.nz;
    const int64_t n = i + (int64_t)(g.nx + 1) * (j + (int64_t)(g.ny + 1) * k);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = in ? u[3 * n + c] : T(0);
}

template <typename T, typename DT, int MODE>
__global__ __launch_bounds__(BX* BY) void k_elastic_fact(Grid3 g, FactCoef<T> coef, const T* __restrict__ fac,
                                                         const uint8_t* __restrict__ mask,
                                                         const T* __restrict__ u, const T* __restrict__ b,
                                                         const DT* __restrict__ dinv, T omega, T shift,
                                                         T* __restrict__ y, int kchunk) {
    __shared__ T su[3][BY][BX + 1];
    __shared__ T sc[24][BY][BX + 1];
    const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
    const int i = blockIdx.x * (BX - 2) - 1 + tx;
    const int j = blockIdx.y * (BY - 2) - 1 + ty;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(kb + kchunk, g.nz + 1);
    const bool elem_col = tx < BX - 1 && ty < BY - 1 && (unsigned)i < (unsigned)g.nx &&
                          (unsigned)j < (unsigned)g.ny;
    const bool owner = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2 && i <= g.nx && j <= g.ny;
    const int64_t ecol = elem_col ? i + (int64_t)g.nx * j : 0;
    const int64_t plane_e = (int64_t)g.nx * g.ny;
    const int64_t plane_n = (int64_t)(g.nx + 1) * (g.ny + 1);
    const int64_t ncol = owner ? i + (int64_t)(g.nx + 1) * j : 0;

    // the element layer kb-1 contributes its upper corners to plane kb
    T lo[4][3];   // corners (0,0),(1,0),(0,1),(1,1) of my element at the lower plane
    {
        T own[3];
        load_node(u, i, j, kb - 1, g, own);
#pragma unroll
        for (int c = 0; c < 3; ++c) su[c][ty][tx] = own[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[0][c] = own[c];
            lo[1][c] = su[c][ty][min(tx + 1, BX - 1)];
            lo[2][c] = su[c][min(ty + 1, BY - 1)][tx];
            lo[3][c] = su[c][min(ty + 1, BY - 1)][min(tx + 1, BX - 1)];
        }
    }
    T carry[3] = {T(0), T(0), T(0)};
    T uown[3] = {lo[0][0], lo[0][1], lo[0][2]};
    // software prefetch ring: the node plane and element factor of the next
    // PF layers are in flight while the current layer is computed
    T pu[PF][3], pf[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d) {
        load_node(u, i, j, kb + d, g, pu[d]);
        const int lay = kb - 1 + d;
        pf[d] = (elem_col && (unsigned)lay < (unsigned)g.nz) ? fac[ecol + plane_e * lay] : T(0);
    }

    auto layer = [&](const int kl, const T (&hi0)[3], const T f) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; ++c) su[c][ty][tx] = hi0[c];
        __syncthreads();
        T hi[4][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            hi[0][c] = hi0[c];
            hi[1][c] = su[c][ty][min(tx + 1, BX - 1)];
            hi[2][c] = su[c][min(ty + 1, BY - 1)][tx];
            hi[3][c] = su[c][min(ty + 1, BY - 1)][min(tx + 1, BX - 1)];
        }
        if (tx < BX - 1 && ty < BY - 1) {
            // lexicographic corners: (cx,cy,cz) -> cx + 2cy + 4cz
            T v[3][8];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                v[c][0] = f * lo[0][c]; v[c][1] = f * lo[1][c]; v[c][2] = f * lo[2][c]; v[c][3] = f * lo[3][c];
                v[c][4] = f * hi[0][c]; v[c][5] = f * hi[1][c]; v[c][6] = f * hi[2][c]; v[c][7] = f * hi[3][c];
            }
            elem_apply(v, coef);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int m = 0; m < 8; ++m) sc[m * 3 + c][ty][tx] = v[c][m];
        }
        __syncthreads();
        if (owner) {
            T low[3], up[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                // node = corner (1,1) of element (tx-1,ty-1), (0,1) of (tx,ty-1),
                // (1,0) of (tx-1,ty), (0,0) of (tx,ty)
                low[c] = sc[3 * 3 + c][ty - 1][tx - 1] + sc[2 * 3 + c][ty - 1][tx] +
                         sc[1 * 3 + c][ty][tx - 1] + sc[0 * 3 + c][ty][tx];
                up[c] = sc[7 * 3 + c][ty - 1][tx - 1] + sc[6 * 3 + c][ty - 1][tx] +
                        sc[5 * 3 + c][ty][tx - 1] + sc[4 * 3 + c][ty][tx];
            }
            if (kl >= kb && kl <= g.nz) {
                const int64_t n = ncol + plane_n * kl;
                const uint8_t mb = mask ? mask[n] : 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T Au = carry[c] + low[c] + shift * uown[c];
                    T out;
                    if (MODE == OP_APPLY) out = Au;
                    else if (MODE == OP_RESID) out = b[3 * n + c] - Au;
                    else out = uown[c] + omega * ld_dinv<T, DT>(dinv, 3 * n + c) * (b[3 * n + c] - Au);
                    y[3 * n + c] = ((mb >> c) & 1) ? T(0) : out;
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) carry[c] = up[c];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) lo[q][c] = hi[q][c];
#pragma unroll
        for (int c = 0; c < 3; ++c) uown[c] = hi0[c];
    };

    for (int k0 = kb - 1; k0 < ke; k0 += PF) {
#pragma unroll
        for (int d = 0; d < PF; ++d) {
            const int kl = k0 + d;
            if (kl >= ke) break;
            // upper node plane kl+1 of element layer kl
            const T hi0[3] = {pu[d][0], pu[d][1], pu[d][2]};
            const T f = pf[d];
            load_node(u, i, j, kl + PF + 1, g, pu[d]);
            const int lay = kl + PF;
            pf[d] = (elem_col && (unsigned)lay < (unsigned)g.nz) ? fac[ecol + plane_e * lay] : T(0);
            layer(kl, hi0, f);
        }
    }
}

}  // namespace

template <typename T>
FactCoef<T> make_fact_coef(double hx, double hy, double hz, double nu) {
    const double h[3] = {hx, hy, hz};
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    double w[3][2];
    for (int d = 0; d < 3; ++d) { w[d][0] = h[d] / 2.0; w[d][1] = h[d] / 6.0; }
    auto akk = [&](int k, int m) {
        if (!((m >> k) & 1)) return 0.0;
        double v = 2.0 / h[k];
        for (int d = 0; d < 3; ++d)
            if (d != k) v *= w[d][(m >> d) & 1];
        return v;
    };
    FactCoef<T> c;
    for (int i = 0; i < 3; ++i)
        for (int m = 0; m < 8; ++m) {
            double s = (lam + 2.0 * mu) * akk(i, m);
            for (int k = 0; k < 3; ++k)
                if (k != i) s += mu * akk(k, m);
            c.D[i][m] = (T)(s / 8.0);
        }
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < 2; ++q) {
            c.cl[d][q] = (T)(lam * w[d][q] / 8.0);
            c.cm[d][q] = (T)(mu * w[d][q] / 8.0);
        }
    return c;
}

template <typename T, typename DT>
void launch_elastic_fact(hipStream_t s, const Grid3& g, OpMode mode, const FactCoef<T>& coef, const T* fac,
                         const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y) {
    const unsigned gx = (unsigned)((g.nx + 1 + (BX - 2) - 1) / (BX - 2));
    const unsigned gy = (unsigned)((g.ny + 1 + (BY - 2) - 1) / (BY - 2));
    // node planes per block: enough blocks to fill the chip, long enough
    // chunks to amortise the one recomputed element layer per block
    int kchunk = 32;
    while (kchunk > 4 && (int64_t)gx * gy * ((g.nz + 1 + kchunk - 1) / kchunk) < 4096) kchunk /= 2;
    const unsigned gz = (unsigned)((g.nz + 1 + kchunk - 1) / kchunk);
    dim3 grid(gx, gy, gz);
    switch (mode) {
        case OP_APPLY:
            k_elastic_fact<T, DT, OP_APPLY><<<grid, BX * BY, 0, s>>>(g, coef, fac, mask, u, b, dinv, omega, shift, y, kchunk);
            break;
        case OP_RESID:
            k_elastic_fact<T, DT, OP_RESID><<<grid, BX * BY, 0, s>>>(g, coef, fac, mask, u, b, dinv, omega, shift, y, kchunk);
            break;
        default:
            k_elastic_fact<T, DT, OP_SMOOTH><<<grid, BX * BY, 0, s>>>(g, coef, fac, mask, u, b, dinv, omega, shift, y, kchunk);
            break;
    }
}

template FactCoef<double> make_fact_coef<double>(double, double, double, double);
template FactCoef<float> make_fact_coef<float>(double, double, double, double);
#define HX_INST_FACT(T, DT)                                                                            \
    template void launch_elastic_fact<T, DT>(hipStream_t, const Grid3&, OpMode, const FactCoef<T>&,   \
                                             const T*, const uint8_t*, const T*, const T*, const DT*, T, \
                                             T, T*);
HX_INST_FACT(double, double)
HX_INST_FACT(float, float)
HX_INST_FACT(float, uint16_t)

}  // namespace hexa
