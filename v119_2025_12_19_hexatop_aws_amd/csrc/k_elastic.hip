// Elasticity operator kernels for the structured hex8 grid (gfx950).
//
// K(rho) u is applied node by node ("gather"): the thread that owns node n
// sums, over the (up to) 8 elements around n, factor_e * (Ke u_e) restricted
// to n's three rows.  Nothing is scattered, so there are no atomics, no
// colour sweeps, and every output is summed in the same order on every run.
// This replaces the element-scatter kernel matvec_unstructured_atomic_kernel!
// (R/src/Solvers/GPUSolver.jl:91-128) together with its expand/contract
// passes (GPUSolver.jl:146-160,322-330): held DOFs are masked in place.
//
// Ke is read through uniform (scalar-cache) loads: every lane of a wave reads
// the same Ke entry at the same time, so the entries live in SGPRs and each
// v_fma takes one scalar operand; LDS bandwidth is left to the data.
#include "hexa_internal.h"

namespace hexa {

namespace {

__device__ constexpr int corner_of(int cx, int cy, int cz) {
    return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
}
__device__ constexpr int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int CY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

template <typename DT>
__device__ __forceinline__ float dinv_to_float(DT v) { return (float)v; }
template <>
__device__ __forceinline__ float dinv_to_float<uint16_t>(uint16_t v) {
    return __uint_as_float(((uint32_t)v) << 16);
}
template <typename T, typename DT>
__device__ __forceinline__ T load_dinv(const DT* p, int64_t i) { return (T)dinv_to_float<DT>(p[i]); }
template <>
__device__ __forceinline__ double load_dinv<double, double>(const double* p, int64_t i) { return p[i]; }

// Contribution of the 4 elements of one element layer (ez = -1 below the
// node, ez = 0 above) to the node's three rows.  un holds the two node
// planes that layer touches: un[0] = plane k+ez, un[1] = plane k+ez+1,
// each a 3x3 block around (i,j).
template <typename T, int EZ>
__device__ __forceinline__ void layer_rows(const T* __restrict__ Ke, const T (&un)[2][9][3],
                                           const T (&f)[4], T& a0, T& a1, T& a2) {
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) {
        const int ex = (sl & 1) - 1, ey = ((sl >> 1) & 1) - 1;
        const int l = corner_of(-ex, -ey, -EZ);
        T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int q = (ex + CX[m] + 1) + 3 * (ey + CY[m] + 1);
            const int pl = CZ[m];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T v = un[pl][q][c];
                s0 += Ke[(3 * l + 0) * 24 + 3 * m + c] * v;
                s1 += Ke[(3 * l + 1) * 24 + 3 * m + c] * v;
                s2 += Ke[(3 * l + 2) * 24 + 3 * m + c] * v;
            }
        }
        a0 += f[sl] * s0;
        a1 += f[sl] * s1;
        a2 += f[sl] * s2;
    }
}

template <typename T>
__device__ __forceinline__ void load_plane(const T* __restrict__ u, int i, int j, int kk, int NX,
                                           int NY, int NZ, T (&pl)[9][3]) {
    const bool kin = (unsigned)kk < (unsigned)NZ;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const int ii = i + dx - 1, jj = j + dy - 1;
            const bool in = kin && (unsigned)ii < (unsigned)NX && (unsigned)jj < (unsigned)NY;
            const int64_t nn = ii + (int64_t)NX * (jj + (int64_t)NY * kk);
#pragma unroll
            for (int c = 0; c < 3; ++c) pl[dx + 3 * dy][c] = in ? u[3 * nn + c] : T(0);
        }
    }
}

template <typename T>
__device__ __forceinline__ void load_factors(const T* __restrict__ fac, const Grid3& g, int i, int j,
                                             int ek, T (&f)[4]) {
    const bool kin = (unsigned)ek < (unsigned)g.nz;
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) {
        const int ei = i + (sl & 1) - 1, ej = j + ((sl >> 1) & 1) - 1;
        const bool ok = kin && (unsigned)ei < (unsigned)g.nx && (unsigned)ej < (unsigned)g.ny;
        f[sl] = ok ? fac[ei + (int64_t)g.nx * (ej + (int64_t)g.ny * ek)] : T(0);
    }
}

template <typename T, typename DT, int MODE>
__global__ __launch_bounds__(256) void k_elastic_dense(Grid3 g, const T* __restrict__ Ke,
                                                       const T* __restrict__ fac,
                                                       const uint8_t* __restrict__ mask,
                                                       const T* __restrict__ u,
                                                       const T* __restrict__ b,
                                                       const DT* __restrict__ dinv, T omega,
                                                       T shift, T* __restrict__ y) {
    const int NX = g.nx + 1, NY = g.ny + 1, NZ = g.nz + 1;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= NX || j >= NY) return;
    const int64_t n = i + (int64_t)NX * (j + (int64_t)NY * k);

    T a0 = 0, a1 = 0, a2 = 0;
    T un[2][9][3];
    T f[4];
    // element layer below the node: node planes k-1 and k
    load_plane(u, i, j, k - 1, NX, NY, NZ, un[0]);
    load_plane(u, i, j, k, NX, NY, NZ, un[1]);
    load_factors(fac, g, i, j, k - 1, f);
    layer_rows<T, -1>(Ke, un, f, a0, a1, a2);
    // element layer above: node planes k and k+1
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) un[0][q][c] = un[1][q][c];
    load_plane(u, i, j, k + 1, NX, NY, NZ, un[1]);
    load_factors(fac, g, i, j, k, f);
    layer_rows<T, 0>(Ke, un, f, a0, a1, a2);

    const uint8_t mb = mask ? mask[n] : 0;
    const T uc[3] = {un[0][4][0], un[0][4][1], un[0][4][2]};
    const T ac[3] = {a0, a1, a2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bool held = (mb >> c) & 1;
        const T Au = ac[c] + shift * uc[c];
        T out;
        if (MODE == OP_APPLY) {
            out = Au;
        } else if (MODE == OP_RESID) {
            out = b[3 * n + c] - Au;
        } else {
            out = uc[c] + omega * load_dinv<T, DT>(dinv, 3 * n + c) * (b[3 * n + c] - Au);
        }
        y[3 * n + c] = held ? T(0) : out;
    }
}

template <typename T>
__global__ void k_stiffness_factor(int64_t ne, const double* __restrict__ rho,
                                   const uint8_t* __restrict__ active, double E0, double Emin,
                                   double penal, T* __restrict__ fac) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double r = rho[e];
        const bool on = active ? active[e] != 0 : true;
        const double p = penal == 1.0 ? r : pow(r, penal);
        fac[e] = on ? (T)(Emin + p * (E0 - Emin)) : T(0);
    }
}

// diag(K(rho)) per DOF, node-centric (deterministic).  x_entry=1 reproduces
// the reference GPU kernel's use of the x-component local diagonal for all
// three components (GPUSolver.jl:130-144); x_entry=0 is the exact
// per-component diagonal of CPUSolver.jl:142-190.
template <typename T>
__global__ void k_diag(Grid3 g, const double* __restrict__ Ke, const T* __restrict__ fac,
                       int x_entry, T* __restrict__ diag) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nn;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % NX), j = (int)((n / NX) % NY), k = (int)(n / ((int64_t)NX * NY));
        double d[3] = {0, 0, 0};
        for (int s = 0; s < 8; ++s) {
            const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
            const int ei = i + ex, ej = j + ey, ek = k + ez;
            if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny ||
                (unsigned)ek >= (unsigned)g.nz)
                continue;
            const double f = (double)fac[ei + (int64_t)g.nx * (ej + (int64_t)g.ny * ek)];
            const int l = corner_of(-ex, -ey, -ez);
            for (int c = 0; c < 3; ++c) {
                const int r = x_entry ? 3 * l : 3 * l + c;
                d[c] += Ke[r * 24 + r] * f;
            }
        }
        for (int c = 0; c < 3; ++c) diag[3 * n + c] = (T)d[c];
    }
}

// Node mask from the per-DOF support flags and the active-element set: a
// node no active element touches is excluded (all components held), which
// mirrors the node renumbering of MeshPruner.prune_system
// (R/src/Mesh/MeshPruner.jl:29-47).
__global__ void k_node_mask(Grid3 g, const uint8_t* __restrict__ fixed_dof,
                            const uint8_t* __restrict__ active, uint8_t* __restrict__ mask) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nn;
         n += (int64_t)gridDim.x * blockDim.x) {
        uint8_t m = 0;
        for (int c = 0; c < 3; ++c) m |= (fixed_dof && fixed_dof[3 * n + c]) ? (1 << c) : 0;
        if (active) {
            const int i = (int)(n % NX), j = (int)((n / NX) % NY), k = (int)(n / ((int64_t)NX * NY));
            bool touched = false;
            for (int s = 0; s < 8 && !touched; ++s) {
                const int ei = i + (s & 1) - 1, ej = j + ((s >> 1) & 1) - 1, ek = k + ((s >> 2) & 1) - 1;
                if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny ||
                    (unsigned)ek >= (unsigned)g.nz)
                    continue;
                touched = active[ei + (int64_t)g.nx * (ej + (int64_t)g.ny * ek)] != 0;
            }
            if (!touched) m = MASK_ALL | 0x8;
        }
        mask[n] = m;
    }
}

// Void pruning on the device (R/src/Solvers/Solver.jl:85-91): an element
// stays active when its Float32 density exceeds the Float32 threshold; the
// per-block counts of active elements go to partials[blockIdx.x].
__global__ __launch_bounds__(256) void k_active_from_rho(int64_t ne, const double* __restrict__ rho, float thr,
                                                         uint8_t* __restrict__ active, double* __restrict__ partials) {
    __shared__ double lds[4];
    double cnt = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t a = (float)rho[e] > thr ? 1 : 0;
        active[e] = a;
        cnt += a;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// Coarse stiffness multiplier (rediscretisation): sum of the 8 children / 8.
// A coarse grid has ceil(n/2) elements per axis; on an odd fine axis the last
// coarse element overhangs the box by one fine layer, whose missing children
// count as void.
template <typename T>
__global__ void k_coarsen_factor(Grid3 f, Grid3 c, const T* __restrict__ ffac, T* __restrict__ cfac) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < c.ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(e % c.nx), J = (int)((e / c.nx) % c.ny), K = (int)(e / ((int64_t)c.nx * c.ny));
        double s = 0;
        for (int q = 0; q < 8; ++q) {
            const int i = 2 * I + CX[q], j = 2 * J + CY[q], k = 2 * K + CZ[q];
            if (i < f.nx && j < f.ny && k < f.nz) s += (double)ffac[i + (int64_t)f.nx * (j + (int64_t)f.ny * k)];
        }
        cfac[e] = (T)(0.125 * s);
    }
}

// Coarse node mask by injection: coarse node (I,J,K) takes the flags of fine
// node (2I,2J,2K), clamped to the box for the overhanging nodes of an odd
// axis (so supports on a max face reach the coarse level).
__global__ void k_coarsen_mask(Grid3 f, Grid3 c, const uint8_t* __restrict__ fmask,
                               uint8_t* __restrict__ cmask) {
    const int CNX = c.nx + 1, CNY = c.ny + 1, FNX = f.nx + 1, FNY = f.ny + 1;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < c.nn;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(n % CNX), J = (int)((n / CNX) % CNY), K = (int)(n / ((int64_t)CNX * CNY));
        const int i = min(2 * I, f.nx), j = min(2 * J, f.ny), k = min(2 * K, f.nz);
        cmask[n] = fmask[i + (int64_t)FNX * (j + (int64_t)FNY * k)];
    }
}

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

template <typename T, typename DT>
void launch_elastic_op(hipStream_t s, const Grid3& g, OpMode mode, const T* Ke, const T* fac,
                       const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega,
                       T shift, T* y) {
    if (elastic_mfma_supported(g)) {
        launch_elastic_mfma<T, DT>(s, g, mode, Ke, fac, mask, u, b, dinv, omega, shift, y);
        return;
    }
    dim3 block(64, 4, 1);
    dim3 grid((unsigned)((g.nx + 1 + 63) / 64), (unsigned)((g.ny + 1 + 3) / 4), (unsigned)(g.nz + 1));
    switch (mode) {
        case OP_APPLY:
            k_elastic_dense<T, DT, OP_APPLY><<<grid, block, 0, s>>>(g, Ke, fac, mask, u, b, dinv, omega,
                                                                   shift, y);
            break;
        case OP_RESID:
            k_elastic_dense<T, DT, OP_RESID><<<grid, block, 0, s>>>(g, Ke, fac, mask, u, b, dinv, omega,
                                                                   shift, y);
            break;
        default:
            k_elastic_dense<T, DT, OP_SMOOTH><<<grid, block, 0, s>>>(g, Ke, fac, mask, u, b, dinv,
                                                                    omega, shift, y);
            break;
    }
}

template <typename T>
void launch_stiffness_factor(hipStream_t s, int64_t ne, const double* rho, const uint8_t* active,
                             double E0, double Emin, double penal, T* fac) {
    k_stiffness_factor<T><<<grid_for(ne), 256, 0, s>>>(ne, rho, active, E0, Emin, penal, fac);
}

template <typename T>
void launch_diag(hipStream_t s, const Grid3& g, const double* Ke, const T* fac, const uint8_t* mask,
                 int x_entry, T* diag) {
    (void)mask;
    k_diag<T><<<grid_for(g.nn), 256, 0, s>>>(g, Ke, fac, x_entry, diag);
}

void launch_node_mask(hipStream_t s, const Grid3& g, const uint8_t* fixed_dof, const uint8_t* active,
                      uint8_t* mask) {
    k_node_mask<<<grid_for(g.nn), 256, 0, s>>>(g, fixed_dof, active, mask);
}

void launch_active_from_rho(hipStream_t s, int64_t ne, const double* rho, float thr, uint8_t* active,
                            double* partials) {
    k_active_from_rho<<<RED_BLOCKS, 256, 0, s>>>(ne, rho, thr, active, partials);
}

template <typename T>
void launch_coarsen_factor(hipStream_t s, const Grid3& fine, const Grid3& coarse, const T* ffac, T* cfac) {
    k_coarsen_factor<T><<<grid_for(coarse.ne), 256, 0, s>>>(fine, coarse, ffac, cfac);
}

void launch_coarsen_mask(hipStream_t s, const Grid3& fine, const Grid3& coarse, const uint8_t* fmask,
                         uint8_t* cmask) {
    k_coarsen_mask<<<grid_for(coarse.nn), 256, 0, s>>>(fine, coarse, fmask, cmask);
}

#define HX_INST_OP(T, DT)                                                                          \
    template void launch_elastic_op<T, DT>(hipStream_t, const Grid3&, OpMode, const T*, const T*, \
                                           const uint8_t*, const T*, const T*, const DT*, T, T, T*);
HX_INST_OP(double, double)
HX_INST_OP(float, float)
HX_INST_OP(float, uint16_t)
template void launch_stiffness_factor<double>(hipStream_t, int64_t, const double*, const uint8_t*,
                                              double, double, double, double*);
template void launch_stiffness_factor<float>(hipStream_t, int64_t, const double*, const uint8_t*,
                                             double, double, double, float*);
template void launch_diag<double>(hipStream_t, const Grid3&, const double*, const double*,
                                  const uint8_t*, int, double*);
template void launch_coarsen_factor<double>(hipStream_t, const Grid3&, const Grid3&, const double*, double*);
template void launch_coarsen_factor<float>(hipStream_t, const Grid3&, const Grid3&, const float*, float*);

}  // namespace hexa
