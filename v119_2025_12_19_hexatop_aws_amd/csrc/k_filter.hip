// Helmholtz (PDE) density filter kernels.
//
// The reference solves (R^2 K + M) x = b with b_n = sum_{e ∋ n} rho_e V/8 and
// returns the element means of x (R/src/Optimization/GPUHelmholtz.jl:59-123),
// applying the 8x8 element matrix by element with eight colour-sorted
// launches through a host-built permutation (:46-53,156-173,230-238).  On the
// structured grid the operator is applied node by node instead (a 27-point
// scalar stencil with the element matrix in scalar registers): one launch,
// coalesced, no colouring, no atomics.
#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

__device__ constexpr int FCX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int FCY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int FCZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ constexpr int fcorner(int cx, int cy, int cz) { return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0)); }

// y = A x, A = sum_e Ae over the elements of the box (scalar DOF per node).
__global__ __launch_bounds__(256) void k_helm_apply(Grid3 g, const double* __restrict__ Ae,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    const int NX = g.nx + 1, NY = g.ny + 1, NZ = g.nz + 1;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= NX || j >= NY) return;
    double xn[27];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int ii = i + dx - 1, jj = j + dy - 1, kk = k + dz - 1;
                const bool in = (unsigned)ii < (unsigned)NX && (unsigned)jj < (unsigned)NY && (unsigned)kk < (unsigned)NZ;
                xn[dx + 3 * dy + 9 * dz] = in ? x[ii + (int64_t)NX * (jj + (int64_t)NY * kk)] : 0.0;
            }
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
        const bool ok = (unsigned)(i + ex) < (unsigned)g.nx && (unsigned)(j + ey) < (unsigned)g.ny &&
                        (unsigned)(k + ez) < (unsigned)g.nz;
        const int l = fcorner(-ex, -ey, -ez);
        double r = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m)
            r += Ae[l * 8 + m] * xn[(ex + FCX[m] + 1) + 3 * (ey + FCY[m] + 1) + 9 * (ez + FCZ[m] + 1)];
        acc += ok ? r : 0.0;
    }
    y[i + (int64_t)NX * (j + (int64_t)NY * k)] = acc;
}

// b_n = w * sum of rho over the elements around n; diag_n = sum of Ae[l][l].
__global__ void k_helm_rhs_diag(Grid3 g, const double* __restrict__ Ae, const double* __restrict__ rho, double w,
                                double* __restrict__ b, double* __restrict__ dinv) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.nn;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(n % NX), j = (int)((n / NX) % NY), k = (int)(n / ((int64_t)NX * NY));
        double s = 0.0, d = 0.0;
        for (int q = 0; q < 8; ++q) {
            const int ex = (q & 1) - 1, ey = ((q >> 1) & 1) - 1, ez = ((q >> 2) & 1) - 1;
            const int ei = i + ex, ej = j + ey, ek = k + ez;
            if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny || (unsigned)ek >= (unsigned)g.nz)
                continue;
            const int l = fcorner(-ex, -ey, -ez);
            if (rho) s += rho[ei + (int64_t)g.nx * (ej + (int64_t)g.ny * ek)];
            d += Ae[l * 8 + l];
        }
        if (b) b[n] = s * w;
        if (dinv) dinv[n] = 1.0 / d;
    }
}

// rho~_e = mean of the 8 nodal values (GPUHelmholtz.jl:99-110); with
// scale != 1 the element sum is multiplied instead (F^T uses V/8).
__global__ void k_helm_extract(Grid3 g, const double* __restrict__ x, double scale, double* __restrict__ out) {
    const int NX = g.nx + 1, NY = g.ny + 1;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < g.ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % g.nx), j = (int)((e / g.nx) % g.ny), k = (int)(e / ((int64_t)g.nx * g.ny));
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c) s += x[(i + FCX[c]) + (int64_t)NX * ((j + FCY[c]) + (int64_t)NY * (k + FCZ[c]))];
        out[e] = s * scale;
    }
}

}  // namespace

void launch_helm_apply(hipStream_t s, const Grid3& g, const double* Ae, const double* x, double* y) {
    dim3 block(64, 4, 1);
    dim3 grid((unsigned)((g.nx + 1 + 63) / 64), (unsigned)((g.ny + 1 + 3) / 4), (unsigned)(g.nz + 1));
    k_helm_apply<<<grid, block, 0, s>>>(g, Ae, x, y);
}
void launch_helm_rhs_diag(hipStream_t s, const Grid3& g, const double* Ae, const double* rho, double w, double* b,
                          double* dinv) {
    k_helm_rhs_diag<<<grid_for(g.nn), 256, 0, s>>>(g, Ae, rho, w, b, dinv);
}
void launch_helm_extract(hipStream_t s, const Grid3& g, const double* x, double scale, double* out) {
    k_helm_extract<<<grid_for(g.ne), 256, 0, s>>>(g, x, scale, out);
}

}  // namespace hexa
