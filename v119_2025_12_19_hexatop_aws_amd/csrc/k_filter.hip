// Helmholtz (PDE) density filter kernels.
//
// The reference solves (R^2 K + M) x = b with b_n = sum_{e ∋ n} rho_e V/8 and
// returns the element means of x (R/src/Optimization/GPUHelmholtz.jl:59-123),
// applying the 8x8 element matrix by element with eight colour-sorted
// launches through a host-built permutation (:46-53,156-173,230-238).  On the
// structured grid the operator is applied node by node instead (a 27-point
// scalar stencil with the element matrix in scalar registers): one launch,
// coalesced, no colouring, no atomics.
#include <type_traits>

#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ constexpr int FCX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int FCY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int FCZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ constexpr int fcorner(int cx, int cy, int cz) { return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0)); }

// y = A x, A = sum_e Ae over the elements of the box (scalar DOF per node)
// plus reg I.  Ae holds the 8x8 element matrix (64 values), the 27-point
// stencil of an interior node (27 values, offsets (dx,dy,dz) at dx+1 +
// 3(dy+1) + 9(dz+1)) and reg (1 value; the reference's CPU filter adds
// 1e-9 I, TopOpt.jl:75, its GPU filter nothing): A has the same coefficients wherever all 8 elements around a node
// exist, so interior nodes take 27 FMAs; nodes on the box faces sum their
// existing elements' rows.  A block of 64 x HB threads owns a 64 x HB tile of
// node columns and marches a chunk of planes along z.  Each plane's tile with
// its one-node halo (66 x (HB+2) values) is loaded once, cooperatively and
// coalesced, into LDS as it is needed (loaded four planes ahead); a thread reads its
// 3 x 3 neighbourhood of the plane from LDS and keeps those of planes k-1,
// k, k+1 in registers (1.5 global loads per node and plane instead of 9).
// With PAP (the fused filter PCG), a block also writes its partial sum of
// x.y over the nodes it computes to pap[block] and block 0 advances the
// iteration count; a halted state skips the launch.
// Node rows per block (one wave each).  Four-wave blocks (four blocks per
// CU) beat eight: at 512x256x256 0.248 against 0.287 ms per apply, at 144^3
// 0.053 against 0.059 ms (one and two waves per block were slower: their
// tiles carry more halo rows per owned row).
constexpr int HB = 4;
constexpr int HT = 64 * HB;                 // threads per block
constexpr int HW = 66, HR = HB + 2;         // staged tile: 66 x (HB+2) nodes
constexpr int HTILE = HW * HR;
constexpr int NQ = (HTILE + HT - 1) / HT;   // tile values staged per thread

template <bool PAP>
__global__ __launch_bounds__(HT) void k_helm_apply(Grid3 g, const double* __restrict__ Ae,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   int kchunk, PcgState* st, double* __restrict__ pap) {
    if (PAP && st->halted) return;
    __shared__ double tile[2][HTILE];
    const int NX = g.nx + 1, NY = g.ny + 1, NZ = g.nz + 1;
    const int tx = threadIdx.x, ty = threadIdx.y, t = tx + 64 * ty;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * HB;
    const int iu = i0 + tx, ju = j0 + ty;
    const int k0 = blockIdx.z * kchunk, k1 = min(k0 + kchunk, NZ);
    // threads past the box edge compute a clamped node and store nothing
    // (no early return: the block's barriers and the PAP reduction need
    // every thread)
    const bool valid = iu < NX && ju < NY;
    const int i = min(iu, NX - 1), j = min(ju, NY - 1);
    const double* S = Ae + 64;
    const double reg = Ae[64 + 27];
    double xy = 0.0;
    const bool inxy = i >= 1 && i < g.nx && j >= 1 && j < g.ny;
    // loads through a buffer descriptor of one plane: a node outside the box
    // is an out-of-range offset and reads 0, so no load carries a per-lane
    // condition (a conditional load compiles to a branch and a wait of its
    // own, which serialises the round trips)
    auto plane_rsrc = [&](int kk) {
        const bool kin = (unsigned)kk < (unsigned)NZ;
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(x + (kin ? (int64_t)NX * NY * kk : 0)),
                                                 (short)0, kin ? NX * NY * 8 : 0, 0x00020000);
    };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, int ii, int jj) {
        const bool in = (unsigned)ii < (unsigned)NX && (unsigned)jj < (unsigned)NY;
        const uint32_t off = in ? (uint32_t)(ii + NX * jj) * 8u : 0x80000000u;
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    };
    // this thread's 3 x 3 neighbourhood of plane kk, straight from memory
    // (the chunk's first three planes)
    auto load9 = [&](int kk, double (&p)[9]) {
        const auto r = plane_rsrc(kk);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) p[dx + 3 * dy] = ld(r, i + dx - 1, j + dy - 1);
    };
    // the tile of plane kk (items t, t + HT, ... of the 66 x (HB+2) tile)
    auto load_tile = [&](int kk, double (&v)[NQ]) {
        const auto r = plane_rsrc(kk);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + q * HT;
            const int ry = e / HW, rx = e - ry * HW;
            v[q] = e < HTILE ? ld(r, i0 - 1 + rx, j0 - 1 + ry) : 0.0;
        }
    };
    auto store_tile = [&](int slot, const double (&v)[NQ]) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (t + q * HT < HTILE) tile[slot][t + q * HT] = v[q];
    };
    auto read9 = [&](int slot, double (&p)[9]) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) p[dx + 3 * dy] = tile[slot][(ty + dy) * HW + tx + dx];
    };
    // Step at plane k (position P of the four-step trip, r = k - k0): a, b, c
    // hold planes k-1, k, k+1.  The tiles of planes k+2 .. k+5 are in flight
    // in gt[(m - k0) % 4] (four planes per block, so that a CU holds enough
    // loads in flight to cover the memory latency: one plane ahead left the
    // stencil at 1.8 TB/s).  The step stores plane k+2's tile to LDS slot
    // P % 2, refills its registers with plane k+6, computes plane k, and after
    // the barrier reads plane k+2's neighbourhood into d.  Four steps rotate
    // the four neighbourhood buffers back to their roles, so no register is
    // copied across the loop.  One barrier per step suffices: a slot is
    // rewritten two steps after it was read, and every thread has passed the
    // barrier in between.
    double gt[4][NQ];
    auto step = [&](auto pos, const double (&a)[9], const double (&b)[9], const double (&c)[9], double (&d)[9],
                    int k) {
        constexpr int P = decltype(pos)::value, G = (P + 2) % 4;
        store_tile(P % 2, gt[G]);
        load_tile(k + 6, gt[G]);
        const double* w[3] = {a, b, c};
        double acc = 0.0;
        if (inxy && k >= 1 && k < g.nz) {
#pragma unroll
            for (int q = 0; q < 27; ++q) acc += S[q] * w[q / 9][q % 9];
        } else {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
                const bool ok = (unsigned)(i + ex) < (unsigned)g.nx && (unsigned)(j + ey) < (unsigned)g.ny &&
                                (unsigned)(k + ez) < (unsigned)g.nz;
                const int l = fcorner(-ex, -ey, -ez);
                double r = 0.0;
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    r += Ae[l * 8 + m] * w[ez + FCZ[m] + 1][(ex + FCX[m] + 1) + 3 * (ey + FCY[m] + 1)];
                acc += ok ? r : 0.0;
            }
        }
        acc = __builtin_fma(reg, b[4], acc);
        const __amdgpu_buffer_rsrc_t ry = __builtin_amdgcn_make_buffer_rsrc(y + (int64_t)NX * NY * k, (short)0,
                                                                            NX * NY * 8, 0x00020000);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, acc), ry,
                                              valid ? (uint32_t)(i + NX * j) * 8u : 0x80000000u, 0, 0);
        if (PAP && valid) xy += b[4] * acc;
        __syncthreads();
        read9(P % 2, d);
    };
    double wa[9], wb[9], wc[9], wd[9];
    load9(k0 - 1, wa);
    load9(k0, wb);
    load9(k0 + 1, wc);
#pragma unroll
    for (int m = 2; m < 6; ++m) load_tile(k0 + m, gt[m % 4]);
    // (a chunk shorter than the four-step trip stops after its last plane:
    // small grids run chunks of one or two planes to fill the GPU)
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    using S3 = std::integral_constant<int, 3>;
    for (int k = k0; k < k1; k += 4) {
        step(S0{}, wa, wb, wc, wd, k);
        if (k + 1 >= k1) break;
        step(S1{}, wb, wc, wd, wa, k + 1);
        if (k + 2 >= k1) break;
        step(S2{}, wc, wd, wa, wb, k + 2);
        if (k + 3 >= k1) break;
        step(S3{}, wd, wa, wb, wc, k + 3);
    }
    if (PAP) {
        __shared__ double lds[HB];
        for (int off = 32; off > 0; off >>= 1) xy += __shfl_xor(xy, off, 64);
        if (tx == 0) lds[ty] = xy;
        __syncthreads();
        const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        if (t == 0) {
            double sum = 0.0;
            for (int w = 0; w < HB; ++w) sum += lds[w];
            pap[blk] = sum;
            if (blk == 0) st->k += 1;
        }
    }
}

// The fused filter PCG (Jacobi, residual checked every iteration) in three
// launches per iteration instead of six: the stencil with the p.Ap partials
// (above), then
//  * k_fupdate: every block sums the p.Ap partials (same order in every
//    block), alpha = rz/pAp, x += alpha p, r -= alpha Ap, partials of r.r
//    and r.(D^-1 r);
//  * k_fcheck: every block sums those, rel = |r|/|b|; unless rel < tol,
//    beta = rz_new/rz and p = D^-1 r + beta p; block 0 records the state.
// The iteration count k is advanced by the stencil's block 0, so both
// kernels read it unchanged; rz alternates between two slots by k's parity,
// so block 0 can store the new value while the other blocks read the old.
constexpr int FB = 256;   // blocks of the fused vector kernels
__device__ __forceinline__ double bsum_bcast(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return t;
}
__device__ __forceinline__ double rz_old(const PcgState* st, int k) { return k == 1 ? st->rz : st->rzb[(k - 1) & 1]; }

__global__ __launch_bounds__(256) void k_fupdate(int64_t n, PcgState* st, const double* __restrict__ pap, int npap,
                                                 double* __restrict__ x, const double* __restrict__ p,
                                                 double* __restrict__ r, const double* __restrict__ Ap,
                                                 const double* __restrict__ dinv, double* __restrict__ part) {
    if (st->halted) return;
    __shared__ double lds[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < npap; i += 256) s += pap[i];
    const double pAp = bsum_bcast(s, lds);
    const double a = rz_old(st, st->k) / pAp;
    double rr = 0.0, rz = 0.0;
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)FB * 256) {
        x[i] += a * p[i];
        const double ri = r[i] - a * Ap[i];
        r[i] = ri;
        rr += ri * ri;
        rz += ri * (ri * dinv[i]);
    }
    rr = bsum_bcast(rr, lds);
    rz = bsum_bcast(rz, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[FB + blockIdx.x] = rz;
        if (blockIdx.x == 0) {
            st->pAp = pAp;
            st->alpha = a;
        }
    }
}

__global__ __launch_bounds__(256) void k_fcheck(int64_t n, PcgState* st, const double* __restrict__ part,
                                                double* __restrict__ p, const double* __restrict__ r,
                                                const double* __restrict__ dinv) {
    if (st->halted) return;
    __shared__ double lds[4];
    const double rr = bsum_bcast(part[threadIdx.x], lds);
    const double rzn = bsum_bcast(part[FB + threadIdx.x], lds);
    const int k = st->k;
    const double rel = sqrt(rr) / st->bnorm;
    const bool conv = rel < st->tol;
    const double beta = rzn / rz_old(st, k);
    if (!conv)
        for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)FB * 256)
            p[i] = dinv[i] * r[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rr = rr;
        st->last_k = k;
        st->rel = rel;
        st->final_rel = rel;
        if (conv) {
            st->halted = PCG_CONVERGED;
        } else {
            st->beta = beta;
            st->rzb[k & 1] = rzn;
        }
    }
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
        if (dinv) dinv[n] = 1.0 / (d + Ae[64 + 27]);
    }
}

// Copies the elements of the box [o, o + sub) of grid g into the sub-grid's
// element array (in), or the sub-grid elements of the box [co, co + csz)
// (sub-grid coordinates) back to their place in g, shifted by o (out).
__global__ void k_box_gather(Grid3 g, Grid3 sub, int ox, int oy, int oz, const double* __restrict__ src,
                             double* __restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < sub.ne;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % sub.nx), j = (int)((e / sub.nx) % sub.ny), k = (int)(e / ((int64_t)sub.nx * sub.ny));
        dst[e] = src[(ox + i) + (int64_t)g.nx * ((oy + j) + (int64_t)g.ny * (oz + k))];
    }
}
__global__ void k_box_scatter(Grid3 g, Grid3 sub, int ox, int oy, int oz, int cx, int cy, int cz, int sx, int sy,
                              int sz, const double* __restrict__ src, double* __restrict__ dst) {
    const int64_t n = (int64_t)sx * sy * sz;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = cx + (int)(e % sx), j = cy + (int)((e / sx) % sy), k = cz + (int)(e / ((int64_t)sx * sy));
        dst[(ox + i) + (int64_t)g.nx * ((oy + j) + (int64_t)g.ny * (oz + k))] =
            src[i + (int64_t)sub.nx * (j + (int64_t)sub.ny * k)];
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

// planes per thread: 16, fewer when that leaves under 2048 blocks, more when
// it leaves over FUSED_PAP_MAX (the fused PCG's partials)
void helm_shape(const Grid3& g, dim3* grid, int* kchunk) {
    const unsigned gx = (unsigned)((g.nx + 1 + 63) / 64), gy = (unsigned)((g.ny + 1 + HB - 1) / HB);
    const int NZ = g.nz + 1;
    int kc = 16;
    while (kc > 1 && (int64_t)gx * gy * ((NZ + kc - 1) / kc) < 2048) kc >>= 1;
    while ((int64_t)gx * gy * ((NZ + kc - 1) / kc) > FUSED_PAP_MAX && kc < NZ) kc <<= 1;
    *grid = dim3(gx, gy, (unsigned)((NZ + kc - 1) / kc));
    *kchunk = kc;
}

void launch_helm_apply(hipStream_t s, const Grid3& g, const double* Ae, const double* x, double* y) {
    dim3 grid;
    int kchunk;
    helm_shape(g, &grid, &kchunk);
    k_helm_apply<false><<<grid, dim3(64, HB, 1), 0, s>>>(g, Ae, x, y, kchunk, nullptr, nullptr);
}

int launch_filter_fused_iteration(hipStream_t s, const Grid3& g, const double* Ae, PcgState* st, double* x,
                                  double* p, double* r, double* Ap, const double* dinv, double* partials) {
    dim3 grid;
    int kchunk;
    helm_shape(g, &grid, &kchunk);
    const int npap = (int)(grid.x * grid.y * grid.z);
    if (npap > FUSED_PAP_MAX) return -1;
    double* part2 = partials + FUSED_PAP_MAX;
    k_helm_apply<true><<<grid, dim3(64, HB, 1), 0, s>>>(g, Ae, p, Ap, kchunk, st, partials);
    k_fupdate<<<FB, 256, 0, s>>>(g.nn, st, partials, npap, x, p, r, Ap, dinv, part2);
    k_fcheck<<<FB, 256, 0, s>>>(g.nn, st, part2, p, r, dinv);
    return 0;
}

void helm_stencil(const double* Ae, double* S) {
    for (int q = 0; q < 27; ++q) S[q] = 0.0;
    // element (ex,ey,ez) in {-1,0}^3 around the node; the node is its corner
    // (-ex,-ey,-ez), the neighbour at offset (dx,dy,dz) its corner (dx-ex, ...)
    for (int s = 0; s < 8; ++s) {
        const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
        const int l = 4 * (-ez) + ((-ey) ? ((-ex) ? 2 : 3) : ((-ex) ? 1 : 0));
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = dx - ex, cy = dy - ey, cz = dz - ez;
                    if (cx < 0 || cx > 1 || cy < 0 || cy > 1 || cz < 0 || cz > 1) continue;
                    const int m = 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0));
                    S[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] += Ae[l * 8 + m];
                }
    }
}
void launch_helm_rhs_diag(hipStream_t s, const Grid3& g, const double* Ae, const double* rho, double w, double* b,
                          double* dinv) {
    k_helm_rhs_diag<<<grid_for(g.nn), 256, 0, s>>>(g, Ae, rho, w, b, dinv);
}
void launch_helm_extract(hipStream_t s, const Grid3& g, const double* x, double scale, double* out) {
    k_helm_extract<<<grid_for(g.ne), 256, 0, s>>>(g, x, scale, out);
}
void launch_box_gather(hipStream_t s, const Grid3& g, const Grid3& sub, int ox, int oy, int oz, const double* src,
                       double* dst) {
    k_box_gather<<<grid_for(sub.ne), 256, 0, s>>>(g, sub, ox, oy, oz, src, dst);
}
void launch_box_scatter(hipStream_t s, const Grid3& g, const Grid3& sub, int ox, int oy, int oz, int cx, int cy,
                        int cz, int sx, int sy, int sz, const double* src, double* dst) {
    k_box_scatter<<<grid_for((int64_t)sx * sy * sz), 256, 0, s>>>(g, sub, ox, oy, oz, cx, cy, cz, sx, sy, sz, src,
                                                                   dst);
}

}  // namespace hexa
