// Dense-Ke elasticity operator on the matrix cores (FP64 MFMA, and FP32 MFMA
// for the FP32 levels of a mixed V-cycle), for element matrices that are not
// exactly factorisable (the reference's f32-widened Ke,
// R/src/Element/Element.jl get_canonical_stiffness in Float32 widened to
// Float64 by the GPU solver's Ke upload).
//
// A block owns a 15 x 15 column of nodes and marches a chunk of node planes
// along z.  For each element layer it takes the 16 x 16 elements around its
// nodes as 16 groups of 16 (one element row each).  A group's element
// outputs Y = Ke U (24 x 16, Ke padded to 32 rows) are 2 x 6
// v_mfma_f64_16x16x4 steps: Ke is the A operand, held in VGPRs for the whole
// launch; U, the group's 16 element DOF vectors, is gathered as the B
// operand from the node planes staged in LDS.  Y, scaled by the element
// factors, goes to LDS, and each owned node's DOF sums its rows of the 8
// elements around the node from there (all threads share the 15 x 15 x 3
// DOFs): a node-centric gather, so no atomics and a fixed summation order.
// Node plane k is complete after element layer k; the part its 4 elements
// below contribute (layer k-1) is carried in registers.
//
// Per element the product costs 32 x 24 FMA on the matrix cores; at
// 16 x 16 elements per 15 x 15 nodes that bounds the launch at about
// 0.84 ms for 33.5 M elements (78.6 TFLOP/s FP64 matrix peak).
#include "hexa_internal.h"

namespace hexa {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step per precision, and the row of the 16 x 16
// result that accumulator register r of lane group lk = lane / 16 holds
// (column lane % 16 in both).  The two precisions differ: checked on gfx950
// with a one-wave probe of each builtin.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef d4 acc;
    __device__ static acc step(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef f4 acc;
    __device__ static acc step(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int lk, int r) { return 4 * lk + r; }
};

template <typename T, typename DT>
__device__ __forceinline__ T to_value(DT v) { return (T)v; }
template <>
__device__ __forceinline__ float to_value<float, uint16_t>(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

// Loads through buffer descriptors of one plane: an offset past the
// descriptor's size (OOB), or a descriptor of size 0, reads 0, so the
// kernel's loads carry no per-lane condition (a guarded load compiles to a
// branch and a wait of its own).
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t OOB = 0x80000000u;
__device__ __forceinline__ Rsrc plane_rsrc(const void* base, int64_t plane_bytes, int k, int nk) {
    const bool in = (unsigned)k < (unsigned)nk;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(static_cast<const char*>(base) + (in ? plane_bytes * k : 0)),
                                             (short)0, in ? (int)plane_bytes : 0, 0x00020000);
}
__device__ __forceinline__ double bload(Rsrc r, uint32_t off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float bload(Rsrc r, uint32_t off, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ uint16_t bload(Rsrc r, uint32_t off, uint16_t) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}

constexpr int MX = 16;              // elements per group (MFMA N) and per tile row
constexpr int MY = 16;              // groups (element rows) per tile
constexpr int OWN = 15;             // owned nodes per tile axis
constexpr int PN = MX + 1;          // staged nodes per plane row
constexpr int PROW = 3 * PN;        // values per staged plane row
constexpr int PLANE = PROW * (MY + 1);   // values per staged node plane
constexpr int NE = MX * MY;         // elements per tile layer
constexpr int NT = 512;             // threads: 8 waves, 2 groups each
constexpr int NOWN = OWN * OWN;     // owned nodes per plane
constexpr int NES = NE + 16;        // Y row stride in values, padded so the 4 rows a
                                    // write instruction spans fall in different LDS banks

__device__ constexpr int MCX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int MCY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int MCZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ constexpr int corner_at(int cx, int cy, int cz) { return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0)); }

template <typename T, typename DT, int MODE>
__global__ __launch_bounds__(NT) void k_elastic_mfma(Grid3 g, const T* __restrict__ Ke, const T* __restrict__ fac,
                                                     const uint8_t* __restrict__ mask, const T* __restrict__ u,
                                                     const T* __restrict__ b, const DT* __restrict__ dinv, T omega,
                                                     T shift, T* __restrict__ y, int kchunk, int tiles_x,
                                                     int tiles_y) {
    typedef Mfma<T> M;
    __shared__ T planes[3 * PLANE];
    __shared__ T Y[24 * NES];   // Y[m * NES + e], e = ey * MX + ex

    const int NX = g.nx + 1, NY = g.ny + 1, NZ = g.nz + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = blockIdx.x;
    const int tx = bid % tiles_x, ty = (bid / tiles_x) % tiles_y, tz = bid / (tiles_x * tiles_y);
    const int i0 = tx * OWN - 1, j0 = ty * OWN - 1;   // global index of local node / element 0
    const int k0 = tz * kchunk, k1 = min(k0 + kchunk, NZ);

    // A operand: lane holds Ke[m][d], m = lane % 16 + 16 t, d = 4 s + lane / 16
    const int lj = lane & 15, lk = lane >> 4;
    T a[2][6];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int m = lj + 16 * t;
            a[t][s] = m < 24 ? Ke[m * 24 + 4 * s + lk] : T(0);
        }
    // B operand: lane holds U[d][e] of element ex = lane % 16, d = 4 s + lane / 16,
    // i.e. component d % 3 of corner d / 3
    int boff[6], bhi[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int d = 4 * s + lk, c = d / 3, comp = d - 3 * c;
        boff[s] = MCY[c] * PROW + (lj + MCX[c]) * 3 + comp;
        bhi[s] = MCZ[c];
    }

    // staging of node plane kp into LDS slot: items tid and tid + NT
    const int64_t nplane = (int64_t)NX * NY;   // nodes per plane
    auto stage_load = [&](int kp, T (&v)[2]) {
        const Rsrc r = plane_rsrc(u, nplane * 3 * (int64_t)sizeof(T), kp, NZ);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int it = tid + q * NT;
            const int rr = it / PROW, w = it - rr * PROW;
            const int nx = w / 3, comp = w - 3 * nx;
            const int i = i0 + nx, j = j0 + rr;
            const bool in = it < PLANE && (unsigned)i < (unsigned)NX && (unsigned)j < (unsigned)NY;
            v[q] = bload(r, in ? (uint32_t)((i + NX * j) * 3 + comp) * (uint32_t)sizeof(T) : OOB, T(0));
        }
    };
    auto stage_store = [&](int slot, const T (&v)[2]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int it = tid + q * NT;
            if (it < PLANE) planes[slot * PLANE + it] = v[q];
        }
    };

    // gather items: (owned node o, component c), item = NOWN c + o, items tid
    // and tid + NT.  Component-major, so that neighbouring lanes read
    // neighbouring elements of one Y row: with the node-major order three
    // lanes in a row read one element of three rows NES apart, which share
    // an LDS bank.
    int gq_n[2], gq_c[2];
    bool gq_in[2];
    int64_t gq_base[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int itm = tid + q * NT;
        const int c = itm / NOWN;
        const int o = itm - NOWN * c;
        gq_c[q] = c;
        gq_n[q] = o;
        const int gx = 1 + o % OWN, gy = 1 + o / OWN;
        gq_in[q] = itm < 3 * NOWN && i0 + gx < NX && j0 + gy < NY;
        gq_base[q] = (i0 + gx) + (int64_t)NX * (j0 + gy);
    }
    T carry[2] = {T(0), T(0)};

    {
        T v[2];
        stage_load(k0 - 1, v);
        stage_store(0, v);
        stage_load(k0, v);
        stage_store(1, v);
    }
    __syncthreads();

    int slot_lo = 0;   // LDS slot of node plane ek; ek + 1 and ek + 2 follow cyclically
    for (int ek = k0 - 1; ek < k1; ++ek) {
        const int slot_hi = slot_lo == 2 ? 0 : slot_lo + 1;
        const int slot_nx = slot_hi == 2 ? 0 : slot_hi + 1;
        T nxt[2];
        stage_load(ek + 2, nxt);
        // element factors of this wave's two groups (0 outside the grid)
        T f[2];
        const Rsrc rf = plane_rsrc(fac, (int64_t)g.nx * g.ny * (int64_t)sizeof(T), ek, g.nz);
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) {
            const int ei = i0 + lj, ej = j0 + 2 * wave + gq;
            const bool in = (unsigned)ei < (unsigned)g.nx && (unsigned)ej < (unsigned)g.ny;
            f[gq] = bload(rf, in ? (uint32_t)(ei + g.nx * ej) * (uint32_t)sizeof(T) : OOB, T(0));
        }
        // flag, right-hand side and inverse diagonal of the items of output plane ek
        uint8_t mb[2] = {0, 0};
        T bv[2] = {T(0), T(0)}, dv[2] = {T(0), T(0)};
        int64_t gdof[2];
        const int ko = ek >= k0 ? ek : -1;   // no output plane in the chunk's first layer
        const Rsrc rm = plane_rsrc(mask, mask ? nplane : 0, ko, NZ);
        const Rsrc rb = plane_rsrc(b, MODE != OP_APPLY ? nplane * 3 * (int64_t)sizeof(T) : 0, ko, NZ);
        const Rsrc rd = plane_rsrc(dinv, MODE == OP_SMOOTH ? nplane * 3 * (int64_t)sizeof(DT) : 0, ko, NZ);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t gn = gq_base[q] + nplane * (ko >= 0 ? ko : 0);
            gdof[q] = 3 * gn + gq_c[q];
            const uint32_t pn = (uint32_t)gq_base[q];   // node index within the plane
            mb[q] = (uint8_t)((__builtin_amdgcn_raw_buffer_load_b8(rm, gq_in[q] ? pn : OOB, 0, 0) >> gq_c[q]) & 1);
            const uint32_t dof = 3 * pn + gq_c[q];
            if (MODE != OP_APPLY) bv[q] = bload(rb, gq_in[q] ? dof * (uint32_t)sizeof(T) : OOB, T(0));
            if (MODE == OP_SMOOTH)
                dv[q] = to_value<T, DT>(bload(rd, gq_in[q] ? dof * (uint32_t)sizeof(DT) : OOB, DT(0)));
        }

        // Y = Ke U for the wave's two groups
        const int lo = slot_lo * PLANE, hi = slot_hi * PLANE;
        typename M::acc acc[2][2];
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) acc[gq][0] = acc[gq][1] = typename M::acc{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int base = (bhi[s] ? hi : lo) + 2 * wave * PROW + boff[s];
            const T b0 = planes[base], b1 = planes[base + PROW];
            acc[0][0] = M::step(a[0][s], b0, acc[0][0]);
            acc[1][0] = M::step(a[0][s], b1, acc[1][0]);
            acc[0][1] = M::step(a[1][s], b0, acc[0][1]);
            acc[1][1] = M::step(a[1][s], b1, acc[1][1]);
        }
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) {
            const int e = (2 * wave + gq) * MX + lj;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = M::row(lk, r);
                Y[m * NES + e] = f[gq] * acc[gq][0][r];
                if (m < 8) Y[(16 + m) * NES + e] = f[gq] * acc[gq][1][r];
            }
        }
        stage_store(slot_nx, nxt);
        __syncthreads();

#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (tid + q * NT < 3 * NOWN) {
                const int o = gq_n[q], c = gq_c[q];
                const int gx = 1 + o % OWN, gy = 1 + o / OWN;
                T bot = T(0), top = T(0);
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int e = (gy - 1 + dy) * MX + (gx - 1 + dx);
                        const int cb = corner_at(1 - dx, 1 - dy, 0), ct = cb + 4;
                        bot += Y[(3 * cb + c) * NES + e];
                        top += Y[(3 * ct + c) * NES + e];
                    }
                if (gq_in[q] && ek >= k0) {
                    const T uc = planes[lo + gy * PROW + 3 * gx + c];
                    const T Au = carry[q] + bot + shift * uc;
                    T v;
                    if (MODE == OP_APPLY) v = Au;
                    else if (MODE == OP_RESID) v = bv[q] - Au;
                    else v = uc + omega * dv[q] * (bv[q] - Au);
                    y[gdof[q]] = mb[q] ? T(0) : v;
                }
                carry[q] = top;
            }
        }
        __syncthreads();
        slot_lo = slot_hi;
    }
}

}  // namespace

bool elastic_mfma_supported(const Grid3& g) {
    return !fallback_kernels_forced() && (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) < (int64_t)1 << 31;
}

template <typename T, typename DT>
void launch_elastic_mfma(hipStream_t s, const Grid3& g, OpMode mode, const T* Ke, const T* fac, const uint8_t* mask,
                         const T* u, const T* b, const DT* dinv, T omega, T shift, T* y) {
    const int tiles_x = (g.nx + 1 + OWN - 1) / OWN, tiles_y = (g.ny + 1 + OWN - 1) / OWN;
    const int NZ = g.nz + 1;
    int kchunk = 32;
    while (kchunk > 4 && (int64_t)tiles_x * tiles_y * ((NZ + kchunk - 1) / kchunk) < 2048) kchunk >>= 1;
    const int tiles = tiles_x * tiles_y * ((NZ + kchunk - 1) / kchunk);
    switch (mode) {
        case OP_APPLY:
            k_elastic_mfma<T, DT, OP_APPLY><<<tiles, NT, 0, s>>>(g, Ke, fac, mask, u, b, dinv, omega, shift, y,
                                                                 kchunk, tiles_x, tiles_y);
            break;
        case OP_RESID:
            k_elastic_mfma<T, DT, OP_RESID><<<tiles, NT, 0, s>>>(g, Ke, fac, mask, u, b, dinv, omega, shift, y,
                                                                 kchunk, tiles_x, tiles_y);
            break;
        default:
            k_elastic_mfma<T, DT, OP_SMOOTH><<<tiles, NT, 0, s>>>(g, Ke, fac, mask, u, b, dinv, omega, shift, y,
                                                                  kchunk, tiles_x, tiles_y);
            break;
    }
}

#define HX_INST_MFMA(T, DT)                                                                                     \
    template void launch_elastic_mfma<T, DT>(hipStream_t, const Grid3&, OpMode, const T*, const T*, const uint8_t*, \
                                             const T*, const T*, const DT*, T, T, T*);
HX_INST_MFMA(double, double)
HX_INST_MFMA(float, float)
HX_INST_MFMA(float, uint16_t)

}  // namespace hexa
