// Row-wave element operator for the exact brick stiffness (gfx950).
//
// Same factored element operator as k_elastic_fact.hip (per-axis sum /
// difference basis: butterfly, modal diagonal + cross-component couplings,
// transposed butterfly), mapped onto the machine differently:
//
//  * x runs along the lanes: a row of LX = 64 (one row per wavefront) or 32
//    (two rows per wavefront, one per lane half) lanes; a lane owns one node
//    column and the element column to its +x side; neighbouring node values
//    and the +x corner results move between lanes with DPP wave shifts, not
//    LDS (at a half boundary the shifted-in value belongs to the other row,
//    and only halo lanes receive it).  The first and last lane of a row are
//    halo: LX - 2 owned columns per LX loaded.
//  * y runs across the NR = 8 RPW rows of a block (8 waves): row r holds
//    element row r and node row r (NR+1 node rows per tile, rows 0 and NR
//    are halo).  The +y node row and the +y corner results are exchanged
//    through LDS (3 values each per node and layer); the top node row is
//    loaded by rows 0-2, one component each.  Both exchanges of a layer
//    share one barrier: the node row staged is the one two planes ahead, and
//    the buffers alternate with the layer parity.
//  * z is marched, with the node planes and element factors of the next three
//    layers in flight (two in the first smoothing sweep).  The butterfly is separable, so the x-y part of a
//    node plane's transform is computed once and used by the element layers
//    below and above it, and the two layers' x-y modal results for a plane
//    are summed before one inverse x-y butterfly: per element about 175
//    flops instead of 210.
//  * memory is read and written through buffer descriptors (see below).
//  * tiles are numbered so that the blocks resident at one time work on one
//    contiguous run of tiles (x fastest, then y, then z-chunk), marching
//    through the same node planes together, and each XCD (workgroups go to
//    the 8 XCDs round-robin) holds a contiguous part of that run, whose
//    shared halo rows then come from its own L2.
//
// Every node sums its contributions in a fixed order: no atomics and the same
// result on every run, whatever the z-chunk length.
#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>

#include "hexa_internal.h"

namespace hexa {
namespace {

#ifndef HEXA_RW_BY
#define HEXA_RW_BY 8
#endif
constexpr int WL = 64;           // lanes per wavefront
constexpr int BY = HEXA_RW_BY;   // wavefronts per block
// Tile layouts: RPW node rows per wavefront, LX = 64 / RPW lanes (node
// columns) per row, NR = BY RPW rows per tile; the first and last column
// and the first row are halo.
template <int RPW>
struct RwTile {
    static constexpr int LX = WL / RPW;
    static constexpr int OWN_X = LX - 2;
    static constexpr int NR = BY * RPW;
    static constexpr int OWN_Y = NR - 1;
};
constexpr int NXCD = 8;
constexpr int DOT_GROUP = 4;     // output planes per (u, y) partial sum
// Cache policy of the output stores: non-temporal (the gfx950 "nt" bit),
// so the streaming output does not evict input rows a neighbouring tile may
// still read from L2 (measured: FP64 product 0.510 -> 0.505 ms, FP32
// 0.283 -> 0.281 ms at 512x256x256; the same bit on the element-factor
// loads changed nothing).
constexpr int ST_NT = 2;

// DPP wave shifts (GFX9 encodings): lane l reads lane l+1 / lane l-1; the
// lane without a source reads 0.
__device__ __forceinline__ int dpp_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true); }

__device__ __forceinline__ float from_next(float v) { return __int_as_float(dpp_next(__float_as_int(v))); }
__device__ __forceinline__ float from_prev(float v) { return __int_as_float(dpp_prev(__float_as_int(v))); }
__device__ __forceinline__ double from_next(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_next((int)(b & 0xffffffffLL)), hi = dpp_next((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double from_prev(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_prev((int)(b & 0xffffffffLL)), hi = dpp_prev((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Buffer (resource-descriptor) memory access: a lane whose offset lies past
// the descriptor's byte count reads 0 and its store is dropped, so halo lanes
// and out-of-box planes need no branches, and every layer issues the same
// memory instructions (the compiler can then keep prefetches in flight
// across the layer's barriers).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t OOB = 0x80000000u;   // > any descriptor's byte count

__device__ __forceinline__ Rsrc rsrc(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void ld3(Rsrc r, uint32_t off, double (&o)[3]) {
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16, 0, 0);
    o[0] = __builtin_bit_cast(double, u32x2{a.x, a.y});
    o[1] = __builtin_bit_cast(double, u32x2{a.z, a.w});
    o[2] = __builtin_bit_cast(double, b);
}
__device__ __forceinline__ void ld3(Rsrc r, uint32_t off, float (&o)[3]) {
    const u32x3 a = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
    o[0] = __uint_as_float(a.x);
    o[1] = __uint_as_float(a.y);
    o[2] = __uint_as_float(a.z);
}
__device__ __forceinline__ void st3(Rsrc r, uint32_t off, const double (&o)[3]) {
    const u32x2 a = __builtin_bit_cast(u32x2, o[0]), b = __builtin_bit_cast(u32x2, o[1]),
                c = __builtin_bit_cast(u32x2, o[2]);
    __builtin_amdgcn_raw_buffer_store_b128(u32x4{a.x, a.y, b.x, b.y}, r, off, 0, ST_NT);
    __builtin_amdgcn_raw_buffer_store_b64(c, r, off + 16, 0, ST_NT);
}
__device__ __forceinline__ void st3(Rsrc r, uint32_t off, const float (&o)[3]) {
    __builtin_amdgcn_raw_buffer_store_b96(u32x3{__float_as_uint(o[0]), __float_as_uint(o[1]), __float_as_uint(o[2])},
                                          r, off, 0, ST_NT);
}
__device__ __forceinline__ double ld1(Rsrc r, uint32_t off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float ld1(Rsrc r, uint32_t off, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <typename T>
__device__ __forceinline__ void ld3_dinv(Rsrc r, uint32_t off, const double*, T (&o)[3]) {
    double d[3];
    ld3(r, off, d);
    for (int c = 0; c < 3; ++c) o[c] = (T)d[c];
}
template <typename T>
__device__ __forceinline__ void ld3_dinv(Rsrc r, uint32_t off, const float*, T (&o)[3]) {
    float d[3];
    ld3(r, off, d);
    for (int c = 0; c < 3; ++c) o[c] = (T)d[c];
}
template <typename T>
__device__ __forceinline__ void ld3_dinv(Rsrc r, uint32_t off, const uint16_t*, T (&o)[3]) {
    for (int c = 0; c < 3; ++c)
        o[c] = (T)__uint_as_float(((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, off + 2 * c, 0, 0)) << 16);
}

// Fused multiply-add, written out: this file is compiled without automatic
// contraction, so every copy of the unrolled layer loop performs the same
// IEEE operations and a plane's result does not depend on which loop
// position (and so which z-chunk length) produced it.
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// A wave-uniform coefficient held in a VGPR.  The FP64 kernel's scalar
// registers cannot hold all of its coefficients; the coupling coefficients
// live in vector registers instead of being spilled and re-read per use.
__device__ __forceinline__ double in_vgpr(double x) {
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
__device__ __forceinline__ float in_vgpr(float x) { return x; }

template <typename T>
__device__ __forceinline__ T ld1_dinv(Rsrc r, uint32_t off, const double*) {
    return (T)ld1(r, off, 0.0);
}
template <typename T>
__device__ __forceinline__ T ld1_dinv(Rsrc r, uint32_t off, const float*) {
    return (T)ld1(r, off, 0.0f);
}
template <typename T>
__device__ __forceinline__ T ld1_dinv(Rsrc r, uint32_t off, const uint16_t*) {
    return (T)__uint_as_float(((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)) << 16);
}

// Occupancy: the FP64 kernels hold 232-256 VGPRs, 2 waves per SIMD (one
// 8-wave block per CU).  The FP32 kernels are held to 128 VGPRs, 4 waves per
// SIMD, so that two blocks share a CU: the first smoothing sweep needed
// 134-138 VGPRs and ran one block per CU (0.547 ms at 512x256x256, 0.449
// held to 128; its BF16-diagonal form spills 12 bytes per lane).  The FP64
// code is the same as with the default bound.
template <typename T, typename DT, int MODE, int RPW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, WL * BY),
                          amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) void k_elastic_rw(Grid3 g, FactCoef<T> coef, const T* __restrict__ fac,
                                                       const uint8_t* __restrict__ mask,
                                                       const T* __restrict__ u, const T* __restrict__ b,
                                                       const DT* __restrict__ dinv, T omega, T shift,
                                                       T* __restrict__ y, int kchunk, int tiles_x, int tiles_y,
                                                       int tiles, double* __restrict__ dot_part, int nb,
                                                       int64_t vstride, int64_t dstride,
                                                       const int32_t* __restrict__ halt, int hs) {
    constexpr int LX = RwTile<RPW>::LX, OWN_X = RwTile<RPW>::OWN_X, NR = RwTile<RPW>::NR,
                  OWN_Y = RwTile<RPW>::OWN_Y;
    // double-buffered by layer parity, so one barrier per layer suffices
    __shared__ T su[2][NR + 1][3][LX];   // node values of each row, plane kl+2
    __shared__ T sc[2][NR][3][LX];       // each row's results for the row above it

    // lane l of wave w: node column lx of tile row ro (RPW rows per wave)
    const int l = threadIdx.x, w = threadIdx.y;
    const int lx = l % LX, ro = w * RPW + l / LX;
    T cl[3][2], cm[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            cl[d][q] = in_vgpr(coef.cl[d][q]);
            cm[d][q] = in_vgpr(coef.cm[d][q]);
        }
    // Tile order.  Workgroup L runs on XCD L % 8; the chip holds WIN blocks
    // at once (8 XCDs x 32 CUs, one FP64 block or two FP32 blocks per CU).
    // Each window of WIN consecutive workgroups takes WIN consecutive tiles,
    // a contiguous share of them per XCD.  Measured at 512x256x256 against
    // one contiguous share per XCD over the whole launch (whose XCDs then
    // work in planes of different z-chunks at once): FP64 product 0.474 ->
    // 0.463 ms, FP64 smoothing sweep 0.768 -> 0.718 ms, FP32 smoothing sweep
    // 0.380 -> 0.353 ms, FP32 residual 0.339 -> 0.308 ms; workgroups in
    // plain tile order (no XCD shares) gave the same sweeps but a slower
    // product at 256x128x128 (0.066 against 0.064 ms).  With nb
    // right-hand sides (vectors vstride apart, partial sums dstride apart)
    // they follow one another in block order, all tiles of one before the
    // next: one launch instead of nb, and the tail of one right-hand side
    // overlaps the start of the next.  (Interleaving them per tile, so that
    // the blocks of a tile share its element factors through the XCD's L2,
    // was measured slower: 0.72 against 0.64 ms for 8 right-hand sides at
    // 256x128x128, against 8 x 0.084 ms in separate launches; it spreads
    // neighbouring tiles apart and loses their halo rows' L2 hits.)
    const int L = blockIdx.x;
    const int total = tiles * nb;
    constexpr int WIN = NXCD * 32 * (sizeof(T) == 4 ? 2 : 1);
    const int xcd = L % NXCD;
    const int wbase = L / WIN * WIN, wn = min(WIN, total - wbase), wl = L - wbase;
    const int wq = wn / NXCD, wr = wn % NXCD;
    const int tb = wbase + xcd * wq + min(xcd, wr) + wl / NXCD;
    const int rb = tb / tiles, t = tb % tiles;
    // a launch queued after its solve (of right-hand side rb) ended does nothing
    if (halt && halt[rb * hs]) return;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    if (rb) {
        if (u) u += rb * vstride;
        if (b) b += rb * vstride;
        y += rb * vstride;
        if (dot_part) dot_part += rb * dstride;
    }

    const int i = tx * OWN_X - 1 + lx;
    const int j = ty * OWN_Y - 1 + ro;
    const int kb = tz * kchunk;
    const int ke = min(kb + kchunk, g.nz + 1);
    const bool col_in = (unsigned)i <= (unsigned)g.nx && (unsigned)j <= (unsigned)g.ny;
    const bool elem = lx < LX - 1 && (unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny;
    const bool owner = lx >= 1 && lx <= LX - 2 && ro >= 1 && col_in;
    const int64_t plane_n = (int64_t)(g.nx + 1) * (g.ny + 1);
    const int64_t plane_e = (int64_t)g.nx * g.ny;

    // loop-invariant lane offsets (bytes); OOB for lanes that must not touch memory
    const int64_t nrow = i + (int64_t)(g.nx + 1) * j;
    const uint32_t uoff = col_in ? (uint32_t)(nrow * 3 * sizeof(T)) : OOB;
    const uint32_t foff = elem ? (uint32_t)((i + (int64_t)g.nx * j) * sizeof(T)) : OOB;
    const uint32_t ooff = owner ? (uint32_t)(nrow * 3 * sizeof(T)) : OOB;
    const uint32_t moff = owner ? (uint32_t)nrow : OOB;
    const uint32_t doff = owner ? (uint32_t)(nrow * 3 * sizeof(DT)) : OOB;
    const int64_t vplane = 3 * plane_n;   // vector entries per node plane
    // the tile's top node row (rows 0..NR-1 hold element rows 0..NR-1 and
    // node rows 1..NR-1 are owned): rows 0-2 load component ro of its nodes
    // as an extra
    const int jx = ty * OWN_Y - 1 + NR;
    const bool x_in = ro < 3 && (unsigned)i <= (unsigned)g.nx && (unsigned)jx <= (unsigned)g.ny;
    const int64_t xent = (i + (int64_t)(g.nx + 1) * jx) * 3 + ro;
    const uint32_t xoff = x_in ? (uint32_t)(xent * sizeof(T)) : OOB;
    const uint32_t xdoff = x_in ? (uint32_t)(xent * sizeof(DT)) : OOB;

    // A node plane's input.  OP_SMOOTH0 (the first smoothing sweep from
    // x = 0) never forms its starting iterate omega D^-1 b: the plane's b and
    // inverse diagonal are loaded and combined when the plane is consumed.
    struct PIn {
        T a[3], d[3], xa, xd;
    };
    const uint32_t udoff = col_in ? (uint32_t)(nrow * 3 * sizeof(DT)) : OOB;
    auto load_in = [&](int k, PIn& o) {
        const bool in = (unsigned)k <= (unsigned)g.nz;
        if (MODE == OP_SMOOTH0) {
            const Rsrc rb = rsrc(b + (in ? vplane * k : 0), in ? vplane * (int64_t)sizeof(T) : 0);
            const Rsrc rd = rsrc(dinv + (in ? vplane * k : 0), in ? vplane * (int64_t)sizeof(DT) : 0);
            ld3(rb, uoff, o.a);
            ld3_dinv(rd, udoff, dinv, o.d);
            o.xa = ld1(rb, xoff, T(0));
            o.xd = ld1_dinv<T>(rd, xdoff, dinv);
        } else {
            const Rsrc ru = rsrc(u + (in ? vplane * k : 0), in ? vplane * (int64_t)sizeof(T) : 0);
            ld3(ru, uoff, o.a);
            o.xa = ld1(ru, xoff, T(0));
        }
    };
    // own node values of a plane, and the extra top-row value
    auto own_value = [&](const PIn& p, T (&o)[3], T& x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = (MODE == OP_SMOOTH0) ? omega * p.d[c] * p.a[c] : p.a[c];
        x = (MODE == OP_SMOOTH0) ? omega * p.xd * p.xa : p.xa;
    };
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    // stage a plane's own row (and the top row) for the wave below
    auto stage = [&](auto par, const T (&o)[3], const T x) {
        constexpr int P = decltype(par)::value;
#pragma unroll
        for (int c = 0; c < 3; ++c) su[P][ro][c][lx] = o[c];
        if (ro < 3) su[P][NR][ro][lx] = x;
    };
    auto load_fac = [&](int lay) -> T {
        const bool in = (unsigned)lay < (unsigned)g.nz;
        return ld1(rsrc(fac + (in ? plane_e * lay : 0), in ? plane_e * (int64_t)sizeof(T) : 0), foff, T(0));
    };
    // x-y modes of the element column's 4 corners in one node plane, from
    // this lane's node value and the value of the node above it (row w+1):
    // P[c][mx + 2 my].
    auto plane_modes = [&](const T (&own)[3], const T (&up)[3], T (&P)[3][4]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T a = own[c], bx = from_next(own[c]), cy = up[c], d = from_next(up[c]);
            const T s0 = a + bx, d0 = bx - a, s1 = cy + d, d1 = d - cy;
            P[c][0] = s0 + s1;
            P[c][1] = d0 + d1;
            P[c][2] = s1 - s0;
            P[c][3] = d1 - d0;
        }
    };

    // One element layer kl (node planes kl, kl+1) and the output of node
    // plane kl.  Plo/Yin: x-y modes of plane kl and the layer below's modal
    // results for plane kl; Yout: the modal results for plane kl+1.
    struct Out {
        Rsrc ry;
        uint32_t mb;
        bool has_out;
        T bv[3], dv[3];
    };
    // this lane's share of (u, y) over its owned outputs, when dot_part is set
    double acc = 0.0;
    // flags, right-hand side and inverse diagonal of output plane kl (the
    // chunk's first layer, kl = kb - 1, has none); issued before the layer's
    // prefetches, so waiting for them leaves the prefetches in flight
    auto load_out = [&](const int kl, Out& o) {
        const bool has_out = kl >= kb;
        const int64_t ko = has_out ? kl : 0;
        o.has_out = has_out;
        o.ry = rsrc(y + vplane * ko, has_out ? vplane * (int64_t)sizeof(T) : 0);
        // (a null mask gives an empty descriptor: all flags read 0)
        o.mb = __builtin_amdgcn_raw_buffer_load_b8(
            rsrc(mask ? mask + plane_n * ko : mask, (mask && has_out) ? plane_n : 0), moff, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c) o.bv[c] = o.dv[c] = T(0);
        if (MODE != OP_APPLY) ld3(rsrc(b + vplane * ko, has_out ? vplane * (int64_t)sizeof(T) : 0), ooff, o.bv);
        if (MODE == OP_SMOOTH || MODE == OP_SMOOTH0)
            ld3_dinv(rsrc(dinv + vplane * ko, has_out ? vplane * (int64_t)sizeof(DT) : 0), doff, dinv, o.dv);
    };
    // Layer kl.  In: Plo, Yin (plane kl), ulo, uhi (own values of planes kl,
    // kl+1), uphi (row w+1 of plane kl+1), unx (own values of plane kl+2).
    // Out: Phi, Yout (plane kl+1), upnx (row w+1 of plane kl+2).  The +y
    // exchange of this layer's results and the staging of plane kl+2's row
    // share one barrier; buffers alternate with the layer parity PAR.
    auto step = [&](auto par, const T (&Plo)[3][4], T (&Phi)[3][4], const T (&Yin)[3][4], T (&Yout)[3][4],
                    const T (&ulo)[3], const T (&uhi)[3], const T (&uphi)[3], const T (&unx)[3], const T xnx,
                    T (&upnx)[3], const T fl, const Out& ot) {
        constexpr int P = decltype(par)::value;
        T N0[3];
        {
            plane_modes(uhi, uphi, Phi);
            // modes v[c][m], m = mx + 2my + 4mz (mode 0 is not used)
            T v[3][8];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (m) v[c][m] = Plo[c][m] + Phi[c][m];
                    v[c][m + 4] = Phi[c][m] - Plo[c][m];
                }
            T N1[3];
            // one output component at a time keeps few modal results live
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T yv[8];
#pragma unroll
                for (int m = 1; m < 8; ++m) yv[m] = coef.D[c][m] * v[c][m];
#pragma unroll
                for (int bb = 0; bb < 3; ++bb) {
                    if (bb == c) continue;
                    const int d = 3 - c - bb;
#pragma unroll
                    for (int qq = 0; qq < 2; ++qq) {
                        const int mq = qq << d;
                        yv[(1 << c) | mq] = fmad(cl[d][qq], v[bb][(1 << bb) | mq], yv[(1 << c) | mq]);
                        yv[(1 << bb) | mq] = fmad(cm[d][qq], v[bb][(1 << c) | mq], yv[(1 << bb) | mq]);
                    }
                }
                // element factor and the z part of the transposed butterfly:
                // plane kl gets (y_m - y_m+4), plane kl+1 gets (y_m + y_m+4)
                T zlo[4];
                Yout[c][0] = fl * yv[4];
                zlo[0] = fmad(-fl, yv[4], Yin[c][0]);
#pragma unroll
                for (int m = 1; m < 4; ++m) {
                    Yout[c][m] = fl * (yv[m] + yv[m + 4]);
                    zlo[m] = fmad(fl, yv[m] - yv[m + 4], Yin[c][m]);
                }
                // x-y part of the transposed butterfly -> corners cx + 2cy
                const T t0 = zlo[0] - zlo[1], t1 = zlo[0] + zlo[1], t2 = zlo[2] - zlo[3], t3 = zlo[2] + zlo[3];
                const T c00 = t0 - t2, c10 = t1 - t3, c01 = t0 + t2, c11 = t1 + t3;
                // node (i, j) takes corner (0,0) of its element and (1,0) of
                // the element of lane l-1; likewise for the row above
                N0[c] = c00 + from_prev(c10);
                N1[c] = c01 + from_prev(c11);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) sc[P][ro][c][lx] = N1[c];
        }
        stage(par, unx, xnx);
        __syncthreads();
        T o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            upnx[c] = su[P][ro + 1][c][lx];
            const T below = sc[P][ro > 0 ? ro - 1 : 0][c][lx];
            const T Au = fmad(shift, ulo[c], N0[c] + (owner ? below : T(0)));
            T v;
            if (MODE == OP_APPLY) v = Au;
            else if (MODE == OP_RESID) v = ot.bv[c] - Au;
            else v = fmad(omega * ot.dv[c], ot.bv[c] - Au, ulo[c]);   // OP_SMOOTH, OP_SMOOTH0
            o[c] = ((ot.mb >> c) & 1) ? T(0) : v;
        }
        st3(ot.ry, ooff, o);
        // partial sums of (u, y) for the operator, of (b, y) for a smoothing
        // sweep (the preconditioned residual's (r, z) when it is the V-cycle's
        // last)
        if (dot_part && ot.has_out && owner) {
            const T(&dl)[3] = MODE == OP_APPLY ? ulo : ot.bv;
            acc += fmad((double)dl[2], (double)o[2], fmad((double)dl[1], (double)o[1], (double)dl[0] * o[0]));
        }
    };

    // Rings indexed by layer: with the chunk's layers counted from 0 at kb-1,
    // layer j keeps the own values of planes kb-1+j.. in uu[j % 3..], its
    // staged rows, x-y modes and modal results in the parity pair [j % 2],
    // and the planes and element factors in flight in rr / ff[j % RING]:
    // three layers ahead, or two in the first smoothing sweep, whose planes
    // carry b and the inverse diagonal (three would cost it occupancy).
    constexpr int RING = MODE == OP_SMOOTH0 ? 2 : 3;
    T uu[3][3], up[2][3], PP[2][3][4], YY[2][3][4], xnx;
    {
        PIn p0, p1;
        load_in(kb - 1, p0);
        load_in(kb, p1);
        own_value(p0, uu[0], xnx);
        stage(P0{}, uu[0], xnx);
        own_value(p1, uu[1], xnx);
        stage(P1{}, uu[1], xnx);
        __syncthreads();
        T up0[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            up0[c] = su[0][ro + 1][c][lx];
            up[0][c] = su[1][ro + 1][c][lx];
        }
        plane_modes(uu[0], up0, PP[0]);
        __syncthreads();   // the staging buffers are rewritten by the first layer
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m) YY[0][c][m] = T(0);
    // prefetch ring: planes kl+2 .. kl+1+RING and factors of layers kl .. kl-1+RING
    PIn rr[RING];
    T ff[RING];
#pragma unroll
    for (int r = 0; r < RING; ++r) load_in(kb + 1 + r, rr[r]);
#pragma unroll
    for (int r = 0; r < RING; ++r) ff[r] = load_fac(kb - 1 + r);

    // (u, y) partial sums: each wave sums its lanes once per group of
    // DOT_GROUP output planes and writes one value per (tile column, wave,
    // group).  Chunks start on group boundaries, so every value, and the
    // fixed-order sum over them, is the same for any chunk length.
    const int tiles_xy = tiles_x * tiles_y;
    const int64_t dot_base = (int64_t)(tx + tiles_x * ty) * BY + w;
    auto flush_dot = [&](const int kl) {
        if (kl >= kb && ((kl + 1) % DOT_GROUP == 0 || kl + 1 == ke)) {
            double v = acc;
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (l == 0) dot_part[(int64_t)(kl / DOT_GROUP) * tiles_xy * BY + dot_base] = v;
            acc = 0.0;
        }
    };
    // layer L = kb - 1 + j (+ a multiple of 6): consumes plane L+2 and the
    // factor of layer L from ring slot j % RING and refills it with plane
    // L+2+RING and the factor of layer L+RING
    auto layer = [&](auto jc, const int L) {
        constexpr int J = decltype(jc)::value;
        constexpr int q = J % 2, s0 = J % 3, s1 = (J + 1) % 3, s2 = (J + 2) % 3, sr = J % RING;
        Out ot;
        load_out(L, ot);
        own_value(rr[sr], uu[s2], xnx);
        const T fl = ff[sr];
        load_in(L + 2 + RING, rr[sr]);
        ff[sr] = load_fac(L + RING);
        step(std::integral_constant<int, q>{}, PP[q], PP[1 - q], YY[q], YY[1 - q], uu[s0], uu[s1], up[q], uu[s2],
             xnx, up[1 - q], fl, ot);
        if (dot_part) flush_dot(L);
    };
    if constexpr (RING == 2) {
        // two layers per trip; the own-value ring is renamed by moves
        // (unrolling to six layers costs this variant spills)
        for (int kl = kb - 1; kl < ke; kl += 2) {
            layer(std::integral_constant<int, 0>{}, kl);
            if (kl + 1 >= ke) break;
            layer(std::integral_constant<int, 1>{}, kl + 1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T t = uu[1][c];
                uu[1][c] = uu[0][c];
                uu[0][c] = uu[2][c];
                uu[2][c] = t;
            }
        }
    } else
    for (int kl = kb - 1; kl < ke; kl += 6) {
        layer(std::integral_constant<int, 0>{}, kl);
        if (kl + 1 >= ke) break;
        layer(std::integral_constant<int, 1>{}, kl + 1);
        if (kl + 2 >= ke) break;
        layer(std::integral_constant<int, 2>{}, kl + 2);
        if (kl + 3 >= ke) break;
        layer(std::integral_constant<int, 3>{}, kl + 3);
        if (kl + 4 >= ke) break;
        layer(std::integral_constant<int, 4>{}, kl + 4);
        if (kl + 5 >= ke) break;
        layer(std::integral_constant<int, 5>{}, kl + 5);
    }
}

}  // namespace

namespace {
struct RwShape {
    int rpw, tiles_x, tiles_y, tiles_z, kchunk;
};

// Chunk lengths (node planes per block along z) measured fastest for a grid
// and kernel (tune_elastic_rw), shared by all handles of the process; grids
// not tuned use the block-count rule below.
enum TuneKind { TK_RW64 = 0, TK_RW32 = 1 };
std::mutex g_tune_mu;
std::map<std::array<int, 5>, int> g_tuned;
int rw_rpw(const Grid3& g, int kind);
int tuned_chunk(const Grid3& g, int kind) {
    const int rpw = rw_rpw(g, kind);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tuned.find({g.nx, g.ny, g.nz, kind, rpw});
    return it == g_tuned.end() ? 0 : it->second;
}

// HEXA_RW_KCHUNK=n (1..64) fixes the chunk length of every launch that does
// not ask for a specific one (used by the tests to cover every chunk length
// and ring phase on small grids).
int chunk_override() {
    const char* e = getenv("HEXA_RW_KCHUNK");
    if (!e || !e[0]) return 0;
    const int v = atoi(e);
    return v >= 1 && v <= 64 ? v : 0;
}

// Rows per wavefront for grid g, a function of the grid alone (the (u, y)
// partial sums are per wavefront, so the layout sets their summation order):
// the layout whose tiles touch fewer 128-byte lines of the node planes per
// plane (one line more than the row's bytes per loaded row, for the
// unaligned start), the one-row layout unless the two-row one saves 3%.
// Two rows of 32 lanes load 17 rows per 15 owned instead of 9 per 7, but
// touch more partial lines per byte and split the x axis into 30-node
// instead of 62-node tiles: at 256x128x128 (257 node columns: 5 tiles of 62
// hold 310) the FP64 product took 0.065 ms against 0.079, at 512x256x256
// 0.489 against 0.476.  HEXA_RW_RPW=1|2 fixes it (tests).
template <int RPW>
double rw_lines(const Grid3& g, int esize) {
    using TL = RwTile<RPW>;
    const double tx = (g.nx + 1 + TL::OWN_X - 1) / TL::OWN_X, ty = (g.ny + 1 + TL::OWN_Y - 1) / TL::OWN_Y;
    return tx * ty * (TL::NR + 1) * (double)((TL::LX * 3 * esize + 127) / 128 + 1);
}
int rw_rpw(const Grid3& g, int kind) {
    const char* e = getenv("HEXA_RW_RPW");
    if (e && (e[0] == '1' || e[0] == '2') && !e[1]) return e[0] - '0';
    const int es = kind == TK_RW64 ? 8 : 4;
    return rw_lines<2>(g, es) < 0.97 * rw_lines<1>(g, es) ? 2 : 1;
}

RwShape rw_shape(const Grid3& g, int kind = TK_RW64, int kchunk = 0) {
    RwShape r;
    r.rpw = rw_rpw(g, kind);
    const int own_x = r.rpw == 2 ? RwTile<2>::OWN_X : RwTile<1>::OWN_X;
    const int own_y = r.rpw == 2 ? RwTile<2>::OWN_Y : RwTile<1>::OWN_Y;
    r.tiles_x = (g.nx + 1 + own_x - 1) / own_x;
    r.tiles_y = (g.ny + 1 + own_y - 1) / own_y;
    // node planes per block: enough blocks to fill the chip, long enough
    // chunks to amortise the one recomputed element layer per block
    if (!kchunk) kchunk = chunk_override();
    r.kchunk = kchunk ? kchunk : tuned_chunk(g, kind);
    if (!r.kchunk) {
        r.kchunk = 32;
        while (r.kchunk > 1 && (int64_t)r.tiles_x * r.tiles_y * ((g.nz + 1 + r.kchunk - 1) / r.kchunk) < 2048)
            r.kchunk /= 2;
    }
    r.tiles_z = (g.nz + 1 + r.kchunk - 1) / r.kchunk;
    return r;
}
}  // namespace

bool elastic_rw_supported(const Grid3& g, size_t elem_bytes) {
    return !fallback_kernels_forced() && 3 * (int64_t)(g.nx + 1) * (g.ny + 1) * (int64_t)elem_bytes <= INT32_MAX;
}

// One partial per (tile column, wave, group of DOT_GROUP output planes) of
// the FP64 kernel (the only one that forms them).
int64_t elastic_rw_partials(const Grid3& g, size_t elem_bytes) {
    const RwShape r = rw_shape(g, elem_bytes == 8 ? TK_RW64 : TK_RW32, DOT_GROUP);
    return (int64_t)r.tiles_x * r.tiles_y * BY * ((g.nz + 1 + DOT_GROUP - 1) / DOT_GROUP);
}

template <typename T, typename DT>
int launch_elastic_rw(hipStream_t s, const Grid3& g, OpMode mode, const FactCoef<T>& coef, const T* fac,
                      const uint8_t* mask, const T* u, const T* b, const DT* dinv, T omega, T shift, T* y,
                      double* dot_part, int nb, const int32_t* halt, int hs) {
    // buffer descriptors address one node plane (byte counts are 31-bit)
    if (!elastic_rw_supported(g, std::max(sizeof(T), sizeof(DT)))) {
        const int64_t vs = batch_stride(3 * g.nn);
        for (int rb = 0; rb < nb; ++rb)
            launch_elastic_fact<T, DT>(s, g, mode, coef, fac, mask, u ? u + rb * vs : u, b ? b + rb * vs : b, dinv,
                                       omega, shift, y + rb * vs);
        return 0;
    }
    RwShape r = rw_shape(g, sizeof(T) == 8 ? TK_RW64 : TK_RW32);
    // partial sums need chunks that start on group boundaries: multiples of
    // DOT_GROUP planes (every tuned length of 4 or more is one)
    if (dot_part && r.kchunk % DOT_GROUP) r = rw_shape(g, sizeof(T) == 8 ? TK_RW64 : TK_RW32,
                                                      ((r.kchunk + DOT_GROUP - 1) / DOT_GROUP) * DOT_GROUP);
    const int tiles = r.tiles_x * r.tiles_y * r.tiles_z;
    const int64_t vs = batch_stride(3 * g.nn), ds = elastic_rw_partials(g, sizeof(T));
    dim3 block(WL, BY);
    const unsigned nblk = (unsigned)tiles * (unsigned)nb;
    auto go = [&](auto md, auto rp) {
        constexpr int M = decltype(md)::value, RP = decltype(rp)::value;
        k_elastic_rw<T, DT, M, RP><<<nblk, block, 0, s>>>(g, coef, fac, mask, u, b, dinv, omega, shift, y, r.kchunk,
                                                         r.tiles_x, r.tiles_y, tiles, dot_part, nb, vs, ds, halt, hs);
    };
    auto by_layout = [&](auto md) {
        if (r.rpw == 2) go(md, std::integral_constant<int, 2>{});
        else go(md, std::integral_constant<int, 1>{});
    };
    switch (mode) {
        case OP_APPLY: by_layout(std::integral_constant<int, OP_APPLY>{}); break;
        case OP_RESID: by_layout(std::integral_constant<int, OP_RESID>{}); break;
        case OP_SMOOTH0: by_layout(std::integral_constant<int, OP_SMOOTH0>{}); break;
        default: by_layout(std::integral_constant<int, OP_SMOOTH>{}); break;
    }
    return dot_part ? (int)ds : 0;
}

// Times K(rho)u (APPLY) for each candidate chunk length on this grid and
// records the fastest for the kernel of precision T.  u is read and y
// overwritten.  Grids already tuned, too
// small to matter, or HEXA_TUNE=0 keep the block-count rule.  The chunk
// length changes the order of the per-block (u, Au) partial sums, so results
// can differ in the last bits between a tuned and an untuned run.
template <typename T>
void tune_elastic_rw(hipStream_t s, const Grid3& g, const FactCoef<T>& coef, const T* fac, const T* u, T* y) {
    const char* te = getenv("HEXA_TUNE");
    if ((te && te[0] == '0') || (int64_t)g.nx * g.ny * g.nz < 32768 || !elastic_rw_supported(g, sizeof(T))) return;
    const int kind = sizeof(T) == 8 ? (int)TK_RW64 : (int)TK_RW32;
    hipEvent_t e0, e1;
    int ncu = 0;   // compute units of the current device
    {
        DeviceOpGuard lock;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
        if (hipEventCreate(&e0) != hipSuccess) return;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return;
        }
    }
    if (!tuned_chunk(g, kind)) {
        // chunks of 2 and 1 planes are tried only where chunks of 4 leave
        // fewer than 4096 blocks (mid-size grids, latency-bound); lengths
        // between 20 and 48 planes where chunks of 32 give 2048 blocks or
        // more: there the block count decides how full the last round of
        // blocks is (512x256x256: 44 planes, 0.491 against 0.524 ms with 32;
        // 256x128x128: 28 planes, 0.069 against 0.079 ms)
        const RwShape r4 = rw_shape(g, kind, 4), r32 = rw_shape(g, kind, 32);
        const bool small = (int64_t)r4.tiles_x * r4.tiles_y * r4.tiles_z < 4096;
        const bool large = (int64_t)r32.tiles_x * r32.tiles_y * r32.tiles_z >= 2048;
        int cand[24], NC = 0;
        auto add = [&](int c) {
            for (int q = 0; q < NC; ++q)
                if (cand[q] == c) return;
            cand[NC++] = c;
        };
        for (int c : {32, 16, 8, 4}) add(c);
        if (small)
            for (int c : {2, 1}) add(c);
        if (large)
            for (int c : {20, 24, 28, 36, 40, 44, 48}) add(c);
        // Lengths that fit the grid into 1-4 full rounds of the blocks the
        // chip holds at once (one FP64 or two FP32 blocks per CU), rounded up
        // to whole groups of partial-sum planes: at 256x128x128 one round
        // of 44-plane chunks (243 blocks) took 0.055 ms against 0.064 ms
        // with the best power of two.
        {
            const int64_t cap = (int64_t)ncu * (sizeof(T) == 8 ? 1 : 2);
            const int64_t txy = (int64_t)r4.tiles_x * r4.tiles_y;
            for (int rounds = 1; rounds <= 4; ++rounds) {
                const int64_t nch = rounds * cap / txy;
                if (nch < 1) continue;
                int kc = (int)((g.nz + 1 + nch - 1) / nch);
                kc = (kc + DOT_GROUP - 1) / DOT_GROUP * DOT_GROUP;
                if (kc >= 4 && kc <= 64) add(kc);
            }
        }
        float t[24] = {};
        auto launch = [&](int kc) {
            const RwShape r = rw_shape(g, kind, kc);
            const int tiles = r.tiles_x * r.tiles_y * r.tiles_z;
            if (r.rpw == 2)
                k_elastic_rw<T, T, OP_APPLY, 2><<<tiles, dim3(WL, BY), 0, s>>>(
                    g, coef, fac, nullptr, u, nullptr, nullptr, T(0), T(0), y, r.kchunk, r.tiles_x, r.tiles_y, tiles,
                    nullptr, 1, 0, 0, nullptr, 0);
            else
                k_elastic_rw<T, T, OP_APPLY, 1><<<tiles, dim3(WL, BY), 0, s>>>(
                    g, coef, fac, nullptr, u, nullptr, nullptr, T(0), T(0), y, r.kchunk, r.tiles_x, r.tiles_y, tiles,
                    nullptr, 1, 0, 0, nullptr, 0);
        };
        // the GPU first reaches its sustained clocks, then every candidate is
        // timed twice (ascending, then descending) and keeps its better time
        for (int w = 0; w < 8; ++w) launch(cand[0]);
        bool ok = true;
        for (int pass = 0; pass < 2 && ok; ++pass)
            for (int q = 0; q < NC && ok; ++q) {
                const int c = pass ? NC - 1 - q : q;
                launch(cand[c]);
                (void)hipEventRecord(e0, s);
                for (int rep = 0; rep < 3; ++rep) launch(cand[c]);
                (void)hipEventRecord(e1, s);
                float ms = 0.0f;
                ok = hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                if (ok && (pass == 0 || ms < t[c])) t[c] = ms;
            }
        int best = 0;   // 1 + index of the fastest candidate
        for (int c = 0; ok && c < NC; ++c)
            if (!best || t[c] < t[best - 1]) best = c + 1;
        best = ok && best ? cand[best - 1] : 0;
        // HEXA_TUNE_LOG=1: each tuned grid's candidate times and choice on stderr
        if (const char* lg = getenv("HEXA_TUNE_LOG"); lg && lg[0] == '1') {
            fprintf(stderr, "[hexa tune] %dx%dx%d %s layout %d:", g.nx, g.ny, g.nz, sizeof(T) == 8 ? "fp64" : "fp32",
                    rw_rpw(g, kind));
            for (int c = 0; c < NC; ++c) fprintf(stderr, " %d:%.4f", cand[c], t[c] / 3.0f);
            fprintf(stderr, " -> %d\n", best);
        }
        if (best) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[{g.nx, g.ny, g.nz, kind, rw_rpw(g, kind)}] = best;
        }
    }
    DeviceOpGuard lock;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}
template void tune_elastic_rw<double>(hipStream_t, const Grid3&, const FactCoef<double>&, const double*,
                                      const double*, double*);
template void tune_elastic_rw<float>(hipStream_t, const Grid3&, const FactCoef<float>&, const float*, const float*,
                                     float*);

#define HX_INST_RW(T, DT)                                                                            \
    template int launch_elastic_rw<T, DT>(hipStream_t, const Grid3&, OpMode, const FactCoef<T>&,     \
                                          const T*, const uint8_t*, const T*, const T*, const DT*, T, \
                                          T, T*, double*, int, const int32_t*, int);
HX_INST_RW(double, double)
HX_INST_RW(float, float)
HX_INST_RW(float, uint16_t)

}  // namespace hexa
