// PCG recurrence kernels with device-resident scalars.
//
// The reference driver (R/src/Solvers/GPUSolver.jl:335-444) brings every
// dot product and norm back to the host (:365,373,380,414) and snapshots
// the best iterate to host memory (:384).  Here alpha, beta, the residual
// check, the divergence tests and the best-iterate snapshot are evaluated on
// the device by single-block kernels reading fixed-order partial sums; the
// host only reads the state back to decide when to stop.  Every kernel is a
// no-op once the state has halted, so iterations queued ahead of the host's
// check cost nothing but their launch.
#include "hexa_internal.h"

namespace hexa {
namespace {

inline unsigned grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)(b < 1 ? 1 : b);
}

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

__device__ __forceinline__ double final_sum(const double* __restrict__ partials, int q, double* lds) {
    double s = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) s += partials[q * RED_BLOCKS + i];
    return block_sum(s, lds);
}

__device__ __forceinline__ double final_sum_n(const double* __restrict__ partials, int n, double* lds) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += RED_THREADS) s += partials[i];
    return block_sum(s, lds);
}

// partials[0] holds ||b||^2 on entry.
__global__ __launch_bounds__(RED_THREADS) void k_init_state(PcgState* st, const double* partials,
                                                            double tol) {
    __shared__ double lds[RED_THREADS / 64];
    const double bb = final_sum(partials, 0, lds);
    if (threadIdx.x == 0) {
        st->bnorm = sqrt(bb);
        st->rz = st->pAp = st->alpha = st->beta = st->rr = st->zAp = 0.0;
        st->rel = st->final_rel = 0.0;
        st->best_rel = INFINITY;
        st->tol = tol;
        st->halted = PCG_RUNNING;
        st->improved = 0;
        st->last_k = 0;
        st->x_iter = 0;
        st->x_done = 0;
        st->k = 0;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_rz(int64_t n, const double* __restrict__ r,
                                                    const double* __restrict__ z,
                                                    double* __restrict__ partials) {
    __shared__ double lds[RED_THREADS / 64];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
        s += r[i] * z[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_THREADS) void k_set_rz(PcgState* st, const double* partials) {
    __shared__ double lds[RED_THREADS / 64];
    const double v = final_sum(partials, 0, lds);
    if (threadIdx.x == 0) st->rz = v;
}

__global__ __launch_bounds__(RED_THREADS) void k_pAp(int64_t n, const PcgState* st,
                                                     const double* __restrict__ p,
                                                     const double* __restrict__ Ap,
                                                     double* __restrict__ partials) {
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
        s += p[i] * Ap[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// alpha = rz / (p, Ap); a breakdown |pAp| < 1e-20 stops the attempt
// (GPUSolver.jl:373-375).
// Batched launches (gridDim.y right-hand sides): block row blockIdx.y works
// on state st + blockIdx.y, vectors vs entries apart and partial sums ps
// apart; one right-hand side (gridDim.y = 1) uses offsets of 0.
__global__ __launch_bounds__(RED_THREADS) void k_alpha(PcgState* st, const double* partials,
                                                       int breakdown_check, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    // the previous iteration's deferred x update (if any) has been applied
    if (threadIdx.x == 0) st->x_done = st->x_iter;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double den = final_sum(partials, 0, lds);
    if (threadIdx.x == 0) {
        st->k += 1;   // one alpha per iteration: the device's iteration count
        st->pAp = den;
        if (breakdown_check && fabs(den) < 1e-20) {
            st->halted = PCG_BREAKDOWN;
            st->alpha = 0.0;
        } else {
            st->alpha = st->rz / den;
        }
    }
}

// First level of a long fixed-order sum: block b sums the b-th of RED_BLOCKS
// equal contiguous ranges of the n partials.
__global__ __launch_bounds__(RED_THREADS) void k_sum_ranges(const PcgState* st, const double* __restrict__ partials,
                                                            int64_t n, double* __restrict__ out, int64_t pin,
                                                            int pout) {
    st += blockIdx.y;
    partials += blockIdx.y * pin;
    out += (int64_t)blockIdx.y * pout;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const int64_t len = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    const int64_t i0 = blockIdx.x * len, i1 = min(n, i0 + len);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += RED_THREADS) s += partials[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// alpha from n partials of (p, Ap) written by the operator kernel.
__global__ __launch_bounds__(RED_THREADS) void k_alpha_n(PcgState* st, const double* partials, int n,
                                                         int breakdown_check, int64_t pin) {
    st += blockIdx.y;
    partials += blockIdx.y * pin;
    if (threadIdx.x == 0) st->x_done = st->x_iter;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double den = final_sum_n(partials, n, lds);
    if (threadIdx.x == 0) {
        st->k += 1;
        st->pAp = den;
        if (breakdown_check && fabs(den) < 1e-20) {
            st->halted = PCG_BREAKDOWN;
            st->alpha = 0.0;
        } else {
            st->alpha = st->rz / den;
        }
    }
}

// The standard (Fletcher-Reeves) beta = (r, z)_new / (r, z) from n partials
// of (r, z) written by the V-cycle's last smoothing sweep; k_beta takes the
// RED_BLOCKS range sums of k_sum_ranges instead.
__device__ __forceinline__ void set_beta(PcgState* st, double rz) {
    st->beta = rz / st->rz;
    st->rz = rz;
}
__global__ __launch_bounds__(RED_THREADS) void k_beta_n(PcgState* st, const double* partials, int n, int64_t pin) {
    st += blockIdx.y;
    partials += blockIdx.y * pin;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double rz = final_sum_n(partials, n, lds);
    if (threadIdx.x == 0) set_beta(st, rz);
}
__global__ __launch_bounds__(RED_THREADS) void k_beta(PcgState* st, const double* partials, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double rz = final_sum(partials, 0, lds);
    if (threadIdx.x == 0) set_beta(st, rz);
}

// x += alpha p; r -= alpha Ap; per-block partial sums of (r,r) and, when
// minv is given, (r, minv r) for the Jacobi preconditioner, in bpart[0..nblk)
// and bpart[nblk..2 nblk) (summed to RED_BLOCKS range sums by
// launch_pcg_update).  r32, when given, receives the new residual in FP32
// (the mixed V-cycle's input).  With x null the x update is deferred to the
// p update of the same iteration (which reads p anyway): the update records
// the iteration st->k in st->x_iter.  One pair of entries per lane (16-byte
// accesses); the lane after the last pair takes an odd last entry.  (Over
// RED_BLOCKS blocks striding through the vector the same update took 18%
// longer at 101.6M entries.)
__global__ __launch_bounds__(RED_THREADS) void k_update(int64_t n, PcgState* st, double* __restrict__ x,
                                                        const double* __restrict__ p, double* __restrict__ r,
                                                        const double* __restrict__ Ap,
                                                        const double* __restrict__ minv,
                                                        double* __restrict__ bpart, int64_t nblk,
                                                        float* __restrict__ r32, int64_t vs) {
    {
        const int rb = blockIdx.y;
        st += rb;
        bpart += (int64_t)rb * 2 * nblk;
        const int64_t o = rb * vs;
        if (x) x += o;
        p += o;
        r += o;
        Ap += o;
        if (r32) r32 += o;
    }
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double a = st->alpha;
    double rr = 0.0, rz = 0.0;
    const int64_t n2 = n / 2;
    const int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x;
    if (i < n2) {
        if (x) {
            double2 xv = reinterpret_cast<double2*>(x)[i];
            const double2 pv = reinterpret_cast<const double2*>(p)[i];
            xv.x += a * pv.x;
            xv.y += a * pv.y;
            reinterpret_cast<double2*>(x)[i] = xv;
        }
        double2 rv = reinterpret_cast<double2*>(r)[i];
        const double2 av = reinterpret_cast<const double2*>(Ap)[i];
        rv.x -= a * av.x;
        rv.y -= a * av.y;
        reinterpret_cast<double2*>(r)[i] = rv;
        if (r32) reinterpret_cast<float2*>(r32)[i] = make_float2((float)rv.x, (float)rv.y);
        rr += rv.x * rv.x;
        rr += rv.y * rv.y;
        if (minv) {
            const double2 mv = reinterpret_cast<const double2*>(minv)[i];
            rz += rv.x * (rv.x * mv.x);
            rz += rv.y * (rv.y * mv.y);
        }
    } else if (i == n2 && (n & 1)) {
        const int64_t e = n - 1;
        if (x) x[e] += a * p[e];
        const double ri = r[e] - a * Ap[e];
        r[e] = ri;
        if (r32) r32[e] = (float)ri;
        rr += ri * ri;
        if (minv) rz += ri * (ri * minv[e]);
    }
    if (!x && i == 0) st->x_iter = st->k;
    rr = block_sum(rr, lds);
    if (threadIdx.x == 0) bpart[blockIdx.x] = rr;
    if (minv) {
        rz = block_sum(rz, lds);
        if (threadIdx.x == 0) bpart[nblk + blockIdx.x] = rz;
    }
}

// Residual check after the update.  GPU semantics (GPUSolver.jl:379-401),
// evaluated when 'check' is set (k == 1 or k % 10 == 0 in the reference):
// best-iterate tracking, rel < tol -> converged, NaN/Inf -> halt,
// rel > 10 or (k > 500 and rel > 100 best) -> diverged.  CPU semantics
// (CPUSolver.jl:233-239): rel < tol every iteration, nothing else.
// For the Jacobi preconditioner beta = (r,z)_new/(r,z) is formed here too.
__global__ __launch_bounds__(RED_THREADS) void k_check(PcgState* st, const double* partials, int check_every,
                                                       int check_first, int cpu_semantics, int jacobi, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    if (st->halted) return;
    const int k = st->k;
    const bool check = (check_every > 0 && k % check_every == 0) || (check_first && k == 1);
    __shared__ double lds[RED_THREADS / 64];
    const double rr = final_sum(partials, 0, lds);
    const double rzn = jacobi ? final_sum(partials, 1, lds) : 0.0;
    if (threadIdx.x != 0) return;
    st->rr = rr;
    st->last_k = k;
    st->improved = 0;
    if (check) {
        const double rel = sqrt(rr) / st->bnorm;
        st->rel = rel;
        st->final_rel = rel;
        if (cpu_semantics) {
            if (rel < st->tol) { st->halted = PCG_CONVERGED; return; }
        } else {
            if (rel < st->best_rel) { st->best_rel = rel; st->improved = 1; }
            if (rel < st->tol) { st->halted = PCG_CONVERGED; return; }
            if (isnan(rel) || isinf(rel)) { st->halted = PCG_NAN; return; }
            if (rel > 10.0 || (k > 500 && rel > st->best_rel * 100.0)) { st->halted = PCG_DIVERGED; return; }
        }
    }
    if (jacobi) {
        st->beta = rzn / st->rz;
        st->rz = rzn;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_flex_dots(int64_t n, const PcgState* st,
                                                           const double* __restrict__ r,
                                                           const double* __restrict__ z,
                                                           const double* __restrict__ Ap,
                                                           double* __restrict__ partials, int64_t vs, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    r += blockIdx.y * vs;
    z += blockIdx.y * vs;
    Ap += blockIdx.y * vs;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    double rz = 0.0, zap = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS) {
        const double zi = z[i];
        rz += r[i] * zi;
        zap += zi * Ap[i];
    }
    rz = block_sum(rz, lds);
    zap = block_sum(zap, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = rz;
        partials[RED_BLOCKS + blockIdx.x] = zap;
    }
}

// Flexible (Polak-Ribiere) beta for a preconditioner that is not exactly a
// fixed SPD matrix (the mixed-precision V-cycle):
// beta = (z_new, r_new - r_old)/(z_old, r_old) = -alpha (z_new, Ap)/(z_old, r_old).
__global__ __launch_bounds__(RED_THREADS) void k_flex_beta(PcgState* st, const double* partials, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    const double rz = final_sum(partials, 0, lds);
    const double zap = final_sum(partials, 1, lds);
    if (threadIdx.x == 0) {
        st->zAp = zap;
        st->beta = -st->alpha * zap / st->rz;
        st->rz = rz;
    }
}

// The same two dots with the correction z in FP32 (the mixed V-cycle's
// output, widened exactly).
__global__ __launch_bounds__(RED_THREADS) void k_flex_dots_z32(int64_t n, const PcgState* st,
                                                               const double* __restrict__ r,
                                                               const float* __restrict__ z,
                                                               const double* __restrict__ Ap,
                                                               double* __restrict__ partials, int64_t vs, int ps) {
    st += blockIdx.y;
    partials += (int64_t)blockIdx.y * ps;
    r += blockIdx.y * vs;
    z += blockIdx.y * vs;
    Ap += blockIdx.y * vs;
    if (st->halted) return;
    __shared__ double lds[RED_THREADS / 64];
    double rz = 0.0, zap = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS) {
        const double zi = (double)z[i];
        rz += r[i] * zi;
        zap += zi * Ap[i];
    }
    rz = block_sum(rz, lds);
    zap = block_sum(zap, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = rz;
        partials[RED_BLOCKS + blockIdx.x] = zap;
    }
}

// p = z + beta p (z in FP32).  With x given, first the x update deferred
// by k_update of this iteration: x += alpha p with the old p, also when this
// iteration's check halted the loop, and once only (the iterations launched
// after a halt neither advance k nor repeat it: the alpha kernel records it
// as done).
__global__ void k_p_update_z32(int64_t n, const PcgState* st, double* __restrict__ p, const float* __restrict__ z,
                               double* __restrict__ x, int64_t vs) {
    st += blockIdx.y;
    p += blockIdx.y * vs;
    z += blockIdx.y * vs;
    if (x) x += blockIdx.y * vs;
    const bool upx = x && st->x_iter == st->k && st->x_done != st->k;
    const bool upp = !st->halted;
    if (!upx && !upp) return;
    const double a = st->alpha, b = st->beta;
    const int64_t n2 = n / 2;
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, ts = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = t0; i < n2; i += ts) {
        double2 pv = reinterpret_cast<double2*>(p)[i];
        if (upx) {
            double2 xv = reinterpret_cast<double2*>(x)[i];
            xv.x += a * pv.x;
            xv.y += a * pv.y;
            reinterpret_cast<double2*>(x)[i] = xv;
        }
        if (upp) {
            const float2 zv = reinterpret_cast<const float2*>(z)[i];
            pv.x = (double)zv.x + b * pv.x;
            pv.y = (double)zv.y + b * pv.y;
            reinterpret_cast<double2*>(p)[i] = pv;
        }
    }
    if ((n & 1) && t0 == 0) {
        const int64_t i = n - 1;
        if (upx) x[i] += a * p[i];
        if (upp) p[i] = (double)z[i] + b * p[i];
    }
}

// p = z + beta p; with z == nullptr, z = minv * r (Jacobi, not stored).
__global__ void k_p_update(int64_t n, const PcgState* st, double* __restrict__ p, const double* __restrict__ z,
                           const double* __restrict__ r, const double* __restrict__ minv, int64_t vs) {
    st += blockIdx.y;
    p += blockIdx.y * vs;
    if (z) z += blockIdx.y * vs;
    if (r) r += blockIdx.y * vs;
    if (st->halted) return;
    const double b = st->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double zi = z ? z[i] : r[i] * minv[i];
        p[i] = zi + b * p[i];
    }
}

__global__ void k_best_copy(int64_t n, const PcgState* st, const double* __restrict__ x,
                            double* __restrict__ best) {
    if (!st->improved) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        best[i] = x[i];
}

}  // namespace

// A running state whose iteration has a deferred x update pending and
// alpha = beta = 0, for timing the recurrence kernels in isolation.
__global__ void k_state_for_timing(PcgState* st) {
    st->halted = PCG_RUNNING;
    st->k = 1;
    st->x_iter = 1;
    st->x_done = 0;
    st->alpha = st->beta = 0.0;
    st->rz = 1.0;
}
void launch_pcg_state_for_timing(hipStream_t s, PcgState* st) { k_state_for_timing<<<1, 1, 0, s>>>(st); }
void launch_pcg_init_state(hipStream_t s, PcgState* st, const double* partials, double tol) {
    k_init_state<<<1, RED_THREADS, 0, s>>>(st, partials, tol);
}
void launch_pcg_rz(hipStream_t s, int64_t n, const double* r, const double* z, double* partials) {
    k_rz<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, r, z, partials);
}
void launch_pcg_set_rz(hipStream_t s, PcgState* st, const double* partials) {
    k_set_rz<<<1, RED_THREADS, 0, s>>>(st, partials);
}
void launch_pcg_pAp(hipStream_t s, int64_t n, PcgState* st, const double* p, const double* Ap,
                    double* partials) {
    k_pAp<<<RED_BLOCKS, RED_THREADS, 0, s>>>(n, st, p, Ap, partials);
}
void launch_pcg_alpha(hipStream_t s, PcgState* st, const double* partials, int breakdown_check) {
    k_alpha<<<1, RED_THREADS, 0, s>>>(st, partials, breakdown_check, 0);
}
void launch_pcg_alpha_n(hipStream_t s, PcgState* st, const double* partials, int64_t n, int breakdown_check,
                        double* scratch, int nb, int64_t pstride) {
    if (n <= RED_BLOCKS) {
        k_alpha_n<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, partials, (int)n, breakdown_check, pstride);
    } else {
        k_sum_ranges<<<dim3(RED_BLOCKS, nb), RED_THREADS, 0, s>>>(st, partials, n, scratch, pstride, RED_BLOCKS);
        k_alpha<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, scratch, breakdown_check, RED_BLOCKS);
    }
}
void launch_pcg_beta_n(hipStream_t s, PcgState* st, const double* partials, int64_t n, double* scratch, int nb,
                       int64_t pstride) {
    if (n <= RED_BLOCKS) {
        k_beta_n<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, partials, (int)n, pstride);
    } else {
        k_sum_ranges<<<dim3(RED_BLOCKS, nb), RED_THREADS, 0, s>>>(st, partials, n, scratch, pstride, RED_BLOCKS);
        k_beta<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, scratch, RED_BLOCKS);
    }
}
int64_t pcg_update_blocks(int64_t n) { return ((n + 1) / 2 + RED_THREADS - 1) / RED_THREADS; }
void launch_pcg_update(hipStream_t s, int64_t n, PcgState* st, double* x, const double* p, double* r,
                       const double* Ap, const double* minv, double* partials, double* upart, float* r32, int nb,
                       int64_t vs, int ps) {
    const int64_t nblk = pcg_update_blocks(n);
    k_update<<<dim3((unsigned)nblk, nb), RED_THREADS, 0, s>>>(n, st, x, p, r, Ap, minv, upart, nblk, r32, vs);
    // the per-block sums to RED_BLOCKS range sums per quantity, in a fixed order
    for (int q = 0; q < (minv ? 2 : 1); ++q)
        k_sum_ranges<<<dim3(RED_BLOCKS, nb), RED_THREADS, 0, s>>>(st, upart + q * nblk, nblk,
                                                                partials + q * RED_BLOCKS, 2 * nblk, ps);
}
void launch_pcg_check(hipStream_t s, PcgState* st, const double* partials, int check_every, int check_first,
                      int cpu_semantics, int jacobi, int nb, int ps) {
    k_check<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, partials, check_every, check_first, cpu_semantics, jacobi, ps);
}
void launch_pcg_flex_dots(hipStream_t s, int64_t n, const PcgState* st, const double* r, const double* z,
                          const double* Ap, double* partials, int nb, int64_t vs, int ps) {
    k_flex_dots<<<dim3(RED_BLOCKS, nb), RED_THREADS, 0, s>>>(n, st, r, z, Ap, partials, vs, ps);
}
void launch_pcg_flex_beta(hipStream_t s, PcgState* st, const double* partials, int nb, int ps) {
    k_flex_beta<<<dim3(1, nb), RED_THREADS, 0, s>>>(st, partials, ps);
}
void launch_pcg_p_update(hipStream_t s, int64_t n, const PcgState* st, double* p, const double* z,
                         const double* r, const double* minv, int nb, int64_t vs) {
    k_p_update<<<dim3(grid_for(n), nb), 256, 0, s>>>(n, st, p, z, r, minv, vs);
}
void launch_pcg_flex_dots_z32(hipStream_t s, int64_t n, const PcgState* st, const double* r, const float* z,
                              const double* Ap, double* partials, int nb, int64_t vs, int ps) {
    k_flex_dots_z32<<<dim3(RED_BLOCKS, nb), RED_THREADS, 0, s>>>(n, st, r, z, Ap, partials, vs, ps);
}
void launch_pcg_p_update_z32(hipStream_t s, int64_t n, const PcgState* st, double* p, const float* z, double* x,
                             int nb, int64_t vs) {
    k_p_update_z32<<<dim3(grid_for((n + 1) / 2), nb), 256, 0, s>>>(n, st, p, z, x, vs);
}
void launch_pcg_best_copy(hipStream_t s, int64_t n, const PcgState* st, const double* x, double* best) {
    k_best_copy<<<grid_for(n), 256, 0, s>>>(n, st, x, best);
}

}  // namespace hexa
