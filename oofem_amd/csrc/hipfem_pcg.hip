// hippcg: the SparseLinearSystemNM of the MI355X path (reference interface
// src/core/sparselinsystemnm.h:56-97), reproducing IMLSolver + CG + the
// diagonal preconditioner (imlsolver.C:102-146, iml/cg.h:22-72,
// diagpre.C:41-68):
//
//   r = b - A x;  stop if |r|/|b| <= tol (|b| = 1 when b = 0)
//   for i = 1..maxit:
//       z = D^-1 r;  rho = r.z;  p = (i == 1) ? z : z + (rho/rho_1) p
//       q = A p;  alpha = rho / p.q;  x += alpha p;  r -= alpha q
//       stop if |r|/|b| <= tol
//
// One iteration is five launches on the context stream:
//   k_pcg_pupdate  p = z + beta p in place
//   k_pcg_spmv     q = A p, per-block p.q partials
//   k_reduce_alpha alpha = rho / sum(p.q)                 (one block)
//   k_pcg_update   x += alpha p; r -= alpha q; z = D^-1 r; partials r.z, r.r
//   k_reduce_beta  convergence test, beta, rho            (one block)
// On request (hippcg_set_fused) p = z + beta p_old is formed inside the
// SpMV gather instead and written by the row owner, saving the first launch;
// on gfx950 that form is slower (the gather reads two vectors).  Summing the
// partials in the last block of the producing kernel (an agent-scope
// release fence per block and a counter) was measured slower still: each
// block's fence writes its XCD's L2 back to memory.
// Dot products are block partials summed in a fixed order, so a solve is
// bitwise reproducible.  Kernels return at once when the device-side "done"
// flag is set, which lets the host enqueue iterations in chunks and read
// the flag once per chunk.
//
// precision 1: FP32 matrix/vectors for an inner PCG (FP64 dot
// accumulation) inside FP64 iterative refinement; the true residual
// r = b - A x and the update x += d are FP64.
// precision 2: one CG on FP32 vectors and the FP32 matrix with FP64 x and
// FP64 dot accumulation; the recursive residual is replaced by the FP64
// true residual whenever it has dropped by a factor delta (reliable
// updates), keeping the Krylov direction, and convergence is decided on
// the true residual.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "hipfem_internal.h"

namespace hipfem {


constexpr int RED_THREADS = 1024;
constexpr int MAX_GRID = 4096;
constexpr int GRP_MAX_GRID = 16384;   // row-group and node-block PCG products: more resident workgroups

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block-wide sum of one value per thread; result valid in thread 0
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; i++) s += sh[i];
    __syncthreads();
    return s;
}

// The PCG scalar updates of the one-block reduction kernels.
__device__ __forceinline__ void step_alpha(PcgState *st, double pq) { st->alpha = st->rho / pq; }

// alpha of one FP32 iteration; a non-positive curvature p.q (the FP32
// operator has lost the definiteness of A on this direction) pauses the
// iteration for an update that restarts the direction
__device__ __forceinline__ void step_alpha_rel(PcgState *st, double pq)
{
    if (pq > 0.0) {
        st->alpha = st->rho / pq;
    } else {
        st->alpha = 0.0;
        st->pause = 1;
        st->restart = 1;
    }
}

// convergence test, beta and rho of a plain iteration (cg.h:50-70)
__device__ __forceinline__ void step_beta(PcgState *st, double rz, double rr, int iter, double *hist, int hist_cap)
{
    const double res = sqrt(rr) / st->normb;
    st->iter = iter;
    st->res = res;
    if (iter < hist_cap) hist[iter] = res;
    if (res <= st->tol) {
        st->done = 1;
    } else if (!(res == res)) {
        st->done = 1;
        st->bad = 1;
    } else {
        st->beta = rz / st->rho;
        st->rho = rz;
    }
}

// convergence / reliable-update test of one FP32 iteration
__device__ __forceinline__ void step_beta_rel(PcgState *st, double rz, double rr, int iter)
{
    const double res = sqrt(rr) / st->normb;
    st->iter = iter;
    st->res = res;
    if (!(res == res)) {
        st->done = 1;
        st->bad = 1;
    } else if (res <= st->tol || res < st->delta * st->rmax) {
        st->pause = 1;   // the host replaces r by the true residual
    } else if (res > 100.0 * st->rmax) {
        st->pause = 1;   // the recursion has drifted: update and restart the direction
        st->restart = 1;
    } else {
        st->beta = rz / st->rho;
        st->rho = rz;
    }
}

// fused form: p_new = z + beta p_old is formed inside the gather (two
// gathers per nonzero) and written by the row owner; q = A p_new
template <int LANES, int U, bool NT, typename VT, typename MT>
__global__ __launch_bounds__(256) void k_pcg_spmv_fused(int n, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col, const MT *__restrict__ val,
                                                        const VT *__restrict__ z, const VT *__restrict__ pold,
                                                        VT *__restrict__ pnew, VT *__restrict__ q,
                                                        const PcgState *__restrict__ st, double *__restrict__ partials,
                                                        int first)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    const VT beta = first ? VT(0) : (VT)st->beta;
    const int rows_per_block = 256 / LANES;
    const int lane = threadIdx.x % LANES;
    double acc = 0.0;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / LANES; row < n; row += gridDim.x * rows_per_block) {
        const int s = rowptr[row], e = rowptr[row + 1];
        VT sum;
        if (first)
            sum = row_partial<LANES, U, NT, VT>(s, e, lane, col, val, [&](int c) { return z[c]; });
        else
            sum = row_partial<LANES, U, NT, VT>(s, e, lane, col, val,
                                                [&](int c) { return fma(beta, pold[c], z[c]); });
        sum = group_sum<LANES>(sum);
        if (lane == 0) {
            const VT pi = first ? z[row] : fma(beta, pold[row], z[row]);
            pnew[row] = pi;
            q[row] = sum;
            acc += (double)pi * (double)sum;
        }
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// separate form: p was updated in place by k_pcg_pupdate; one gather per nonzero
template <int LANES, int U, bool NT, typename VT, typename MT>
__global__ __launch_bounds__(256) void k_pcg_spmv_p(int n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col, const MT *__restrict__ val,
                                                    const VT *__restrict__ p, VT *__restrict__ q,
                                                    const PcgState *__restrict__ st, double *__restrict__ partials)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    const int rows_per_block = 256 / LANES;
    const int lane = threadIdx.x % LANES;
    double acc = 0.0;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / LANES; row < n; row += gridDim.x * rows_per_block) {
        const int s = rowptr[row], e = rowptr[row + 1];
        VT sum = row_partial<LANES, U, NT, VT>(s, e, lane, col, val, [&](int c) { return p[c]; });
        sum = group_sum<LANES>(sum);
        if (lane == 0) {
            q[row] = sum;
            acc += (double)p[row] * (double)sum;
        }
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// streaming form of k_pcg_spmv_p (block-contiguous value/column streams)
template <int RPB, int U, typename VT, typename MT>
__global__ __launch_bounds__(256) void k_pcg_spmv_p_stream(int n, const int32_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ col, const MT *__restrict__ val,
                                                           const VT *__restrict__ p, VT *__restrict__ q,
                                                           const PcgState *__restrict__ st, double *__restrict__ partials)
{
    __shared__ VT prod[STREAM_CAP];
    __shared__ int rp[RPB + 1];
    __shared__ double sh[4];
    if (halted(st)) return;
    constexpr int LPR = 256 / RPB;
    double acc = 0.0;
    for (int r0 = blockIdx.x * RPB; r0 < n; r0 += gridDim.x * RPB) {
        const VT sum = stream_block_row<RPB, U, VT>(r0, n, rowptr, col, val, [&](int c) { return p[c]; }, prod, rp);
        const int row = r0 + threadIdx.x / LPR;
        if (threadIdx.x % LPR == 0 && row < n) {
            q[row] = sum;
            acc += (double)p[row] * (double)sum;
        }
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// node-block form of k_pcg_spmv_p (bcol per 3x3 block, one gather per block).
// FP: p_new = z + beta p_old is formed inside the gather (p = p_old, first:
// p_new = z) and written by the row owner to pnew, as k_pcg_pupdate would.
template <typename VT, typename MT, bool FP>
__global__ __launch_bounds__(256) void k_pcg_spmv_p_nb(int nbr, const int32_t *__restrict__ bptr,
                                                       const int32_t *__restrict__ bcol, const MT *__restrict__ val,
                                                       const VT *__restrict__ p, const VT *__restrict__ z,
                                                       VT *__restrict__ pnew, VT *__restrict__ q,
                                                       const PcgState *__restrict__ st, double *__restrict__ partials,
                                                       int first)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    const VT beta = (FP && !first) ? (VT)st->beta : VT(0);
    auto pv = [&](int i) -> VT {
        if constexpr (FP)
            return first ? z[i] : fma(beta, p[i], z[i]);
        else
            return p[i];
    };
    const int lane = threadIdx.x % 32;
    double acc = 0.0;
    for (int I = blockIdx.x * 8 + threadIdx.x / 32; I < nbr; I += gridDim.x * 8) {
        const int bo = bptr[I], nbk = bptr[I + 1] - bo;
        VT a[3];
        nb_row<VT>(bo, nbk, 9 * bo, lane, bcol, val,
                   [&](int bc, VT &x0, VT &x1, VT &x2) {
                       x0 = pv(3 * bc);
                       x1 = pv(3 * bc + 1);
                       x2 = pv(3 * bc + 2);
                   },
                   a);
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const VT pi = pv(3 * I + r);
                if constexpr (FP) pnew[3 * I + r] = pi;
                q[3 * I + r] = a[r];
                acc += (double)pi * (double)a[r];
            }
        }
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// row-group form of k_pcg_spmv_p (partially constrained node patterns);
// FP as in k_pcg_spmv_p_nb
template <typename VT, typename MT, bool FP>
__global__ __launch_bounds__(256) void k_pcg_spmv_p_grp(int ngrp, const int32_t *__restrict__ grow,
                                                        const int32_t *__restrict__ gch,
                                                        const uint32_t *__restrict__ ccol,
                                                        const int32_t *__restrict__ rowptr, const MT *__restrict__ val,
                                                        const VT *__restrict__ p, const VT *__restrict__ z,
                                                        VT *__restrict__ pnew, VT *__restrict__ q,
                                                        const PcgState *__restrict__ st, double *__restrict__ partials,
                                                        int first)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    const VT beta = (FP && !first) ? (VT)st->beta : VT(0);
    auto pv = [&](int i) -> VT {
        if constexpr (FP)
            return first ? z[i] : fma(beta, p[i], z[i]);
        else
            return p[i];
    };
    const int lane = threadIdx.x % GRP_LANES;
    double acc = 0.0;
    for (int g = blockIdx.x * GRP_PER_BLOCK + threadIdx.x / GRP_LANES; g < ngrp; g += gridDim.x * GRP_PER_BLOCK) {
        const int r0 = grow[g], nr = grow[g + 1] - r0, c0 = gch[g], nch = gch[g + 1] - c0;
        VT a[3];
        grp_row<VT>(r0, nr, c0, nch, lane, rowptr, ccol, val, pv, a);
        if (lane == 0)
            for (int r = 0; r < nr; r++) {
                const VT pi = pv(r0 + r);
                if constexpr (FP) pnew[r0 + r] = pi;
                q[r0 + r] = a[r];
                acc += (double)pi * (double)a[r];
            }
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// p = z (first iteration) or p = z + beta p, in place
template <typename VT>
__global__ __launch_bounds__(256) void k_pcg_pupdate(int n, const VT *__restrict__ z, VT *__restrict__ p,
                                                     const PcgState *__restrict__ st, int first)
{
    if (halted(st)) return;
    const VT beta = first ? VT(0) : (VT)st->beta;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        p[i] = first ? z[i] : fma(beta, p[i], z[i]);
}

template <typename XT, typename VT>
__global__ __launch_bounds__(256) void k_pcg_update(int n, XT *__restrict__ x, VT *__restrict__ r, VT *__restrict__ z,
                                                    const VT *__restrict__ p, const VT *__restrict__ q,
                                                    const VT *__restrict__ dinv, const PcgState *__restrict__ st,
                                                    double *__restrict__ partials)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    const double alpha = st->alpha;
    const VT a = (VT)alpha;
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] = fma((XT)a, (XT)p[i], x[i]);
        const VT ri = fma(-a, q[i], r[i]);
        r[i] = ri;
        const VT zi = dinv[i] * ri;
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += (double)ri * (double)ri;
    }
    const double s1 = block_sum<256>(rz, sh);
    const double s2 = block_sum<256>(rr, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}

// r = b - q (q = A x, or x == 0), z = D^-1 r; partials b.b, r.r, r.z
template <typename VT>
__global__ __launch_bounds__(256) void k_pcg_init(int n, const VT *__restrict__ b, const VT *__restrict__ q, int x_zero,
                                                  VT *__restrict__ r, VT *__restrict__ z, const VT *__restrict__ dinv,
                                                  double *__restrict__ partials)
{
    __shared__ double sh[4];
    double bb = 0.0, rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const VT bi = b[i];
        const VT ri = x_zero ? bi : bi - q[i];
        r[i] = ri;
        const VT zi = dinv[i] * ri;
        z[i] = zi;
        bb += (double)bi * (double)bi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    const double s1 = block_sum<256>(bb, sh);
    const double s2 = block_sum<256>(rr, sh);
    const double s3 = block_sum<256>(rz, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
        partials[2 * gridDim.x + blockIdx.x] = s3;
    }
}

// thread t sums p[t], p[t + RED_THREADS], ... in that order; the loads of
// eight terms are issued before their adds (the same sum, bit for bit, with
// one memory latency per eight terms instead of one per term)
__device__ double reduce_partials(const double *__restrict__ p, int nb, double *sh)
{
    double s = 0.0;
    for (int i0 = threadIdx.x; i0 < nb; i0 += 8 * RED_THREADS) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int i = i0 + j * RED_THREADS;
            v[j] = i < nb ? p[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (i0 + j * RED_THREADS < nb) s += v[j];
    }
    return block_sum<RED_THREADS>(s, sh);
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_init(PcgState *st, const double *__restrict__ partials, int nb,
                                                             double tol, double *__restrict__ hist, int hist_cap)
{
    __shared__ double sh[RED_THREADS / 64];
    const double bb = reduce_partials(partials, nb, sh);
    const double rr = reduce_partials(partials + nb, nb, sh);
    const double rz = reduce_partials(partials + 2 * nb, nb, sh);
    if (threadIdx.x == 0) {
        double normb = sqrt(bb);
        if (normb == 0.0) normb = 1;
        st->normb = normb;
        st->tol = tol;
        st->res = sqrt(rr) / normb;
        st->rho = rz;
        st->alpha = 0.0;
        st->beta = 0.0;
        st->iter = 0;
        st->bad = 0;
        st->pause = 0;
        st->restart = 0;
        st->done = (st->res <= tol) ? 1 : 0;
        if (hist_cap > 0) hist[0] = st->res;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_alpha(PcgState *st, const double *__restrict__ partials, int nb)
{
    __shared__ double sh[RED_THREADS / 64];
    if (halted(st)) return;
    const double pq = reduce_partials(partials, nb, sh);
    if (threadIdx.x == 0) step_alpha(st, pq);
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_beta(PcgState *st, const double *__restrict__ partials, int nb,
                                                             int iter, double *__restrict__ hist, int hist_cap)
{
    __shared__ double sh[RED_THREADS / 64];
    if (halted(st)) return;
    const double rz = reduce_partials(partials, nb, sh);
    const double rr = reduce_partials(partials + nb, nb, sh);
    if (threadIdx.x == 0) step_beta(st, rz, rr, iter, hist, hist_cap);
}

__global__ void k_dinv(int n, const int32_t *__restrict__ diag, const double *__restrict__ val, int use_diag,
                       double *__restrict__ dinv, float *__restrict__ dinv32, int *__restrict__ zero_row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 1.0;
    if (use_diag) {
        d = diag[i] >= 0 ? val[diag[i]] : 0.0;
        if (d == 0.0) atomicMin(zero_row, i);
    }
    dinv[i] = 1. / d;
    if (dinv32) dinv32[i] = (float)(1. / d);
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_d2f(int n, const double *__restrict__ a, float *__restrict__ b)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = (float)a[i];
}

__global__ void k_zero_f(int n, float *__restrict__ a)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = 0.0f;
}

__global__ void k_add_f2d(int n, const float *__restrict__ d, double *__restrict__ x)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) x[i] += (double)d[i];
}

// r = b - q (FP64), partial r.r, b.b
__global__ __launch_bounds__(256) void k_residual(int n, const double *__restrict__ b, const double *__restrict__ q,
                                                  double *__restrict__ r, double *__restrict__ partials, int x_zero)
{
    __shared__ double sh[4];
    double rr = 0.0, bb = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double ri = x_zero ? b[i] : b[i] - q[i];
        r[i] = ri;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    const double s1 = block_sum<256>(rr, sh);
    const double s2 = block_sum<256>(bb, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1;
        partials[gridDim.x + blockIdx.x] = s2;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce2(const double *__restrict__ partials, int nb, double *out)
{
    __shared__ double sh[RED_THREADS / 64];
    const double a = reduce_partials(partials, nb, sh);
    const double b = reduce_partials(partials + nb, nb, sh);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// ------------------------------------------------------ reliable updates
// Mixed-precision CG with reliable updates (Sleijpen & van der Vorst 1996;
// the GPU form of Clark et al. 2010): the iteration runs on FP32 vectors and
// an FP32 operator with FP64 dot accumulation and an FP64 solution x; when
// the recursive residual has dropped by the factor delta since the last
// update (or meets the tolerance), the FP64 true residual r = b - A x
// replaces it and the search direction is kept.

// true residual norm after k_residual; init also sets |b| and the counters
__global__ __launch_bounds__(RED_THREADS) void k_reduce_true(PcgState *st, const double *__restrict__ partials, int nb,
                                                             double tol, double delta, int init)
{
    __shared__ double sh[RED_THREADS / 64];
    if (st->done && !init) return;
    const double rr = reduce_partials(partials, nb, sh);
    const double bb = reduce_partials(partials + nb, nb, sh);
    if (threadIdx.x == 0) {
        if (init) {
            double normb = sqrt(bb);
            if (normb == 0.0) normb = 1;
            st->normb = normb;
            st->tol = tol;
            st->delta = delta;
            st->alpha = st->beta = 0.0;
            st->iter = 0;
            st->bad = 0;
            st->done = 0;
            st->restart = 0;
        }
        const double res = sqrt(rr) / st->normb;
        st->res = res;
        if (res <= st->tol) st->done = 1;
        if (!(res == res)) {
            st->done = 1;
            st->bad = 1;
        }
    }
}

// r32 = (float) r, z32 = D^-1 r32; partials r32.z32
__global__ __launch_bounds__(256) void k_rel_refresh(int n, const double *__restrict__ r, float *__restrict__ r32,
                                                     float *__restrict__ z32, const float *__restrict__ dinv32,
                                                     const PcgState *__restrict__ st, double *__restrict__ partials)
{
    __shared__ double sh[4];
    if (st->done) return;
    double rz = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float ri = (float)r[i];
        const float zi = dinv32[i] * ri;
        r32[i] = ri;
        z32[i] = zi;
        rz += (double)ri * (double)zi;
    }
    const double t = block_sum<256>(rz, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// q32 = (float) q64 and the partials of p.q after the element operator
__global__ __launch_bounds__(256) void k_dot_pq(int n, const float *__restrict__ p, const double *__restrict__ q64,
                                                float *__restrict__ q, const PcgState *__restrict__ st,
                                                double *__restrict__ partials)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double qi = q64[i];
        q[i] = (float)qi;
        acc += (double)p[i] * qi;
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// the element operator's slot sum fused with k_dot_pq: q32 = (float) of the
// FP64 sum over the colours touching each equation (ascending) of the FP32
// slot values, and the partials of p.q in that FP64 sum; the same grid and
// order as k_mf_slot_sum followed by k_dot_pq, so the same bits.  NC > 0:
// exactly NC colours, every slot read at once (the slots of colours that
// touch no element at an equation hold zero, and adding +0 to a sum that
// starts at +0 changes no bit)
template <int NC>
__global__ __launch_bounds__(256) void k_slot_sum_dot(int n, int ncolors, const uint16_t *__restrict__ eqmask,
                                                      const float *__restrict__ yslot, const float *__restrict__ p,
                                                      float *__restrict__ q, const PcgState *__restrict__ st,
                                                      double *__restrict__ partials)
{
    __shared__ double sh[4];
    if (halted(st)) return;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double y = 0.0;
        if constexpr (NC > 0) {
            float v[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) v[c] = yslot[(size_t)c * n + i];
#pragma unroll
            for (int c = 0; c < NC; c++) y += (double)v[c];
        } else {
            const unsigned m = eqmask[i];
            for (int c = 0; c < ncolors; c++)
                if (m >> c & 1u) y += (double)yslot[(size_t)c * n + i];
        }
        q[i] = (float)y;
        acc += (double)p[i] * y;
    }
    const double t = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// after a refresh: beta from the true residual (the direction is kept), rmax
__global__ __launch_bounds__(RED_THREADS) void k_reduce_refresh(PcgState *st, const double *__restrict__ partials,
                                                                int nb, int first)
{
    __shared__ double sh[RED_THREADS / 64];
    if (st->done) return;
    const double rz = reduce_partials(partials, nb, sh);
    if (threadIdx.x == 0) {
        st->beta = (first || st->restart) ? 0.0 : rz / st->rho;
        st->rho = rz;
        st->rmax = st->res;
        st->pause = 0;
        st->restart = 0;
    }
}

// one-block reduction forms of step_alpha_rel / step_beta_rel
__global__ __launch_bounds__(RED_THREADS) void k_reduce_alpha_rel(PcgState *st, const double *__restrict__ partials,
                                                                  int nb)
{
    __shared__ double sh[RED_THREADS / 64];
    if (halted(st)) return;
    const double pq = reduce_partials(partials, nb, sh);
    if (threadIdx.x == 0) step_alpha_rel(st, pq);
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_beta_rel(PcgState *st, const double *__restrict__ partials,
                                                                 int nb, int iter)
{
    __shared__ double sh[RED_THREADS / 64];
    if (halted(st)) return;
    const double rz = reduce_partials(partials, nb, sh);
    const double rr = reduce_partials(partials + nb, nb, sh);
    if (threadIdx.x == 0) step_beta_rel(st, rz, rr, iter);
}

static int vec_grid(int n) { return std::max(1, std::min((n + 255) / 256, MAX_GRID)); }
static int spmv_grid(int n, int lanes) { return std::max(1, std::min((n + 256 / lanes - 1) / (256 / lanes), MAX_GRID)); }

template <typename VT, typename MT>
static void launch_pcg_spmv_fused(const SpmvShape &sh, int grid, hipStream_t s, int n, const int32_t *rp,
                                  const int32_t *ci, const MT *v, const VT *z, const VT *po, VT *pn, VT *q,
                                  const PcgState *st, double *part, int first)
{
    dispatch_shape(sh, [&](auto L, auto U, auto NT) {
        k_pcg_spmv_fused<decltype(L)::value, decltype(U)::value, decltype(NT)::value, VT, MT>
            <<<grid, 256, 0, s>>>(n, rp, ci, v, z, po, pn, q, st, part, first);
    });
}

template <typename VT, typename MT>
static void launch_pcg_spmv_p(const SpmvShape &sh, int grid, hipStream_t s, int n, const int32_t *rp,
                              const int32_t *ci, const MT *v, const VT *p, VT *q, const PcgState *st, double *part)
{
    if (sh.stream)
        dispatch_stream(sh, [&](auto R, auto U) {
            k_pcg_spmv_p_stream<decltype(R)::value, decltype(U)::value, VT, MT>
                <<<grid, 256, 0, s>>>(n, rp, ci, v, p, q, st, part);
        });
    else
        dispatch_shape(sh, [&](auto L, auto U, auto NT) {
            k_pcg_spmv_p<decltype(L)::value, decltype(U)::value, decltype(NT)::value, VT, MT>
                <<<grid, 256, 0, s>>>(n, rp, ci, v, p, q, st, part);
        });
}

}  // namespace hipfem

using namespace hipfem;

struct hippcg_solver {
    hipfem_ctx *ctx = nullptr;
    double tol = 1.e-5;
    int maxit = 200;
    int precond = 0;
    int precision = 0;
    double inner_tol = 1e-4;
    int max_outer = 100;
    int last_outer = 0;
    int check_every = 16;
    int fused = 0;   // 1: p formed in the SpMV gather; 0: separate in-place p update (faster on gfx950)
    double delta = 0.1;   // precision 2: reliable update when |r| < delta |r at last update|
    int last_updates = 0;
    int last_mf = 0;      // the last precision-2 solve used the element-by-element operator
    int last_restarts = 0;
    // preconditioner cache (IMLSolver re-inits on a new matrix or version, imlsolver.C:112-114)
    const hipcsr_mtx *lhs = nullptr;
    long lhs_version = -1;
    int lhs_precond = -1;
    int n = 0;
    double *dinv = nullptr;
    float *dinv32 = nullptr;
    // work vectors
    double *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr, *hx = nullptr, *hb = nullptr;
    float *fr = nullptr, *fz = nullptr, *fp0 = nullptr, *fp1 = nullptr, *fq = nullptr, *fx = nullptr, *fb = nullptr;
    PcgState *st = nullptr;
    int *zero_row = nullptr;
    double *scal = nullptr;     // two doubles of device scratch
    // optional per-launch timing of the PCG SpMV (HIP events on the context stream)
    int profiling = 0;
    std::vector<hipEvent_t> ev;           // 2 events per iteration slot
    int prof_n64 = 0, prof_n32 = 0;
    double prof_ms64 = 0.0, prof_ms32 = 0.0;
    // optional residual history of the last plain solve: ||r_k|| / ||b||
    double *d_hist = nullptr;
    int hist_cap = 0;
    int hist_len = 0;
};

static int ensure_events(hippcg_solver *s, size_t slots)
{
    while (s->ev.size() < 2 * slots) {
        hipEvent_t e;
        HIPFEM_CHECK(hipEventCreate(&e));
        s->ev.push_back(e);
    }
    return 0;
}

static int solver_free(hippcg_solver *s)
{
    void *ps[] = {s->dinv, s->dinv32, s->r, s->z, s->p0, s->p1, s->q, s->hx, s->hb,
                  s->fr, s->fz, s->fp0, s->fp1, s->fq, s->fx, s->fb, s->st, s->zero_row, s->scal};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
    s->dinv = nullptr; s->dinv32 = nullptr;
    s->r = s->z = s->p0 = s->p1 = s->q = s->hx = s->hb = nullptr;
    s->fr = s->fz = s->fp0 = s->fp1 = s->fq = s->fx = s->fb = nullptr;
    s->st = nullptr;
    s->zero_row = nullptr;
    s->scal = nullptr;
    s->n = 0;
    s->lhs = nullptr;
    s->lhs_version = -1;
    return 0;
}

static int solver_alloc(hippcg_solver *s, int n)
{
    if (s->n == n && s->r) return 0;
    int rc = solver_free(s);
    if (rc) return rc;
    const size_t nd = sizeof(double) * std::max(1, n), nf = sizeof(float) * std::max(1, n);
    for (double **p : {&s->dinv, &s->r, &s->z, &s->p0, &s->p1, &s->q, &s->hx, &s->hb}) HIPFEM_CHECK(hipMalloc(p, nd));
    for (float **p : {&s->dinv32, &s->fr, &s->fz, &s->fp0, &s->fp1, &s->fq, &s->fx, &s->fb}) HIPFEM_CHECK(hipMalloc(p, nf));
    HIPFEM_CHECK(hipMalloc(&s->st, sizeof(PcgState)));
    HIPFEM_CHECK(hipMalloc(&s->zero_row, sizeof(int)));
    HIPFEM_CHECK(hipMalloc(&s->scal, 2 * sizeof(double)));
    s->n = n;
    return 0;
}

static int init_precond(hippcg_solver *s, hipcsr_mtx *A)
{
    if (s->lhs == A && s->lhs_version == A->version && s->lhs_precond == s->precond) return 0;
    hipStream_t st = s->ctx->stream;
    const int big = INT32_MAX;
    HIPFEM_CHECK(hipMemcpyAsync(s->zero_row, &big, sizeof(int), hipMemcpyHostToDevice, st));
    if (A->n) {
        k_dinv<<<(A->n + 255) / 256, 256, 0, st>>>(A->n, A->d_diag, A->d_val, s->precond == 1, s->dinv, s->dinv32,
                                                   s->zero_row);
        HIPFEM_CHECK(hipGetLastError());
    }
    int zr = big;
    HIPFEM_CHECK(hipMemcpyAsync(&zr, s->zero_row, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    // DiagPreconditioner::init diagpre.C:49-55: zero diagonal is an error
    HIPFEM_REQUIRE(zr == big, HIPFEM_ENUMERIC, "DiagPreconditioner: zero diagonal detected in equation %d", zr + 1);
    s->lhs = A;
    s->lhs_version = A->version;
    s->lhs_precond = s->precond;
    return 0;
}

static int read_state_at(hipfem_ctx *ctx, const PcgState *dst, PcgState *out)
{
    hipStream_t st = ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(ctx->h_pinned, dst, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    memcpy(out, ctx->h_pinned, sizeof(PcgState));
    return 0;
}

static int read_state(hippcg_solver *s, PcgState *out)
{
    hipStream_t st = s->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(s->ctx->h_pinned, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    memcpy(out, s->ctx->h_pinned, sizeof(PcgState));
    return 0;
}

// Where a plain PCG run keeps its device state: the solver's own buffers
// (state, the context's partial sums, optional history and launch timing),
// or caller-provided memory (the allocation-free entry point: part != null,
// no history, no timing).
struct PcgRunBufs {
    PcgState *st = nullptr;
    double *part = nullptr;   // null: the context's partials, grown as needed
    size_t part_cap = 0;      // doubles available at part
    bool prof = false;
    double *hist = nullptr;
    int hist_cap = 0;
};

static PcgRunBufs solver_bufs(hippcg_solver *s)
{
    PcgRunBufs b;
    b.st = s->st;
    b.prof = s->profiling != 0;
    b.hist = s->d_hist;
    b.hist_cap = s->hist_cap;
    return b;
}

// Generic PCG on device vectors.  VT: vector/arith type, MT: matrix values.
// b, x: length n; x is the initial guess (x_zero: known to be 0).
template <typename VT, typename XT, typename MT>
static int pcg_run(hippcg_solver *s, const PcgRunBufs &bufs, hipcsr_mtx *A, const MT *val, const VT *b, XT *x,
                   int x_zero, double tol, int maxit, VT *r, VT *z, VT *p0, VT *p1, VT *q, const VT *dinv, int *iters,
                   double *res, int *converged)
{
    hipStream_t st = s->ctx->stream;
    PcgState *const pst = bufs.st;
    const int n = A->n;
    const SpmvShape sh = A->shape;
    const int gv = vec_grid(n);
    const bool nodeblock = sh.nodeblock && A->nodeblocked;
    const bool grouped = sh.nodeblock && !A->nodeblocked && A->grouped;
    const int gs = nodeblock    ? std::max(1, std::min((A->nbr + 7) / 8, GRP_MAX_GRID))
                   : grouped   ? std::max(1, std::min((A->ngrp + GRP_PER_BLOCK - 1) / GRP_PER_BLOCK, GRP_MAX_GRID))
                   : sh.stream ? std::max(1, std::min((n + sh.rpb - 1) / sh.rpb, MAX_GRID))
                               : spmv_grid(n, sh.lanes);
    // p = z + beta p formed in the product's gather (the lane, node-block
    // and row-group forms), or in a pass of its own
    const bool fused = s->fused && !sh.stream;
    const size_t need = 3 * (size_t)std::max(gv, gs);
    double *part = bufs.part;
    int rc = 0;
    if (!part) {
        rc = ensure_partials(s->ctx, need);
        if (rc) return rc;
        part = s->ctx->d_partials;
    } else {
        HIPFEM_REQUIRE(need <= bufs.part_cap, HIPFEM_ETOOBIG, "pcg: %zu partial sums exceed the %zu provided", need,
                       bufs.part_cap);
    }
    if (!x_zero) {
        // q = A x with x in XT (FP64 only for the plain path)
        if constexpr (std::is_same<VT, double>::value && std::is_same<XT, double>::value) {
            rc = csr_spmv_dev(A, x, q);
            if (rc) return rc;
        } else {
            set_error("pcg: non-zero initial guess needs the FP64 path");
            return HIPFEM_EINVAL;
        }
    }
    k_pcg_init<VT><<<gv, 256, 0, st>>>(n, b, q, x_zero, r, z, dinv, part);
    k_reduce_init<<<1, RED_THREADS, 0, st>>>(pst, part, gv, tol, bufs.hist, bufs.hist_cap);
    HIPFEM_CHECK(hipGetLastError());
    PcgState h;
    rc = read_state_at(s->ctx, pst, &h);
    if (rc) return rc;
    if (h.done) {
        *iters = 0;
        *res = h.res;
        *converged = 1;
        if (bufs.hist) s->hist_len = std::min(1, bufs.hist_cap);
        return 0;
    }
    int it = 1;
    VT *pold = p0, *pnew = p1;
    int chunk = s->check_every;
    const bool prof = bufs.prof;
    while (it <= maxit) {
        const int stop = std::min(maxit, it + chunk - 1);
        if (prof) {
            rc = ensure_events(s, (size_t)stop);
            if (rc) return rc;
        }
        for (; it <= stop; it++) {
            const int first = (it == 1);
            if (!fused) k_pcg_pupdate<VT><<<gv, 256, 0, st>>>(n, z, p0, pst, first);
            if (prof) HIPFEM_CHECK(hipEventRecord(s->ev[2 * (it - 1)], st));
            // plain PCG never pauses, so the two direction buffers can swap
            // on the host: after "done" every launch returns at once
            VT *pin = fused ? pold : p0;
            if (nodeblock) {
                if (fused)
                    k_pcg_spmv_p_nb<VT, MT, true><<<gs, 256, 0, st>>>(A->nbr, A->d_bptr, A->d_bcol, val, pin, z, pnew,
                                                                      q, pst, part, first);
                else
                    k_pcg_spmv_p_nb<VT, MT, false><<<gs, 256, 0, st>>>(A->nbr, A->d_bptr, A->d_bcol, val, pin, z,
                                                                       nullptr, q, pst, part, first);
            } else if (grouped) {
                if (fused)
                    k_pcg_spmv_p_grp<VT, MT, true><<<gs, 256, 0, st>>>(A->ngrp, A->d_grow, A->d_gch, A->d_ccol,
                                                                       A->d_rowptr, val, pin, z, pnew, q,
                                                                       pst, part, first);
                else
                    k_pcg_spmv_p_grp<VT, MT, false><<<gs, 256, 0, st>>>(A->ngrp, A->d_grow, A->d_gch, A->d_ccol,
                                                                        A->d_rowptr, val, pin, z, nullptr,
                                                                        q, pst, part, first);
            } else if (fused) {
                launch_pcg_spmv_fused<VT, MT>(sh, gs, st, n, A->d_rowptr, A->d_colind, val, z, pold, pnew, q, pst,
                                              part, first);
            } else {
                launch_pcg_spmv_p<VT, MT>(sh, gs, st, n, A->d_rowptr, A->d_colind, val, p0, q, pst, part);
            }
            if (prof) HIPFEM_CHECK(hipEventRecord(s->ev[2 * (it - 1) + 1], st));
            k_reduce_alpha<<<1, RED_THREADS, 0, st>>>(pst, part, gs);
            k_pcg_update<XT, VT><<<gv, 256, 0, st>>>(n, x, r, z, fused ? pnew : p0, q, dinv, pst, part);
            k_reduce_beta<<<1, RED_THREADS, 0, st>>>(pst, part, gv, it, bufs.hist, bufs.hist_cap);
            if (fused) std::swap(pold, pnew);
        }
        HIPFEM_CHECK(hipGetLastError());
        rc = read_state_at(s->ctx, pst, &h);
        if (rc) return rc;
        if (h.done) break;
        chunk = std::min(chunk * 2, 64);
    }
    *iters = h.iter;
    *res = h.res;
    *converged = (h.done && !h.bad && h.res <= tol) ? 1 : 0;
    if (bufs.hist) s->hist_len = std::min(h.iter + 1, bufs.hist_cap);
    if (prof) {
        // launches after convergence returned at once: time iterations 1..h.iter only
        HIPFEM_CHECK(hipStreamSynchronize(st));
        double tot = 0.0;
        for (int k = 0; k < h.iter; k++) {
            float ms = 0.f;
            HIPFEM_CHECK(hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
            tot += ms;
        }
        if (std::is_same<VT, double>::value) {
            s->prof_n64 += h.iter;
            s->prof_ms64 += tot;
        } else {
            s->prof_n32 += h.iter;
            s->prof_ms32 += tot;
        }
    }
    HIPFEM_REQUIRE(!h.bad, HIPFEM_ENUMERIC, "pcg: residual became NaN at iteration %d", h.iter);
    return 0;
}

// precision 2: CG on FP32 vectors and the FP32 copy of A, FP64 x and dot
// accumulation, reliable updates with the FP64 true residual b - A x.
static int pcg_reliable(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx, int maxit, int *iters,
                        double *res, int *converged)
{
    hipStream_t st = s->ctx->stream;
    const int n = A->n;
    // inner operator: the element-by-element FP32 form of A when A is the
    // unchanged result of a batched assembly, else the FP32 copy of A
    const bool mf = model_mf_available(A);
    int rc = mf ? 0 : csr_refresh_fp32(A);
    if (rc) return rc;
    s->last_mf = mf ? 1 : 0;
    const SpmvShape sh = A->shape;
    const int gv = vec_grid(n);
    const bool nodeblock = sh.nodeblock && A->nodeblocked;
    const bool grouped = sh.nodeblock && !A->nodeblocked && A->grouped;
    const int gs = nodeblock    ? std::max(1, std::min((A->nbr + 7) / 8, GRP_MAX_GRID))
                   : grouped   ? std::max(1, std::min((A->ngrp + GRP_PER_BLOCK - 1) / GRP_PER_BLOCK, GRP_MAX_GRID))
                   : sh.stream ? std::max(1, std::min((n + sh.rpb - 1) / sh.rpb, MAX_GRID))
                               : spmv_grid(n, sh.lanes);
    rc = ensure_partials(s->ctx, 3 * (size_t)std::max(gv, gs));
    if (rc) return rc;
    double *part = s->ctx->d_partials;
    // r = b - A x in FP64, then the FP32 residual; init = 1 sets |b|
    auto true_residual = [&](int init) -> int {
        int e = csr_spmv_dev(A, dx, s->q);
        if (e) return e;
        k_residual<<<gv, 256, 0, st>>>(n, db, s->q, s->r, part, 0);
        k_reduce_true<<<1, RED_THREADS, 0, st>>>(s->st, part, gv, s->tol, s->delta, init);
        k_rel_refresh<<<gv, 256, 0, st>>>(n, s->r, s->fr, s->fz, s->dinv32, s->st, part);
        k_reduce_refresh<<<1, RED_THREADS, 0, st>>>(s->st, part, gv, init);
        HIPFEM_CHECK(hipGetLastError());
        return 0;
    };
    rc = true_residual(1);
    if (rc) return rc;
    PcgState h;
    rc = read_state(s, &h);
    if (rc) return rc;
    int it = 1, chunk = s->check_every;
    bool first = true;
    const bool prof = s->profiling != 0;
    s->last_updates = 0;
    s->last_restarts = 0;
    int last_restart_iter = -1;
    while (!h.done && it <= maxit) {
        const int stop = std::min(maxit, it + chunk - 1);
        if (prof) {
            rc = ensure_events(s, (size_t)stop);
            if (rc) return rc;
        }
        for (int k = it; k <= stop; k++) {
            k_pcg_pupdate<float><<<gv, 256, 0, st>>>(n, s->fz, s->fp0, s->st, first ? 1 : 0);
            first = false;
            if (prof) HIPFEM_CHECK(hipEventRecord(s->ev[2 * (k - 1)], st));
            if (mf) {
                const float *ys = nullptr;
                const uint16_t *em = nullptr;
                int nc = 0;
                rc = model_mf_slots_f32(A->mf_model, s->fp0, s->st, &ys, &em, &nc);
                if (rc) return rc;
                if (ys) {
                    if (nc == 8)
                        k_slot_sum_dot<8><<<gv, 256, 0, st>>>(n, nc, em, ys, s->fp0, s->fq, s->st, part);
                    else
                        k_slot_sum_dot<0><<<gv, 256, 0, st>>>(n, nc, em, ys, s->fp0, s->fq, s->st, part);
                } else {
                    rc = model_mf_apply_f32(A->mf_model, s->fp0, s->q, s->st);
                    if (rc) return rc;
                    k_dot_pq<<<gv, 256, 0, st>>>(n, s->fp0, s->q, s->fq, s->st, part);
                }
            } else if (nodeblock) {
                k_pcg_spmv_p_nb<float, float, false><<<gs, 256, 0, st>>>(A->nbr, A->d_bptr, A->d_bcol, A->d_val32,
                                                                         s->fp0, nullptr, nullptr, s->fq, s->st, part,
                                                                         0);
            } else if (grouped) {
                k_pcg_spmv_p_grp<float, float, false><<<gs, 256, 0, st>>>(A->ngrp, A->d_grow, A->d_gch, A->d_ccol,
                                                                          A->d_rowptr, A->d_val32, s->fp0,
                                                                          nullptr, nullptr, s->fq, s->st, part, 0);
            } else {
                launch_pcg_spmv_p<float, float>(sh, gs, st, n, A->d_rowptr, A->d_colind, A->d_val32, s->fp0, s->fq,
                                                s->st, part);
            }
            if (prof) HIPFEM_CHECK(hipEventRecord(s->ev[2 * (k - 1) + 1], st));
            k_reduce_alpha_rel<<<1, RED_THREADS, 0, st>>>(s->st, part, mf ? gv : gs);
            k_pcg_update<double, float><<<gv, 256, 0, st>>>(n, dx, s->fr, s->fz, s->fp0, s->fq, s->dinv32, s->st,
                                                            part);
            k_reduce_beta_rel<<<1, RED_THREADS, 0, st>>>(s->st, part, gv, k);
        }
        HIPFEM_CHECK(hipGetLastError());
        rc = read_state(s, &h);
        if (rc) return rc;
        if (h.done) break;
        if (h.pause) {
            // iterations after the pause returned at once; resume after the
            // last completed one (a pause inside k_reduce_alpha_rel leaves
            // that iteration without an update: it is done again)
            const bool restart = h.restart != 0;
            // a restart that makes no progress (the FP32 operator is not
            // positive on the preconditioned residual) ends the solve
            if (restart && h.iter == last_restart_iter) break;
            if (restart) last_restart_iter = h.iter;
            if (s->last_updates > maxit) break;
            rc = true_residual(0);
            if (rc) return rc;
            rc = read_state(s, &h);
            if (rc) return rc;
            s->last_updates++;
            if (restart) {
                first = true;
                s->last_restarts++;
            }
            it = h.iter + 1;
            continue;
        }
        it = stop + 1;
        chunk = std::min(chunk * 2, 64);
    }
    *iters = h.iter;
    *res = h.res;
    *converged = (h.done && !h.bad && h.res <= s->tol) ? 1 : 0;
    if (prof) {
        HIPFEM_CHECK(hipStreamSynchronize(st));
        double tot = 0.0;
        for (int k = 0; k < h.iter; k++) {
            float ms = 0.f;
            HIPFEM_CHECK(hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
            tot += ms;
        }
        s->prof_n32 += h.iter;
        s->prof_ms32 += tot;
    }
    HIPFEM_REQUIRE(!h.bad, HIPFEM_ENUMERIC, "pcg: residual became NaN at iteration %d", h.iter);
    return 0;
}

static int solve_dev(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx, int *reason, int *iters,
                     double *achieved)
{
    HIPFEM_REQUIRE(A->has_structure, HIPFEM_EINVAL, "solve: matrix has no structure");
    HIPFEM_REQUIRE(!A->staged, HIPFEM_EINVAL, "solve: assemble() contributions not flushed (call assembleEnd)");
    int rc = solver_alloc(s, A->n);
    if (rc) return rc;
    rc = init_precond(s, A);
    if (rc) return rc;
    int it = 0, conv = 0;
    double res = 0.0;
    s->last_outer = 0;
    s->prof_n64 = s->prof_n32 = 0;
    s->prof_ms64 = s->prof_ms32 = 0.0;
    if (A->n == 0) {
        *reason = HIPFEM_CR_CONVERGED;
        *iters = 0;
        *achieved = 0.0;
        return 0;
    }
    if (s->precision == 0) {
        rc = pcg_run<double, double, double>(s, solver_bufs(s), A, A->d_val, db, dx, 0, s->tol, s->maxit, s->r, s->z,
                                              s->p0, s->p1, s->q, s->dinv, &it, &res, &conv);
        if (rc) return rc;
    } else if (s->precision == 2) {
        rc = pcg_reliable(s, A, db, dx, s->maxit, &it, &res, &conv);
        if (rc) return rc;
        s->last_outer = s->last_updates;
    } else {
        // FP64 iterative refinement around an FP32 inner PCG
        rc = csr_refresh_fp32(A);
        if (rc) return rc;
        hipStream_t st = s->ctx->stream;
        const int n = A->n, gv = vec_grid(n);
        rc = ensure_partials(s->ctx, 3 * (size_t)std::max(gv, MAX_GRID));
        if (rc) return rc;
        double *out = s->scal;
        int total = 0;
        for (int outer = 0; outer < s->max_outer; outer++) {
            rc = csr_spmv_dev(A, dx, s->q);
            if (rc) return rc;
            k_residual<<<gv, 256, 0, st>>>(n, db, s->q, s->r, s->ctx->d_partials, 0);
            k_reduce2<<<1, RED_THREADS, 0, st>>>(s->ctx->d_partials, gv, out);
            HIPFEM_CHECK(hipGetLastError());
            double h2[2];
            HIPFEM_CHECK(hipMemcpyAsync(s->ctx->h_pinned, out, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPFEM_CHECK(hipStreamSynchronize(st));
            memcpy(h2, s->ctx->h_pinned, sizeof h2);
            double normb = sqrt(h2[1]);
            if (normb == 0.0) normb = 1;
            res = sqrt(h2[0]) / normb;
            s->last_outer = outer;
            if (res <= s->tol) { conv = 1; break; }
            if (total >= s->maxit) break;
            k_d2f<<<gv, 256, 0, st>>>(n, s->r, s->fb);
            k_zero_f<<<gv, 256, 0, st>>>(n, s->fx);
            // inner solve to inner_tol relative, but never below what the outer
            // tolerance needs from this correction
            const double itol = std::max(s->inner_tol, 0.5 * s->tol / res);
            int iit = 0, iconv = 0;
            double ires = 0;
            rc = pcg_run<float, float, float>(s, solver_bufs(s), A, A->d_val32, s->fb, s->fx, 1, itol, s->maxit - total,
                                              s->fr, s->fz, s->fp0, s->fp1, s->fq, s->dinv32, &iit, &ires, &iconv);
            if (rc) return rc;
            total += iit;
            k_add_f2d<<<gv, 256, 0, st>>>(n, s->fx, dx);
            HIPFEM_CHECK(hipGetLastError());
            s->last_outer = outer + 1;
        }
        it = total;
    }
    *reason = conv ? HIPFEM_CR_CONVERGED : HIPFEM_CR_DIVERGED_ITS;
    *iters = it;
    *achieved = res;
    return 0;
}

// ------------------------------------------------- allocation-free plain PCG
// Device memory of hippcg_solve_dev_noalloc, one copy per GPU, held until the
// process ends: the diagonal inverse and five work vectors of up to
// HIPPCG_NOALLOC_MAX_N equations, the block partial sums, the solver state
// and the zero-diagonal flag.  Host code never uses the array's name as an
// address: it takes the device's address once from hipGetSymbolAddress.
constexpr size_t NA_N = HIPPCG_NOALLOC_MAX_N;
constexpr size_t NA_PART = 3 * (size_t)std::max(MAX_GRID, GRP_MAX_GRID);
constexpr size_t NA_OFF_PART = 6 * NA_N * sizeof(double);
constexpr size_t NA_OFF_STATE = NA_OFF_PART + NA_PART * sizeof(double);
constexpr size_t NA_OFF_FLAG = NA_OFF_STATE + 256;
constexpr size_t NA_BYTES = NA_OFF_FLAG + 256;
static_assert(sizeof(PcgState) <= 256, "PcgState exceeds its scratch slot");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static std::mutex na_mutex;
static unsigned char *na_ptr[64] = {};

// address of this device's scratch (the current device must be `device`)
static int noalloc_scratch(int device, unsigned char **out)
{
    HIPFEM_REQUIRE(device >= 0 && device < 64, HIPFEM_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[device]) {
        void *ptr = nullptr;
        HIPFEM_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        HIPFEM_REQUIRE(ptr != nullptr, HIPFEM_EHIP, "no device address for the solver scratch");
        na_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    *out = na_ptr[device];
    return 0;
}

extern "C" int hippcg_solve_dev_noalloc(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx, int *reason,
                                        int *iters, double *achieved)
{
    HIPFEM_REQUIRE(s && A && db && dx && reason && iters && achieved, HIPFEM_EINVAL, "solve_noalloc: null argument");
    HIPFEM_REQUIRE(A->ctx == s->ctx, HIPFEM_EINVAL, "solve_noalloc: matrix and solver on different contexts");
    HIPFEM_REQUIRE(A->has_structure, HIPFEM_EINVAL, "solve_noalloc: matrix has no structure");
    HIPFEM_REQUIRE(!A->staged, HIPFEM_EINVAL, "solve_noalloc: assemble() contributions not flushed (call assembleEnd)");
    HIPFEM_REQUIRE(s->precision == 0, HIPFEM_EINVAL, "solve_noalloc: plain FP64 PCG (precision 0) only");
    HIPFEM_REQUIRE((size_t)A->n <= NA_N, HIPFEM_ETOOBIG, "solve_noalloc: %d equations exceed the limit of %zu", A->n,
                   NA_N);
    DeviceGuard g(s->ctx->device);
    if (A->n == 0) {
        *reason = HIPFEM_CR_CONVERGED;
        *iters = 0;
        *achieved = 0.0;
        return 0;
    }
    unsigned char *base = nullptr;
    int rc = noalloc_scratch(s->ctx->device, &base);
    if (rc) return rc;
    double *v = reinterpret_cast<double *>(base);
    double *dinv = v, *r = v + NA_N, *z = v + 2 * NA_N, *p0 = v + 3 * NA_N, *p1 = v + 4 * NA_N, *q = v + 5 * NA_N;
    PcgRunBufs bufs;
    bufs.part = reinterpret_cast<double *>(base + NA_OFF_PART);
    bufs.part_cap = NA_PART;
    bufs.st = reinterpret_cast<PcgState *>(base + NA_OFF_STATE);
    int *zero_row = reinterpret_cast<int *>(base + NA_OFF_FLAG);
    hipStream_t st = s->ctx->stream;
    // DiagPreconditioner::init on every call: the scratch holds whatever the
    // last call on this GPU left there
    k_set_int<<<1, 64, 0, st>>>(zero_row, INT32_MAX);
    if (A->n) {
        k_dinv<<<(A->n + 255) / 256, 256, 0, st>>>(A->n, A->d_diag, A->d_val, s->precond == 1, dinv, nullptr, zero_row);
        HIPFEM_CHECK(hipGetLastError());
    }
    HIPFEM_CHECK(hipMemcpyAsync(s->ctx->h_pinned, zero_row, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    int zr = 0;
    memcpy(&zr, s->ctx->h_pinned, sizeof(int));
    HIPFEM_REQUIRE(zr == INT32_MAX, HIPFEM_ENUMERIC, "DiagPreconditioner: zero diagonal detected in equation %d", zr + 1);
    int it = 0, conv = 0;
    double res = 0.0;
    rc = pcg_run<double, double, double>(s, bufs, A, A->d_val, db, dx, 0, s->tol, s->maxit, r, z, p0, p1, q, dinv,
                                          &it, &res, &conv);
    if (rc) return rc;
    *reason = conv ? HIPFEM_CR_CONVERGED : HIPFEM_CR_DIVERGED_ITS;
    *iters = it;
    *achieved = res;
    return 0;
}

extern "C" int hippcg_create(hipfem_ctx *ctx, hippcg_solver **out)
{
    HIPFEM_REQUIRE(ctx && out, HIPFEM_EINVAL, "hippcg_create: null argument");
    hippcg_solver *s = new hippcg_solver();
    s->ctx = ctx;
    ctx_retain(ctx);
    *out = s;
    return 0;
}

extern "C" void hippcg_destroy(hippcg_solver *s)
{
    if (!s) return;
    hipfem_ctx *c = s->ctx;
    {
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        solver_free(s);
        if (s->d_hist) (void)hipFree(s->d_hist);
        for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
        delete s;
    }
    ctx_release(c);
}

extern "C" int hippcg_set_options(hippcg_solver *s, double lstol, int lsiter, int lsprecond)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    HIPFEM_REQUIRE(lsprecond == 0 || lsprecond == 1, HIPFEM_EINVAL, "unknown preconditioner type %d", lsprecond);
    HIPFEM_REQUIRE(lsiter >= 0, HIPFEM_EINVAL, "lsiter < 0");
    s->tol = lstol;
    s->maxit = lsiter;
    s->precond = lsprecond;
    return 0;
}

extern "C" int hippcg_set_precision(hippcg_solver *s, int precision, double inner_tol, int max_outer)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    HIPFEM_REQUIRE(precision >= 0 && precision <= 2, HIPFEM_EINVAL, "precision must be 0, 1 or 2");
    s->precision = precision;
    if (inner_tol > 0) s->inner_tol = inner_tol;
    if (max_outer > 0) s->max_outer = max_outer;
    return 0;
}

extern "C" int hippcg_last_outer(hippcg_solver *s) { return s ? s->last_outer : 0; }

extern "C" int hippcg_last_matrix_free(hippcg_solver *s) { return s ? s->last_mf : 0; }

extern "C" int hippcg_set_reliable(hippcg_solver *s, double delta)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    HIPFEM_REQUIRE(delta > 0.0 && delta < 1.0, HIPFEM_EINVAL, "reliable-update factor must be in (0, 1)");
    s->delta = delta;
    return 0;
}

extern "C" int hippcg_set_fused(hippcg_solver *s, int fused)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    s->fused = fused ? 1 : 0;
    return 0;
}

extern "C" int hippcg_set_history(hippcg_solver *s, int cap)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    HIPFEM_REQUIRE(cap >= 0, HIPFEM_EINVAL, "history capacity %d < 0", cap);
    DeviceGuard g(s->ctx->device);
    HIPFEM_CHECK(hipStreamSynchronize(s->ctx->stream));
    if (s->d_hist) HIPFEM_CHECK(hipFree(s->d_hist));
    s->d_hist = nullptr;
    s->hist_cap = 0;
    s->hist_len = 0;
    if (cap > 0) {
        HIPFEM_CHECK(hipMalloc(&s->d_hist, sizeof(double) * (size_t)cap));
        s->hist_cap = cap;
    }
    return 0;
}

extern "C" int hippcg_get_history(hippcg_solver *s, double *out, int cap, int *len)
{
    HIPFEM_REQUIRE(s && len, HIPFEM_EINVAL, "null argument");
    *len = s->hist_len;
    const int n = std::min(cap, s->hist_len);
    if (n > 0 && out) {
        DeviceGuard g(s->ctx->device);
        HIPFEM_CHECK(hipMemcpyAsync(out, s->d_hist, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s->ctx->stream));
        HIPFEM_CHECK(hipStreamSynchronize(s->ctx->stream));
    }
    return 0;
}

extern "C" int hippcg_set_profiling(hippcg_solver *s, int on)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    s->profiling = on ? 1 : 0;
    return 0;
}

extern "C" int hippcg_profile(hippcg_solver *s, int *n64, double *ms64, int *n32, double *ms32)
{
    HIPFEM_REQUIRE(s, HIPFEM_EINVAL, "null solver");
    if (n64) *n64 = s->prof_n64;
    if (ms64) *ms64 = s->prof_ms64;
    if (n32) *n32 = s->prof_n32;
    if (ms32) *ms32 = s->prof_ms32;
    return 0;
}

extern "C" int hippcg_solve_dev(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx, int *reason, int *iters,
                                double *achieved)
{
    HIPFEM_REQUIRE(s && A && db && dx && reason && iters && achieved, HIPFEM_EINVAL, "solve: null argument");
    HIPFEM_REQUIRE(s->ctx == A->ctx, HIPFEM_EINVAL, "solve: solver and matrix belong to different contexts");
    DeviceGuard g(s->ctx->device);
    return solve_dev(s, A, db, dx, reason, iters, achieved);
}

// SparseLinearSystemNM::solve(A, B, X) (sparselinsystemnm.C:47-70): the
// columns of B (n x nrhs, column-major, host) in turn, X holding the initial
// guesses and receiving the solutions, stopping at the first column that
// does not converge.  With precision 2, the diagonal preconditioner and zero
// initial guesses the columns go six at a time through the batched solver
// (one block: its bulk phase over the whole matrix reads every matrix entry
// once per iteration for six columns; FP32 PCG with FP64 reliable updates,
// the same method as precision 2).  *iters and *achieved are the largest
// over the columns.
extern "C" int hippcg_solve_multi(hippcg_solver *s, hipcsr_mtx *A, int nrhs, const double *B, double *X,
                                  int *reason, int *iters, double *achieved)
{
    HIPFEM_REQUIRE(s && A && reason && iters && achieved && nrhs >= 0 && (nrhs == 0 || (B && X)), HIPFEM_EINVAL,
                   "solve_multi: bad arguments");
    HIPFEM_REQUIRE(A->has_structure, HIPFEM_EINVAL, "solve_multi: matrix has no structure");
    const int n = A->n;
    *reason = HIPFEM_CR_CONVERGED;
    *iters = 0;
    *achieved = 0.0;
    bool zero = true;
    for (size_t i = 0; i < (size_t)n * nrhs && zero; i++) zero = X[i] == 0.0;
    if (s->precision != 2 || s->precond != 1 || !zero || n == 0) {
        for (int j = 0; j < nrhs; j++) {
            int it = 0;
            double ach = 0.0;
            int rc = hippcg_solve(s, A, B + (size_t)j * n, X + (size_t)j * n, reason, &it, &ach);
            if (rc) return rc;
            *iters = std::max(*iters, it);
            *achieved = std::max(*achieved, ach);
            if (*reason != HIPFEM_CR_CONVERGED) return 0;
        }
        return 0;
    }
    DeviceGuard g(s->ctx->device);
    const int32_t rows[2] = {0, n};
    hfem_batch *bt = nullptr;
    int rc = hfem_batch_create(A, 1, rows, &bt);
    if (rc) return rc;
    double *db = nullptr, *dx = nullptr;
    if (hipMalloc(&db, sizeof(double) * 6 * (size_t)n) != hipSuccess ||
        hipMalloc(&dx, sizeof(double) * 6 * (size_t)n) != hipSuccess) {
        if (db) (void)hipFree(db);
        hfem_batch_destroy(bt);
        set_error("solve_multi: out of device memory");
        return HIPFEM_ENOMEM;
    }
    rc = hfem_batch_set_options(bt, s->tol, s->delta, s->maxit);
    if (!rc) rc = hfem_batch_set_split(bt, 0);
    std::vector<double> hb(6 * (size_t)n), hx(6 * (size_t)n);
    hipStream_t st = s->ctx->stream;
    for (int j0 = 0; j0 < nrhs && !rc; j0 += 6) {
        const int m = std::min(6, nrhs - j0);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 6; k++) hb[(size_t)i * 6 + k] = k < m ? B[(size_t)(j0 + k) * n + i] : 0.0;
        int it = 0, upd = 0;
        double res = 0.0;
        if (hipMemcpyAsync(db, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("solve_multi: upload failed");
            rc = HIPFEM_EHIP;
            break;
        }
        rc = hfem_batch_solve_dev(bt, db, dx, &it, &upd, &res);
        if (rc) break;
        if (hipMemcpyAsync(hx.data(), dx, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("solve_multi: download failed");
            rc = HIPFEM_EHIP;
            break;
        }
        for (int k = 0; k < m; k++)
            for (int i = 0; i < n; i++) X[(size_t)(j0 + k) * n + i] = hx[(size_t)i * 6 + k];
        *iters = std::max(*iters, it);
        *achieved = std::max(*achieved, std::fabs(res));
        if (res < 0.0) {
            *reason = HIPFEM_CR_DIVERGED_ITS;
            break;
        }
    }
    (void)hipFree(db);
    (void)hipFree(dx);
    hfem_batch_destroy(bt);
    return rc;
}

extern "C" int hippcg_solve(hippcg_solver *s, hipcsr_mtx *A, const double *b, double *x, int *reason, int *iters,
                            double *achieved)
{
    HIPFEM_REQUIRE(s && A && b && x && reason && iters && achieved, HIPFEM_EINVAL, "solve: null argument");
    HIPFEM_REQUIRE(s->ctx == A->ctx, HIPFEM_EINVAL, "solve: solver and matrix belong to different contexts");
    DeviceGuard g(s->ctx->device);
    int rc = solver_alloc(s, A->n);
    if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    if (A->n) {
        HIPFEM_CHECK(hipMemcpyAsync(s->hb, b, sizeof(double) * A->n, hipMemcpyHostToDevice, st));
        HIPFEM_CHECK(hipMemcpyAsync(s->hx, x, sizeof(double) * A->n, hipMemcpyHostToDevice, st));
    }
    rc = solve_dev(s, A, s->hb, s->hx, reason, iters, achieved);
    if (!rc && A->n) {
        hipError_t e = hipMemcpyAsync(x, s->hx, sizeof(double) * A->n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_error("solve: copy back failed: %s", hipGetErrorString(e));
            rc = HIPFEM_EHIP;
        }
    }
    return rc;
}
