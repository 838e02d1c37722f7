// hfem_batch: many independent SPD systems solved in one persistent launch
// (the RVE batch: block-diagonal stiffness from a batched assembly, six
// right-hand sides per block, one per unit macro strain).
//
// Layout: the blocks are consecutive row ranges [rows[b], rows[b+1]) of one
// CSR matrix whose entries stay inside their block.  Vectors of the six
// right-hand sides are interleaved per row ([row][6]) so that a column's six
// entries are one 24-byte (FP32) or 48-byte (FP64) gather.
//
// Solver per block (one workgroup owns a block from start to finish; blocks
// are taken from a device queue in descending size, so large blocks start
// first): Jacobi PCG on FP32 values and vectors with FP64 dot accumulation
// and FP64 solution; whenever a column's recursive residual has dropped by
// delta since its last update (or meets the tolerance), every column's
// residual is replaced by the FP64 true residual b - A x (FP64 values) and
// the search directions are kept (reliable updates, as in hippcg precision
// 2).  Convergence is decided on the true residual of every column.
//
// Shape: LPR lanes per row (8, 16, 32 or 64) and the workgroup size (256,
// 512 or 1024) are template parameters; hfem_batch_autotune times them on
// median blocks and keeps the fastest.
//
// Large blocks: one workgroup streams a block at a fraction of one CU's
// bandwidth, so a block much larger than the batch's mean share would set
// the whole launch's length.  Blocks above a split size (stored entries)
// are therefore solved first, together, in a bulk-synchronous phase: every
// launch covers all their rows in tiles of TILE_ROWS rows (one workgroup
// each, the whole GPU; the products use a 3x3 node-block form when the
// pattern has one), per-tile partial dots are summed per block by a
// reduction launch, and the same per-column scalars, reliable updates and
// stopping rule as the persistent kernel live in a BulkBlock record per
// block.  Finished blocks drop out of the tile list when the host next looks
// (every BULK_CHECK iterations); until then their tiles return at once.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hipfem_internal.h"

namespace hipfem {

constexpr int NRHS = 6;

struct BatchCtl {
    int next;          // queue head
    int pad[15];
};

// per-workgroup block-wide sum of NV doubles (one value per thread each);
// sh holds NV * (BS / 64) doubles; results in out[NV] for every thread
template <int BS, int NV>
__device__ __forceinline__ void block_sums(double v[NV], double *sh, double *out)
{
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (l == 0) sh[k * (BS / 64) + w] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int i = 0; i < BS / 64; i++) s += sh[threadIdx.x * (BS / 64) + i];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// y[row][k] = sum_j A(row, j) x[j][k] for the rows of one block, LPR lanes per row
template <int BS, int LPR, typename MT, typename XT, typename YT>
__device__ __forceinline__ void block_spmv(int r0, int r1, const int32_t *__restrict__ rowptr,
                                           const int32_t *__restrict__ col, const MT *__restrict__ val,
                                           const XT *__restrict__ x, YT *__restrict__ y)
{
    const int lane = threadIdx.x % LPR;
    for (int row = r0 + threadIdx.x / LPR; row < r1; row += BS / LPR) {
        const int s = rowptr[row], e = rowptr[row + 1];
        float acc32[NRHS];
        double acc64[NRHS];
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            acc32[k] = 0.f;
            acc64[k] = 0.0;
        }
        for (int i = s + lane; i < e; i += LPR) {
            const int c = col[i];
            const MT a = val[i];
            const XT *xc = x + (size_t)c * NRHS;
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                if constexpr (sizeof(MT) == 4)
                    acc32[k] = fmaf(a, (float)xc[k], acc32[k]);
                else
                    acc64[k] = fma((double)a, (double)xc[k], acc64[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            if constexpr (sizeof(MT) == 4) {
                float v = acc32[k];
#pragma unroll
                for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                acc32[k] = v;
            } else {
                double v = acc64[k];
#pragma unroll
                for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                acc64[k] = v;
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < NRHS; k++)
                y[(size_t)row * NRHS + k] = sizeof(MT) == 4 ? (YT)acc32[k] : (YT)acc64[k];
        }
    }
}

template <int BS, int LPR>
__global__ __launch_bounds__(BS) void k_batch_pcg(int nblocks, const int *__restrict__ order,
                                                  const int32_t *__restrict__ rows, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col, const double *__restrict__ val64,
                                                  const float *__restrict__ val32, const float *__restrict__ dinv,
                                                  const double *__restrict__ b, double *__restrict__ x,
                                                  float *__restrict__ r, float *__restrict__ z, float *__restrict__ p,
                                                  float *__restrict__ q, double *__restrict__ r64, double tol,
                                                  double delta, int maxit, BatchCtl *ctl, int *__restrict__ iters,
                                                  double *__restrict__ resid, int *__restrict__ updates)
{
    __shared__ double sh[2 * NRHS * (BS / 64)];
    __shared__ double red[2 * NRHS];
    __shared__ int s_blk;
    for (;;) {
        if (threadIdx.x == 0) s_blk = atomicAdd(&ctl->next, 1);
        __syncthreads();
        const int qi = s_blk;
        __syncthreads();
        if (qi >= nblocks) return;
        const int blk = order[qi];
        const int r0 = rows[blk], r1 = rows[blk + 1];
        // x = 0: r = b, z = D^-1 r, rho = r.z, |b|
        double part[2 * NRHS];
#pragma unroll
        for (int k = 0; k < 2 * NRHS; k++) part[k] = 0.0;
        for (int i = r0 + threadIdx.x; i < r1; i += BS) {
            const float d = dinv[i];
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                const size_t o = (size_t)i * NRHS + k;
                const double bi = b[o];
                x[o] = 0.0;
                const float ri = (float)bi;
                r[o] = ri;
                const float zi = d * ri;
                z[o] = zi;
                p[o] = zi;
                part[k] += (double)ri * (double)zi;
                part[NRHS + k] += bi * bi;
            }
        }
        block_sums<BS, 2 * NRHS>(part, sh, red);
        double rho[NRHS], normb[NRHS], rmax[NRHS], res[NRHS];
        bool conv[NRHS];
        bool all = true;
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            rho[k] = red[k];
            normb[k] = sqrt(red[NRHS + k]);
            if (normb[k] == 0.0) normb[k] = 1.0;
            res[k] = sqrt(red[NRHS + k]) / normb[k];
            rmax[k] = res[k];
            conv[k] = res[k] <= tol;
            all = all && conv[k];
        }
        int it = 0, nupd = 0;
        while (!all && it < maxit) {
            it++;
            // q = A p (FP32), p.q
            block_spmv<BS, LPR, float, float, float>(r0, r1, rowptr, col, val32, p, q);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 2 * NRHS; k++) part[k] = 0.0;
            for (int i = r0 + threadIdx.x; i < r1; i += BS)
#pragma unroll
                for (int k = 0; k < NRHS; k++) {
                    const size_t o = (size_t)i * NRHS + k;
                    part[k] += (double)p[o] * (double)q[o];
                }
            block_sums<BS, NRHS>(part, sh, red);
            float alpha[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; k++) alpha[k] = (!conv[k] && red[k] > 0.0) ? (float)(rho[k] / red[k]) : 0.f;
            // x += alpha p (FP64), r -= alpha q, z = D^-1 r; r.z, r.r
#pragma unroll
            for (int k = 0; k < 2 * NRHS; k++) part[k] = 0.0;
            for (int i = r0 + threadIdx.x; i < r1; i += BS) {
                const float d = dinv[i];
#pragma unroll
                for (int k = 0; k < NRHS; k++) {
                    const size_t o = (size_t)i * NRHS + k;
                    x[o] += (double)alpha[k] * (double)p[o];
                    const float ri = fmaf(-alpha[k], q[o], r[o]);
                    r[o] = ri;
                    const float zi = d * ri;
                    z[o] = zi;
                    part[k] += (double)ri * (double)zi;
                    part[NRHS + k] += (double)ri * (double)ri;
                }
            }
            block_sums<BS, 2 * NRHS>(part, sh, red);
            bool refresh = false;
            float beta[NRHS];
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                res[k] = sqrt(red[NRHS + k]) / normb[k];
                if (!conv[k] && (res[k] <= tol || res[k] < delta * rmax[k])) refresh = true;
                beta[k] = conv[k] ? 0.f : (float)(red[k] / rho[k]);
            }
            if (refresh) {
                // reliable update: r64 = b - A x (FP64 values), every column
                nupd++;
                block_spmv<BS, LPR, double, double, double>(r0, r1, rowptr, col, val64, x, r64);
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 2 * NRHS; k++) part[k] = 0.0;
                for (int i = r0 + threadIdx.x; i < r1; i += BS) {
                    const float d = dinv[i];
#pragma unroll
                    for (int k = 0; k < NRHS; k++) {
                        const size_t o = (size_t)i * NRHS + k;
                        const double rt = b[o] - r64[o];
                        const float ri = (float)rt;
                        r[o] = ri;
                        const float zi = d * ri;
                        z[o] = zi;
                        part[k] += (double)ri * (double)zi;
                        part[NRHS + k] += rt * rt;
                    }
                }
                block_sums<BS, 2 * NRHS>(part, sh, red);
                all = true;
#pragma unroll
                for (int k = 0; k < NRHS; k++) {
                    res[k] = sqrt(red[NRHS + k]) / normb[k];
                    conv[k] = res[k] <= tol;
                    all = all && conv[k];
                    beta[k] = conv[k] ? 0.f : (float)(red[k] / rho[k]);
                    rho[k] = red[k];
                    rmax[k] = res[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < NRHS; k++)
                    if (!conv[k]) rho[k] = red[k];
            }
            if (all) break;
            // p = z + beta p (converged columns keep p = z + 0 p: alpha is 0 for them)
            for (int i = r0 + threadIdx.x; i < r1; i += BS)
#pragma unroll
                for (int k = 0; k < NRHS; k++) {
                    const size_t o = (size_t)i * NRHS + k;
                    p[o] = fmaf(beta[k], p[o], z[o]);
                }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            iters[blk] = it;
            updates[blk] = nupd;
            double mx = 0.0;
            for (int k = 0; k < NRHS; k++) mx = fmax(mx, res[k]);
            resid[blk] = all ? mx : -mx;   // negative: not converged within maxit
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ bulk phase
constexpr int BULK_TILE = 256;   // threads per workgroup
constexpr int TILE_ROWS = 255;   // rows per tile: 85 whole nodes, one row per thread
constexpr int BULK_CHECK = 16;   // iterations between host looks at the done flags
constexpr int NPART = 2 * NRHS;  // partial sums per tile

struct BulkBlock {
    double rho[NRHS], normb[NRHS], rmax[NRHS], res[NRHS];
    float alpha[NRHS], beta[NRHS];
    int conv;      // bit k: column k meets the tolerance
    int it, nupd;
    int refresh;   // 1: this iteration replaces the residuals by b - A x
    int done;      // 1: every column converged, or maxit reached
    int blk;       // block number
};

struct BulkArgs {
    const int4 *tiles;       // {first row, end row, bulk index, 0}; node-aligned in the node-block form
    const int *act;          // bulk indices of the blocks in the tile list
    const int *act_t0;       // their tile ranges [act_t0[a], act_t0[a + 1])
    BulkBlock *st;
    const int32_t *rowptr, *col;
    const int32_t *bptr, *bcol;  // node-block view (node rows, block columns) when the pattern has one
    const float4 *vwide;         // wide form: FP32 values block-major, 12 floats per block
    const double *val64;
    const float *val32, *dinv;
    const double *b;
    double *x;
    float *r, *z, *p, *q;
    double *part;            // [tile][NPART]
    double tol, delta;
    int maxit;
    int *iters, *updates;
    double *resid;
};

// x = 0, r = b, z = p = D^-1 r; partials r.z and b.b
__global__ __launch_bounds__(BULK_TILE) void k_bulk_init(BulkArgs A)
{
    __shared__ double sh[NPART * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    double v[NPART];
#pragma unroll
    for (int k = 0; k < NPART; k++) v[k] = 0.0;
    const int i = t.x + threadIdx.x;
    if (i < t.y) {
        const float d = A.dinv[i];
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            const size_t o = (size_t)i * NRHS + k;
            const double bi = A.b[o];
            A.x[o] = 0.0;
            const float ri = (float)bi;
            A.r[o] = ri;
            const float zi = d * ri;
            A.z[o] = zi;
            A.p[o] = zi;
            v[k] = (double)ri * (double)zi;
            v[NRHS + k] = bi * bi;
        }
    }
    block_sums<BULK_TILE, NPART>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// q = A p (FP32) for the tile's rows, LPR lanes per row; partials p.q
template <int LPR>
__global__ __launch_bounds__(BULK_TILE) void k_bulk_spmv(BulkArgs A)
{
    __shared__ double sh[NRHS * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    if (A.st[t.z].done) return;
    double v[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; k++) v[k] = 0.0;
    const int lane = threadIdx.x % LPR;
    for (int row = t.x + threadIdx.x / LPR; row < t.y; row += BULK_TILE / LPR) {
        const int s = A.rowptr[row], e = A.rowptr[row + 1];
        float acc[NRHS];
#pragma unroll
        for (int k = 0; k < NRHS; k++) acc[k] = 0.f;
        for (int i = s + lane; i < e; i += LPR) {
            const float a = A.val32[i];
            const float2 *xc = reinterpret_cast<const float2 *>(A.p + (size_t)A.col[i] * NRHS);
            const float2 x0 = xc[0], x1 = xc[1], x2 = xc[2];
            acc[0] = fmaf(a, x0.x, acc[0]);
            acc[1] = fmaf(a, x0.y, acc[1]);
            acc[2] = fmaf(a, x1.x, acc[2]);
            acc[3] = fmaf(a, x1.y, acc[3]);
            acc[4] = fmaf(a, x2.x, acc[4]);
            acc[5] = fmaf(a, x2.y, acc[5]);
        }
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            float w = acc[k];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
            acc[k] = w;
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                const size_t o = (size_t)row * NRHS + k;
                A.q[o] = acc[k];
                v[k] += (double)A.p[o] * (double)acc[k];
            }
        }
    }
    block_sums<BULK_TILE, NRHS>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// x += alpha p (FP64), r -= alpha q, z = D^-1 r; partials r.z and r.r
__global__ __launch_bounds__(BULK_TILE) void k_bulk_update(BulkArgs A)
{
    __shared__ double sh[NPART * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    const BulkBlock &S = A.st[t.z];
    if (S.done) return;
    double v[NPART];
#pragma unroll
    for (int k = 0; k < NPART; k++) v[k] = 0.0;
    const int i = t.x + threadIdx.x;
    if (i < t.y) {
        const float d = A.dinv[i];
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            const size_t o = (size_t)i * NRHS + k;
            const float al = S.alpha[k];
            A.x[o] += (double)al * (double)A.p[o];
            const float ri = fmaf(-al, A.q[o], A.r[o]);
            A.r[o] = ri;
            const float zi = d * ri;
            A.z[o] = zi;
            v[k] = (double)ri * (double)zi;
            v[NRHS + k] = (double)ri * (double)ri;
        }
    }
    block_sums<BULK_TILE, NPART>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// reliable update for the tiles of blocks that asked for one:
// r = b - A x (FP64 values and solution), z = D^-1 r; partials r.z and |r|^2
template <int LPR>
__global__ __launch_bounds__(BULK_TILE) void k_bulk_refresh(BulkArgs A)
{
    __shared__ double sh[NPART * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    const BulkBlock &S = A.st[t.z];
    if (S.done || !S.refresh) return;
    double v[NPART];
#pragma unroll
    for (int k = 0; k < NPART; k++) v[k] = 0.0;
    const int lane = threadIdx.x % LPR;
    for (int row = t.x + threadIdx.x / LPR; row < t.y; row += BULK_TILE / LPR) {
        const int s = A.rowptr[row], e = A.rowptr[row + 1];
        double acc[NRHS];
#pragma unroll
        for (int k = 0; k < NRHS; k++) acc[k] = 0.0;
        for (int i = s + lane; i < e; i += LPR) {
            const double a = A.val64[i];
            const double2 *xc = reinterpret_cast<const double2 *>(A.x + (size_t)A.col[i] * NRHS);
            const double2 x0 = xc[0], x1 = xc[1], x2 = xc[2];
            acc[0] = fma(a, x0.x, acc[0]);
            acc[1] = fma(a, x0.y, acc[1]);
            acc[2] = fma(a, x1.x, acc[2]);
            acc[3] = fma(a, x1.y, acc[3]);
            acc[4] = fma(a, x2.x, acc[4]);
            acc[5] = fma(a, x2.y, acc[5]);
        }
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            double w = acc[k];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
            acc[k] = w;
        }
        if (lane == 0) {
            const float d = A.dinv[row];
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                const size_t o = (size_t)row * NRHS + k;
                const double rt = A.b[o] - acc[k];
                const float ri = (float)rt;
                A.r[o] = ri;
                const float zi = d * ri;
                A.z[o] = zi;
                v[k] += (double)ri * (double)zi;
                v[NRHS + k] += rt * rt;
            }
        }
    }
    block_sums<BULK_TILE, NPART>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// Node-block forms of the two products, for patterns whose rows come in
// fully free node triples (every RVE: boundary nodes are prescribed whole).
// Node row I holds nbk 3x3 blocks; block k's column node is bcol[bo + k] and
// its row r values sit at val[9 bo + 3 r nbk + 3 k + (0..2)] (the CSR arrays
// themselves).  The six right-hand sides of a node are 18 consecutive
// entries of the interleaved vectors, so a block gathers 72 (FP32) or 144
// (FP64) contiguous bytes and the column index is read once per 9 values.
// LPR lanes share a node row; after the lane reduction lane j < 18 owns
// output (row j / 6, column j % 6).
template <int LPR, typename VT, typename MT>
__device__ __forceinline__ void nb_row6(int bo, int nbk, int lane, const int32_t *__restrict__ bcol,
                                        const MT *__restrict__ val, const VT *__restrict__ x, VT acc[3][NRHS])
{
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < NRHS; k++) acc[r][k] = VT(0);
    const MT *v0 = val + 9 * (size_t)bo;
    for (int k = lane; k < nbk; k += LPR) {
        const VT *xb = x + (size_t)bcol[bo + k] * (3 * NRHS);
        VT xv[3][NRHS];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < NRHS; j += 2) {
                if constexpr (sizeof(VT) == 4) {
                    const float2 t = *reinterpret_cast<const float2 *>(xb + c * NRHS + j);
                    xv[c][j] = t.x;
                    xv[c][j + 1] = t.y;
                } else {
                    const double2 t = *reinterpret_cast<const double2 *>(xb + c * NRHS + j);
                    xv[c][j] = t.x;
                    xv[c][j + 1] = t.y;
                }
            }
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const MT *vr = v0 + 3 * (size_t)r * nbk + 3 * k;
            const VT a0 = (VT)vr[0], a1 = (VT)vr[1], a2 = (VT)vr[2];
#pragma unroll
            for (int j = 0; j < NRHS; j++) acc[r][j] = fma(a2, xv[2][j], fma(a1, xv[1][j], fma(a0, xv[0][j], acc[r][j])));
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int j = 0; j < NRHS; j++) {
            VT w = acc[r][j];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
            acc[r][j] = w;
        }
}

// node-block product with the FP32 values of every 3x3 block stored
// together, padded to 12 floats (three 16-byte loads per block instead of
// nine 4-byte ones; vw from k_bulk_widen)
template <int LPR>
__device__ __forceinline__ void nb_row6_wide(int bo, int nbk, int lane, const int32_t *__restrict__ bcol,
                                             const float4 *__restrict__ vw, const float *__restrict__ x,
                                             float acc[3][NRHS])
{
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < NRHS; k++) acc[r][k] = 0.f;
    for (int k = lane; k < nbk; k += LPR) {
        const float *xb = x + (size_t)bcol[bo + k] * (3 * NRHS);
        const float4 *v = vw + 3 * ((size_t)bo + k);
        const float4 v0 = v[0], v1 = v[1], v2 = v[2];
        const float a[3][3] = {{v0.x, v0.y, v0.z}, {v0.w, v1.x, v1.y}, {v1.z, v1.w, v2.x}};
        float xv[3][NRHS];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < NRHS; j += 2) {
                const float2 t = *reinterpret_cast<const float2 *>(xb + c * NRHS + j);
                xv[c][j] = t.x;
                xv[c][j + 1] = t.y;
            }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int j = 0; j < NRHS; j++)
                acc[r][j] = fmaf(a[r][2], xv[2][j], fmaf(a[r][1], xv[1][j], fmaf(a[r][0], xv[0][j], acc[r][j])));
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int j = 0; j < NRHS; j++) {
            float w = acc[r][j];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
            acc[r][j] = w;
        }
}

// value of output j (0..17) of a reduced node row, selected without
// dynamic register indexing
template <typename VT>
__device__ __forceinline__ VT pick18(const VT acc[3][NRHS], int j)
{
    VT v = VT(0);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < NRHS; c++) v = (j == r * NRHS + c) ? acc[r][c] : v;
    return v;
}

// q = A p (FP32) in the node-block form; partials p.q
template <int LPR>
__global__ __launch_bounds__(BULK_TILE) void k_bulk_spmv_nb(BulkArgs A)
{
    static_assert(LPR >= 4, "18 outputs per node row need at least 4 lanes (stores spread over the lanes)");
    __shared__ double sh[NRHS * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    if (A.st[t.z].done) return;
    double v[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; k++) v[k] = 0.0;
    const int lane = threadIdx.x % LPR;
    for (int I = t.x / 3 + threadIdx.x / LPR; I < t.y / 3; I += BULK_TILE / LPR) {
        const int bo = A.bptr[I], nbk = A.bptr[I + 1] - bo;
        float acc[3][NRHS];
        nb_row6<LPR>(bo, nbk, lane, A.bcol, A.val32, A.p, acc);
        for (int j = lane; j < 3 * NRHS; j += LPR) {
            const size_t o = (size_t)I * (3 * NRHS) + j;
            const float qj = pick18(acc, j);
            A.q[o] = qj;
            const double pq = (double)A.p[o] * (double)qj;
#pragma unroll
            for (int k = 0; k < NRHS; k++) v[k] += (j % NRHS == k) ? pq : 0.0;
        }
    }
    block_sums<BULK_TILE, NRHS>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// q = A p (FP32) in the wide node-block form; partials p.q
template <int LPR>
__global__ __launch_bounds__(BULK_TILE) void k_bulk_spmv_nbw(BulkArgs A)
{
    __shared__ double sh[NRHS * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    if (A.st[t.z].done) return;
    double v[NRHS];
#pragma unroll
    for (int k = 0; k < NRHS; k++) v[k] = 0.0;
    const int lane = threadIdx.x % LPR;
    for (int I = t.x / 3 + threadIdx.x / LPR; I < t.y / 3; I += BULK_TILE / LPR) {
        const int bo = A.bptr[I], nbk = A.bptr[I + 1] - bo;
        float acc[3][NRHS];
        nb_row6_wide<LPR>(bo, nbk, lane, A.bcol, A.vwide, A.p, acc);
        for (int j = lane; j < 3 * NRHS; j += LPR) {
            const size_t o = (size_t)I * (3 * NRHS) + j;
            const float qj = pick18(acc, j);
            A.q[o] = qj;
            const double pq = (double)A.p[o] * (double)qj;
#pragma unroll
            for (int k = 0; k < NRHS; k++) v[k] += (j % NRHS == k) ? pq : 0.0;
        }
    }
    block_sums<BULK_TILE, NRHS>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// block-major padded copy of the FP32 values of a node-blocked CSR: block
// bo + k of node row I gets row r, column c at vw[12 (bo + k) + 3 r + c]
__global__ void k_bulk_widen(int nbr, const int32_t *__restrict__ bptr, const float *__restrict__ val32,
                             float *__restrict__ vw)
{
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < nbr; I += gridDim.x * blockDim.x) {
        const int bo = bptr[I], nbk = bptr[I + 1] - bo;
        for (int k = 0; k < nbk; k++)
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++)
                    vw[12 * ((size_t)bo + k) + 3 * r + c] = val32[9 * (size_t)bo + 3 * (size_t)r * nbk + 3 * k + c];
                if (r == 2)
                    for (int c = 9; c < 12; c++) vw[12 * ((size_t)bo + k) + c] = 0.f;
            }
    }
}

// reliable update in the node-block form: r = b - A x (FP64), z = D^-1 r
template <int LPR>
__global__ __launch_bounds__(BULK_TILE) void k_bulk_refresh_nb(BulkArgs A)
{
    __shared__ double sh[NPART * (BULK_TILE / 64)];
    const int4 t = A.tiles[blockIdx.x];
    const BulkBlock &S = A.st[t.z];
    if (S.done || !S.refresh) return;
    double v[NPART];
#pragma unroll
    for (int k = 0; k < NPART; k++) v[k] = 0.0;
    const int lane = threadIdx.x % LPR;
    for (int I = t.x / 3 + threadIdx.x / LPR; I < t.y / 3; I += BULK_TILE / LPR) {
        const int bo = A.bptr[I], nbk = A.bptr[I + 1] - bo;
        double acc[3][NRHS];
        nb_row6<LPR>(bo, nbk, lane, A.bcol, A.val64, A.x, acc);
        for (int j = lane; j < 3 * NRHS; j += LPR) {
            const size_t o = (size_t)I * (3 * NRHS) + j;
            const double rt = A.b[o] - pick18(acc, j);
            const float ri = (float)rt;
            A.r[o] = ri;
            const float zi = A.dinv[3 * I + j / NRHS] * ri;
            A.z[o] = zi;
            const double rz = (double)ri * (double)zi, rr = rt * rt;
#pragma unroll
            for (int k = 0; k < NRHS; k++) {
                v[k] += (j % NRHS == k) ? rz : 0.0;
                v[NRHS + k] += (j % NRHS == k) ? rr : 0.0;
            }
        }
    }
    block_sums<BULK_TILE, NPART>(v, sh, A.part + (size_t)blockIdx.x * NPART);
}

// p = z + beta p
__global__ __launch_bounds__(BULK_TILE) void k_bulk_pupdate(BulkArgs A)
{
    const int4 t = A.tiles[blockIdx.x];
    const BulkBlock &S = A.st[t.z];
    if (S.done) return;
    const int i = t.x + threadIdx.x;
    if (i < t.y) {
#pragma unroll
        for (int k = 0; k < NRHS; k++) {
            const size_t o = (size_t)i * NRHS + k;
            A.p[o] = fmaf(S.beta[k], A.p[o], A.z[o]);
        }
    }
}

enum BulkPhase { BP_INIT, BP_ALPHA, BP_BETA, BP_REFRESH };

__device__ void bulk_finish(const BulkArgs &A, BulkBlock &S, bool all)
{
    double mx = 0.0;
    for (int k = 0; k < NRHS; k++) mx = fmax(mx, S.res[k]);
    A.iters[S.blk] = S.it;
    A.updates[S.blk] = S.nupd;
    A.resid[S.blk] = all ? mx : -mx;
    S.done = 1;
}

// one workgroup per block of the tile list: sum the block's tile partials
// and advance its scalars (the same rules as k_batch_pcg)
template <int PHASE>
__global__ __launch_bounds__(256) void k_bulk_reduce(BulkArgs A)
{
    __shared__ double sh[NPART * 4];
    __shared__ double red[NPART];
    BulkBlock &S = A.st[A.act[blockIdx.x]];
    if (S.done) return;
    if (PHASE == BP_REFRESH && !S.refresh) return;
    constexpr int NV = PHASE == BP_ALPHA ? NRHS : NPART;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    for (int ti = A.act_t0[blockIdx.x] + threadIdx.x; ti < A.act_t0[blockIdx.x + 1]; ti += 256)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += A.part[(size_t)ti * NPART + k];
    block_sums<256, NV>(v, sh, red);
    if (threadIdx.x != 0) return;
    if (PHASE == BP_INIT) {
        bool all = true;
        S.conv = 0;
        for (int k = 0; k < NRHS; k++) {
            S.rho[k] = red[k];
            const double nb = sqrt(red[NRHS + k]);
            S.normb[k] = nb == 0.0 ? 1.0 : nb;
            S.res[k] = nb / S.normb[k];
            S.rmax[k] = S.res[k];
            if (S.res[k] <= A.tol) S.conv |= 1 << k;
            all = all && (S.conv >> k & 1);
        }
        S.it = 0;
        S.nupd = 0;
        S.refresh = 0;
        if (all || A.maxit <= 0) bulk_finish(A, S, all);
    } else if (PHASE == BP_ALPHA) {
        S.it++;
        for (int k = 0; k < NRHS; k++)
            S.alpha[k] = (!(S.conv >> k & 1) && red[k] > 0.0) ? (float)(S.rho[k] / red[k]) : 0.f;
    } else if (PHASE == BP_BETA) {
        bool refresh = false;
        for (int k = 0; k < NRHS; k++) {
            S.res[k] = sqrt(red[NRHS + k]) / S.normb[k];
            if (!(S.conv >> k & 1) && (S.res[k] <= A.tol || S.res[k] < A.delta * S.rmax[k])) refresh = true;
        }
        S.refresh = refresh;
        if (!refresh) {
            for (int k = 0; k < NRHS; k++) {
                const bool c = S.conv >> k & 1;
                S.beta[k] = c ? 0.f : (float)(red[k] / S.rho[k]);
                if (!c) S.rho[k] = red[k];
            }
            if (S.it >= A.maxit) bulk_finish(A, S, false);
        }
    } else {
        S.nupd++;
        bool all = true;
        S.conv = 0;
        for (int k = 0; k < NRHS; k++) {
            S.res[k] = sqrt(red[NRHS + k]) / S.normb[k];
            const bool c = S.res[k] <= A.tol;
            if (c) S.conv |= 1 << k;
            all = all && c;
            S.beta[k] = c ? 0.f : (float)(red[k] / S.rho[k]);
            S.rho[k] = red[k];
            S.rmax[k] = S.res[k];
        }
        S.refresh = 0;
        if (all || S.it >= A.maxit) bulk_finish(A, S, all);
    }
}

typedef void (*bulk_kernel_t)(BulkArgs);
struct BulkShape {
    int lpr;
    bulk_kernel_t spmv, refresh, spmv_nb, refresh_nb, spmv_nbw;
};
static const BulkShape kBulkShapes[] = {
    {4, k_bulk_spmv<4>, k_bulk_refresh<4>, k_bulk_spmv_nb<4>, k_bulk_refresh_nb<4>, k_bulk_spmv_nbw<4>},
    {8, k_bulk_spmv<8>, k_bulk_refresh<8>, k_bulk_spmv_nb<8>, k_bulk_refresh_nb<8>, k_bulk_spmv_nbw<8>},
    {16, k_bulk_spmv<16>, k_bulk_refresh<16>, k_bulk_spmv_nb<16>, k_bulk_refresh_nb<16>, k_bulk_spmv_nbw<16>},
};
constexpr int NBULK = sizeof(kBulkShapes) / sizeof(kBulkShapes[0]);

__global__ void k_batch_dinv(int n, const int32_t *__restrict__ diag, const double *__restrict__ val,
                             float *__restrict__ dinv, int *__restrict__ zero_row)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i] >= 0 ? val[diag[i]] : 0.0;
    if (d == 0.0) atomicMin(zero_row, i);
    dinv[i] = (float)(1.0 / d);
}

typedef void (*batch_kernel_t)(int, const int *, const int32_t *, const int32_t *, const int32_t *, const double *,
                               const float *, const float *, const double *, double *, float *, float *, float *,
                               float *, double *, double, double, int, BatchCtl *, int *, double *, int *);

struct BatchShape {
    int bs, lpr;
    batch_kernel_t fn;
};

static const BatchShape kShapes[] = {
    {256, 8, k_batch_pcg<256, 8>},     {256, 16, k_batch_pcg<256, 16>},   {256, 32, k_batch_pcg<256, 32>},
    {512, 8, k_batch_pcg<512, 8>},     {512, 16, k_batch_pcg<512, 16>},   {512, 32, k_batch_pcg<512, 32>},
    {1024, 16, k_batch_pcg<1024, 16>}, {1024, 32, k_batch_pcg<1024, 32>}, {1024, 64, k_batch_pcg<1024, 64>},
};
constexpr int NSHAPES = sizeof(kShapes) / sizeof(kShapes[0]);

}  // namespace hipfem

using namespace hipfem;

struct hfem_batch {
    hipfem_ctx *ctx = nullptr;
    hipcsr_mtx *A = nullptr;
    int nblocks = 0, n = 0;
    std::vector<int32_t> h_rows;
    std::vector<long long> h_nnz;   // stored entries per block
    std::vector<int> h_sorted;      // blocks by descending size
    int32_t *d_rows = nullptr;
    int *d_order = nullptr;         // queue of the persistent kernel (blocks at or below the split)
    float *d_dinv = nullptr;
    double *d_x = nullptr, *d_r64 = nullptr;
    float *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_q = nullptr;
    BatchCtl *d_ctl = nullptr;
    int *d_iters = nullptr, *d_upd = nullptr, *d_zero = nullptr;
    double *d_res = nullptr;
    int shape = 4;              // index into kShapes (512 threads, 16 lanes per row)
    int grid = 0;               // persistent workgroups (0: one per block up to the device's capacity)
    double tol = 1e-8, delta = 0.1;
    int maxit = 20000;
    // partition between the bulk phase and the persistent kernel
    long long split = -1;       // stored entries above which a block goes to the bulk phase (-1: automatic)
    long long split_used = -1;
    int part_shape = -1;        // shape the automatic split was computed for
    bool part_dirty = true;
    std::vector<int> h_bulk;    // bulk blocks (descending size)
    int nsmall = 0;
    // bulk phase
    int bulk_shape = 1;         // index into kBulkShapes
    bool nodeblock_ok = false;  // node-block kernels apply to this matrix
    bool nodeblock = true;      // use them when they apply
    bool wide = false;          // node-block products from the block-major padded FP32 copy
    float *d_vwide = nullptr;   // that copy (12 floats per block), built on first use
    long vwide_version = -1;
    int tiles_cap = 0;
    BulkBlock *d_bst = nullptr;
    int4 *d_tiles = nullptr;
    int *d_act = nullptr, *d_act_t0 = nullptr;
    double *d_part = nullptr;
    std::vector<BulkBlock> h_bst;
    std::vector<int4> h_tiles;
    std::vector<int> h_act, h_act_t0;
    double tile_bytes = 0.0;    // algorithmic bytes of one bulk product over the current tile list
    // HIP-event timing of the bulk products (hfem_batch_set_profiling)
    bool profiling = false;
    long long prof_n = 0;
    double prof_ms = 0.0, prof_bytes = 0.0;
    std::vector<hipEvent_t> ev;  // 2 * BULK_CHECK events, reused
    int ev_used = 0;
};

static int batch_free(hfem_batch *bt)
{
    void *ps[] = {bt->d_rows, bt->d_order, bt->d_dinv, bt->d_x, bt->d_r64, bt->d_r, bt->d_z, bt->d_p, bt->d_q,
                  bt->d_ctl, bt->d_iters, bt->d_upd, bt->d_zero, bt->d_res, bt->d_bst, bt->d_tiles, bt->d_act,
                  bt->d_act_t0, bt->d_part, bt->d_vwide};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
    for (hipEvent_t e : bt->ev) HIPFEM_CHECK(hipEventDestroy(e));
    bt->ev.clear();
    return 0;
}

extern "C" int hfem_batch_create(hipcsr_mtx *A, int nblocks, const int32_t *rows, hfem_batch **out)
{
    HIPFEM_REQUIRE(A && rows && out && nblocks > 0, HIPFEM_EINVAL, "hfem_batch_create: bad arguments");
    HIPFEM_REQUIRE(A->has_structure, HIPFEM_EINVAL, "hfem_batch_create: matrix has no structure");
    HIPFEM_REQUIRE(rows[0] == 0 && rows[nblocks] == A->n, HIPFEM_EINVAL,
                   "hfem_batch_create: block rows must start at 0 and end at the matrix size");
    for (int b = 0; b < nblocks; b++) HIPFEM_REQUIRE(rows[b + 1] > rows[b], HIPFEM_EINVAL, "empty block %d", b);
    // every entry must stay inside its block (independent systems)
    for (int b = 0; b < nblocks; b++)
        for (int i = rows[b]; i < rows[b + 1]; i++)
            for (int k = A->h_rowptr[i]; k < A->h_rowptr[i + 1]; k++)
                HIPFEM_REQUIRE(A->h_colind[k] >= rows[b] && A->h_colind[k] < rows[b + 1], HIPFEM_EINVAL,
                               "entry (%d, %d) couples block %d to another block", i + 1, A->h_colind[k] + 1, b);
    DeviceGuard g(A->ctx->device);
    hfem_batch *bt = new hfem_batch();
    bt->ctx = A->ctx;
    bt->A = A;
    bt->nblocks = nblocks;
    bt->n = A->n;
    bt->h_rows.assign(rows, rows + nblocks + 1);
    bt->h_nnz.resize(nblocks);
    for (int b = 0; b < nblocks; b++) bt->h_nnz[b] = (long long)A->h_rowptr[rows[b + 1]] - A->h_rowptr[rows[b]];
    // descending size: the largest blocks are taken first
    bt->h_sorted.resize(nblocks);
    for (int b = 0; b < nblocks; b++) bt->h_sorted[b] = b;
    std::stable_sort(bt->h_sorted.begin(), bt->h_sorted.end(),
                     [&](int a, int c) { return bt->h_nnz[a] > bt->h_nnz[c]; });
    const size_t nv = (size_t)std::max(1, bt->n) * NRHS;
    // tiles of TILE_ROWS rows never straddle blocks: at most n / TILE_ROWS + nblocks
    bt->tiles_cap = bt->n / TILE_ROWS + nblocks + 1;
    // node-block form: the pattern has a node-block view and every block
    // starts on a node boundary
    bt->nodeblock_ok = A->nodeblocked;
    for (int b = 0; b <= nblocks && bt->nodeblock_ok; b++) bt->nodeblock_ok = rows[b] % 3 == 0;
    auto fail = [&](int rc) { batch_free(bt); delete bt; return rc; };
    if (hipMalloc(&bt->d_rows, sizeof(int32_t) * (nblocks + 1)) != hipSuccess ||
        hipMalloc(&bt->d_order, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_dinv, sizeof(float) * std::max(1, bt->n)) != hipSuccess ||
        hipMalloc(&bt->d_x, sizeof(double) * nv) != hipSuccess || hipMalloc(&bt->d_r64, sizeof(double) * nv) != hipSuccess ||
        hipMalloc(&bt->d_r, sizeof(float) * nv) != hipSuccess || hipMalloc(&bt->d_z, sizeof(float) * nv) != hipSuccess ||
        hipMalloc(&bt->d_p, sizeof(float) * nv) != hipSuccess || hipMalloc(&bt->d_q, sizeof(float) * nv) != hipSuccess ||
        hipMalloc(&bt->d_ctl, sizeof(BatchCtl)) != hipSuccess || hipMalloc(&bt->d_iters, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_upd, sizeof(int) * nblocks) != hipSuccess || hipMalloc(&bt->d_zero, sizeof(int)) != hipSuccess ||
        hipMalloc(&bt->d_res, sizeof(double) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_bst, sizeof(BulkBlock) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_tiles, sizeof(int4) * bt->tiles_cap) != hipSuccess ||
        hipMalloc(&bt->d_act, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_act_t0, sizeof(int) * (nblocks + 1)) != hipSuccess ||
        hipMalloc(&bt->d_part, sizeof(double) * NPART * bt->tiles_cap) != hipSuccess) {
        set_error("hfem_batch_create: out of device memory");
        return fail(HIPFEM_ENOMEM);
    }
    hipStream_t st = bt->ctx->stream;
    if (hipMemcpyAsync(bt->d_rows, rows, sizeof(int32_t) * (nblocks + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("hfem_batch_create: upload failed");
        return fail(HIPFEM_EHIP);
    }
    ctx_retain(bt->ctx);
    *out = bt;
    return 0;
}

extern "C" void hfem_batch_destroy(hfem_batch *bt)
{
    if (!bt) return;
    hipfem_ctx *c = bt->ctx;
    {
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        batch_free(bt);
        delete bt;
    }
    ctx_release(c);
}

extern "C" int hfem_batch_set_options(hfem_batch *bt, double tol, double delta, int maxit)
{
    HIPFEM_REQUIRE(bt && tol > 0 && delta > 0 && delta < 1 && maxit > 0, HIPFEM_EINVAL, "batch options: bad values");
    bt->tol = tol;
    bt->delta = delta;
    bt->maxit = maxit;
    return 0;
}

extern "C" int hfem_batch_nshapes(void) { return NSHAPES; }

extern "C" int hfem_batch_shape_info(int shape, int *block_size, int *lanes_per_row)
{
    HIPFEM_REQUIRE(shape >= 0 && shape < NSHAPES, HIPFEM_ERANGE, "batch shape %d out of range", shape);
    *block_size = kShapes[shape].bs;
    *lanes_per_row = kShapes[shape].lpr;
    return 0;
}

extern "C" int hfem_batch_set_shape(hfem_batch *bt, int shape, int grid)
{
    HIPFEM_REQUIRE(bt && shape >= 0 && shape < NSHAPES && grid >= 0, HIPFEM_EINVAL, "batch shape: bad values");
    bt->shape = shape;
    bt->grid = grid;
    return 0;
}

extern "C" int hfem_batch_nbulk_shapes(void) { return NBULK; }

extern "C" int hfem_batch_bulk_shape_info(int shape, int *lanes_per_row)
{
    HIPFEM_REQUIRE(shape >= 0 && shape < NBULK, HIPFEM_ERANGE, "bulk shape %d out of range", shape);
    *lanes_per_row = kBulkShapes[shape].lpr;
    return 0;
}

extern "C" int hfem_batch_set_bulk_shape(hfem_batch *bt, int shape)
{
    HIPFEM_REQUIRE(bt && shape >= 0 && shape < NBULK, HIPFEM_EINVAL, "bulk shape: bad values");
    bt->bulk_shape = shape;
    return 0;
}

extern "C" int hfem_batch_set_bulk_form(hfem_batch *bt, int form)
{
    HIPFEM_REQUIRE(bt && form >= 0 && form <= 2, HIPFEM_EINVAL, "bulk form must be 0, 1 or 2");
    bt->nodeblock = form != 0;
    bt->wide = form == 2;
    return 0;
}

extern "C" int hfem_batch_bulk_form(const hfem_batch *bt)
{
    if (!(bt && bt->nodeblock && bt->nodeblock_ok && bt->A->nodeblocked)) return 0;
    return bt->wide ? 2 : 1;
}

// HIP-event timing of the bulk products: on resets the counters; the
// profile is the number of timed products, their summed duration and their
// summed algorithmic bytes
extern "C" int hfem_batch_set_profiling(hfem_batch *bt, int on)
{
    HIPFEM_REQUIRE(bt, HIPFEM_EINVAL, "batch profiling: null handle");
    DeviceGuard g(bt->ctx->device);
    if (on && bt->ev.empty()) {
        bt->ev.resize(2 * BULK_CHECK);
        for (auto &e : bt->ev) HIPFEM_CHECK(hipEventCreate(&e));
    }
    bt->profiling = on != 0;
    bt->prof_n = 0;
    bt->prof_ms = bt->prof_bytes = 0.0;
    bt->ev_used = 0;
    return 0;
}

extern "C" int hfem_batch_profile(const hfem_batch *bt, long long *launches, double *ms, double *bytes)
{
    HIPFEM_REQUIRE(bt && launches && ms && bytes, HIPFEM_EINVAL, "batch profile: null argument");
    *launches = bt->prof_n;
    *ms = bt->prof_ms;
    *bytes = bt->prof_bytes;
    return 0;
}

extern "C" int hfem_batch_set_split(hfem_batch *bt, long long split)
{
    HIPFEM_REQUIRE(bt && split >= -1, HIPFEM_EINVAL, "batch split: bad value");
    bt->split = split;
    bt->part_dirty = true;
    return 0;
}

static int device_cus(int *ncu)
{
    int dev = 0;
    HIPFEM_CHECK(hipGetDevice(&dev));
    HIPFEM_CHECK(hipDeviceGetAttribute(ncu, hipDeviceAttributeMultiprocessorCount, dev));
    return 0;
}

// blocks above the split go to the bulk phase; the automatic split is an
// even share of the batch's entries over the workgroups the persistent
// kernel keeps resident (a block above it would outlast the rest)
static int batch_partition(hfem_batch *bt)
{
    if (!bt->part_dirty && (bt->split >= 0 || bt->part_shape == bt->shape)) return 0;
    long long thr = bt->split;
    if (thr < 0) {
        int ncu = 0;
        int rc = device_cus(&ncu);
        if (rc) return rc;
        long long total = 0;
        for (long long v : bt->h_nnz) total += v;
        const long long resident = (long long)ncu * std::max(1, 2048 / kShapes[bt->shape].bs);
        thr = total / resident;
    }
    bt->h_bulk.clear();
    std::vector<int> small;
    for (int b : bt->h_sorted) (bt->h_nnz[b] > thr ? bt->h_bulk : small).push_back(b);
    bt->nsmall = (int)small.size();
    if (!small.empty())
        HIPFEM_CHECK(hipMemcpy(bt->d_order, small.data(), sizeof(int) * small.size(), hipMemcpyHostToDevice));
    bt->split_used = thr;
    bt->part_shape = bt->shape;
    bt->part_dirty = false;
    return 0;
}

extern "C" int hfem_batch_partition_info(hfem_batch *bt, int *nbulk, long long *split_used)
{
    HIPFEM_REQUIRE(bt, HIPFEM_EINVAL, "batch partition: null handle");
    DeviceGuard g(bt->ctx->device);
    int rc = batch_partition(bt);
    if (rc) return rc;
    if (nbulk) *nbulk = (int)bt->h_bulk.size();
    if (split_used) *split_used = bt->split_used;
    return 0;
}

// dinv, FP32 values, partition; checks for zero diagonals
static int batch_prepare(hfem_batch *bt)
{
    hipcsr_mtx *A = bt->A;
    hipStream_t st = bt->ctx->stream;
    int rc = csr_refresh_fp32(A);
    if (rc) return rc;
    const int big = INT32_MAX;
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_zero, &big, sizeof(int), hipMemcpyHostToDevice, st));
    k_batch_dinv<<<(bt->n + 255) / 256, 256, 0, st>>>(bt->n, A->d_diag, A->d_val, bt->d_dinv, bt->d_zero);
    HIPFEM_CHECK(hipGetLastError());
    int zr = big;
    HIPFEM_CHECK(hipMemcpyAsync(&zr, bt->d_zero, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    HIPFEM_REQUIRE(zr == big, HIPFEM_ENUMERIC, "batch: zero diagonal in equation %d", zr + 1);
    return batch_partition(bt);
}

// tile list for the bulk blocks whose index is in `act` (the stream must be
// idle: the host arrays are reused)
static int bulk_tiles(hfem_batch *bt, const std::vector<int> &act)
{
    bt->h_tiles.clear();
    bt->h_act = act;
    bt->h_act_t0.clear();
    // algorithmic bytes of one product over these blocks: values, block
    // columns and node-row pointers (node-block form) or columns and row
    // pointers (CSR rows), p read once and q written once (six FP32
    // columns per row)
    const bool nb = bt->nodeblock && bt->nodeblock_ok && bt->A->nodeblocked;
    bt->tile_bytes = 0.0;
    for (int a : act) {
        const int blk = bt->h_bulk[a];
        const double nnz = (double)bt->h_nnz[blk], rows = (double)(bt->h_rows[blk + 1] - bt->h_rows[blk]);
        bt->tile_bytes += !nb ? 8.0 * nnz + 4.0 * rows + 48.0 * rows
                          : (bt->wide ? 16.0 / 3.0 : 4.0) * nnz + 4.0 * nnz / 9.0 + 4.0 * rows / 3.0 + 48.0 * rows;
    }
    for (int a : act) {
        const int blk = bt->h_bulk[a];
        bt->h_act_t0.push_back((int)bt->h_tiles.size());
        for (int r = bt->h_rows[blk]; r < bt->h_rows[blk + 1]; r += TILE_ROWS)
            bt->h_tiles.push_back(make_int4(r, std::min(r + TILE_ROWS, bt->h_rows[blk + 1]), a, 0));
    }
    bt->h_act_t0.push_back((int)bt->h_tiles.size());
    HIPFEM_REQUIRE((int)bt->h_tiles.size() <= bt->tiles_cap, HIPFEM_EINVAL, "bulk tiles exceed their capacity");
    hipStream_t st = bt->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_tiles, bt->h_tiles.data(), sizeof(int4) * bt->h_tiles.size(),
                                hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_act, bt->h_act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_act_t0, bt->h_act_t0.data(), sizeof(int) * bt->h_act_t0.size(),
                                hipMemcpyHostToDevice, st));
    return 0;
}

// the bulk phase: every bulk block to convergence (or maxit)
static int bulk_solve(hfem_batch *bt, const double *db, double *dx, int maxit)
{
    const int nbulk = (int)bt->h_bulk.size();
    if (nbulk == 0) return 0;
    hipcsr_mtx *M = bt->A;
    hipStream_t st = bt->ctx->stream;
    bt->h_bst.assign(nbulk, BulkBlock{});
    for (int a = 0; a < nbulk; a++) bt->h_bst[a].blk = bt->h_bulk[a];
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_bst, bt->h_bst.data(), sizeof(BulkBlock) * nbulk, hipMemcpyHostToDevice, st));
    std::vector<int> act(nbulk);
    for (int a = 0; a < nbulk; a++) act[a] = a;
    int rc = bulk_tiles(bt, act);
    if (rc) return rc;
    BulkArgs A;
    A.tiles = bt->d_tiles;
    A.act = bt->d_act;
    A.act_t0 = bt->d_act_t0;
    A.st = bt->d_bst;
    A.rowptr = M->d_rowptr;
    A.col = M->d_colind;
    A.bptr = M->d_bptr;
    A.bcol = M->d_bcol;
    A.val64 = M->d_val;
    A.val32 = M->d_val32;
    A.dinv = bt->d_dinv;
    A.b = db;
    A.x = dx;
    A.r = bt->d_r;
    A.z = bt->d_z;
    A.p = bt->d_p;
    A.q = bt->d_q;
    A.part = bt->d_part;
    A.tol = bt->tol;
    A.delta = bt->delta;
    A.maxit = maxit;
    A.iters = bt->d_iters;
    A.updates = bt->d_upd;
    A.resid = bt->d_res;
    const BulkShape &sh = kBulkShapes[bt->bulk_shape];
    const bool nb = bt->nodeblock && bt->nodeblock_ok && M->nodeblocked;
    const bool wide = nb && bt->wide;
    if (wide && (!bt->d_vwide || bt->vwide_version != M->val32_version)) {
        if (!bt->d_vwide) HIPFEM_CHECK(hipMalloc(&bt->d_vwide, sizeof(float) * 12 * (size_t)std::max<int64_t>(1, M->nnz / 9)));
        k_bulk_widen<<<std::max(1, std::min((M->nbr + 255) / 256, 8192)), 256, 0, st>>>(M->nbr, M->d_bptr, M->d_val32,
                                                                                        bt->d_vwide);
        HIPFEM_CHECK(hipGetLastError());
        bt->vwide_version = M->val32_version;
    }
    A.vwide = reinterpret_cast<const float4 *>(bt->d_vwide);
    const bulk_kernel_t spmv = wide ? sh.spmv_nbw : nb ? sh.spmv_nb : sh.spmv;
    const bulk_kernel_t refresh = nb ? sh.refresh_nb : sh.refresh;
    int ntiles = (int)bt->h_tiles.size(), nact = nbulk;
    k_bulk_init<<<ntiles, BULK_TILE, 0, st>>>(A);
    k_bulk_reduce<BP_INIT><<<nact, 256, 0, st>>>(A);
    HIPFEM_CHECK(hipGetLastError());
    // every block is done once maxit iterations have run: the loop ends by then
    for (int it = 0;; it++) {
        if (it % BULK_CHECK == 0 || it >= maxit) {
            HIPFEM_CHECK(hipMemcpyAsync(bt->h_bst.data(), bt->d_bst, sizeof(BulkBlock) * nbulk,
                                        hipMemcpyDeviceToHost, st));
            HIPFEM_CHECK(hipStreamSynchronize(st));
            // the products since the last look ran over the current tile list
            for (int k = 0; k < bt->ev_used; k++) {
                float ms = 0.f;
                HIPFEM_CHECK(hipEventElapsedTime(&ms, bt->ev[2 * k], bt->ev[2 * k + 1]));
                bt->prof_ms += ms;
                bt->prof_bytes += bt->tile_bytes;
                bt->prof_n++;
            }
            bt->ev_used = 0;
            std::vector<int> still;
            for (int a : act)
                if (!bt->h_bst[a].done) still.push_back(a);
            if (still.empty()) break;
            HIPFEM_REQUIRE(it <= maxit, HIPFEM_ENUMERIC, "bulk phase: %d blocks not done after maxit",
                           (int)still.size());
            if (still.size() != act.size()) {
                act.swap(still);
                rc = bulk_tiles(bt, act);
                if (rc) return rc;
                ntiles = (int)bt->h_tiles.size();
                nact = (int)act.size();
            }
        }
        const bool timed = bt->profiling && bt->ev_used < BULK_CHECK;
        if (timed) HIPFEM_CHECK(hipEventRecord(bt->ev[2 * bt->ev_used], st));
        hipLaunchKernelGGL(spmv, dim3(ntiles), dim3(BULK_TILE), 0, st, A);
        if (timed) HIPFEM_CHECK(hipEventRecord(bt->ev[2 * bt->ev_used++ + 1], st));
        k_bulk_reduce<BP_ALPHA><<<nact, 256, 0, st>>>(A);
        k_bulk_update<<<ntiles, BULK_TILE, 0, st>>>(A);
        k_bulk_reduce<BP_BETA><<<nact, 256, 0, st>>>(A);
        hipLaunchKernelGGL(refresh, dim3(ntiles), dim3(BULK_TILE), 0, st, A);
        k_bulk_reduce<BP_REFRESH><<<nact, 256, 0, st>>>(A);
        k_bulk_pupdate<<<ntiles, BULK_TILE, 0, st>>>(A);
        HIPFEM_CHECK(hipGetLastError());
    }
    return 0;
}

// the persistent kernel over `count` queue entries starting at `qfirst`
static int persistent_solve(hfem_batch *bt, const double *db, double *dx, int qfirst, int count, int maxit)
{
    if (count <= 0) return 0;
    hipcsr_mtx *A = bt->A;
    hipStream_t st = bt->ctx->stream;
    BatchCtl c0;
    memset(&c0, 0, sizeof c0);
    c0.next = qfirst;
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_ctl, &c0, sizeof(BatchCtl), hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    const BatchShape &sh = kShapes[bt->shape];
    int ncu = 0;
    int rc = device_cus(&ncu);
    if (rc) return rc;
    const int per_cu = std::max(1, 2048 / sh.bs);
    const int grid = bt->grid > 0 ? bt->grid : std::max(1, std::min(count, ncu * per_cu));
    hipLaunchKernelGGL(sh.fn, dim3(grid), dim3(sh.bs), 0, st, qfirst + count, bt->d_order, bt->d_rows, A->d_rowptr,
                       A->d_colind, A->d_val, A->d_val32, bt->d_dinv, db, dx, bt->d_r, bt->d_z, bt->d_p, bt->d_q,
                       bt->d_r64, bt->tol, bt->delta, maxit, bt->d_ctl, bt->d_iters, bt->d_res, bt->d_upd);
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}

// solve every block for the six right-hand sides b ([row][6], device, FP64):
// the bulk phase for the blocks above the split, then the persistent kernel
// for the rest; solutions in x ([row][6], device); per block: iterations,
// reliable updates and the final max relative true residual (negative when
// a block did not converge within maxit)
extern "C" int hfem_batch_solve_dev(hfem_batch *bt, const double *db, double *dx, int *iters, int *updates,
                                    double *resid)
{
    HIPFEM_REQUIRE(bt && db && dx, HIPFEM_EINVAL, "batch solve: null argument");
    HIPFEM_REQUIRE(bt->A->has_structure && bt->A->n == bt->n, HIPFEM_EINVAL, "batch solve: matrix changed size");
    DeviceGuard g(bt->ctx->device);
    int rc = batch_prepare(bt);
    if (rc) return rc;
    rc = bulk_solve(bt, db, dx, bt->maxit);
    if (rc) return rc;
    rc = persistent_solve(bt, db, dx, 0, bt->nsmall, bt->maxit);
    if (rc) return rc;
    hipStream_t st = bt->ctx->stream;
    if (iters) HIPFEM_CHECK(hipMemcpyAsync(iters, bt->d_iters, sizeof(int) * bt->nblocks, hipMemcpyDeviceToHost, st));
    if (updates) HIPFEM_CHECK(hipMemcpyAsync(updates, bt->d_upd, sizeof(int) * bt->nblocks, hipMemcpyDeviceToHost, st));
    if (resid) HIPFEM_CHECK(hipMemcpyAsync(resid, bt->d_res, sizeof(double) * bt->nblocks, hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    return 0;
}

constexpr int TUNE_ITERS = 40;

// time every persistent-kernel shape on `sample` blocks of median size
// among those at or below the split (a contiguous stretch of the queue
// around its middle), each run capped at TUNE_ITERS iterations, and keep
// the fastest; ms[NSHAPES] receives the timings (all 0 when no block is at
// or below the split)
extern "C" int hfem_batch_autotune(hfem_batch *bt, const double *db, double *dx, int sample, double *ms, int *best)
{
    HIPFEM_REQUIRE(bt && db && dx && sample > 0, HIPFEM_EINVAL, "autotune: bad arguments");
    DeviceGuard g(bt->ctx->device);
    int rc = batch_prepare(bt);
    if (rc) return rc;
    if (ms)
        for (int s = 0; s < NSHAPES; s++) ms[s] = 0.0;
    if (best) *best = bt->shape;
    if (bt->nsmall == 0) return 0;
    hipEvent_t e0, e1;
    HIPFEM_CHECK(hipEventCreate(&e0));
    HIPFEM_CHECK(hipEventCreate(&e1));
    int bi = bt->shape;
    float bms = 1e30f;
    const int ns = std::min(sample, bt->nsmall);
    const int q0 = std::max(0, bt->nsmall / 2 - ns / 2);
    const int cap = std::min(bt->maxit, TUNE_ITERS);
    for (int s = 0; s < NSHAPES && rc == 0; s++) {
        bt->shape = s;
        rc = (hipEventRecord(e0, bt->ctx->stream) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (!rc) rc = persistent_solve(bt, db, dx, q0, ns, cap);
        if (!rc) rc = (hipEventRecord(e1, bt->ctx->stream) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (!rc) rc = (hipEventSynchronize(e1) == hipSuccess) ? 0 : HIPFEM_EHIP;
        float t = 0.f;
        if (!rc) rc = (hipEventElapsedTime(&t, e0, e1) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (ms) ms[s] = t;
        if (t < bms) {
            bms = t;
            bi = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        set_error("autotune: a timed run failed");
        return rc;
    }
    bt->shape = bi;
    bt->grid = 0;
    if (best) *best = bi;
    return 0;
}

// time every bulk shape on the blocks above the split, each run capped at
// TUNE_ITERS iterations, and keep the fastest; ms[NBULK] receives the
// timings (all 0 when no block is above the split)
extern "C" int hfem_batch_autotune_bulk(hfem_batch *bt, const double *db, double *dx, double *ms, int *best)
{
    HIPFEM_REQUIRE(bt && db && dx, HIPFEM_EINVAL, "bulk autotune: bad arguments");
    DeviceGuard g(bt->ctx->device);
    int rc = batch_prepare(bt);
    if (rc) return rc;
    if (ms)
        for (int s = 0; s < NBULK; s++) ms[s] = 0.0;
    if (best) *best = bt->bulk_shape;
    if (bt->h_bulk.empty()) return 0;
    hipEvent_t e0, e1;
    HIPFEM_CHECK(hipEventCreate(&e0));
    HIPFEM_CHECK(hipEventCreate(&e1));
    int bi = bt->bulk_shape;
    float bms = 1e30f;
    const int cap = std::min(bt->maxit, TUNE_ITERS);
    for (int s = 0; s < NBULK && rc == 0; s++) {
        bt->bulk_shape = s;
        rc = (hipEventRecord(e0, bt->ctx->stream) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (!rc) rc = bulk_solve(bt, db, dx, cap);
        if (!rc) rc = (hipEventRecord(e1, bt->ctx->stream) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (!rc) rc = (hipEventSynchronize(e1) == hipSuccess) ? 0 : HIPFEM_EHIP;
        float t = 0.f;
        if (!rc) rc = (hipEventElapsedTime(&t, e0, e1) == hipSuccess) ? 0 : HIPFEM_EHIP;
        if (ms) ms[s] = t;
        if (t < bms) {
            bms = t;
            bi = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        set_error("bulk autotune: a timed run failed");
        return rc;
    }
    bt->bulk_shape = bi;
    if (best) *best = bi;
    return 0;
}
