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
// the largest blocks and keeps the fastest.

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
    int32_t *d_rows = nullptr;
    int *d_order = nullptr;
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
};

static int batch_free(hfem_batch *bt)
{
    void *ps[] = {bt->d_rows, bt->d_order, bt->d_dinv, bt->d_x, bt->d_r64, bt->d_r, bt->d_z, bt->d_p, bt->d_q,
                  bt->d_ctl, bt->d_iters, bt->d_upd, bt->d_zero, bt->d_res};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
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
    // descending size: the largest blocks are taken first
    std::vector<int> order(nblocks);
    for (int b = 0; b < nblocks; b++) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return A->h_rowptr[rows[a + 1]] - A->h_rowptr[rows[a]] > A->h_rowptr[rows[c + 1]] - A->h_rowptr[rows[c]];
    });
    const size_t nv = (size_t)std::max(1, bt->n) * NRHS;
    auto fail = [&](int rc) { batch_free(bt); delete bt; return rc; };
    if (hipMalloc(&bt->d_rows, sizeof(int32_t) * (nblocks + 1)) != hipSuccess ||
        hipMalloc(&bt->d_order, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_dinv, sizeof(float) * std::max(1, bt->n)) != hipSuccess ||
        hipMalloc(&bt->d_x, sizeof(double) * nv) != hipSuccess || hipMalloc(&bt->d_r64, sizeof(double) * nv) != hipSuccess ||
        hipMalloc(&bt->d_r, sizeof(float) * nv) != hipSuccess || hipMalloc(&bt->d_z, sizeof(float) * nv) != hipSuccess ||
        hipMalloc(&bt->d_p, sizeof(float) * nv) != hipSuccess || hipMalloc(&bt->d_q, sizeof(float) * nv) != hipSuccess ||
        hipMalloc(&bt->d_ctl, sizeof(BatchCtl)) != hipSuccess || hipMalloc(&bt->d_iters, sizeof(int) * nblocks) != hipSuccess ||
        hipMalloc(&bt->d_upd, sizeof(int) * nblocks) != hipSuccess || hipMalloc(&bt->d_zero, sizeof(int)) != hipSuccess ||
        hipMalloc(&bt->d_res, sizeof(double) * nblocks) != hipSuccess) {
        set_error("hfem_batch_create: out of device memory");
        return fail(HIPFEM_ENOMEM);
    }
    hipStream_t st = bt->ctx->stream;
    if (hipMemcpyAsync(bt->d_rows, rows, sizeof(int32_t) * (nblocks + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(bt->d_order, order.data(), sizeof(int) * nblocks, hipMemcpyHostToDevice, st) != hipSuccess ||
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

// solve every block for the six right-hand sides b ([row][6], device, FP64);
// solutions in x ([row][6], device); per block: iterations, reliable
// updates and the final max relative true residual (negative when a block
// did not converge within maxit)
static int batch_solve(hfem_batch *bt, const double *db, double *dx, int *iters, int *updates, double *resid,
                       int nb_limit, int qfirst = 0)
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
    BatchCtl c0;
    memset(&c0, 0, sizeof c0);
    c0.next = qfirst;
    HIPFEM_CHECK(hipMemcpyAsync(bt->d_ctl, &c0, sizeof(BatchCtl), hipMemcpyHostToDevice, st));
    const BatchShape &sh = kShapes[bt->shape];
    const int nb = nb_limit > 0 ? std::min(qfirst + nb_limit, bt->nblocks) : bt->nblocks;
    int dev = 0, ncu = 0;
    HIPFEM_CHECK(hipGetDevice(&dev));
    HIPFEM_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = std::max(1, 2048 / sh.bs);
    const int grid = bt->grid > 0 ? bt->grid : std::max(1, std::min(nb - qfirst, ncu * per_cu));
    hipLaunchKernelGGL(sh.fn, dim3(grid), dim3(sh.bs), 0, st, nb, bt->d_order, bt->d_rows, A->d_rowptr, A->d_colind,
                       A->d_val, A->d_val32, bt->d_dinv, db, dx, bt->d_r, bt->d_z, bt->d_p, bt->d_q, bt->d_r64, bt->tol,
                       bt->delta, bt->maxit, bt->d_ctl, bt->d_iters, bt->d_res, bt->d_upd);
    HIPFEM_CHECK(hipGetLastError());
    if (iters) HIPFEM_CHECK(hipMemcpyAsync(iters, bt->d_iters, sizeof(int) * bt->nblocks, hipMemcpyDeviceToHost, st));
    if (updates) HIPFEM_CHECK(hipMemcpyAsync(updates, bt->d_upd, sizeof(int) * bt->nblocks, hipMemcpyDeviceToHost, st));
    if (resid) HIPFEM_CHECK(hipMemcpyAsync(resid, bt->d_res, sizeof(double) * bt->nblocks, hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hfem_batch_solve_dev(hfem_batch *bt, const double *db, double *dx, int *iters, int *updates,
                                    double *resid)
{
    HIPFEM_REQUIRE(bt && db && dx, HIPFEM_EINVAL, "batch solve: null argument");
    HIPFEM_REQUIRE(bt->A->has_structure && bt->A->n == bt->n, HIPFEM_EINVAL, "batch solve: matrix changed size");
    DeviceGuard g(bt->ctx->device);
    return batch_solve(bt, db, dx, iters, updates, resid, 0);
}

// time every shape on `sample` blocks of median size (a contiguous stretch
// of the size-ordered queue around its middle), each run capped at
// TUNE_ITERS iterations, and keep the fastest; ms[NSHAPES] receives the
// timings
extern "C" int hfem_batch_autotune(hfem_batch *bt, const double *db, double *dx, int sample, double *ms, int *best)
{
    HIPFEM_REQUIRE(bt && db && dx && sample > 0, HIPFEM_EINVAL, "autotune: bad arguments");
    DeviceGuard g(bt->ctx->device);
    hipEvent_t e0, e1;
    HIPFEM_CHECK(hipEventCreate(&e0));
    HIPFEM_CHECK(hipEventCreate(&e1));
    const int keep_shape = bt->shape, keep_grid = bt->grid;
    int bi = keep_shape;
    float bms = 1e30f;
    const int ns = std::min(sample, bt->nblocks);
    const int q0 = std::max(0, bt->nblocks / 2 - ns / 2);
    constexpr int TUNE_ITERS = 40;
    const int keep_maxit = bt->maxit;
    bt->maxit = std::min(keep_maxit, TUNE_ITERS);
    for (int s = 0; s < NSHAPES; s++) {
        bt->shape = s;
        bt->grid = 0;
        HIPFEM_CHECK(hipEventRecord(e0, bt->ctx->stream));
        int rc = batch_solve(bt, db, dx, nullptr, nullptr, nullptr, ns, q0);
        if (rc) {
            bt->maxit = keep_maxit;
            return rc;
        }
        HIPFEM_CHECK(hipEventRecord(e1, bt->ctx->stream));
        HIPFEM_CHECK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPFEM_CHECK(hipEventElapsedTime(&t, e0, e1));
        if (ms) ms[s] = t;
        if (t < bms) {
            bms = t;
            bi = s;
        }
    }
    (void)keep_grid;
    bt->maxit = keep_maxit;
    bt->shape = bi;
    bt->grid = 0;
    if (best) *best = bi;
    HIPFEM_CHECK(hipEventDestroy(e0));
    HIPFEM_CHECK(hipEventDestroy(e1));
    (void)keep_shape;
    return 0;
}
