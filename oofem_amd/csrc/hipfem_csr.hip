// hipcsr: the SparseMtrx of the MI355X path (reference interface
// src/core/sparsemtrx.h:60-281; storage replaces CompCol/SymCompCol,
// compcol.C:128-316, symcompcol.C:98-222).
//
// Storage: full (unsymmetric) CSR, int32 row pointers and columns, FP64
// values, sorted columns.  The matrix lives in HBM; a host copy of the
// pattern serves assemble()/at() lookups.  SpMV: a group of LANES lanes per
// row (LANES in {4, 8, 16, 32, 64}, picked from the mean row length) reads
// the row's values and columns with unit stride, gathers x, and reduces with
// cross-lane shuffles; one wave holds 64/LANES rows.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "hipfem_internal.h"

namespace hipfem {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int ensure_partials(hipfem_ctx *ctx, size_t n)
{
    if (ctx->partials_cap >= n) return 0;
    if (ctx->d_partials) HIPFEM_CHECK(hipFree(ctx->d_partials));
    ctx->d_partials = nullptr;
    size_t cap = std::max<size_t>(n, 8192);
    HIPFEM_CHECK(hipMalloc(&ctx->d_partials, cap * sizeof(double)));
    ctx->partials_cap = cap;
    return 0;
}

// ------------------------------------------------------------------ kernels
template <int LANES, int U, bool NT, typename VT>
__global__ __launch_bounds__(256) void k_spmv(int n, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ col,
                                              const VT *__restrict__ val,
                                              const double *__restrict__ x, double *__restrict__ y)
{
    const int rows_per_block = 256 / LANES;
    const int lane = threadIdx.x % LANES;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / LANES; row < n;
         row += gridDim.x * rows_per_block) {
        const int s = rowptr[row], e = rowptr[row + 1];
        double sum = row_partial<LANES, U, NT, double>(s, e, lane, col, val, [&](int c) { return x[c]; });
        sum = group_sum<LANES>(sum);
        if (lane == 0) y[row] = sum;
    }
}

template <int RPB, int U, typename VT>
__global__ __launch_bounds__(256) void k_spmv_stream(int n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, const VT *__restrict__ val,
                                                     const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double prod[STREAM_CAP];
    __shared__ int rp[RPB + 1];
    constexpr int LPR = 256 / RPB;
    for (int r0 = blockIdx.x * RPB; r0 < n; r0 += gridDim.x * RPB) {
        const double sum = stream_block_row<RPB, U, double>(r0, n, rowptr, col, val, [&](int c) { return x[c]; },
                                                            prod, rp);
        const int row = r0 + threadIdx.x / LPR;
        if (threadIdx.x % LPR == 0 && row < n) y[row] = sum;
    }
}

__global__ void k_diag_pos(int n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                           int32_t *__restrict__ diag)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int lo = rowptr[r], hi = rowptr[r + 1] - 1, pos = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = col[mid];
        if (c == r) { pos = mid; break; }
        if (c < r) lo = mid + 1; else hi = mid - 1;
    }
    diag[r] = pos;
}

__global__ void k_axpy_vals(int64_t nnz, double a, const double *__restrict__ src, double *__restrict__ dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] += a * src[i];
}

__global__ void k_scale_vals(int64_t nnz, double a, double *__restrict__ v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        v[i] *= a;
}

__global__ void k_to_fp32(int64_t nnz, const double *__restrict__ v, float *__restrict__ o)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = (float)v[i];
}

__global__ void k_add_diag(int n, double a, const double *__restrict__ d, const int32_t *__restrict__ diag,
                           double *__restrict__ v)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && diag[r] >= 0) v[diag[r]] += a * d[r];
}

// y = A^T x through the transpose pattern: tmap[k] is the position in val of
// the k-th transpose entry, so each output row is a plain gather-sum.
__global__ void k_spmv_t(int n, const int32_t *__restrict__ tptr, const int32_t *__restrict__ tcol,
                         const int32_t *__restrict__ tmap, const double *__restrict__ val,
                         const double *__restrict__ x, double *__restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int k = tptr[r]; k < tptr[r + 1]; k++) s += val[tmap[k]] * x[tcol[k]];
    y[r] = s;
}

static int grid_for(int64_t work, int block = 256, int64_t cap = 8192)
{
    int64_t g = (work + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min(g, cap));
}

int spmv_lanes(const hipcsr_mtx *m) { return m->shape.lanes; }

// default shape from the mean row length: 16 lanes for FEM rows of 30-100
// entries, unrolled so one pass covers a typical row
static SpmvShape default_shape(double avg_row)
{
    SpmvShape s;
    s.lanes = avg_row <= 12 ? 8 : (avg_row <= 120 ? 16 : 32);
    const int passes = (int)std::ceil(avg_row / s.lanes);
    s.unroll = passes <= 1 ? 1 : passes <= 2 ? 2 : passes <= 4 ? 4 : 6;
    s.nt = 0;  // non-temporal streams measured slower on gfx950 for this access pattern
    return s;
}

// node-block SpMV: one 32-lane half wave per node row (see nb_row)
template <typename MT>
__global__ __launch_bounds__(256) void k_spmv_nb(int nbr, const int32_t *__restrict__ bptr,
                                                 const int32_t *__restrict__ bcol, const MT *__restrict__ val,
                                                 const double *__restrict__ x, double *__restrict__ y)
{
    const int lane = threadIdx.x % 32;
    for (int I = blockIdx.x * 8 + threadIdx.x / 32; I < nbr; I += gridDim.x * 8) {
        const int bo = bptr[I], nbk = bptr[I + 1] - bo;
        double acc[3];
        nb_row<double>(bo, nbk, 9 * bo, lane, bcol, val,
                       [&](int bc, double &a, double &b, double &c) {
                           const double *xp = x + 3 * bc;
                           a = xp[0];
                           b = xp[1];
                           c = xp[2];
                       },
                       acc);
        if (lane == 0) {
            y[3 * I] = acc[0];
            y[3 * I + 1] = acc[1];
            y[3 * I + 2] = acc[2];
        }
    }
}

// row-group SpMV: GRP_LANES lanes per row group (see grp_row)
template <typename MT>
__global__ __launch_bounds__(256) void k_spmv_grp(int ngrp, const int32_t *__restrict__ grow,
                                                  const int32_t *__restrict__ gch, const uint32_t *__restrict__ ccol,
                                                  const int32_t *__restrict__ rowptr,
                                                  const MT *__restrict__ val, const double *__restrict__ x,
                                                  double *__restrict__ y)
{
    const int lane = threadIdx.x % GRP_LANES;
    for (int g = blockIdx.x * GRP_PER_BLOCK + threadIdx.x / GRP_LANES; g < ngrp; g += gridDim.x * GRP_PER_BLOCK) {
        const int r0 = grow[g], nr = grow[g + 1] - r0, c0 = gch[g], nch = gch[g + 1] - c0;
        double acc[3];
        grp_row<double>(r0, nr, c0, nch, lane, rowptr, ccol, val, [&](int c) { return x[c]; }, acc);
        if (lane == 0)
            for (int r = 0; r < nr; r++) y[r0 + r] = acc[r];
    }
}

template <typename VT>
static int launch_spmv(const hipcsr_mtx *m, const SpmvShape &sh, hipStream_t st, int n, const int32_t *rp,
                       const int32_t *ci, const VT *v, const double *x, double *y)
{
    bool ok;
    if (sh.nodeblock && m->nodeblocked) {
        const int grid = std::max(1, std::min((m->nbr + 7) / 8, 1 << 20));
        k_spmv_nb<VT><<<grid, 256, 0, st>>>(m->nbr, m->d_bptr, m->d_bcol, v, x, y);
        ok = true;
    } else if (sh.nodeblock && m->grouped) {
        const int grid = std::max(1, std::min((m->ngrp + GRP_PER_BLOCK - 1) / GRP_PER_BLOCK, 1 << 20));
        k_spmv_grp<VT><<<grid, 256, 0, st>>>(m->ngrp, m->d_grow, m->d_gch, m->d_ccol, rp, v, x, y);
        ok = true;
    } else if (sh.stream) {
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + sh.rpb - 1) / sh.rpb, 1 << 20));
        ok = dispatch_stream(sh, [&](auto R, auto U) {
            k_spmv_stream<decltype(R)::value, decltype(U)::value, VT><<<grid, 256, 0, st>>>(n, rp, ci, v, x, y);
        });
    } else {
        const int rows_per_block = 256 / sh.lanes;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, 1 << 20));
        ok = dispatch_shape(sh, [&](auto L, auto U, auto NT) {
            k_spmv<decltype(L)::value, decltype(U)::value, decltype(NT)::value, VT>
                <<<grid, 256, 0, st>>>(n, rp, ci, v, x, y);
        });
    }
    HIPFEM_REQUIRE(ok, HIPFEM_EINVAL, "unsupported SpMV shape (%d lanes, unroll %d)", sh.lanes, sh.unroll);
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}

int csr_spmv_dev(hipcsr_mtx *m, const double *dx, double *dy)
{
    return launch_spmv<double>(m, m->shape, m->ctx->stream, m->n, m->d_rowptr, m->d_colind, m->d_val, dx, dy);
}

int csr_free(hipcsr_mtx *m)
{
    for (void *p : {(void *)m->d_rowptr, (void *)m->d_colind, (void *)m->d_diag, (void *)m->d_val,
                    (void *)m->d_val32, (void *)m->d_tptr, (void *)m->d_tcol, (void *)m->d_tmap,
                    (void *)m->d_bptr, (void *)m->d_bcol, (void *)m->d_grow, (void *)m->d_gch,
                    (void *)m->d_ccol})
        if (p) HIPFEM_CHECK(hipFree(p));
    m->d_rowptr = m->d_colind = m->d_diag = m->d_tptr = m->d_tcol = m->d_tmap = nullptr;
    m->d_bptr = m->d_bcol = nullptr;
    m->d_grow = m->d_gch = nullptr;
    m->d_ccol = nullptr;
    m->grouped = false;
    m->ngrp = 0;
    m->d_val = nullptr;
    m->d_val32 = nullptr;
    m->has_structure = false;
    m->nodeblocked = false;
    return 0;
}

// Node-blocked pattern test: n = 3 nbr; rows 3I..3I+2 share one column list
// made of whole triples.  Fills bptr (node-row block offsets) and bcol.
static bool node_blocked(int n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                         std::vector<int32_t> &bptr, std::vector<int32_t> &bcol)
{
    if (n == 0 || n % 3) return false;
    const int nbr = n / 3;
    bptr.assign((size_t)nbr + 1, 0);
    for (int I = 0; I < nbr; I++) {
        const int r0 = 3 * I, len = rp[r0 + 1] - rp[r0];
        if (len % 3) return false;
        for (int r = 1; r < 3; r++) {
            if (rp[r0 + r + 1] - rp[r0 + r] != len) return false;
            if (!std::equal(ci.begin() + rp[r0], ci.begin() + rp[r0 + 1], ci.begin() + rp[r0 + r])) return false;
        }
        for (int t = 0; t < len; t += 3) {
            const int c = ci[rp[r0] + t];
            if (c % 3 || ci[rp[r0] + t + 1] != c + 1 || ci[rp[r0] + t + 2] != c + 2) return false;
        }
        bptr[I + 1] = bptr[I] + len / 3;
    }
    bcol.resize((size_t)bptr[nbr]);
    for (int I = 0; I < nbr; I++)
        for (int k = 0; k < bptr[I + 1] - bptr[I]; k++) bcol[bptr[I] + k] = ci[rp[3 * I] + 3 * k] / 3;
    return true;
}

// Row-group view: consecutive rows with identical column lists (at most 3)
// form a group; each group's column list is cut into chunks of up to 3
// consecutive columns.  Worth using when groups and chunks are mostly full
// (FEM patterns with some partially constrained nodes).
static bool row_groups(int n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                       std::vector<int32_t> &grow, std::vector<int32_t> &gch, std::vector<uint32_t> &ccol)
{
    if (n == 0 || n >= (int)GRP_COL_MASK) return false;
    grow.assign(1, 0);
    gch.assign(1, 0);
    ccol.clear();
    int r = 0;
    int64_t group_cols = 0;
    while (r < n) {
        const int len = rp[r + 1] - rp[r];
        group_cols += len;
        int nr = 1;
        while (nr < 3 && r + nr < n && rp[r + nr + 1] - rp[r + nr] == len &&
               std::equal(ci.begin() + rp[r], ci.begin() + rp[r + 1], ci.begin() + rp[r + nr]))
            nr++;
        const int32_t *c = ci.data() + rp[r];
        for (int i = 0; i < len;) {
            int L = 1;
            while (L < 3 && i + L < len && c[i + L] == c[i] + L) L++;
            ccol.push_back((uint32_t)c[i] | ((uint32_t)(L - 1) << 30));
            i += L;
        }
        r += nr;
        grow.push_back(r);
        gch.push_back((int32_t)ccol.size());
    }
    const double rows_per_group = (double)n / (double)(grow.size() - 1);
    const double cols_per_chunk = ccol.empty() ? 0.0 : (double)group_cols / (double)ccol.size();
    return rows_per_group >= 2.5 && cols_per_chunk >= 2.5;
}

int csr_alloc(hipcsr_mtx *m, int n, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind)
{
    int rc = csr_free(m);
    if (rc) return rc;
    m->n = n;
    m->nnz = (int64_t)colind.size();
    m->h_rowptr = rowptr;
    m->h_colind = colind;
    m->avg_row = n ? (double)m->nnz / n : 0.0;
    for (int k = 0; k < 4; k++) {
        const int rpb = 8 << k;
        int mx = 0;
        for (int r0 = 0; r0 < n; r0 += rpb) mx = std::max(mx, rowptr[std::min(r0 + rpb, n)] - rowptr[r0]);
        m->stream_max[k] = mx;
    }
    m->shape = default_shape(m->avg_row);
    std::vector<int32_t> bptr, bcol;
    m->nodeblocked = node_blocked(n, rowptr, colind, bptr, bcol);
    m->nbr = m->nodeblocked ? n / 3 : 0;
    std::vector<int32_t> grow, gch;
    std::vector<uint32_t> ccol;
    m->grouped = !m->nodeblocked && row_groups(n, rowptr, colind, grow, gch, ccol);
    m->ngrp = m->grouped ? (int)grow.size() - 1 : 0;
    m->shape.nodeblock = (m->nodeblocked || m->grouped) ? 1 : 0;
    HIPFEM_CHECK(hipSetDevice(m->ctx->device));
    if (m->grouped) {
        HIPFEM_CHECK(hipMalloc(&m->d_grow, sizeof(int32_t) * grow.size()));
        HIPFEM_CHECK(hipMalloc(&m->d_gch, sizeof(int32_t) * gch.size()));
        HIPFEM_CHECK(hipMalloc(&m->d_ccol, sizeof(uint32_t) * std::max<size_t>(1, ccol.size())));
        hipStream_t s0 = m->ctx->stream;
        HIPFEM_CHECK(hipMemcpyAsync(m->d_grow, grow.data(), sizeof(int32_t) * grow.size(), hipMemcpyHostToDevice, s0));
        HIPFEM_CHECK(hipMemcpyAsync(m->d_gch, gch.data(), sizeof(int32_t) * gch.size(), hipMemcpyHostToDevice, s0));
        if (!ccol.empty()) {
            HIPFEM_CHECK(hipMemcpyAsync(m->d_ccol, ccol.data(), sizeof(uint32_t) * ccol.size(), hipMemcpyHostToDevice, s0));
        }
        // the host vectors die with this scope: finish the copies first
        HIPFEM_CHECK(hipStreamSynchronize(s0));
    }
    if (m->nodeblocked) {
        HIPFEM_CHECK(hipMalloc(&m->d_bptr, sizeof(int32_t) * (m->nbr + 1)));
        HIPFEM_CHECK(hipMalloc(&m->d_bcol, sizeof(int32_t) * std::max<size_t>(1, bcol.size())));
        HIPFEM_CHECK(hipMemcpyAsync(m->d_bptr, bptr.data(), sizeof(int32_t) * (m->nbr + 1), hipMemcpyHostToDevice,
                                    m->ctx->stream));
        if (!bcol.empty())
            HIPFEM_CHECK(hipMemcpyAsync(m->d_bcol, bcol.data(), sizeof(int32_t) * bcol.size(), hipMemcpyHostToDevice,
                                        m->ctx->stream));
    }
    HIPFEM_CHECK(hipMalloc(&m->d_rowptr, sizeof(int32_t) * (n + 1)));
    HIPFEM_CHECK(hipMalloc(&m->d_colind, sizeof(int32_t) * std::max<int64_t>(1, m->nnz)));
    HIPFEM_CHECK(hipMalloc(&m->d_diag, sizeof(int32_t) * std::max(1, n)));
    HIPFEM_CHECK(hipMalloc(&m->d_val, sizeof(double) * std::max<int64_t>(1, m->nnz)));
    hipStream_t st = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(m->d_rowptr, rowptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    if (m->nnz)
        HIPFEM_CHECK(hipMemcpyAsync(m->d_colind, colind.data(), sizeof(int32_t) * m->nnz, hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemsetAsync(m->d_val, 0, sizeof(double) * std::max<int64_t>(1, m->nnz), st));
    if (n) {
        k_diag_pos<<<(n + 255) / 256, 256, 0, st>>>(n, m->d_rowptr, m->d_colind, m->d_diag);
        HIPFEM_CHECK(hipGetLastError());
    }
    HIPFEM_CHECK(hipStreamSynchronize(st));
    m->has_structure = true;
    m->pattern_id = next_pattern_id();
    m->h_stage.clear();
    m->staged = false;
    m->version++;
    return 0;
}

int csr_refresh_fp32(hipcsr_mtx *m)
{
    if (m->d_val32 && m->val32_version == m->version) return 0;
    if (!m->d_val32) HIPFEM_CHECK(hipMalloc(&m->d_val32, sizeof(float) * std::max<int64_t>(1, m->nnz)));
    k_to_fp32<<<grid_for(m->nnz), 256, 0, m->ctx->stream>>>(m->nnz, m->d_val, m->d_val32);
    HIPFEM_CHECK(hipGetLastError());
    m->val32_version = m->version;
    return 0;
}

int spmv_fp32_dev(hipcsr_mtx *m, const double *dx, double *dy)
{
    return launch_spmv<float>(m, m->shape, m->ctx->stream, m->n, m->d_rowptr, m->d_colind, m->d_val32, dx, dy);
}

static int64_t find_pos(const hipcsr_mtx *m, int r, int c)
{
    const int32_t *b = m->h_colind.data() + m->h_rowptr[r];
    const int32_t *e = m->h_colind.data() + m->h_rowptr[r + 1];
    const int32_t *p = std::lower_bound(b, e, c);
    if (p == e || *p != c) return -1;
    return p - m->h_colind.data();
}

long next_pattern_id()
{
    static long id = 0;
    return ++id;
}

static int flush_stage(hipcsr_mtx *m)
{
    if (!m->staged) return 0;
    double *tmp = nullptr;
    HIPFEM_CHECK(hipMalloc(&tmp, sizeof(double) * m->nnz));
    HIPFEM_CHECK(hipMemcpyAsync(tmp, m->h_stage.data(), sizeof(double) * m->nnz, hipMemcpyHostToDevice, m->ctx->stream));
    k_axpy_vals<<<grid_for(m->nnz), 256, 0, m->ctx->stream>>>(m->nnz, 1.0, tmp, m->d_val);
    HIPFEM_CHECK(hipGetLastError());
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIPFEM_CHECK(hipFree(tmp));
    m->h_stage.clear();
    m->h_stage.shrink_to_fit();
    m->staged = false;
    m->version++;
    return 0;
}

}  // namespace hipfem

using namespace hipfem;

// ================================================================== context
extern "C" int hipfem_ctx_create(int device, hipfem_ctx **out)
{
    HIPFEM_REQUIRE(out, HIPFEM_EINVAL, "hipfem_ctx_create: out is NULL");
    int ndev = 0;
    HIPFEM_CHECK(hipGetDeviceCount(&ndev));
    HIPFEM_REQUIRE(device >= 0 && device < ndev, HIPFEM_EINVAL, "hipfem_ctx_create: device %d of %d", device, ndev);
    DeviceGuard g(device);
    hipfem_ctx *c = new hipfem_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
        return HIPFEM_EHIP;
    }
    e = hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("hipHostMalloc: %s", hipGetErrorString(e));
        return HIPFEM_EHIP;
    }
    *out = c;
    return 0;
}

static void ctx_free(hipfem_ctx *c)
{
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->d_partials) (void)hipFree(c->d_partials);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

namespace hipfem {
void ctx_retain(hipfem_ctx *c) { c->refs++; }
void ctx_release(hipfem_ctx *c)
{
    if (--c->refs == 0 && c->destroy_pending) ctx_free(c);
}
}  // namespace hipfem

extern "C" void hipfem_ctx_destroy(hipfem_ctx *c)
{
    if (!c) return;
    if (c->refs > 0) {
        c->destroy_pending = true;   // freed by the last object's destroy
        return;
    }
    ctx_free(c);
}

extern "C" int hipfem_ctx_synchronize(hipfem_ctx *c)
{
    HIPFEM_REQUIRE(c, HIPFEM_EINVAL, "null context");
    HIPFEM_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" void *hipfem_ctx_stream(hipfem_ctx *c) { return c ? (void *)c->stream : nullptr; }

// ------------------------------------------------------------ device vectors
namespace hipfem {
// y[i * incy] = a x[i * incx] + b y[i * incy] (x may be NULL: a x = 0)
__global__ void k_axpby_strided(int64_t n, double a, const double *__restrict__ x, int64_t incx, double b,
                                double *__restrict__ y, int64_t incy)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double xv = x ? a * x[i * incx] : 0.0;
        y[i * incy] = b == 0.0 ? xv : xv + b * y[i * incy];
    }
}

__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (x ? a * x[i] : 0.0) + (b != 0.0 ? b * y[i] : 0.0);  // b = 0 never reads y
}

__global__ __launch_bounds__(256) void k_dot_part(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                                                  double *__restrict__ part)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += x[i] * y[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// max |x_i|: per-block maxima, then one block over them
__global__ void k_amax_part(int64_t n, const double *__restrict__ x, double *__restrict__ part)
{
    __shared__ double sh[4];
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmax(m, fabs(x[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void k_amax_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double sh[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, part[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void k_dot_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = sh[0] + sh[1] + sh[2] + sh[3];
}
}  // namespace hipfem

extern "C" int hipfem_dev_alloc(hipfem_ctx *c, int64_t n, double **out)
{
    HIPFEM_REQUIRE(c && out && n >= 0, HIPFEM_EINVAL, "dev_alloc: bad arguments");
    DeviceGuard g(c->device);
    HIPFEM_CHECK(hipMalloc((void **)out, sizeof(double) * std::max<int64_t>(1, n)));
    return 0;
}

extern "C" int hipfem_dev_free(hipfem_ctx *c, double *p)
{
    HIPFEM_REQUIRE(c, HIPFEM_EINVAL, "dev_free: null context");
    DeviceGuard g(c->device);
    HIPFEM_CHECK(hipStreamSynchronize(c->stream));
    if (p) HIPFEM_CHECK(hipFree(p));
    return 0;
}

extern "C" int hipfem_dev_upload(hipfem_ctx *c, double *dst, const double *src, int64_t n)
{
    HIPFEM_REQUIRE(c && (n == 0 || (dst && src)), HIPFEM_EINVAL, "dev_upload: bad arguments");
    DeviceGuard g(c->device);
    if (n) HIPFEM_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPFEM_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hipfem_dev_download(hipfem_ctx *c, double *dst, const double *src, int64_t n)
{
    HIPFEM_REQUIRE(c && (n == 0 || (dst && src)), HIPFEM_EINVAL, "dev_download: bad arguments");
    DeviceGuard g(c->device);
    if (n) HIPFEM_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPFEM_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hipfem_dev_axpby(hipfem_ctx *c, int64_t n, double a, const double *x, double b, double *y)
{
    HIPFEM_REQUIRE(c && (n == 0 || y), HIPFEM_EINVAL, "dev_axpby: bad arguments");
    DeviceGuard g(c->device);
    if (n) {
        k_axpby<<<grid_for(n), 256, 0, c->stream>>>(n, a, x, b, y);
        HIPFEM_CHECK(hipGetLastError());
    }
    return 0;
}

// y = x with 16-byte accesses, four in flight per thread (bandwidth reference)
__global__ __launch_bounds__(256) void k_copy_wide(int64_t n2, const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const double2 a = x[i], b = x[i + stride], c = x[i + 2 * stride], d = x[i + 3 * stride];
        y[i] = a;
        y[i + stride] = b;
        y[i + 2 * stride] = c;
        y[i + 3 * stride] = d;
    }
    for (; i < n2; i += stride) y[i] = x[i];
}

extern "C" int hipfem_dev_copy(hipfem_ctx *c, int64_t n, const double *x, double *y, int method)
{
    HIPFEM_REQUIRE(c && (n == 0 || (x && y)) && n >= 0 && (method == 0 || method == 1), HIPFEM_EINVAL,
                   "dev_copy: bad arguments");
    DeviceGuard g(c->device);
    if (n == 0) return 0;
    if (method == 0) {
        HIPFEM_CHECK(hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    } else {
        HIPFEM_REQUIRE(n % 2 == 0 && ((uintptr_t)x % 16) == 0 && ((uintptr_t)y % 16) == 0, HIPFEM_EINVAL,
                       "dev_copy: the wide copy needs an even length and 16-byte aligned vectors");
        int dev = 0, ncu = 0;
        HIPFEM_CHECK(hipGetDevice(&dev));
        HIPFEM_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        k_copy_wide<<<std::max(1, ncu * 8), 256, 0, c->stream>>>(n / 2, reinterpret_cast<const double2 *>(x),
                                                                  reinterpret_cast<double2 *>(y));
        HIPFEM_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" int hipfem_dev_dot(hipfem_ctx *c, int64_t n, const double *x, const double *y, double *out)
{
    HIPFEM_REQUIRE(c && out && (n == 0 || (x && y)), HIPFEM_EINVAL, "dev_dot: bad arguments");
    DeviceGuard g(c->device);
    const int nb = grid_for(n, 256, 1024);
    int rc = ensure_partials(c, (size_t)nb + 1);
    if (rc) return rc;
    if (n) {
        k_dot_part<<<nb, 256, 0, c->stream>>>(n, x, y, c->d_partials);
        k_dot_final<<<1, 256, 0, c->stream>>>(nb, c->d_partials, c->d_partials + nb);
        HIPFEM_CHECK(hipGetLastError());
        HIPFEM_CHECK(hipMemcpyAsync(c->h_pinned, c->d_partials + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPFEM_CHECK(hipStreamSynchronize(c->stream));
        *out = c->h_pinned[0];
    } else {
        *out = 0.0;
    }
    return 0;
}
extern "C" int hipfem_dev_axpby_strided(hipfem_ctx *c, int64_t n, double a, const double *x, int64_t incx, double b,
                                        double *y, int64_t incy)
{
    HIPFEM_REQUIRE(c && n >= 0 && incx > 0 && incy > 0 && (n == 0 || y), HIPFEM_EINVAL, "dev_axpby_strided: bad arguments");
    DeviceGuard g(c->device);
    if (n) {
        k_axpby_strided<<<grid_for(n), 256, 0, c->stream>>>(n, a, x, incx, b, y, incy);
        HIPFEM_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" int hipfem_dev_amax(hipfem_ctx *c, int64_t n, const double *x, double *out)
{
    HIPFEM_REQUIRE(c && out && (n == 0 || x), HIPFEM_EINVAL, "dev_amax: bad arguments");
    DeviceGuard g(c->device);
    const int nb = grid_for(n, 256, 1024);
    int rc = ensure_partials(c, (size_t)nb + 1);
    if (rc) return rc;
    if (n) {
        k_amax_part<<<nb, 256, 0, c->stream>>>(n, x, c->d_partials);
        k_amax_final<<<1, 256, 0, c->stream>>>(nb, c->d_partials, c->d_partials + nb);
        HIPFEM_CHECK(hipGetLastError());
        HIPFEM_CHECK(hipMemcpyAsync(c->h_pinned, c->d_partials + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPFEM_CHECK(hipStreamSynchronize(c->stream));
        *out = c->h_pinned[0];
    } else {
        *out = 0.0;
    }
    return 0;
}
extern "C" const char *hipfem_last_error(void) { return g_err.c_str(); }
extern "C" const char *hipfem_version(void) { return "hipfem 0.1 gfx950"; }

// =================================================================== hipcsr
extern "C" int hipcsr_create(hipfem_ctx *ctx, hipcsr_mtx **out)
{
    HIPFEM_REQUIRE(ctx && out, HIPFEM_EINVAL, "hipcsr_create: null argument");
    hipcsr_mtx *m = new hipcsr_mtx();
    m->ctx = ctx;
    ctx_retain(ctx);
    *out = m;
    return 0;
}

extern "C" void hipcsr_destroy(hipcsr_mtx *m)
{
    if (!m) return;
    hipfem_ctx *c = m->ctx;
    {
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        if (m->mf_model) model_unlink(m->mf_model, m);
        csr_free(m);
        delete m;
    }
    ctx_release(c);
}

extern "C" int hipcsr_set_matrix_free(hipcsr_mtx *m, int on)
{
    HIPFEM_REQUIRE(m, HIPFEM_EINVAL, "null matrix");
    m->mf_enabled = on != 0;
    return 0;
}

extern "C" int hipcsr_is_matrix_free(const hipcsr_mtx *m) { return (m && model_mf_available(m)) ? 1 : 0; }

__global__ void k_d2f_vec(int n, const double *__restrict__ a, float *__restrict__ b)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = (float)a[i];
}

extern "C" int hipcsr_times_matrix_free(hipcsr_mtx *m, const double *x, double *y)
{
    HIPFEM_REQUIRE(m && x && y, HIPFEM_EINVAL, "times_matrix_free: null argument");
    HIPFEM_REQUIRE(model_mf_available(m), HIPFEM_EINVAL,
                   "times_matrix_free: the matrix has no current element-by-element view");
    DeviceGuard g(m->ctx->device);
    hipStream_t st = m->ctx->stream;
    const int n = m->n;
    double *dd = nullptr;
    float *fx = nullptr;
    HIPFEM_CHECK(hipMalloc(&dd, sizeof(double) * std::max(1, n)));
    HIPFEM_CHECK(hipMalloc(&fx, sizeof(float) * std::max(1, n)));
    int rc = 0;
    const int grid = std::max(1, std::min((n + 255) / 256, 4096));
    if (n) {
        HIPFEM_CHECK(hipMemcpyAsync(dd, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
        k_d2f_vec<<<grid, 256, 0, st>>>(n, dd, fx);
        rc = model_mf_apply_f32(m->mf_model, fx, dd, nullptr);
        if (!rc) {
            HIPFEM_CHECK(hipMemcpyAsync(y, dd, sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIPFEM_CHECK(hipStreamSynchronize(st));
        }
    }
    HIPFEM_CHECK(hipFree(dd));
    HIPFEM_CHECK(hipFree(fx));
    return rc;
}

extern "C" int hipcsr_clone(const hipcsr_mtx *src, hipcsr_mtx **out)
{
    HIPFEM_REQUIRE(src && out, HIPFEM_EINVAL, "hipcsr_clone: null argument");
    DeviceGuard g(src->ctx->device);
    hipcsr_mtx *m = new hipcsr_mtx();
    m->ctx = src->ctx;
    if (src->has_structure) {
        int rc = csr_alloc(m, src->n, src->h_rowptr, src->h_colind);
        if (rc) { delete m; return rc; }
        HIPFEM_CHECK(hipMemcpyAsync(m->d_val, src->d_val, sizeof(double) * src->nnz, hipMemcpyDeviceToDevice, m->ctx->stream));
        HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
        m->h_stage = src->h_stage;
        m->staged = src->staged;
    }
    m->version = src->version;
    ctx_retain(m->ctx);
    *out = m;
    return 0;
}

extern "C" int hipcsr_build_structure(hipcsr_mtx *m, int neq, int nelem, int ndofe, const int *loc)
{
    HIPFEM_REQUIRE(m && neq >= 0 && nelem >= 0 && ndofe > 0 && (loc || nelem == 0), HIPFEM_EINVAL,
                   "hipcsr_build_structure: bad arguments");
    DeviceGuard g(m->ctx->device);
    // rows as sorted unique column lists (CompCol::buildInternalStructure semantics)
    std::vector<std::vector<int32_t>> rows((size_t)neq);
    for (int e = 0; e < nelem; e++) {
        const int *l = loc + (size_t)e * ndofe;
        for (int i = 0; i < ndofe; i++) {
            const int ii = l[i];
            if (ii <= 0) continue;
            HIPFEM_REQUIRE(ii <= neq, HIPFEM_ERANGE, "location %d > neq %d (element %d)", ii, neq, e + 1);
            for (int j = 0; j < ndofe; j++)
                if (l[j] > 0) rows[ii - 1].push_back(l[j] - 1);
        }
    }
    std::vector<int32_t> rowptr((size_t)neq + 1, 0), colind;
    int64_t total = 0;
    for (int r = 0; r < neq; r++) {
        auto &v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        total += (int64_t)v.size();
        HIPFEM_REQUIRE(total < INT32_MAX, HIPFEM_ETOOBIG, "pattern exceeds int32 indices");
        rowptr[r + 1] = (int32_t)total;
    }
    colind.reserve((size_t)total);
    for (auto &v : rows) colind.insert(colind.end(), v.begin(), v.end());
    return csr_alloc(m, neq, rowptr, colind);
}

extern "C" int hipcsr_build_structure_csr(hipcsr_mtx *m, int n, const int64_t *rowptr, const int32_t *colind)
{
    HIPFEM_REQUIRE(m && n >= 0 && rowptr && (colind || rowptr[n] == 0), HIPFEM_EINVAL,
                   "hipcsr_build_structure_csr: bad arguments");
    HIPFEM_REQUIRE(rowptr[n] < INT32_MAX, HIPFEM_ETOOBIG, "pattern exceeds int32 indices");
    DeviceGuard g(m->ctx->device);
    std::vector<int32_t> rp((size_t)n + 1), ci((size_t)rowptr[n]);
    for (int i = 0; i <= n; i++) rp[i] = (int32_t)rowptr[i];
    for (int64_t k = 0; k < rowptr[n]; k++) {
        HIPFEM_REQUIRE(colind[k] >= 0 && colind[k] < n, HIPFEM_ERANGE, "column %d out of range", colind[k]);
        ci[k] = colind[k];
    }
    for (int r = 0; r < n; r++)
        for (int k = rp[r] + 1; k < rp[r + 1]; k++)
            HIPFEM_REQUIRE(ci[k - 1] < ci[k], HIPFEM_EINVAL, "columns of row %d not sorted/unique", r);
    return csr_alloc(m, n, rp, ci);
}

static int stage_add(hipcsr_mtx *m, int r, int c, double v)
{
    if (!m->staged) {
        m->h_stage.assign((size_t)m->nnz, 0.0);
        m->staged = true;
    }
    const int64_t p = find_pos(m, r, c);
    HIPFEM_REQUIRE(p >= 0, HIPFEM_ERANGE, "assemble: entry (%d,%d) is not in the pattern", r + 1, c + 1);
    m->h_stage[p] += v;
    return 0;
}

extern "C" int hipcsr_assemble(hipcsr_mtx *m, int n, const int *loc, const double *mat)
{
    return hipcsr_assemble_rc(m, n, loc, n, loc, mat);
}

extern "C" int hipcsr_assemble_rc(hipcsr_mtx *m, int nr, const int *rloc, int nc, const int *cloc, const double *mat)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "assemble: matrix has no structure");
    HIPFEM_REQUIRE(nr >= 0 && nc >= 0 && (nr == 0 || (rloc && mat)) && (nc == 0 || cloc), HIPFEM_EINVAL,
                   "assemble: bad arguments");
    for (int i = 0; i < nr; i++) {
        const int ii = rloc[i];
        if (ii <= 0) continue;
        HIPFEM_REQUIRE(ii <= m->n, HIPFEM_ERANGE, "assemble: row %d > %d", ii, m->n);
        for (int j = 0; j < nc; j++) {
            const int jj = cloc[j];
            if (jj <= 0) continue;
            HIPFEM_REQUIRE(jj <= m->n, HIPFEM_ERANGE, "assemble: column %d > %d", jj, m->n);
            int rc = stage_add(m, ii - 1, jj - 1, mat[(size_t)i * nc + j]);
            if (rc) return rc;
        }
    }
    return 0;
}

extern "C" int hipcsr_assemble_begin(hipcsr_mtx *m)
{
    HIPFEM_REQUIRE(m, HIPFEM_EINVAL, "null matrix");
    return 0;
}

extern "C" int hipcsr_assemble_end(hipcsr_mtx *m)
{
    HIPFEM_REQUIRE(m, HIPFEM_EINVAL, "null matrix");
    DeviceGuard g(m->ctx->device);
    int rc = flush_stage(m);
    if (rc) return rc;
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    return 0;
}

extern "C" int hipcsr_zero(hipcsr_mtx *m)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "zero: matrix has no structure");
    DeviceGuard g(m->ctx->device);
    if (m->nnz) {
        // fill kernel rather than a memset node, so callers may capture it
        k_scale_vals<<<grid_for(m->nnz), 256, 0, m->ctx->stream>>>(m->nnz, 0.0, m->d_val);
        HIPFEM_CHECK(hipGetLastError());
    }
    m->h_stage.clear();
    m->staged = false;
    m->version++;
    return 0;
}

extern "C" int hipcsr_scale(hipcsr_mtx *m, double a)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "times(double): no structure");
    DeviceGuard g(m->ctx->device);
    int rc = flush_stage(m);
    if (rc) return rc;
    k_scale_vals<<<grid_for(m->nnz), 256, 0, m->ctx->stream>>>(m->nnz, a, m->d_val);
    HIPFEM_CHECK(hipGetLastError());
    m->version++;
    return 0;
}

extern "C" int hipcsr_add(hipcsr_mtx *m, double a, const hipcsr_mtx *o)
{
    HIPFEM_REQUIRE(m && o && m->has_structure && o->has_structure, HIPFEM_EINVAL, "add: no structure");
    HIPFEM_REQUIRE(m->n == o->n && m->nnz == o->nnz && m->h_rowptr == o->h_rowptr && m->h_colind == o->h_colind,
                   HIPFEM_EINVAL, "add: patterns differ");
    DeviceGuard g(m->ctx->device);
    int rc = flush_stage(m);
    if (rc) return rc;
    HIPFEM_REQUIRE(!o->staged, HIPFEM_EINVAL, "add: operand has unflushed assemble() contributions");
    HIPFEM_CHECK(hipStreamSynchronize(o->ctx->stream));
    k_axpy_vals<<<grid_for(m->nnz), 256, 0, m->ctx->stream>>>(m->nnz, a, o->d_val, m->d_val);
    HIPFEM_CHECK(hipGetLastError());
    m->version++;
    return 0;
}

extern "C" int hipcsr_add_diagonal(hipcsr_mtx *m, double a, const double *d)
{
    HIPFEM_REQUIRE(m && m->has_structure && d, HIPFEM_EINVAL, "addDiagonal: bad arguments");
    DeviceGuard g(m->ctx->device);
    for (int i = 0; i < m->n; i++)
        HIPFEM_REQUIRE(find_pos(m, i, i) >= 0 || d[i] == 0.0, HIPFEM_ERANGE, "addDiagonal: (%d,%d) not allocated", i + 1, i + 1);
    int rc = flush_stage(m);
    if (rc) return rc;
    double *dd = nullptr;
    HIPFEM_CHECK(hipMalloc(&dd, sizeof(double) * std::max(1, m->n)));
    HIPFEM_CHECK(hipMemcpyAsync(dd, d, sizeof(double) * m->n, hipMemcpyHostToDevice, m->ctx->stream));
    if (m->n) {
        k_add_diag<<<(m->n + 255) / 256, 256, 0, m->ctx->stream>>>(m->n, a, dd, m->d_diag, m->d_val);
        HIPFEM_CHECK(hipGetLastError());
    }
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIPFEM_CHECK(hipFree(dd));
    m->version++;
    return 0;
}

extern "C" int hipcsr_times_dev(hipcsr_mtx *m, const double *dx, double *dy)
{
    HIPFEM_REQUIRE(m && m->has_structure && dx && dy, HIPFEM_EINVAL, "times: bad arguments");
    HIPFEM_REQUIRE(!m->staged, HIPFEM_EINVAL, "times: assemble() contributions not flushed (call assembleEnd)");
    DeviceGuard g(m->ctx->device);
    if (m->n == 0) return 0;
    return csr_spmv_dev(m, dx, dy);
}

extern "C" int hipcsr_times(hipcsr_mtx *m, const double *x, double *y)
{
    HIPFEM_REQUIRE(m && m->has_structure && x && y, HIPFEM_EINVAL, "times: bad arguments");
    HIPFEM_REQUIRE(!m->staged, HIPFEM_EINVAL, "times: assemble() contributions not flushed (call assembleEnd)");
    DeviceGuard g(m->ctx->device);
    if (m->n == 0) return 0;
    double *d = nullptr;
    HIPFEM_CHECK(hipMalloc(&d, sizeof(double) * 2 * m->n));
    hipStream_t st = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(d, x, sizeof(double) * m->n, hipMemcpyHostToDevice, st));
    int rc = csr_spmv_dev(m, d, d + m->n);
    if (rc) { (void)hipFree(d); return rc; }
    HIPFEM_CHECK(hipMemcpyAsync(y, d + m->n, sizeof(double) * m->n, hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    HIPFEM_CHECK(hipFree(d));
    return 0;
}

static int build_transpose(hipcsr_mtx *m)
{
    if (m->d_tptr) return 0;
    const int n = m->n;
    std::vector<int32_t> cnt((size_t)n + 1, 0);
    for (int64_t k = 0; k < m->nnz; k++) cnt[m->h_colind[k] + 1]++;
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1), tcol((size_t)m->nnz), tmap((size_t)m->nnz);
    for (int r = 0; r < n; r++)
        for (int k = m->h_rowptr[r]; k < m->h_rowptr[r + 1]; k++) {
            const int c = m->h_colind[k];
            tcol[fill[c]] = r;
            tmap[fill[c]++] = k;
        }
    HIPFEM_CHECK(hipMalloc(&m->d_tptr, sizeof(int32_t) * (n + 1)));
    HIPFEM_CHECK(hipMalloc(&m->d_tcol, sizeof(int32_t) * std::max<int64_t>(1, m->nnz)));
    HIPFEM_CHECK(hipMalloc(&m->d_tmap, sizeof(int32_t) * std::max<int64_t>(1, m->nnz)));
    hipStream_t st = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(m->d_tptr, cnt.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemcpyAsync(m->d_tcol, tcol.data(), sizeof(int32_t) * m->nnz, hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemcpyAsync(m->d_tmap, tmap.data(), sizeof(int32_t) * m->nnz, hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hipcsr_times_t(hipcsr_mtx *m, const double *x, double *y)
{
    HIPFEM_REQUIRE(m && m->has_structure && x && y, HIPFEM_EINVAL, "timesT: bad arguments");
    HIPFEM_REQUIRE(!m->staged, HIPFEM_EINVAL, "timesT: assemble() contributions not flushed (call assembleEnd)");
    DeviceGuard g(m->ctx->device);
    if (m->n == 0) return 0;
    int rc = build_transpose(m);
    if (rc) return rc;
    double *d = nullptr;
    HIPFEM_CHECK(hipMalloc(&d, sizeof(double) * 2 * m->n));
    hipStream_t st = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(d, x, sizeof(double) * m->n, hipMemcpyHostToDevice, st));
    k_spmv_t<<<(m->n + 255) / 256, 256, 0, st>>>(m->n, m->d_tptr, m->d_tcol, m->d_tmap, m->d_val, d, d + m->n);
    HIPFEM_CHECK(hipGetLastError());
    HIPFEM_CHECK(hipMemcpyAsync(y, d + m->n, sizeof(double) * m->n, hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    HIPFEM_CHECK(hipFree(d));
    return 0;
}

extern "C" int hipcsr_times_mat(hipcsr_mtx *m, int ncol, const double *B, double *Y)
{
    HIPFEM_REQUIRE(m && m->has_structure && ncol >= 0 && (ncol == 0 || (B && Y)), HIPFEM_EINVAL,
                   "times(FloatMatrix): bad arguments");
    std::vector<double> x((size_t)m->n), y((size_t)m->n);
    for (int c = 0; c < ncol; c++) {
        for (int i = 0; i < m->n; i++) x[i] = B[(size_t)i * ncol + c];
        int rc = hipcsr_times(m, x.data(), y.data());
        if (rc) return rc;
        for (int i = 0; i < m->n; i++) Y[(size_t)i * ncol + c] = y[i];
    }
    return 0;
}

extern "C" int hipcsr_get(hipcsr_mtx *m, int i, int j, double *v)
{
    HIPFEM_REQUIRE(m && m->has_structure && v, HIPFEM_EINVAL, "at: bad arguments");
    HIPFEM_REQUIRE(i >= 1 && j >= 1 && i <= m->n && j <= m->n, HIPFEM_ERANGE, "at(%d,%d) outside %dx%d", i, j, m->n, m->n);
    DeviceGuard g(m->ctx->device);
    const int64_t p = find_pos(m, i - 1, j - 1);
    if (p < 0) { *v = 0.0; return 0; }
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    double d = 0.0;
    HIPFEM_CHECK(hipMemcpy(&d, m->d_val + p, sizeof(double), hipMemcpyDeviceToHost));
    if (m->staged) d += m->h_stage[p];
    *v = d;
    return 0;
}

extern "C" int hipcsr_find(const hipcsr_mtx *m, int i, int j, int64_t *pos)
{
    HIPFEM_REQUIRE(m && m->has_structure && pos, HIPFEM_EINVAL, "find: bad arguments");
    HIPFEM_REQUIRE(i >= 1 && j >= 1 && i <= m->n && j <= m->n, HIPFEM_ERANGE, "at(%d,%d) outside %dx%d", i, j, m->n, m->n);
    *pos = find_pos(m, i - 1, j - 1);
    return 0;
}

extern "C" int hipcsr_set(hipcsr_mtx *m, int i, int j, double v)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "at: no structure");
    HIPFEM_REQUIRE(i >= 1 && j >= 1 && i <= m->n && j <= m->n, HIPFEM_ERANGE, "at(%d,%d) outside %dx%d", i, j, m->n, m->n);
    DeviceGuard g(m->ctx->device);
    const int64_t p = find_pos(m, i - 1, j - 1);
    HIPFEM_REQUIRE(p >= 0, HIPFEM_ERANGE, "at(%d,%d): entry not in the pattern", i, j);
    if (m->staged) m->h_stage[p] = 0.0;
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(m->d_val + p, &v, sizeof(double), hipMemcpyHostToDevice));
    m->version++;
    return 0;
}

extern "C" int hipcsr_is_allocated(hipcsr_mtx *m, int i, int j)
{
    if (!m || !m->has_structure || i < 1 || j < 1 || i > m->n || j > m->n) return 0;
    return find_pos(m, i - 1, j - 1) >= 0;
}

extern "C" int hipcsr_set_spmv_shape(hipcsr_mtx *m, int lanes, int unroll, int nontemporal)
{
    HIPFEM_REQUIRE(m, HIPFEM_EINVAL, "null matrix");
    SpmvShape s;
    s.lanes = lanes;
    s.unroll = unroll;
    s.nt = nontemporal ? 1 : 0;
    s.stream = 0;
    const bool ok = (lanes == 8 || lanes == 16 || lanes == 32) && (unroll == 1 || unroll == 2 || unroll == 4 || unroll == 6);
    HIPFEM_REQUIRE(ok, HIPFEM_EINVAL, "unsupported SpMV shape (%d lanes, unroll %d)", lanes, unroll);
    m->shape = s;
    return 0;
}

extern "C" int hipcsr_set_spmv_stream(hipcsr_mtx *m, int rows_per_block, int unroll)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "set_spmv_stream: no structure");
    if (rows_per_block == 0) {
        m->shape.stream = 0;
        return 0;
    }
    int k = -1;
    for (int i = 0; i < 4; i++)
        if ((8 << i) == rows_per_block) k = i;
    HIPFEM_REQUIRE(k >= 0 && (unroll == 2 || unroll == 4 || unroll == 6), HIPFEM_EINVAL,
                   "unsupported streaming shape (%d rows per block, unroll %d)", rows_per_block, unroll);
    HIPFEM_REQUIRE(m->stream_max[k] <= STREAM_CAP, HIPFEM_EINVAL,
                   "streaming SpMV: %d rows per block hold up to %d entries (> %d)", rows_per_block, m->stream_max[k],
                   STREAM_CAP);
    m->shape.stream = 1;
    m->shape.nodeblock = 0;
    m->shape.rpb = rows_per_block;
    m->shape.unroll = unroll;
    return 0;
}

extern "C" int hipcsr_set_spmv_nodeblock(hipcsr_mtx *m, int on)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "set_spmv_nodeblock: no structure");
    HIPFEM_REQUIRE(!on || m->nodeblocked || m->grouped, HIPFEM_EINVAL,
                   "node-block SpMV needs a node-structured pattern (rows of a node sharing one column list)");
    m->shape.nodeblock = on ? 1 : 0;
    if (on) m->shape.stream = 0;
    return 0;
}

extern "C" int hipcsr_nodeblock_sizes(const hipcsr_mtx *m, int64_t *groups, int64_t *blocks)
{
    HIPFEM_REQUIRE(m && groups && blocks, HIPFEM_EINVAL, "null argument");
    *groups = m->nodeblocked ? m->nbr : m->grouped ? m->ngrp : 0;
    *blocks = 0;
    if (m->nodeblocked) *blocks = m->nnz / 9;
    if (m->grouped) {
        int32_t nch = 0;
        HIPFEM_CHECK(hipMemcpy(&nch, m->d_gch + m->ngrp, sizeof(int32_t), hipMemcpyDeviceToHost));
        *blocks = nch;
    }
    return 0;
}

extern "C" int hipcsr_is_nodeblocked(const hipcsr_mtx *m)
{
    return !m ? 0 : m->nodeblocked ? 1 : m->grouped ? 2 : 0;
}

extern "C" int hipcsr_get_spmv_shape(const hipcsr_mtx *m, int *lanes, int *unroll, int *nontemporal)
{
    HIPFEM_REQUIRE(m, HIPFEM_EINVAL, "null matrix");
    if (lanes) *lanes = m->shape.lanes;
    if (unroll) *unroll = m->shape.unroll;
    if (nontemporal) *nontemporal = m->shape.nt;
    return 0;
}

extern "C" int hipcsr_nrows(const hipcsr_mtx *m) { return m ? m->n : 0; }
extern "C" int64_t hipcsr_nnz(const hipcsr_mtx *m) { return m ? m->nnz : 0; }
extern "C" long hipcsr_version(const hipcsr_mtx *m) { return m ? m->version : 0; }

extern "C" int hipcsr_download(hipcsr_mtx *m, int64_t *rowptr, int32_t *colind, double *val)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "download: no structure");
    DeviceGuard g(m->ctx->device);
    if (rowptr)
        for (int i = 0; i <= m->n; i++) rowptr[i] = m->h_rowptr[i];
    if (colind && m->nnz) memcpy(colind, m->h_colind.data(), sizeof(int32_t) * m->nnz);
    if (val && m->nnz) {
        HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
        HIPFEM_CHECK(hipMemcpy(val, m->d_val, sizeof(double) * m->nnz, hipMemcpyDeviceToHost));
        if (m->staged)
            for (int64_t k = 0; k < m->nnz; k++) val[k] += m->h_stage[k];
    }
    return 0;
}

extern "C" int hipcsr_upload_values(hipcsr_mtx *m, const double *val)
{
    HIPFEM_REQUIRE(m && m->has_structure && (val || !m->nnz), HIPFEM_EINVAL, "upload: bad arguments");
    DeviceGuard g(m->ctx->device);
    if (m->nnz) {
        HIPFEM_CHECK(hipMemcpyAsync(m->d_val, val, sizeof(double) * m->nnz, hipMemcpyHostToDevice, m->ctx->stream));
        HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    }
    m->h_stage.clear();
    m->staged = false;
    m->version++;
    return 0;
}

extern "C" int hipcsr_device_arrays(hipcsr_mtx *m, const int32_t **rowptr, const int32_t **colind, double **val)
{
    HIPFEM_REQUIRE(m && m->has_structure, HIPFEM_EINVAL, "device_arrays: no structure");
    if (rowptr) *rowptr = m->d_rowptr;
    if (colind) *colind = m->d_colind;
    if (val) *val = m->d_val;
    return 0;
}
