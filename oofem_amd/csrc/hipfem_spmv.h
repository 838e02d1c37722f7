// Row kernel shared by the plain SpMV and the PCG's fused SpMV.
//
// A group of LANES consecutive lanes owns one CSR row.  Each pass over the
// row issues U independent (column, value) load pairs per lane before any
// gather, then U gathers, so a lane keeps 2U streaming loads and U gathers in
// flight instead of one dependent chain per nonzero.  With NT the value and
// column streams are loaded non-temporally: they are read once per SpMV and
// should not displace the gathered vectors from L2 / MALL.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hipfem {

template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// sum_k val[k] * g(col[k]) over k in [s, e), partial per lane (not reduced)
template <int LANES, int U, bool NT, typename AT, typename MT, typename G>
__device__ __forceinline__ AT row_partial(int s, int e, int lane, const int32_t *__restrict__ col,
                                          const MT *__restrict__ val, G gather)
{
    AT sum = AT(0);
    for (int base = s + lane; base < e; base += U * LANES) {
        int c[U];
        MT v[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int k = base + j * LANES;
            if (k < e) {
                c[j] = ld_stream<NT>(col + k);
                v[j] = ld_stream<NT>(val + k);
            } else {
                c[j] = -1;
                v[j] = MT(0);
            }
        }
        AT g[U];
#pragma unroll
        for (int j = 0; j < U; j++) g[j] = (c[j] >= 0) ? gather(c[j]) : AT(0);
#pragma unroll
        for (int j = 0; j < U; j++) sum += (AT)v[j] * g[j];
    }
    return sum;
}

template <int LANES, typename T>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
    return v;
}

// kernel shape of an SpMV launch; a property of the matrix (tunable)
struct SpmvShape {
    int lanes = 16;   // 8, 16, 32
    int unroll = 4;   // 1, 2, 4, 6
    int nt = 0;       // non-temporal streams
    int stream = 0;   // 1: block-contiguous streaming form (rows_per_block below)
    int rpb = 16;     // rows per block of the streaming form
    int nodeblock = 0;  // 1: node-block form (needs a node-blocked pattern)
};

// Node-block form.  When every row triple (3I, 3I+1, 3I+2) shares one column
// list made of whole triples (3J, 3J+1, 3J+2) -- all free nodes fully free
// and numbered consecutively, as in the reference's node-order numbering --
// the CSR value array of node row I is a sequence of 3x3 blocks laid out
// [r][k][c]: val[rowptr[3I] + r*3*nbk + 3*k + c] with nbk = number of block
// columns.  The kernel then reads one block column index per 9 values
// (bcol) instead of one column per value, and gathers x once per block.
// A 32-lane half wave owns one node row; lane k owns block column k.
template <typename AT, typename MT, typename G>
__device__ __forceinline__ void nb_row(int bo, int nbk, int vbase, int lane, const int32_t *__restrict__ bcol,
                                       const MT *__restrict__ val, G gather3, AT acc[3])
{
    acc[0] = acc[1] = acc[2] = AT(0);
    for (int k = lane; k < nbk; k += 32) {
        const int bc = bcol[bo + k];
        AT x0, x1, x2;
        gather3(bc, x0, x1, x2);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const MT *vp = val + vbase + r * 3 * nbk + 3 * k;
            acc[r] += (AT)vp[0] * x0 + (AT)vp[1] * x1 + (AT)vp[2] * x2;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) acc[r] = group_sum<32>(acc[r]);
}

// Row-group form, for patterns with partially constrained nodes: rows
// [r0, r0 + nr) (nr <= 3) share one column list, which is cut into chunks of
// up to 3 consecutive columns.  Chunk k of the group starts at column
// ccol & GRP_COL_MASK and holds (ccol >> 30) + 1 columns; its offset inside
// every row of the group is the sum of the lengths of the chunks before it,
// which the group's lanes form by a prefix sum over the 16 lanes of a DPP
// row (four row shifts, then the row total from lane 15 as the carry into
// the next 16 chunks), so no offset is stored.  Lane k of a GRP_LANES-lane
// group takes chunks k, k + GRP_LANES, ...; for a fully free node row this
// is the node-block access pattern.  16 lanes (four groups per wave, up to
// two chunks per lane for a hexahedral node row of 27 chunks) measured 11 %
// faster on gfx950 than 32 lanes (one chunk per lane) and than 8 lanes, for
// FP64 and FP32 values alike.
constexpr uint32_t GRP_COL_MASK = 0x3FFFFFFFu;
constexpr int GRP_LANES = 16;
constexpr int GRP_PER_BLOCK = 256 / GRP_LANES;   // groups per 256-thread block

// inclusive prefix sum of v over the 16 lanes of a DPP row (lanes without a
// source in the row add 0); every lane of the row must be active
__device__ __forceinline__ int row16_inclusive_sum(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);   // row_shr:8
    return v;
}

template <typename AT, typename MT, typename G>
__device__ __forceinline__ void grp_row(int r0, int nr, int c0, int nch, int lane, const int32_t *__restrict__ rowptr,
                                        const uint32_t *__restrict__ ccol, const MT *__restrict__ val, G gather,
                                        AT acc[3])
{
    static_assert(GRP_LANES == 16, "the chunk offsets are summed over one 16-lane DPP row");
    acc[0] = acc[1] = acc[2] = AT(0);
    const int rp0 = rowptr[r0], rp1 = nr > 1 ? rowptr[r0 + 1] : 0, rp2 = nr > 2 ? rowptr[r0 + 2] : 0;
    int carry = 0;   // columns in the chunks of the earlier rounds
    // the round count is the same for the group's 16 lanes, so all of them
    // take part in the row sums
    for (int base = 0; base < nch; base += GRP_LANES) {
        const int k = base + lane;
        const bool valid = k < nch;
        const uint32_t cc = valid ? ccol[c0 + k] : 0u;
        const int L = valid ? (int)(cc >> 30) + 1 : 0;
        const int incl = row16_inclusive_sum(L);
        const int off = carry + incl - L;
        carry += __shfl(incl, GRP_LANES - 1, GRP_LANES);
        if (!valid) continue;
        const int col = (int)(cc & GRP_COL_MASK);
        if (nr == 3 && L == 3) {
            // full 3x3 block (the common case): nine independent loads
            const AT x0 = gather(col), x1 = gather(col + 1), x2 = gather(col + 2);
            const MT *v0 = val + rp0 + off, *v1 = val + rp1 + off, *v2 = val + rp2 + off;
            const MT a0 = v0[0], a1 = v0[1], a2 = v0[2], b0 = v1[0], b1 = v1[1], b2 = v1[2];
            const MT c0v = v2[0], c1v = v2[1], c2v = v2[2];
            acc[0] += (AT)a0 * x0 + (AT)a1 * x1 + (AT)a2 * x2;
            acc[1] += (AT)b0 * x0 + (AT)b1 * x1 + (AT)b2 * x2;
            acc[2] += (AT)c0v * x0 + (AT)c1v * x1 + (AT)c2v * x2;
            continue;
        }
        const AT x0 = gather(col), x1 = L > 1 ? gather(col + 1) : AT(0), x2 = L > 2 ? gather(col + 2) : AT(0);
        const int rps[3] = {rp0, rp1, rp2};
#pragma unroll
        for (int r = 0; r < 3; r++) {
            if (r >= nr) break;
            const MT *vp = val + rps[r] + off;
            AT a = (AT)vp[0] * x0;
            if (L > 1) a += (AT)vp[1] * x1;
            if (L > 2) a += (AT)vp[2] * x2;
            acc[r] += a;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) acc[r] = group_sum<GRP_LANES>(acc[r]);
}

// Streaming form: a 256-thread block owns rows [r0, r0 + RPB).  Their values
// and columns are one contiguous range of the CSR arrays, which the block
// reads with unit stride across all 256 lanes (U loads in flight per lane),
// multiplying each value by its gathered vector entry into LDS.  A second
// phase sums each row's products with 256/RPB lanes per row.  The host
// guarantees that no block's range exceeds STREAM_CAP entries.
constexpr int STREAM_CAP = 4096;

template <int RPB, int U, typename AT, typename MT, typename G>
__device__ __forceinline__ AT stream_block_row(int r0, int n, const int32_t *__restrict__ rowptr,
                                               const int32_t *__restrict__ col, const MT *__restrict__ val,
                                               G gather, AT *prod, int *rp)
{
    constexpr int LPR = 256 / RPB;  // lanes per row in the reduction phase
    const int t = threadIdx.x;
    const int rend = min(r0 + RPB, n);
    const int s0 = rowptr[r0];
    const int len = rowptr[rend] - s0;
    if (t <= RPB) rp[t] = rowptr[min(r0 + t, rend)] - s0;
    for (int base = t; base < len; base += U * 256) {
        int c[U];
        MT v[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int i = base + j * 256;
            if (i < len) {
                c[j] = col[s0 + i];
                v[j] = val[s0 + i];
            } else {
                c[j] = -1;
                v[j] = MT(0);
            }
        }
#pragma unroll
        for (int j = 0; j < U; j++)
            if (c[j] >= 0) prod[base + j * 256] = (AT)v[j] * gather(c[j]);
    }
    __syncthreads();
    const int r = t / LPR, lane = t % LPR;
    AT sum = AT(0);
    if (r0 + r < rend)
        for (int i = rp[r] + lane; i < rp[r + 1]; i += LPR) sum += prod[i];
    sum = group_sum<LPR>(sum);
    __syncthreads();  // prod / rp are reused by the next row block
    return sum;       // valid in lane 0 of each row group; row = r0 + t / LPR
}

}  // namespace hipfem
