// Internal declarations shared by the HIP translation units of libhipfem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hipfem.h"
#include "hipfem_spmv.h"

#include <type_traits>

namespace hipfem {
// Calls f(lanes, unroll, nt) with compile-time constants for the supported
// SpMV shapes; returns false for an unsupported shape.
template <typename F>
bool dispatch_shape(const SpmvShape &s, F &&f)
{
    auto with_nt = [&](auto L, auto U) {
        if (s.nt)
            f(L, U, std::true_type{});
        else
            f(L, U, std::false_type{});
        return true;
    };
    auto with_u = [&](auto L) {
        switch (s.unroll) {
        case 1: return with_nt(L, std::integral_constant<int, 1>{});
        case 2: return with_nt(L, std::integral_constant<int, 2>{});
        case 4: return with_nt(L, std::integral_constant<int, 4>{});
        case 6: return with_nt(L, std::integral_constant<int, 6>{});
        }
        return false;
    };
    switch (s.lanes) {
    case 8: return with_u(std::integral_constant<int, 8>{});
    case 16: return with_u(std::integral_constant<int, 16>{});
    case 32: return with_u(std::integral_constant<int, 32>{});
    }
    return false;
}

// Same for the streaming form: f(rows_per_block, unroll).
template <typename F>
bool dispatch_stream(const SpmvShape &s, F &&f)
{
    auto with_u = [&](auto R) {
        switch (s.unroll) {
        case 2: f(R, std::integral_constant<int, 2>{}); return true;
        case 4: f(R, std::integral_constant<int, 4>{}); return true;
        case 6: f(R, std::integral_constant<int, 6>{}); return true;
        }
        return false;
    };
    switch (s.rpb) {
    case 8: return with_u(std::integral_constant<int, 8>{});
    case 16: return with_u(std::integral_constant<int, 16>{});
    case 32: return with_u(std::integral_constant<int, 32>{});
    case 64: return with_u(std::integral_constant<int, 64>{});
    }
    return false;
}
}  // namespace hipfem

namespace hipfem {

void set_error(const char *fmt, ...);

// Every device launch of a context goes to ctx->stream (default priority).
// Small reduction scratch is owned by the context; the PCG keeps its own
// vectors.
// Makes the context's device current for the scope of an API call.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace hipfem

#define HIPFEM_CHECK(expr)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            hipfem::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,  \
                              __LINE__);                                                         \
            return HIPFEM_EHIP;                                                                  \
        }                                                                                        \
    } while (0)

#define HIPFEM_REQUIRE(cond, code, ...)       \
    do {                                      \
        if (!(cond)) {                        \
            hipfem::set_error(__VA_ARGS__);   \
            return (code);                    \
        }                                     \
    } while (0)

struct hipfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *d_partials = nullptr;   // per-block partial sums (reduction scratch)
    size_t partials_cap = 0;        // in doubles
    double *h_pinned = nullptr;     // small pinned host buffer for scalar readback
    // Objects created on the context hold a reference; hipfem_ctx_destroy on a
    // context with live objects only marks it, and the last object's destroy
    // frees it.
    int refs = 0;
    bool destroy_pending = false;
};

namespace hipfem {
void ctx_retain(hipfem_ctx *c);
void ctx_release(hipfem_ctx *c);
}  // namespace hipfem

struct hfem_model;

namespace hipfem {
// Device-resident scalars of one PCG solve.
struct PcgState {
    double rho, alpha, beta, normb, res, tol;
    double rmax, delta;   // reliable updates: residual at the last update, trigger factor
    int iter, done, bad, pause;
    int restart;          // reliable updates: restart the direction after the next update
};

// kernels of an iteration return at once once the solve is done, or while it
// is paused for a reliable update (the host then recomputes the residual)
__device__ __forceinline__ bool halted(const PcgState *st) { return (st->done | st->pause) != 0; }
}  // namespace hipfem

struct hipcsr_mtx {
    hipfem_ctx *ctx = nullptr;
    int n = 0;
    int64_t nnz = 0;
    bool has_structure = false;
    std::vector<int32_t> h_rowptr, h_colind;   // host copy of the pattern
    int32_t *d_rowptr = nullptr, *d_colind = nullptr, *d_diag = nullptr;
    double *d_val = nullptr;
    float *d_val32 = nullptr;                  // FP32 copy for mixed precision
    long val32_version = -1;
    std::vector<double> h_stage;               // host-side assemble() staging
    bool staged = false;
    long version = 0;
    long pattern_id = 0;                       // changes whenever the pattern is rebuilt
    // transpose pattern for timesT, built on first use
    int32_t *d_tptr = nullptr, *d_tcol = nullptr, *d_tmap = nullptr;
    double avg_row = 0.0;
    hipfem::SpmvShape shape;                   // SpMV kernel shape (hipcsr_set_spmv_shape)
    int stream_max[4] = {0, 0, 0, 0};          // largest block range for 8/16/32/64 rows per block
    // node-block view of the same CSR arrays (valid when the pattern is node-blocked)
    bool nodeblocked = false;
    int nbr = 0;                               // node rows = n / 3
    int32_t *d_bptr = nullptr, *d_bcol = nullptr;
    // row-group view (patterns with partially constrained nodes), see grp_row
    bool grouped = false;
    int ngrp = 0;
    int32_t *d_grow = nullptr, *d_gch = nullptr;
    uint32_t *d_ccol = nullptr;
    // element-by-element view: set when the values were assembled by a batched
    // evaluator and not changed since (mf_version == version); the FP32 PCG
    // of precision 2 then applies the element operator instead of the stored
    // values
    hfem_model *mf_model = nullptr;
    long mf_version = -1;
    bool mf_enabled = true;
};

namespace hipfem {
int ensure_partials(hipfem_ctx *ctx, size_t n);
int csr_alloc(hipcsr_mtx *m, int n, const std::vector<int32_t> &rowptr, const std::vector<int32_t> &colind);
int csr_free(hipcsr_mtx *m);
int csr_spmv_dev(hipcsr_mtx *m, const double *dx, double *dy);
int csr_refresh_fp32(hipcsr_mtx *m);
// launch helper: rows-per-lane group width chosen from the average row length
int spmv_lanes(const hipcsr_mtx *m);
int spmv_fp32_dev(hipcsr_mtx *m, const double *dx, double *dy);
long next_pattern_id();
// element-by-element operator of a linked evaluator: y = K_ff x (FP32
// element arithmetic, FP64 sum over elements, free equations only); returns
// at once inside kernels when the PCG state is halted
bool model_mf_available(const hipcsr_mtx *m);
int model_mf_apply_f32(hfem_model *md, const float *x, double *y, const PcgState *st);
// the slot form's element pass alone: every element's FP32 forces into its
// colour's slot vector; yslot / eqmask / ncolors describe the slots for a
// caller that sums them itself.  Sets *yslot = nullptr (and launches
// nothing) when the model does not use the slot form.
int model_mf_slots_f32(hfem_model *md, const float *x, const PcgState *st, const float **yslot,
                       const uint16_t **eqmask, int *ncolors);
void model_unlink(hfem_model *md, hipcsr_mtx *m);
}  // namespace hipfem

struct hfem_mlp;
namespace hipfem {
// surrogate network: owner context, material {E, nu, sig0, H} and the
// input / stress / tangent scale factors
void mlp_info(const hfem_mlp *m, hipfem_ctx **ctx, double mat4[4], float scales[3]);
// BF16 MFMA inference of n scaled rows (device arrays) on the context stream
int mlp_infer_bf16_dev(hfem_mlp *m, int n, const float *dX, float *dY);
}  // namespace hipfem
