// hfem_mlp: learned constitutive surrogate for MisesMat.
//
// Network: 13 inputs -> 3 x 128 hidden (SiLU) -> 27 outputs, FP32 master
// weights in one flat buffer (W1 128x13, b1, W2 128x128, b2, W3 128x128, b3,
// W4 27x128, b4; row = output unit).
//   inputs  (eps_{n+1} - epsp_n) (6), epsp_n (6), kappa_n (1); the return
//           map of misesmat.C:181-263 depends on the strain only through
//           eps_{n+1} - epsp_n, so this difference is the strain input;
//           all scaled by E / sig0 (the yield strain is 1)
//   outputs sigma_{n+1} (6) and the upper triangle of the consistent
//           tangent (21, row-major), each as its departure from the elastic
//           trial: o_sigma = (D_e (eps - epsp_n) - sigma) / sig0 and
//           o_D = (D_e - D) / E, so both vanish on the elastic branch and the
//           network learns only the plastic correction
// Training data: random strain paths driven through the plain return
// mapping (mises_return / mises_tangent, the device form of MisesMat):
// each path starts virgin and takes 1..maxsteps Gaussian increments of
// every Voigt strain component; every step is one sample.
// Training: mean-squared error over the 27 scaled outputs, Adam, batches
// drawn with replacement by a counter-based hash; forward and backward are
// the FP32 kernels below (tiled GEMM with fused bias/SiLU epilogues).
// Inference: BF16 operands on MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation, weights staged in LDS in operand order.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hipfem_internal.h"
#include "hipfem_material.h"

namespace hipfem {

constexpr int MLP_IN = 13, MLP_H = 128, MLP_OUT = 27;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_H;
constexpr int OFF_W4 = OFF_B3 + MLP_H;
constexpr int OFF_B4 = OFF_W4 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = OFF_B4 + MLP_OUT;
constexpr int MLP_SPLITK = 32;   // weight-gradient slabs (workspace MLP_SPLITK x 128 x 128)

// --------------------------------------------------------------- hashing RNG
__host__ __device__ inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__host__ __device__ inline double unit_open(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__host__ __device__ inline int path_length(uint64_t seed, int64_t p, int maxsteps)
{
    return 1 + (int)(mix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(p + 1))) % (uint64_t)maxsteps);
}

// --------------------------------------------------------- training samples
__global__ void k_mlp_paths(int64_t npaths, const int64_t *__restrict__ off, int64_t nsamples, uint64_t seed,
                            double E, double nu, double sig0, double H, double stdev, float in_scale,
                            float sig_scale, float tan_scale, float *__restrict__ X, float *__restrict__ Y)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npaths) return;
    const int64_t o = off[p], L = off[p + 1] - o;
    double eps[6] = {0, 0, 0, 0, 0, 0}, epsp[6] = {0, 0, 0, 0, 0, 0}, kappa = 0.0;
    uint64_t ctr = mix64(seed * 0x632BE59BD9B4E019ull + (uint64_t)p);
    for (int64_t s = 0; s < L && o + s < nsamples; s++) {
#pragma unroll
        for (int i = 0; i < 6; i += 2) {   // Box-Muller pairs
            const uint64_t h1 = mix64(ctr++), h2 = mix64(ctr++);
            const double r = sqrt(-2.0 * log(unit_open(h1))), t = 6.283185307179586 * unit_open(h2);
            eps[i] += stdev * r * cos(t);
            eps[i + 1] += stdev * r * sin(t);
        }
        MisesOut m;
        mises_return(E, nu, sig0, H, eps, epsp, kappa, m);
        double D[36];
        mises_tangent(E, nu, sig0, H, kappa, m.kappa, m.sdev, D);
        float *x = X + (o + s) * MLP_IN;
        float *y = Y + (o + s) * MLP_OUT;
        double el[6], st[6];
#pragma unroll
        for (int i = 0; i < 6; i++) el[i] = eps[i] - epsp[i];
        elastic_trial(E, nu, el, st);
#pragma unroll
        for (int i = 0; i < 6; i++) {
            x[i] = (float)(el[i] * in_scale);
            x[6 + i] = (float)(epsp[i] * in_scale);
            y[i] = (float)((st[i] - m.sig[i]) * sig_scale);
        }
        x[12] = (float)(kappa * in_scale);
        int k = 0;
        for (int r = 0; r < 6; r++)
            for (int c = r; c < 6; c++, k++) y[6 + k] = (float)((elastic_upper(E, nu, k) - D[6 * r + c]) * tan_scale);
#pragma unroll
        for (int i = 0; i < 6; i++) epsp[i] = m.epsp[i];
        kappa = m.kappa;
    }
}

// ------------------------------------------------------------- FP32 kernels
__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + expf(-z)); }

// C[M x N] = sum_k A(m,k) B(k,n) with A(m,k) = A[m*sam + k*sak] and
// B(k,n) = B[k*sbk + n*sbn]; 64x64 tiles, 16-deep K slices, 4x4 outputs per
// thread; every output is one thread's fixed-order sum (deterministic).
// EPI 0: C = acc; 1: Z = acc + bias, C = silu(Z); 2: C = acc + bias.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A, int64_t sam,
                                              int64_t sak, const float *__restrict__ B, int64_t sbk, int64_t sbn,
                                              const float *__restrict__ bias, float *__restrict__ C, int64_t ldc,
                                              float *__restrict__ Z)
{
    __shared__ float As[16][65];
    __shared__ float Bs[16][65];
    const int tm = threadIdx.x / 16, tn = threadIdx.x % 16;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    // split K: slice z of gridDim.z sums k in [kb, ke) into its own C slab
    const int kc = (((K + gridDim.z - 1) / gridDim.z) + 15) / 16 * 16;
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    C += (int64_t)blockIdx.z * M * ldc;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += 16) {
        for (int t = threadIdx.x; t < 1024; t += 256) {
            int mm, kk;
            if (sam == 1) { mm = t % 64; kk = t / 64; } else { kk = t % 16; mm = t / 16; }
            const int m = m0 + mm, k = k0 + kk;
            As[kk][mm] = (m < M && k < ke) ? A[m * sam + k * sak] : 0.f;
        }
        for (int t = threadIdx.x; t < 1024; t += 256) {
            int nn, kk;
            if (sbn == 1) { nn = t % 64; kk = t / 64; } else { kk = t % 16; nn = t / 16; }
            const int n = n0 + nn, k = k0 + kk;
            Bs[kk][nn] = (n < N && k < ke) ? B[k * sbk + n * sbn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][tm * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tn * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + tm * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + tn * 4 + j;
            if (n >= N) continue;
            float v = acc[i][j];
            if (EPI == 1) {
                v += bias[n];
                Z[m * ldc + n] = v;
                v = v * sigmoidf_(v);
            } else if (EPI == 2) {
                v += bias[n];
            }
            C[m * ldc + n] = v;
        }
    }
}

// out = sum over the S split-K slabs of ws (fixed order)
__global__ void k_sum_slabs(int S, int64_t n, const float *__restrict__ ws, float *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < S; z++) s += ws[(int64_t)z * n + i];
        out[i] = s;
    }
}

// dZ = dH * silu'(Z), in place on dH
__global__ void k_silu_back(int64_t n, const float *__restrict__ Z, float *__restrict__ dH)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float z = Z[i], s = sigmoidf_(z);
        dH[i] *= s * (1.f + z * (1.f - s));
    }
}

// column sums of a [rows x cols] row-major matrix (bias gradients), fixed order
__global__ void k_colsum(int rows, int cols, const float *__restrict__ A, float *__restrict__ out)
{
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += 256) s += A[(int64_t)r * cols + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// dY = 2 (Yp - Y) / (rows * 27); per-block partial sums of the squared error
__global__ __launch_bounds__(256) void k_mse(int64_t n, const float *__restrict__ Yp, const float *__restrict__ Y,
                                             float scale, float *__restrict__ dY, double *__restrict__ partials)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float d = Yp[i] - Y[i];
        dY[i] = 2.f * scale * d;
        s += (double)d * (double)d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void k_sum_partials(int nb, const double *__restrict__ partials, double scale, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += partials[i];
    *out = s * scale;
}

// batch rows drawn with replacement: row = hash(seed, step, i) mod nsamples
__global__ void k_gather_batch(int64_t nsamples, const float *__restrict__ X, const float *__restrict__ Y, int B,
                               uint64_t seed, int64_t step, float *__restrict__ Xb, float *__restrict__ Yb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const int64_t r = (int64_t)(mix64(seed ^ mix64((uint64_t)step * 0x9E3779B97F4A7C15ull + (uint64_t)i)) %
                                (uint64_t)nsamples);
    for (int k = 0; k < MLP_IN; k++) Xb[(int64_t)i * MLP_IN + k] = X[r * MLP_IN + k];
    for (int k = 0; k < MLP_OUT; k++) Yb[(int64_t)i * MLP_OUT + k] = Y[r * MLP_OUT + k];
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// --------------------------------------------------------- BF16 inference
typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

// operand-order weight blocks: layer (KS k-steps of 16, NT n-tiles of 32);
// block (nt, ks) holds 64 lanes x 8 values: lane l -> output j = 32 nt + l%32,
// inputs k = 16 ks + 8 (l/32) + v
constexpr int BF_L1 = 4 * 1 * 512, BF_L2 = 4 * 8 * 512, BF_L3 = 4 * 8 * 512, BF_L4 = 1 * 8 * 512;
constexpr int BF_TOTAL = BF_L1 + BF_L2 + BF_L3 + BF_L4;   // bf16 values (76 KiB)
constexpr int ACT_LD = MLP_H + 8;                          // padded LDS row (bf16)

__device__ __forceinline__ float bf2f(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// one 32x32 output tile of 64 lanes: value r of lane l is row
// 8 (r/4) + 4 (l/32) + r%4, column l%32
__device__ __forceinline__ int tile_row(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// the activation tile of a wave is private to it: LDS operations of one
// wave complete in order, so a wave-level barrier (no block barrier) orders
// its stores before its own loads of the next layer
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// z silu(z) = z / (1 + e^-z) with the hardware exp2 and reciprocal (the
// result is rounded to BF16, 8 bits, right after), then to BF16 by
// v_cvt_pk_bf16_f32 (round to nearest even, as f2bf)
__device__ __forceinline__ uint16_t silu_bf16(float z)
{
    const float h = z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
    const __bf16 b = (__bf16)h;
    uint16_t u;
    memcpy(&u, &b, 2);
    return u;
}

constexpr int BF_WAVES = 8;   // waves per workgroup sharing the LDS weights

__global__ __launch_bounds__(64 * BF_WAVES) void k_mlp_bf16(int n, const float *__restrict__ X,
                                                             float *__restrict__ Y, const uint16_t *__restrict__ wpk,
                                                             const float *__restrict__ params)
{
    extern __shared__ uint16_t lds[];
    uint16_t *sW = lds;                                   // BF_TOTAL
    uint16_t *sA = lds + BF_TOTAL;                        // BF_WAVES waves x 32 x ACT_LD
    for (int i = threadIdx.x; i < BF_TOTAL / 8; i += 64 * BF_WAVES)
        reinterpret_cast<uint4 *>(sW)[i] = reinterpret_cast<const uint4 *>(wpk)[i];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    uint16_t *act = sA + wv * 32 * ACT_LD;
    const uint16_t *W1 = sW, *W2 = sW + BF_L1, *W3 = W2 + BF_L2, *W4 = W3 + BF_L3;
    const int ntile = (n + 31) / 32;
    __syncthreads();
    for (int base = blockIdx.x * BF_WAVES; base < ntile; base += gridDim.x * BF_WAVES) {
        const int tile = base + wv;
        const int row0 = tile * 32;
        // layer 1: A operand straight from X (13 inputs, zero padded to 16)
        bf16x8 a;
        {
            const int i = row0 + (l & 31);
            uint16_t v8[8];
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const int k = 8 * (l >> 5) + v;
                v8[v] = (i < n && k < MLP_IN) ? f2bf(X[(int64_t)i * MLP_IN + k]) : (uint16_t)0;
            }
            memcpy(&a, v8, 16);
        }
        f32x16 acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            acc[nt] = (f32x16){0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(W1 + (nt * 64 + l) * 8);
            acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nt], 0, 0, 0);
        }
        // hidden layers: bias + SiLU into LDS, read back as the next A operand
        for (int layer = 0; layer < 3; layer++) {
            const float *bias = params + (layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : OFF_B3);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const int j = nt * 32 + (l & 31);
                const float bj = bias[j];
#pragma unroll
                for (int r = 0; r < 16; r++) act[tile_row(l, r) * ACT_LD + j] = silu_bf16(acc[nt][r] + bj);
            }
            wave_lds_sync();
            const uint16_t *Wn = layer == 0 ? W2 : layer == 1 ? W3 : W4;
#pragma unroll
            for (int nt = 0; nt < 4; nt++) acc[nt] = (f32x16){0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (layer < 2) {
#pragma unroll
                for (int ks = 0; ks < 8; ks++) {
                    const bf16x8 av =
                        *reinterpret_cast<const bf16x8 *>(act + (l & 31) * ACT_LD + ks * 16 + 8 * (l >> 5));
#pragma unroll
                    for (int nt = 0; nt < 4; nt++) {
                        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Wn + ((nt * 8 + ks) * 64 + l) * 8);
                        acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b, acc[nt], 0, 0, 0);
                    }
                }
            } else {
#pragma unroll
                for (int ks = 0; ks < 8; ks++) {
                    const bf16x8 av =
                        *reinterpret_cast<const bf16x8 *>(act + (l & 31) * ACT_LD + ks * 16 + 8 * (l >> 5));
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Wn + (ks * 64 + l) * 8);
                    acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b, acc[0], 0, 0, 0);
                }
            }
            wave_lds_sync();
        }
        // output layer: 27 of 32 columns
        const int j = l & 31;
        if (j < MLP_OUT) {
            const float bj = params[OFF_B4 + j];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int i = row0 + tile_row(l, r);
                if (i < n) Y[(int64_t)i * MLP_OUT + j] = acc[0][r] + bj;
            }
        }
    }
}

// pack W (out x in, FP32) of one layer into operand-order BF16 blocks
__global__ void k_pack_bf16(const float *__restrict__ W, int nout, int nin, int NT, int KS, uint16_t *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NT * KS * 512) return;
    const int v = idx & 7, l = (idx >> 3) & 63, blk = idx >> 9;
    const int nt = blk / KS, ks = blk % KS;
    const int j = nt * 32 + (l & 31), k = ks * 16 + 8 * (l >> 5) + v;
    out[idx] = (j < nout && k < nin) ? f2bf(W[j * nin + k]) : (uint16_t)0;
}

}  // namespace hipfem

using namespace hipfem;

struct hfem_mlp {
    hipfem_ctx *ctx = nullptr;
    double E = 0, nu = 0, sig0 = 0, H = 0;
    float in_scale = 1, sig_scale = 1, tan_scale = 1;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    int64_t adam_t = 0;
    // training set
    int64_t nsamples = 0;
    float *d_X = nullptr, *d_Y = nullptr;
    // work buffers for a batch of up to cap rows
    int cap = 0;
    float *d_xb = nullptr, *d_yb = nullptr, *d_z[3] = {nullptr, nullptr, nullptr}, *d_h[3] = {nullptr, nullptr, nullptr};
    float *d_yp = nullptr, *d_dy = nullptr, *d_dh = nullptr;
    float *d_ws = nullptr;       // split-K workspace of the weight gradients
    double *d_loss = nullptr;
    // BF16 operand-order weights (valid when packed_version == param_version)
    uint16_t *d_wpk = nullptr;
    long param_version = 0, packed_version = -1;
};

static int mlp_ensure_batch(hfem_mlp *m, int B)
{
    if (B <= m->cap) return 0;
    void *ps[] = {m->d_xb, m->d_yb, m->d_z[0], m->d_z[1], m->d_z[2], m->d_h[0], m->d_h[1], m->d_h[2],
                  m->d_yp, m->d_dy, m->d_dh};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
    const size_t b = (size_t)B;
    HIPFEM_CHECK(hipMalloc(&m->d_xb, sizeof(float) * b * MLP_IN));
    HIPFEM_CHECK(hipMalloc(&m->d_yb, sizeof(float) * b * MLP_OUT));
    for (int i = 0; i < 3; i++) {
        HIPFEM_CHECK(hipMalloc(&m->d_z[i], sizeof(float) * b * MLP_H));
        HIPFEM_CHECK(hipMalloc(&m->d_h[i], sizeof(float) * b * MLP_H));
    }
    HIPFEM_CHECK(hipMalloc(&m->d_yp, sizeof(float) * b * MLP_OUT));
    HIPFEM_CHECK(hipMalloc(&m->d_dy, sizeof(float) * b * MLP_OUT));
    HIPFEM_CHECK(hipMalloc(&m->d_dh, sizeof(float) * b * MLP_H));
    if (!m->d_ws) HIPFEM_CHECK(hipMalloc(&m->d_ws, sizeof(float) * MLP_SPLITK * MLP_H * MLP_H));
    m->cap = B;
    return 0;
}

static dim3 gemm_grid(int M, int N) { return dim3((N + 63) / 64, (M + 63) / 64); }

// forward pass of B rows of xb: z/h of the hidden layers kept for backward
static int mlp_forward(hfem_mlp *m, int B, const float *xb, float *yp)
{
    hipStream_t s = m->ctx->stream;
    const float *P = m->d_p;
    k_gemm<1><<<gemm_grid(B, MLP_H), 256, 0, s>>>(B, MLP_H, MLP_IN, xb, MLP_IN, 1, P + OFF_W1, 1, MLP_IN, P + OFF_B1,
                                                  m->d_h[0], MLP_H, m->d_z[0]);
    k_gemm<1><<<gemm_grid(B, MLP_H), 256, 0, s>>>(B, MLP_H, MLP_H, m->d_h[0], MLP_H, 1, P + OFF_W2, 1, MLP_H,
                                                  P + OFF_B2, m->d_h[1], MLP_H, m->d_z[1]);
    k_gemm<1><<<gemm_grid(B, MLP_H), 256, 0, s>>>(B, MLP_H, MLP_H, m->d_h[1], MLP_H, 1, P + OFF_W3, 1, MLP_H,
                                                  P + OFF_B3, m->d_h[2], MLP_H, m->d_z[2]);
    k_gemm<2><<<gemm_grid(B, MLP_OUT), 256, 0, s>>>(B, MLP_OUT, MLP_H, m->d_h[2], MLP_H, 1, P + OFF_W4, 1, MLP_H,
                                                    P + OFF_B4, yp, MLP_OUT, nullptr);
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}

// gradients of the mean-squared error of B rows into d_g; loss into d_loss
static int mlp_backward(hfem_mlp *m, int B, const float *xb, const float *yb)
{
    hipStream_t s = m->ctx->stream;
    const float *P = m->d_p;
    float *G = m->d_g;
    int rc = ensure_partials(m->ctx, 1024);
    if (rc) return rc;
    const int64_t ny = (int64_t)B * MLP_OUT;
    const int nbm = (int)std::min<int64_t>(1024, (ny + 255) / 256);
    k_mse<<<nbm, 256, 0, s>>>(ny, m->d_yp, yb, 1.f / (float)ny, m->d_dy, m->ctx->d_partials);
    k_sum_partials<<<1, 64, 0, s>>>(nbm, m->ctx->d_partials, 1.0 / (double)ny, m->d_loss);
    // weight gradients reduce over the batch: split K into S slabs of at
    // least 256 rows, then sum the slabs in order
    const int S = std::max(1, std::min(MLP_SPLITK, B / 256));
    auto wgrad = [&](int M, int N, const float *A, int64_t sam, int64_t sak, const float *Bm, int64_t sbk,
                     int64_t sbn, float *out) {
        dim3 g = gemm_grid(M, N);
        g.z = S;
        k_gemm<0><<<g, 256, 0, s>>>(M, N, B, A, sam, sak, Bm, sbk, sbn, nullptr, m->d_ws, N, nullptr);
        const int64_t mn = (int64_t)M * N;
        k_sum_slabs<<<(int)std::min<int64_t>(1024, (mn + 255) / 256), 256, 0, s>>>(S, mn, m->d_ws, out);
    };
    // layer 4: dW4 = dY^T H3, db4 = colsum dY, dH3 = dY W4
    wgrad(MLP_OUT, MLP_H, m->d_dy, 1, MLP_OUT, m->d_h[2], MLP_H, 1, G + OFF_W4);
    k_colsum<<<MLP_OUT, 256, 0, s>>>(B, MLP_OUT, m->d_dy, G + OFF_B4);
    k_gemm<0><<<gemm_grid(B, MLP_H), 256, 0, s>>>(B, MLP_H, MLP_OUT, m->d_dy, MLP_OUT, 1, P + OFF_W4, MLP_H, 1,
                                                  nullptr, m->d_dh, MLP_H, nullptr);
    const int64_t nh = (int64_t)B * MLP_H;
    const int nbe = (int)std::min<int64_t>(4096, (nh + 255) / 256);
    const int offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
    for (int L = 2; L >= 0; L--) {
        k_silu_back<<<nbe, 256, 0, s>>>(nh, m->d_z[L], m->d_dh);   // dZ_L
        const float *in = L == 0 ? xb : m->d_h[L - 1];
        const int nin = L == 0 ? MLP_IN : MLP_H;
        wgrad(MLP_H, nin, m->d_dh, 1, MLP_H, in, nin, 1, G + offW[L]);
        k_colsum<<<MLP_H, 256, 0, s>>>(B, MLP_H, m->d_dh, G + offB[L]);
        if (L > 0) {
            // dH_{L-1} = dZ_L W_L: uses d_dh as input, so write into the z
            // buffer of layer L (no longer needed) and swap
            k_gemm<0><<<gemm_grid(B, MLP_H), 256, 0, s>>>(B, MLP_H, MLP_H, m->d_dh, MLP_H, 1, P + offW[L], MLP_H, 1,
                                                          nullptr, m->d_z[L], MLP_H, nullptr);
            std::swap(m->d_z[L], m->d_dh);
        }
    }
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}

static int mlp_free(hfem_mlp *m)
{
    void *ps[] = {m->d_p, m->d_g, m->d_m, m->d_v, m->d_X, m->d_Y, m->d_xb, m->d_yb, m->d_z[0], m->d_z[1],
                  m->d_z[2], m->d_h[0], m->d_h[1], m->d_h[2], m->d_yp, m->d_dy, m->d_dh, m->d_loss, m->d_wpk,
                  m->d_ws};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
    return 0;
}

extern "C" int hfem_mlp_nparams(void) { return MLP_NPARAM; }

extern "C" int hfem_mlp_create(hipfem_ctx *ctx, const double *mat4, uint64_t seed, hfem_mlp **out)
{
    HIPFEM_REQUIRE(ctx && mat4 && out, HIPFEM_EINVAL, "hfem_mlp_create: null argument");
    HIPFEM_REQUIRE(mat4[0] > 0 && mat4[2] > 0, HIPFEM_EINVAL, "hfem_mlp_create: E and sig0 must be positive");
    DeviceGuard g(ctx->device);
    hfem_mlp *m = new hfem_mlp();
    m->ctx = ctx;
    m->E = mat4[0];
    m->nu = mat4[1];
    m->sig0 = mat4[2];
    m->H = mat4[3];
    m->in_scale = (float)(m->E / m->sig0);
    m->sig_scale = (float)(1.0 / m->sig0);
    m->tan_scale = (float)(1.0 / m->E);
    auto fail = [&](int rc) { mlp_free(m); delete m; return rc; };
    // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    std::vector<float> p(MLP_NPARAM);
    const int fan[4] = {MLP_IN, MLP_H, MLP_H, MLP_H};
    const int offs[5] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4, MLP_NPARAM};
    for (int L = 0; L < 4; L++) {
        const double a = 1.0 / std::sqrt((double)fan[L]);
        for (int i = offs[L]; i < offs[L + 1]; i++)
            p[i] = (float)(a * (2.0 * unit_open(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)i)) - 1.0));
    }
    const size_t nb = sizeof(float) * MLP_NPARAM;
    if (hipMalloc(&m->d_p, nb) != hipSuccess || hipMalloc(&m->d_g, nb) != hipSuccess ||
        hipMalloc(&m->d_m, nb) != hipSuccess || hipMalloc(&m->d_v, nb) != hipSuccess ||
        hipMalloc(&m->d_loss, sizeof(double)) != hipSuccess) {
        set_error("hfem_mlp_create: out of device memory");
        return fail(HIPFEM_ENOMEM);
    }
    if (hipMemcpy(m->d_p, p.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->d_m, 0, nb) != hipSuccess || hipMemset(m->d_v, 0, nb) != hipSuccess ||
        hipMemset(m->d_g, 0, nb) != hipSuccess) {
        set_error("hfem_mlp_create: upload failed");
        return fail(HIPFEM_EHIP);
    }
    ctx_retain(ctx);
    *out = m;
    return 0;
}

extern "C" void hfem_mlp_destroy(hfem_mlp *m)
{
    if (!m) return;
    hipfem_ctx *c = m->ctx;
    {
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        mlp_free(m);
        delete m;
    }
    ctx_release(c);
}

extern "C" int hfem_mlp_scales(const hfem_mlp *m, double *scales3)
{
    HIPFEM_REQUIRE(m && scales3, HIPFEM_EINVAL, "null argument");
    scales3[0] = m->in_scale;
    scales3[1] = m->sig_scale;
    scales3[2] = m->tan_scale;
    return 0;
}

extern "C" int hfem_mlp_get_params(hfem_mlp *m, float *p)
{
    HIPFEM_REQUIRE(m && p, HIPFEM_EINVAL, "null argument");
    DeviceGuard g(m->ctx->device);
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(p, m->d_p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hfem_mlp_set_params(hfem_mlp *m, const float *p)
{
    HIPFEM_REQUIRE(m && p, HIPFEM_EINVAL, "null argument");
    DeviceGuard g(m->ctx->device);
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(m->d_p, p, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    m->param_version++;
    return 0;
}

extern "C" int hfem_mlp_generate(hfem_mlp *m, int64_t nsamples, uint64_t seed, double stdev, int maxsteps)
{
    HIPFEM_REQUIRE(m && nsamples > 0 && stdev > 0 && maxsteps > 0, HIPFEM_EINVAL, "generate: bad arguments");
    DeviceGuard g(m->ctx->device);
    hipStream_t s = m->ctx->stream;
    // path offsets: path p holds 1..maxsteps samples (same length rule on the device)
    std::vector<int64_t> off(1, 0);
    while (off.back() < nsamples) off.push_back(off.back() + path_length(seed, (int64_t)off.size() - 1, maxsteps));
    const int64_t npaths = (int64_t)off.size() - 1;
    int64_t *d_off = nullptr;
    HIPFEM_CHECK(hipStreamSynchronize(s));
    if (m->d_X) HIPFEM_CHECK(hipFree(m->d_X));
    if (m->d_Y) HIPFEM_CHECK(hipFree(m->d_Y));
    m->d_X = m->d_Y = nullptr;
    m->nsamples = 0;
    HIPFEM_CHECK(hipMalloc(&m->d_X, sizeof(float) * (size_t)nsamples * MLP_IN));
    HIPFEM_CHECK(hipMalloc(&m->d_Y, sizeof(float) * (size_t)nsamples * MLP_OUT));
    HIPFEM_CHECK(hipMalloc(&d_off, sizeof(int64_t) * off.size()));
    HIPFEM_CHECK(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, s));
    k_mlp_paths<<<(unsigned)((npaths + 127) / 128), 128, 0, s>>>(npaths, d_off, nsamples, seed, m->E, m->nu, m->sig0,
                                                                 m->H, stdev, m->in_scale, m->sig_scale,
                                                                 m->tan_scale, m->d_X, m->d_Y);
    HIPFEM_CHECK(hipGetLastError());
    HIPFEM_CHECK(hipStreamSynchronize(s));
    HIPFEM_CHECK(hipFree(d_off));
    m->nsamples = nsamples;
    return 0;
}

extern "C" int hfem_mlp_get_samples(hfem_mlp *m, int64_t first, int64_t count, float *X, float *Y)
{
    HIPFEM_REQUIRE(m && first >= 0 && count >= 0 && first + count <= m->nsamples, HIPFEM_ERANGE,
                   "get_samples: range outside the training set");
    DeviceGuard g(m->ctx->device);
    HIPFEM_CHECK(hipStreamSynchronize(m->ctx->stream));
    if (X) HIPFEM_CHECK(hipMemcpy(X, m->d_X + first * MLP_IN, sizeof(float) * count * MLP_IN, hipMemcpyDeviceToHost));
    if (Y) HIPFEM_CHECK(hipMemcpy(Y, m->d_Y + first * MLP_OUT, sizeof(float) * count * MLP_OUT, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hfem_mlp_train(hfem_mlp *m, int steps, int batch, double lr, double lr_final, uint64_t seed,
                              double *final_loss)
{
    HIPFEM_REQUIRE(m && steps >= 0 && batch > 0 && lr > 0, HIPFEM_EINVAL, "train: bad arguments");
    HIPFEM_REQUIRE(m->nsamples > 0, HIPFEM_EINVAL, "train: no training set (call hfem_mlp_generate)");
    DeviceGuard g(m->ctx->device);
    hipStream_t s = m->ctx->stream;
    int rc = mlp_ensure_batch(m, batch);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < steps; it++) {
        k_gather_batch<<<(batch + 255) / 256, 256, 0, s>>>(m->nsamples, m->d_X, m->d_Y, batch, seed, it, m->d_xb,
                                                           m->d_yb);
        rc = mlp_forward(m, batch, m->d_xb, m->d_yp);
        if (rc) return rc;
        rc = mlp_backward(m, batch, m->d_xb, m->d_yb);
        if (rc) return rc;
        m->adam_t++;
        // cosine decay from lr to lr_final
        const double c = 0.5 * (1.0 + std::cos(3.141592653589793 * (double)it / std::max(1, steps)));
        const float lrt = (float)(lr_final + (lr - lr_final) * c);
        const float bc1 = 1.f - std::pow(b1, (float)m->adam_t), bc2 = 1.f - std::pow(b2, (float)m->adam_t);
        k_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, s>>>(MLP_NPARAM, m->d_p, m->d_g, m->d_m, m->d_v, lrt, b1, b2, 1e-8f,
                                                         bc1, bc2);
        HIPFEM_CHECK(hipGetLastError());
    }
    m->param_version++;
    if (final_loss) {
        HIPFEM_CHECK(hipMemcpyAsync(final_loss, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPFEM_CHECK(hipStreamSynchronize(s));
    return 0;
}

// host-array helpers: run the FP32 forward (and backward) on n given rows
extern "C" int hfem_mlp_forward(hfem_mlp *m, int n, const float *X, float *Y)
{
    HIPFEM_REQUIRE(m && n >= 0 && X && Y, HIPFEM_EINVAL, "forward: bad arguments");
    DeviceGuard g(m->ctx->device);
    int rc = mlp_ensure_batch(m, std::max(n, 1));
    if (rc) return rc;
    hipStream_t s = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(m->d_xb, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, s));
    rc = mlp_forward(m, n, m->d_xb, m->d_yp);
    if (rc) return rc;
    HIPFEM_CHECK(hipMemcpyAsync(Y, m->d_yp, sizeof(float) * n * MLP_OUT, hipMemcpyDeviceToHost, s));
    HIPFEM_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int hfem_mlp_gradients(hfem_mlp *m, int n, const float *X, const float *Y, float *grad, double *loss)
{
    HIPFEM_REQUIRE(m && n > 0 && X && Y && grad, HIPFEM_EINVAL, "gradients: bad arguments");
    DeviceGuard g(m->ctx->device);
    int rc = mlp_ensure_batch(m, n);
    if (rc) return rc;
    hipStream_t s = m->ctx->stream;
    HIPFEM_CHECK(hipMemcpyAsync(m->d_xb, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, s));
    HIPFEM_CHECK(hipMemcpyAsync(m->d_yb, Y, sizeof(float) * n * MLP_OUT, hipMemcpyHostToDevice, s));
    rc = mlp_forward(m, n, m->d_xb, m->d_yp);
    if (rc) return rc;
    rc = mlp_backward(m, n, m->d_xb, m->d_yb);
    if (rc) return rc;
    HIPFEM_CHECK(hipMemcpyAsync(grad, m->d_g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s));
    if (loss) HIPFEM_CHECK(hipMemcpyAsync(loss, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPFEM_CHECK(hipStreamSynchronize(s));
    return 0;
}

namespace hipfem {
void mlp_info(const hfem_mlp *m, hipfem_ctx **ctx, double mat4[4], float scales[3])
{
    *ctx = m->ctx;
    mat4[0] = m->E;
    mat4[1] = m->nu;
    mat4[2] = m->sig0;
    mat4[3] = m->H;
    scales[0] = m->in_scale;
    scales[1] = m->sig_scale;
    scales[2] = m->tan_scale;
}

// BF16 inference of n rows of device X into device Y on the context stream
int mlp_infer_bf16_dev(hfem_mlp *m, int n, const float *dX, float *dY)
{
    hipStream_t s = m->ctx->stream;
    if (m->packed_version != m->param_version || !m->d_wpk) {
        if (!m->d_wpk) HIPFEM_CHECK(hipMalloc(&m->d_wpk, sizeof(uint16_t) * BF_TOTAL));
        uint16_t *w = m->d_wpk;
        k_pack_bf16<<<(BF_L1 + 255) / 256, 256, 0, s>>>(m->d_p + OFF_W1, MLP_H, MLP_IN, 4, 1, w);
        k_pack_bf16<<<(BF_L2 + 255) / 256, 256, 0, s>>>(m->d_p + OFF_W2, MLP_H, MLP_H, 4, 8, w + BF_L1);
        k_pack_bf16<<<(BF_L3 + 255) / 256, 256, 0, s>>>(m->d_p + OFF_W3, MLP_H, MLP_H, 4, 8, w + BF_L1 + BF_L2);
        k_pack_bf16<<<(BF_L4 + 255) / 256, 256, 0, s>>>(m->d_p + OFF_W4, MLP_OUT, MLP_H, 1, 8,
                                                        w + BF_L1 + BF_L2 + BF_L3);
        HIPFEM_CHECK(hipGetLastError());
        m->packed_version = m->param_version;
    }
    if (n <= 0) return 0;
    const size_t lds = sizeof(uint16_t) * (BF_TOTAL + BF_WAVES * 32 * ACT_LD);
    int dev = 0, maxlds = 0;
    HIPFEM_CHECK(hipGetDevice(&dev));
    HIPFEM_CHECK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIPFEM_REQUIRE((size_t)maxlds >= lds, HIPFEM_EINVAL,
                   "surrogate inference needs %zu bytes of LDS per workgroup, the device offers %d", lds, maxlds);
    const int ntile = (n + 31) / 32;
    // one workgroup per CU (the LDS holds one), each looping over its tiles
    const int grid = std::max(1, std::min((ntile + BF_WAVES - 1) / BF_WAVES, 256));
    k_mlp_bf16<<<grid, 64 * BF_WAVES, lds, s>>>(n, dX, dY, m->d_wpk, m->d_p);
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}
}  // namespace hipfem

extern "C" int hfem_mlp_infer_bf16(hfem_mlp *m, int n, const float *X, float *Y)
{
    HIPFEM_REQUIRE(m && n >= 0 && X && Y, HIPFEM_EINVAL, "infer: bad arguments");
    DeviceGuard g(m->ctx->device);
    hipStream_t s = m->ctx->stream;
    float *dX = nullptr, *dY = nullptr;
    HIPFEM_CHECK(hipMalloc(&dX, sizeof(float) * std::max(1, n) * MLP_IN));
    HIPFEM_CHECK(hipMalloc(&dY, sizeof(float) * std::max(1, n) * MLP_OUT));
    int rc = 0;
    if (n) HIPFEM_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, s));
    rc = mlp_infer_bf16_dev(m, n, dX, dY);
    if (!rc && n) {
        HIPFEM_CHECK(hipMemcpyAsync(Y, dY, sizeof(float) * n * MLP_OUT, hipMemcpyDeviceToHost, s));
    }
    HIPFEM_CHECK(hipStreamSynchronize(s));
    HIPFEM_CHECK(hipFree(dX));
    HIPFEM_CHECK(hipFree(dY));
    return rc;
}
