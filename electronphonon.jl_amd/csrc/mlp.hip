// Band-energy surrogate: a small MLP (periodic Fourier features of k -> 2 hidden SiLU layers -> band
// energies) trained on the GPU on exact eigenvalues, then run in BF16 on the matrix cores to prune
// fine-grid k-points whose predicted bands all lie farther than a safety margin outside the energy
// window. The exact filter (src/filter.jl:40-100) stays the arbiter: survivors are re-solved exactly.
// No counterpart exists in the reference; this is the engine's own pre-filter.
//
// Training kernels (FP32): features, a tiled row-major GEMM with transpose flags (forward, data and
// weight gradients), SiLU forward/backward, MSE loss gradient, column sums (bias gradients), Adam.
// Inference kernel (BF16 operands, FP32 accumulation): the whole network for 16 points per wave on
// 16x16x32 bf16 MFMAs with the activations kept in LDS, fused with the window test.
#include "common.hpp"

namespace ephip {

// f[n, j] = cos(2 pi m_j . k_n), f[n, nf2 + j] = sin(2 pi m_j . k_n); argument reduced mod 1 in FP64.
__global__ void mlp_features_kernel(int64_t n, const double *__restrict__ xk, int nf2, const int *__restrict__ mvec,
                                    float *__restrict__ f) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nf2) return;
    const int64_t p = e / nf2;
    const int j = (int)(e % nf2);
    const double t = mvec[3 * j] * xk[3 * p] + mvec[3 * j + 1] * xk[3 * p + 1] + mvec[3 * j + 2] * xk[3 * p + 2];
    float c, s;
    cis2pi(t, c, s);
    f[p * 2 * nf2 + j] = c;
    f[p * 2 * nf2 + nf2 + j] = s;
}

// C[m, n] = sum_k opA[m, k] opB[k, n] (+ bias[n]) (+ beta C[m, n]), row-major operands:
// opA[m, k] = ta ? A[k*lda + m] : A[m*lda + k], opB[k, n] = tb ? B[n*ldb + k] : B[k*ldb + n].
// 64 x 64 tile per 256-thread workgroup, 4 x 4 outputs per thread, K in stages of 16.
// Split-K: grid z-slice z covers k in [z*kchunk, min(K, (z+1)*kchunk)) and writes C + z*cstride.
__global__ __launch_bounds__(256) void gemm_f32_kernel(int ta, int tb, int M, int N, int K, const float *__restrict__ A,
                                                       int64_t lda, const float *__restrict__ B, int64_t ldb,
                                                       float *__restrict__ C, int64_t ldc,
                                                       const float *__restrict__ bias, float beta, int kchunk,
                                                       int64_t cstride) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    C += blockIdx.z * cstride;
    float acc[4][4] = {};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        for (int e = tid; e < 64 * 16; e += 256) {
            const int mm = e % 64, kk = e / 64;
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < kend) v = ta ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
            As[kk][mm] = v;
            const int n = n0 + mm;
            float w = 0.f;
            if (n < N && k < kend) w = tb ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
            Bs[kk][mm] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
            if (m < M && n < N) {
                float v = acc[i][j];
                if (bias) v += bias[n];
                if (beta != 0.f) v += beta * C[(int64_t)m * ldc + n];
                C[(int64_t)m * ldc + n] = v;
            }
        }
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }

__global__ void silu_fwd_kernel(int64_t n, const float *__restrict__ z, float *__restrict__ a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) a[e] = silu(z[e]);
}

// gZ = gA * silu'(Z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void silu_bwd_kernel(int64_t n, const float *__restrict__ z, const float *__restrict__ ga,
                                float *__restrict__ gz) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float s = 1.f / (1.f + __expf(-z[e]));
    gz[e] = ga[e] * s * (1.f + z[e] * (1.f - s));
}

// gY = 2 (y - t) / (n_pts * nout); partial[block] = sum of (y - t)^2 over the block's elements
__global__ void mse_grad_kernel(int64_t n, const float *__restrict__ y, const float *__restrict__ t, float scale,
                                float *__restrict__ gy, double *__restrict__ partial) {
    __shared__ double red[256];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (e < n) {
        const float d = y[e] - t[e];
        gy[e] = 2.f * d * scale;
        d2 = (double)d * d;
    }
    red[threadIdx.x] = d2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[c] = sum_r X[r*ldx + c], one thread per column
__global__ void colsum_kernel(int rows, int cols, const float *__restrict__ x, int64_t ldx, float *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += x[(int64_t)r * ldx + c];
    out[c] = s;
}

__global__ void adam_kernel(int64_t n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void f32_to_bf16_kernel(int64_t n, const float *__restrict__ x, __bf16 *__restrict__ y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = static_cast<__bf16>(x[e]);
}

// Fused BF16 inference + window pruning. One wave = 16 points; LDS holds the wave's activations as
// bf16 [16][HMAX + 8]. Layer l: C (16 x out) = A (16 x in) W^T on 16x16x32 bf16 MFMAs (lane l holds
// A[l%16][8(l/16)+e], B[8(l/16)+e][l%16], C[4(l/16)+r][l%16]); bias + SiLU in FP32, rounded to bf16
// for the next layer. Output layer in FP32: keep[p] = any band within [wmin - margin, wmax + margin].
// Requirements (checked on the host): nf = 2*nf2, hidden and nout multiples of 32 / 16 as below.
constexpr int MLP_HMAX = 128;
__global__ __launch_bounds__(256) void mlp_bf16_prune_kernel(int64_t npts, const double *__restrict__ xk, int nf2,
                                                             const int *__restrict__ mvec, int hidden, int nout,
                                                             const __bf16 *__restrict__ W1, const float *__restrict__ b1,
                                                             const __bf16 *__restrict__ W2, const float *__restrict__ b2,
                                                             const __bf16 *__restrict__ W3, const float *__restrict__ b3,
                                                             float wmin, float wmax, float margin,
                                                             float *__restrict__ y_out, uint8_t *__restrict__ keep) {
    __shared__ __attribute__((aligned(16))) __bf16 act[4][2][16][MLP_HMAX + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (p0 >= npts) return;
    const int nf = 2 * nf2;
    // features of the wave's 16 points into act[wave][0]
    for (int e = lane; e < 16 * nf2; e += 64) {
        const int i = e / nf2, j = e % nf2;
        const int64_t p = p0 + i;
        float c = 0.f, s = 0.f;
        if (p < npts) {
            const double t = mvec[3 * j] * xk[3 * p] + mvec[3 * j + 1] * xk[3 * p + 1] + mvec[3 * j + 2] * xk[3 * p + 2];
            cis2pi(t, c, s);
        }
        act[wave][0][i][j] = static_cast<__bf16>(c);
        act[wave][0][i][nf2 + j] = static_cast<__bf16>(s);
    }
    wave_lds_sync();
    const int li = lane & 15, lk = 8 * (lane >> 4), q = lane >> 4;
    int cur = 0;
    // two hidden layers
    for (int layer = 0; layer < 2; layer++) {
        const int in = layer == 0 ? nf : hidden;
        const __bf16 *W = layer == 0 ? W1 : W2;
        const float *bb = layer == 0 ? b1 : b2;
        for (int cb = 0; cb < hidden / 16; cb++) {
            f32x4 acc{0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < in; k0 += 32) {
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&act[wave][cur][li][k0 + lk]);
                const bf16x8 w = *reinterpret_cast<const bf16x8 *>(&W[(int64_t)(16 * cb + li) * in + k0 + lk]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc, 0, 0, 0);
            }
            const float bias = bb[16 * cb + li];
#pragma unroll
            for (int r = 0; r < 4; r++) act[wave][cur ^ 1][4 * q + r][16 * cb + li] = static_cast<__bf16>(silu(acc[r] + bias));
        }
        wave_lds_sync();
        cur ^= 1;
    }
    // output layer + window test: point 4q + r is kept if any band of any column block is within
    // [wmin - margin, wmax + margin]; the OR over the 16 bands of a block is a ballot over its lanes.
    for (int cb = 0; cb < nout / 16; cb++) {
        f32x4 acc{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < hidden; k0 += 32) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&act[wave][cur][li][k0 + lk]);
            const bf16x8 w = *reinterpret_cast<const bf16x8 *>(&W3[(int64_t)(16 * cb + li) * hidden + k0 + lk]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc, 0, 0, 0);
        }
        const float bias = b3[16 * cb + li];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t p = p0 + 4 * q + r;
            const float y = acc[r] + bias;
            if (p < npts && y_out) y_out[p * nout + 16 * cb + li] = y;
            const bool in = (y >= wmin - margin) && (y <= wmax + margin);
            const unsigned long long b = __ballot(in);
            if (li == 0 && ((b >> (16 * q)) & 0xffffull) != 0ull && p < npts) keep[p] = 1;
        }
    }
}

__global__ void zero_u8_kernel(int64_t n, uint8_t *__restrict__ x) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = 0;
}

}  // namespace ephip

using namespace ephip;

static inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int ephip_mlp_features(int64_t npts, const double *xk, int32_t nf2, const int32_t *mvec, float *f,
                                  void *stream) {
    EPHIP_REQUIRE(npts >= 0 && nf2 >= 1, "mlp_features: bad sizes");
    if (npts == 0) return EPHIP_OK;
    EPHIP_REQUIRE(xk && mvec && f, "mlp_features: null device pointer");
    mlp_features_kernel<<<nblk(npts * nf2), 256, 0, as_stream(stream)>>>(npts, xk, nf2, mvec, f);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gemm_f32(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float *A, int64_t lda,
                              const float *B, int64_t ldb, float *C, int64_t ldc, const float *bias, float beta,
                              void *stream) {
    EPHIP_REQUIRE(M >= 0 && N >= 0 && K >= 0, "gemm_f32: negative size");
    EPHIP_REQUIRE((M + 63) / 64 <= 65535, "gemm_f32: M too large");
    if (M == 0 || N == 0) return EPHIP_OK;
    EPHIP_REQUIRE(A && B && C, "gemm_f32: null device pointer");
    EPHIP_REQUIRE(lda >= (ta ? M : K) && ldb >= (tb ? K : N) && ldc >= N, "gemm_f32: leading dimension too small");
    dim3 g((N + 63) / 64, (M + 63) / 64);
    gemm_f32_kernel<<<g, 256, 0, as_stream(stream)>>>(ta, tb, M, N, K, A, lda, B, ldb, C, ldc, bias, beta,
                                                      K > 0 ? K : 1, 0);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gemm_f32_splitk(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float *A,
                                     int64_t lda, const float *B, int64_t ldb, float *Cpart, int64_t ldc,
                                     int32_t nsplit, void *stream) {
    EPHIP_REQUIRE(M >= 0 && N >= 0 && K >= 0 && nsplit >= 1 && nsplit <= 65535, "gemm_f32_splitk: bad sizes");
    EPHIP_REQUIRE((M + 63) / 64 <= 65535, "gemm_f32_splitk: M too large");
    if (M == 0 || N == 0) return EPHIP_OK;
    EPHIP_REQUIRE(A && B && Cpart, "gemm_f32_splitk: null device pointer");
    EPHIP_REQUIRE(lda >= (ta ? M : K) && ldb >= (tb ? K : N) && ldc >= N, "gemm_f32_splitk: leading dimension too small");
    const int kchunk = ((K + nsplit - 1) / nsplit + 15) / 16 * 16;
    dim3 g((N + 63) / 64, (M + 63) / 64, nsplit);
    gemm_f32_kernel<<<g, 256, 0, as_stream(stream)>>>(ta, tb, M, N, K, A, lda, B, ldb, Cpart, ldc, nullptr, 0.f,
                                                      kchunk > 0 ? kchunk : 16, (int64_t)M * ldc);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_silu_fwd(int64_t n, const float *z, float *a, void *stream) {
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(z && a, "silu_fwd: null device pointer");
    silu_fwd_kernel<<<nblk(n), 256, 0, as_stream(stream)>>>(n, z, a);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_silu_bwd(int64_t n, const float *z, const float *ga, float *gz, void *stream) {
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(z && ga && gz, "silu_bwd: null device pointer");
    silu_bwd_kernel<<<nblk(n), 256, 0, as_stream(stream)>>>(n, z, ga, gz);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_mse_grad(int64_t n, const float *y, const float *t, float scale, float *gy, double *partial,
                              void *stream) {
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(y && t && gy && partial, "mse_grad: null device pointer");
    mse_grad_kernel<<<nblk(n), 256, 0, as_stream(stream)>>>(n, y, t, scale, gy, partial);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_colsum(int32_t rows, int32_t cols, const float *x, int64_t ldx, float *out, void *stream) {
    EPHIP_REQUIRE(rows >= 0 && cols >= 0 && ldx >= cols, "colsum: bad sizes");
    if (cols == 0) return EPHIP_OK;
    EPHIP_REQUIRE(x && out, "colsum: null device pointer");
    colsum_kernel<<<nblk(cols, 64), 64, 0, as_stream(stream)>>>(rows, cols, x, ldx, out);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_adam_step(int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1,
                               float beta2, float eps, int32_t step, void *stream) {
    EPHIP_REQUIRE(step >= 1, "adam_step: step counts from 1");
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(p && g && m && v, "adam_step: null device pointer");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    adam_kernel<<<nblk(n), 256, 0, as_stream(stream)>>>(n, p, g, m, v, lr, beta1, beta2, eps, c1, c2);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_f32_to_bf16(int64_t n, const float *x, void *y, void *stream) {
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(x && y, "f32_to_bf16: null device pointer");
    f32_to_bf16_kernel<<<nblk(n), 256, 0, as_stream(stream)>>>(n, x, static_cast<__bf16 *>(y));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_mlp_bf16_prune(int64_t npts, const double *xk, int32_t nf2, const int32_t *mvec, int32_t hidden,
                                    int32_t nout, const void *W1, const float *b1, const void *W2, const float *b2,
                                    const void *W3, const float *b3, float wmin, float wmax, float margin,
                                    float *y_out, uint8_t *keep, void *stream) {
    EPHIP_REQUIRE(npts >= 0, "mlp_bf16_prune: negative size");
    EPHIP_REQUIRE(nf2 >= 16 && (2 * nf2) % 32 == 0 && 2 * nf2 <= MLP_HMAX,
                  "mlp_bf16_prune: 2*nf2=%d must be a multiple of 32 in 32..%d", 2 * nf2, MLP_HMAX);
    EPHIP_REQUIRE(hidden % 32 == 0 && hidden >= 32 && hidden <= MLP_HMAX,
                  "mlp_bf16_prune: hidden=%d must be a multiple of 32 in 32..%d", hidden, MLP_HMAX);
    EPHIP_REQUIRE(nout % 16 == 0 && nout >= 16 && nout <= MLP_HMAX, "mlp_bf16_prune: nout=%d must be a multiple of 16",
                  nout);
    EPHIP_REQUIRE(margin >= 0.f, "mlp_bf16_prune: negative margin");
    if (npts == 0) return EPHIP_OK;
    EPHIP_REQUIRE(xk && mvec && W1 && b1 && W2 && b2 && W3 && b3 && keep, "mlp_bf16_prune: null device pointer");
    hipStream_t s = as_stream(stream);
    zero_u8_kernel<<<nblk(npts), 256, 0, s>>>(npts, keep);
    mlp_bf16_prune_kernel<<<nblk(npts, 64), 256, 0, s>>>(
        npts, xk, nf2, mvec, hidden, nout, static_cast<const __bf16 *>(W1), b1, static_cast<const __bf16 *>(W2), b2,
        static_cast<const __bf16 *>(W3), b3, wmin, wmax, margin, y_out, keep);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
