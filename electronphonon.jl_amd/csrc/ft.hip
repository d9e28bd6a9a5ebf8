// Phase-sum Fourier transform on MFMA:  out[d, j] = sum_r op[d, r] * exp(2 pi i R_r . x_j).
//
// This is the contraction of the reference's fourier_batched! (src/wannier/batched_fourier_core.jl:
// 117-139, a ZGEMM of op_r against a phase matrix built by fourier_phase!, :24-30). Here the phase
// matrix is never stored: each workgroup generates its (R x column) phase tile in LDS from the
// FP64 R.x argument reduced mod 1, and the complex product runs as a real GEMM on the matrix cores:
//
//   [out_re]   [op_re  -op_im] [P_re]        rows (d, re|im) interleaved, K = (r, re|im)
//   [out_im] = [op_im   op_re] [P_im]
//
// so the real A' operand (2 ndata x 2 nr) is formed on the fly from the complex op tile and
// no flop is wasted (8 real flops per complex multiply-add, as in ZGEMM). The MFMA lane layouts
// used here were checked on gfx950: f32 16x16x4 and bf16 16x16x32 return C[4(l/16)+r][l%16] in
// lane l, register r; f64 16x16x4 returns C[4r+l/16][l%16].
//
// Tiles: 256 threads = 4 waves in a 2 x 2 arrangement; a workgroup owns TM = 64 complex rows and
// TN columns; K advances KC = 16 complex R per LDS stage. FP32 and FP64 use the 16x16x4 MFMA
// (lane l holds A[l%16][l/16], B[l/16][l%16]); the split-BF16 form uses
// 16x16x32 bf16 MFMAs with hi/lo operand splitting and three products per real product.
#include "common.hpp"

namespace ephip {

template <class T>
__device__ __forceinline__ typename std::conditional<sizeof(T) == 4, f32x4, f64x4>::type
mfma16x16x4(T a, T b, typename std::conditional<sizeof(T) == 4, f32x4, f64x4>::type c) {
    if constexpr (sizeof(T) == 4)
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int FT_TM = 64;  // complex rows per workgroup tile
constexpr int FT_KC = 16;  // complex R per LDS stage
constexpr int FT_APAD = 8;
constexpr int FT_BPAD = 16;

// Separable phases. exp(2 pi i R.x) = prod_a exp(2 pi i R_a x_a): a workgroup tabulates the three
// per-axis factors for its columns once, for every integer R_a in the range irvec spans on axis a
// (FP64 argument reduced mod 1, then sincospi), and a phase costs two FP32 complex products (a few
// ulp) instead of an FP64 dot product and a sincos per (R, column). Used when the three spans add
// up to at most FT_TAB_MAX rows; otherwise the phases are generated directly.
constexpr int FT_TAB_MAX = 24;

struct FtAxes {
    int lo0, lo1, lo2;  // smallest R_a on each axis
    int b1, b2;         // first table row of axes 1 and 2 (axis 0 starts at row 0)
    int total;          // table rows in use
};

// Per-axis range of irvec[0..nr) (nr > 0), computed by the whole 256-thread workgroup.
__device__ FtAxes ft_axes(const int *__restrict__ irvec, int nr, int *red) {
    int mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int r = threadIdx.x; r < nr; r += blockDim.x)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int v = irvec[3 * r + a];
            mn[a] = min(mn[a], v);
            mx[a] = max(mx[a], v);
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], off, 64));
        }
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            red[6 * wave + a] = mn[a];
            red[6 * wave + 3 + a] = mx[a];
        }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; a++) {
        mn[a] = red[a];
        mx[a] = red[3 + a];
        for (int w = 1; w < nw; w++) {
            mn[a] = min(mn[a], red[6 * w + a]);
            mx[a] = max(mx[a], red[6 * w + 3 + a]);
        }
    }
    FtAxes ax;
    ax.lo0 = mn[0];
    ax.lo1 = mn[1];
    ax.lo2 = mn[2];
    ax.b1 = mx[0] - mn[0] + 1;
    ax.b2 = ax.b1 + mx[1] - mn[1] + 1;
    ax.total = ax.b2 + mx[2] - mn[2] + 1;
    return ax;
}

// tab[row][j] for the workgroup's TN columns (zero for columns past nk).
template <int TN>
__device__ void ft_fill_table(float2 (*tab)[TN], const FtAxes &ax, const double *__restrict__ xk,
                              const int *__restrict__ col_index, int64_t j0, int nk) {
    for (int e = threadIdx.x; e < ax.total * TN; e += blockDim.x) {
        const int row = e / TN, j = e % TN;
        const int a = row < ax.b1 ? 0 : (row < ax.b2 ? 1 : 2);
        const int ra = a == 0 ? ax.lo0 + row : (a == 1 ? ax.lo1 + row - ax.b1 : ax.lo2 + row - ax.b2);
        float c = 0.f, s = 0.f;
        if (j0 + j < nk) {
            const int64_t pt = col_index ? col_index[j0 + j] : j0 + j;
            cis2pi((double)ra * xk[3 * pt + a], c, s);
        }
        tab[row][j] = make_float2(c, s);
    }
}

// Phase of R = (R[0], R[1], R[2]) for table column j (tab: rows of ld columns).
__device__ __forceinline__ void ft_table_phase(const float2 *tab, int ld, const FtAxes &ax,
                                               const int *__restrict__ R, int j, float &c, float &s) {
    const float2 p0 = tab[(R[0] - ax.lo0) * ld + j];
    const float2 p1 = tab[(ax.b1 + R[1] - ax.lo1) * ld + j];
    const float2 p2 = tab[(ax.b2 + R[2] - ax.lo2) * ld + j];
    const float r01 = p0.x * p1.x - p0.y * p1.y, i01 = p0.x * p1.y + p0.y * p1.x;
    c = r01 * p2.x - i01 * p2.y;
    s = r01 * p2.y + i01 * p2.x;
}

// Optional batch of independent contractions, one per grid z-slice u ("unit"): unit u reads its
// operand at op + u*op_stride, writes at out + u*out_stride, has ncol[u] columns (or nk), and column
// j uses point xk[col_index[u*col_stride + j]] (or xk[j]). A zero-initialised FtBatch is one plain call.
struct FtBatch {
    int64_t op_stride = 0;
    int64_t out_stride = 0;
    const int *col_index = nullptr;
    int64_t col_stride = 0;
    const int *ncol = nullptr;
};

template <class T, int TN>
__global__ __launch_bounds__(256) void ft_mfma_kernel(int64_t ndata, int nr, const cplx<T> *__restrict__ op,
                                                      int64_t ldop, const int *__restrict__ irvec, int nk,
                                                      const double *__restrict__ xk, cplx<T> *__restrict__ out,
                                                      int64_t ldout, FtBatch bt) {
    using acc_t = typename std::conditional<sizeof(T) == 4, f32x4, f64x4>::type;
    constexpr int NCB = TN / 32;  // 16-column blocks per wave
    __shared__ cplx<T> As[FT_KC][FT_TM + FT_APAD];
    __shared__ T Bs[FT_KC][2][TN + FT_BPAD];
    constexpr bool kSep = sizeof(T) == 4;  // separable phases for FP32 (FP64 keeps direct sincos)
    __shared__ float2 tab[kSep ? FT_TAB_MAX : 1][kSep ? TN : 1];
    __shared__ int red[24];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t d0 = (int64_t)blockIdx.y * FT_TM;
    const int64_t j0 = (int64_t)blockIdx.x * TN;
    {
        const int u = blockIdx.z;
        if (bt.ncol) nk = bt.ncol[u];
        if (j0 >= nk) return;  // uniform for the block: nothing of this unit in this column tile
        op += u * bt.op_stride;
        out += u * bt.out_stride;
        if (bt.col_index) bt.col_index += u * bt.col_stride;
    }
    bool use_tab = false;
    FtAxes ax{};
    if constexpr (kSep) {
        ax = ft_axes(irvec, nr, red);
        use_tab = ax.total <= FT_TAB_MAX;
        if (use_tab) ft_fill_table<TN>(tab, ax, xk, bt.col_index, j0, nk);
    }

    // The phase column this thread generates is fixed for the whole kernel; its R rows are uniform
    // per wavefront (TN is a multiple of 64).
    const int pj = tid % TN;
    const int prl0 = __builtin_amdgcn_readfirstlane(tid / TN);
    constexpr int PSTEP = 256 / TN;
    double x0 = 0, x1 = 0, x2 = 0;
    const bool pvalid = (j0 + pj) < nk;
    if (pvalid) {
        const int64_t pt = bt.col_index ? bt.col_index[j0 + pj] : j0 + pj;
        x0 = xk[3 * pt];
        x1 = xk[3 * pt + 1];
        x2 = xk[3 * pt + 2];
    }

    acc_t acc[4][NCB];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    for (int r0 = 0; r0 < nr; r0 += FT_KC) {
        // op tile: FT_TM x FT_KC complex, coalesced along d
        for (int e = tid; e < FT_TM * FT_KC; e += 256) {
            const int dl = e % FT_TM, rl = e / FT_TM;
            const int64_t d = d0 + dl;
            const int r = r0 + rl;
            cplx<T> v{0, 0};
            if (d < ndata && r < nr) v = op[d + ldop * r];
            As[rl][dl] = v;
        }
        // phase tile: FT_KC x TN
        if (use_tab) {
            for (int rl = prl0; rl < FT_KC; rl += PSTEP) {
                const int r = r0 + rl;
                float c = 0.f, s = 0.f;
                if (r < nr) ft_table_phase(&tab[0][0], TN, ax, irvec + 3 * r, pj, c, s);
                Bs[rl][0][pj] = c;
                Bs[rl][1][pj] = s;
            }
        } else {
            for (int rl = prl0; rl < FT_KC; rl += PSTEP) {
                const int r = r0 + rl;
                T c = 0, s = 0;
                if (pvalid && r < nr) {
                    const double t = irvec[3 * r] * x0 + irvec[3 * r + 1] * x1 + irvec[3 * r + 2] * x2;
                    cis2pi(t, c, s);
                }
                Bs[rl][0][pj] = c;
                Bs[rl][1][pj] = s;
            }
        }
        __syncthreads();
        // The f32 MFMA returns row 4(l/16)+r of the 16x16 block in lane l, register r; the f64 MFMA
        // returns row 4r + l/16. Feeding the f64 A operand with rows permuted by
        // i -> 4(i%4) + i/4 makes both land in the same (lane, register) slots.
        const int il = lane & 15;
        const int irow = (sizeof(T) == 8) ? (4 * (il & 3) + (il >> 2)) : il;
#pragma unroll
        for (int ks = 0; ks < FT_KC / 2; ks++) {
            const int kk = lane >> 4;
            const int rl = 2 * ks + (kk >> 1);
            const int cc = kk & 1;
            T a[4], b[NCB];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const int rr = wm * 64 + 16 * rb + irow;
                const cplx<T> v = As[rl][rr >> 1];
                a[rb] = (rr & 1) == 0 ? (cc == 0 ? v.re : -v.im) : (cc == 0 ? v.im : v.re);
            }
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) b[cb] = Bs[rl][cc][wn * (TN / 2) + 16 * cb + (lane & 15)];
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) acc[rb][cb] = mfma16x16x4<T>(a[rb], b[cb], acc[rb][cb]);
        }
        __syncthreads();
    }

    const int q = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int64_t d = d0 + (wm * 64 + 16 * rb) / 2 + 2 * q;
            const int64_t j = j0 + wn * (TN / 2) + 16 * cb + (lane & 15);
            if (j < nk) {
                if (d < ndata) out[d + ldout * j] = cplx<T>{acc[rb][cb][0], acc[rb][cb][1]};
                if (d + 1 < ndata) out[d + 1 + ldout * j] = cplx<T>{acc[rb][cb][2], acc[rb][cb][3]};
            }
        }
}

// Double-buffered form of ft_mfma_kernel: two LDS stages; the next stage's operand loads are issued
// before the current stage's MFMAs and its phase entries are generated one per K step between them,
// so global latency and the sincos work overlap the matrix cores; one barrier per stage. KC (complex R
// per stage) is a tuning parameter.
template <class T, int TN, int KC>
__global__ __launch_bounds__(256) void ft_mfma_db_kernel(int64_t ndata, int nr, const cplx<T> *__restrict__ op,
                                                         int64_t ldop, const int *__restrict__ irvec, int nk,
                                                         const double *__restrict__ xk, cplx<T> *__restrict__ out,
                                                         int64_t ldout, FtBatch bt) {
    using acc_t = typename std::conditional<sizeof(T) == 4, f32x4, f64x4>::type;
    constexpr int NCB = TN / 32;
    constexpr int AE = FT_TM * KC / 256;  // op elements staged per thread per stage
    constexpr int PSTEP = 256 / TN;
    constexpr int BE = KC / PSTEP;        // phase entries per thread per stage (<= KC/2)
    static_assert(AE >= 1 && BE >= 1 && BE <= KC / 2, "tile parameters");
    __shared__ cplx<T> As[2][KC][FT_TM + FT_APAD];
    __shared__ T Bs[2][KC][2][TN + FT_BPAD];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t d0 = (int64_t)blockIdx.y * FT_TM;
    const int64_t j0 = (int64_t)blockIdx.x * TN;
    {
        const int u = blockIdx.z;
        if (bt.ncol) nk = bt.ncol[u];
        if (j0 >= nk) return;
        op += u * bt.op_stride;
        out += u * bt.out_stride;
        if (bt.col_index) bt.col_index += u * bt.col_stride;
    }
    const int pj = tid % TN;
    const int prl0 = tid / TN;
    double x0 = 0, x1 = 0, x2 = 0;
    const bool pvalid = (j0 + pj) < nk;
    if (pvalid) {
        const int64_t pt = bt.col_index ? bt.col_index[j0 + pj] : j0 + pj;
        x0 = xk[3 * pt];
        x1 = xk[3 * pt + 1];
        x2 = xk[3 * pt + 2];
    }
    acc_t acc[4][NCB];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    cplx<T> areg[AE];
    T creg[BE], sreg[BE];
    auto load_a = [&](int r0) {
#pragma unroll
        for (int i = 0; i < AE; i++) {
            const int e = tid + 256 * i, dl = e % FT_TM, rl = e / FT_TM;
            const int64_t d = d0 + dl;
            const int r = r0 + rl;
            areg[i] = (d < ndata && r < nr) ? op[d + ldop * r] : cplx<T>{0, 0};
        }
    };
    auto phase_one = [&](int r0, int i) {
        const int r = r0 + prl0 + PSTEP * i;
        T c = 0, s = 0;
        if (pvalid && r < nr) {
            const double t = irvec[3 * r] * x0 + irvec[3 * r + 1] * x1 + irvec[3 * r + 2] * x2;
            cis2pi(t, c, s);
        }
        creg[i] = c;
        sreg[i] = s;
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < AE; i++) {
            const int e = tid + 256 * i;
            As[buf][e / FT_TM][e % FT_TM] = areg[i];
        }
#pragma unroll
        for (int i = 0; i < BE; i++) {
            Bs[buf][prl0 + PSTEP * i][0][pj] = creg[i];
            Bs[buf][prl0 + PSTEP * i][1][pj] = sreg[i];
        }
    };

    load_a(0);
#pragma unroll
    for (int i = 0; i < BE; i++) phase_one(0, i);
    stash(0);
    __syncthreads();

    const int il = lane & 15;
    const int irow = (sizeof(T) == 8) ? (4 * (il & 3) + (il >> 2)) : il;  // see ft_mfma_kernel
    const int nchunks = (nr + KC - 1) / KC;
    for (int ch = 0; ch < nchunks; ch++) {
        const int buf = ch & 1;
        const int rn = (ch + 1) * KC;
        const bool more = ch + 1 < nchunks;
        if (more) load_a(rn);
#pragma unroll
        for (int ks = 0; ks < KC / 2; ks++) {
            if (ks < BE && more) phase_one(rn, ks);
            const int kk = lane >> 4;
            const int rl = 2 * ks + (kk >> 1);
            const int cc = kk & 1;
            T a[4], b[NCB];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const int rr = wm * 64 + 16 * rb + irow;
                const cplx<T> v = As[buf][rl][rr >> 1];
                a[rb] = (rr & 1) == 0 ? (cc == 0 ? v.re : -v.im) : (cc == 0 ? v.im : v.re);
            }
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) b[cb] = Bs[buf][rl][cc][wn * (TN / 2) + 16 * cb + (lane & 15)];
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) acc[rb][cb] = mfma16x16x4<T>(a[rb], b[cb], acc[rb][cb]);
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
    }

    const int q = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int64_t d = d0 + (wm * 64 + 16 * rb) / 2 + 2 * q;
            const int64_t j = j0 + wn * (TN / 2) + 16 * cb + (lane & 15);
            if (j < nk) {
                if (d < ndata) out[d + ldout * j] = cplx<T>{acc[rb][cb][0], acc[rb][cb][1]};
                if (d + 1 < ndata) out[d + 1 + ldout * j] = cplx<T>{acc[rb][cb][2], acc[rb][cb][3]};
            }
        }
}

// Kernel variant per contraction kind, set by the tuner (ephip_ft_set_variant). Kinds:
// 0 = FP32 wide (nk > 64), 1 = FP32 narrow (nk <= 64), 2 = FP64. Variants: 0 = single-buffered,
// 1 = double-buffered KC 16, 2 = double-buffered KC 8.
static int g_ft_variant[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

template <class T, int TN>
static void launch_ft(int variant, dim3 g, hipStream_t s, int64_t ndata, int nr, const cplx<T> *op, int64_t ldop,
                      const int *irvec, int nk, const double *xk, cplx<T> *out, int64_t ldout, FtBatch bt) {
    if (variant == 1)
        ft_mfma_db_kernel<T, TN, 16><<<g, 256, 0, s>>>(ndata, nr, op, ldop, irvec, nk, xk, out, ldout, bt);
    else if (variant == 2)
        ft_mfma_db_kernel<T, TN, 8><<<g, 256, 0, s>>>(ndata, nr, op, ldop, irvec, nk, xk, out, ldout, bt);
    else
        ft_mfma_kernel<T, TN><<<g, 256, 0, s>>>(ndata, nr, op, ldop, irvec, nk, xk, out, ldout, bt);
}

// ---------------------------------------------------------------------------------------------
// Split-BF16 form. Every real operand x is written as x_hi + x_lo with x_hi = bf16(x) and
// x_lo = bf16(x - x_hi); the product keeps hi*hi + hi*lo + lo*hi (the lo*lo term is below FP32
// resolution of the sum) and accumulates in FP32 on the 16x16x32 bf16 MFMA. One K step covers
// 32 real K = 16 complex R, i.e. one whole LDS stage. Lane l holds A[l%16][8(l/16)+e] and
// B[8(l/16)+e][l%16], e = 0..7.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
    hi = static_cast<__bf16>(x);
    lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

template <int TN>
__global__ __launch_bounds__(256) void ft_bf16x3_kernel(int64_t ndata, int nr, const cplx<float> *__restrict__ op,
                                                        int64_t ldop, const int *__restrict__ irvec, int nk,
                                                        const double *__restrict__ xk, cplx<float> *__restrict__ out,
                                                        int64_t ldout, FtBatch bt) {
    constexpr int NCB = TN / 32;
    constexpr int KR = 2 * FT_KC;  // real K per stage = 32
    // Pre-split operands, stored row-major along K so each lane reads its 8 K values as 16 bytes.
    __shared__ __attribute__((aligned(16))) __bf16 Ahi[2 * FT_TM][KR + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Alo[2 * FT_TM][KR + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bhi[TN][KR + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Blo[TN][KR + 8];
    __shared__ float2 tab[FT_TAB_MAX][TN];
    __shared__ int red[24];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t d0 = (int64_t)blockIdx.y * FT_TM;
    const int64_t j0 = (int64_t)blockIdx.x * TN;
    {
        const int u = blockIdx.z;
        if (bt.ncol) nk = bt.ncol[u];
        if (j0 >= nk) return;  // uniform for the block: nothing of this unit in this column tile
        op += u * bt.op_stride;
        out += u * bt.out_stride;
        if (bt.col_index) bt.col_index += u * bt.col_stride;
    }
    const FtAxes ax = ft_axes(irvec, nr, red);
    const bool use_tab = ax.total <= FT_TAB_MAX;
    if (use_tab) ft_fill_table<TN>(tab, ax, xk, bt.col_index, j0, nk);

    const int pj = tid % TN;
    const int prl0 = __builtin_amdgcn_readfirstlane(tid / TN);
    constexpr int PSTEP = 256 / TN;
    double x0 = 0, x1 = 0, x2 = 0;
    const bool pvalid = (j0 + pj) < nk;
    if (pvalid) {
        const int64_t pt = bt.col_index ? bt.col_index[j0 + pj] : j0 + pj;
        x0 = xk[3 * pt];
        x1 = xk[3 * pt + 1];
        x2 = xk[3 * pt + 2];
    }

    f32x4 acc[4][NCB];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < NCB; j++) acc[i][j] = f32x4{0, 0, 0, 0};

    for (int r0 = 0; r0 < nr; r0 += FT_KC) {
        // A' rows (d, re|im), K = (r, re|im): [re -im; im re]
        for (int e = tid; e < FT_TM * FT_KC; e += 256) {
            const int dl = e % FT_TM, rl = e / FT_TM;
            const int64_t d = d0 + dl;
            const int r = r0 + rl;
            cplx<float> v{0.f, 0.f};
            if (d < ndata && r < nr) v = op[d + ldop * r];
            __bf16 hr, lr, hi, li;
            split_bf16(v.re, hr, lr);
            split_bf16(v.im, hi, li);
            // rows (d, re) = [re, -im] and (d, im) = [im, re] along K = (r, re|im); negation is exact
            *reinterpret_cast<bf16x2 *>(&Ahi[2 * dl][2 * rl]) = bf16x2{hr, -hi};
            *reinterpret_cast<bf16x2 *>(&Alo[2 * dl][2 * rl]) = bf16x2{lr, -li};
            *reinterpret_cast<bf16x2 *>(&Ahi[2 * dl + 1][2 * rl]) = bf16x2{hi, hr};
            *reinterpret_cast<bf16x2 *>(&Alo[2 * dl + 1][2 * rl]) = bf16x2{li, lr};
        }
        for (int rl = prl0; rl < FT_KC; rl += PSTEP) {
            const int r = r0 + rl;
            float c = 0.f, s = 0.f;
            if (r < nr) {
                if (use_tab) {
                    ft_table_phase(&tab[0][0], TN, ax, irvec + 3 * r, pj, c, s);
                } else if (pvalid) {
                    const double t = irvec[3 * r] * x0 + irvec[3 * r + 1] * x1 + irvec[3 * r + 2] * x2;
                    cis2pi(t, c, s);
                }
            }
            __bf16 hc, lc, hs, ls;
            split_bf16(c, hc, lc);
            split_bf16(s, hs, ls);
            // (cos, sin) pairs are adjacent along K: one 32-bit LDS store per hi and lo
            *reinterpret_cast<bf16x2 *>(&Bhi[pj][2 * rl]) = bf16x2{hc, hs};
            *reinterpret_cast<bf16x2 *>(&Blo[pj][2 * rl]) = bf16x2{lc, ls};
        }
        __syncthreads();
        const int kb = 8 * (lane >> 4);
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const int rr = wm * 64 + 16 * rb + (lane & 15);
            ah[rb] = *reinterpret_cast<const bf16x8 *>(&Ahi[rr][kb]);
            al[rb] = *reinterpret_cast<const bf16x8 *>(&Alo[rr][kb]);
        }
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int jl = wn * (TN / 2) + 16 * cb + (lane & 15);
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&Bhi[jl][kb]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&Blo[jl][kb]);
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], bh, acc[rb][cb], 0, 0, 0);
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bl, acc[rb][cb], 0, 0, 0);
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], bh, acc[rb][cb], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    const int q = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int64_t d = d0 + (wm * 64 + 16 * rb) / 2 + 2 * q;
            const int64_t j = j0 + wn * (TN / 2) + 16 * cb + (lane & 15);
            if (j < nk) {
                if (d < ndata) out[d + ldout * j] = cplx<float>{acc[rb][cb][0], acc[rb][cb][1]};
                if (d + 1 < ndata) out[d + 1 + ldout * j] = cplx<float>{acc[rb][cb][2], acc[rb][cb][3]};
            }
        }
}

// ---------------------------------------------------------------------------------------------
// Pre-packed split-BF16 form. Both operands are split once into bf16 hi/lo and stored in the order
// the 16x16x32 MFMA consumes them, so the kernel spends its VALU budget on nothing but addresses:
//
//  * operand A (ndata x nr complex, per unit) -> apk, stage-major: for stage s (R = 16 s .. 16 s + 15,
//    zero past nr) and row d, 64 bf16 at apk[((u*nrs + s)*ndata + d)*64]: re_hi[16] im_hi[16]
//    re_lo[16] im_lo[16]. A lane's fragment (row l%16, K chunk l/16) is 16 contiguous bytes.
//  * phases e^{2 pi i R.x_p} of a point table -> bpk: for stage s and point p, 64 bf16 at
//    bpk[(s*npts + p)*64]: hi = im[16] re[16], then lo likewise.
//
// With K = (re of 16 R, im of 16 R) the complex product of one stage is two real products,
//   out_re += [A_re A_im] [P_re; -P_im],   out_im += [A_re A_im] [P_im; P_re],
// so A is never duplicated. [P_im; P_re] is a point's hi (or lo) block as stored; [P_re; -P_im] is the
// same block read two 8-element chunks further on, with the sign bits of lanes 32..63 flipped. Each
// term is hi*hi + hi*lo + lo*hi (three MFMAs, FP32 accumulation).
// Columns are staged per stage through double-buffered LDS with a 144-byte pitch (conflict-free
// 16-byte reads), A fragments go straight from global memory to registers.
// ---------------------------------------------------------------------------------------------
constexpr int FTP_BLD = 72;  // LDS pitch of one column's 64 bf16 (144 bytes)

// Workgroup of 4 waves as WR (rows) x 4/WR (columns); a wave owns RB 16-row blocks x CB 16-column
// blocks. Rows past ndata and columns past the unit's count read a valid row/column (clamped) and
// are not stored, so no load is predicated. Offsets are 32-bit against per-stage uniform bases.
// PF: the next stage's A fragments are loaded into a second register set before the current
// stage's MFMAs (costs 8 RB registers; without it the loads are issued at the top of each stage).
// Flip the sign bits selected by mask in each 32-bit pair of bf16 values.
__device__ __forceinline__ bf16x8 bf16_flip(bf16x8 v, uint32_t mask) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w = __builtin_bit_cast(u32x4, v);
    w ^= mask;
    return __builtin_bit_cast(bf16x8, w);
}

template <int RB, int CB, int WR>
struct FtpShape {
    static constexpr int WC = 4 / WR;
    static constexpr int TM = WR * RB * 16;       // complex rows per workgroup
    static constexpr int TN = WC * CB * 16;       // columns per workgroup
    static constexpr int PPT = TN * 8 / 256;      // 16-byte phase pieces per thread and stage
    static_assert(TN * 8 % 256 == 0, "phase pieces must split evenly over the workgroup");
};

// Main loop: accumulates the workgroup tile (rows d0.., columns j0..) of one unit into accr/acci.
// apk and col_index point at the unit's operand and column list; Bs is 2 x TN x FTP_BLD bf16 of LDS.
// MID: the next stage's phase pieces are written to LDS halfway through the MFMA sequence instead of
// after it, so the LDS write latency overlaps the second half.
template <int RB, int CB, int WR, bool PF, bool MID = false>
__device__ __forceinline__ void ft_bf16p_mainloop(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                  const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                  const int *__restrict__ col_index, int64_t j0, int64_t d0,
                                                  __bf16 *Bs, f32x4 (&accr)[RB][CB], f32x4 (&acci)[RB][CB]) {
    using S = FtpShape<RB, CB, WR>;
    constexpr int TN = S::TN, PPT = S::PPT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave % WR, wc = wave / WR;
    const int c = lane >> 4, il = lane & 15;

    uint32_t bsrc[PPT], bdst[PPT];
#pragma unroll
    for (int t = 0; t < PPT; t++) {
        const int e = tid + 256 * t, jl = e / 8, pc = e % 8;
        const int64_t j = min(j0 + jl, (int64_t)nk - 1);
        bsrc[t] = (uint32_t)(col_index ? col_index[j] : j) * 64u + 8u * pc;
        bdst[t] = jl * FTP_BLD + 8 * pc;
    }
    const int64_t bstage = npts * 64;
    // [P_re; -P_im] fragment: chunk (c + 2) % 4 of the stored [P_im; P_re], negated on lanes 32..63
    const int c1 = (c + 2) & 3;
    const uint32_t nmask = c >= 2 ? 0x80008000u : 0u;
    uint32_t aoff[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int64_t d = min(d0 + wr * (RB * 16) + rb * 16 + il, ndata - 1);
        aoff[rb] = (uint32_t)d * 64u + 8u * c;
    }
    const int64_t astage = ndata * 64;

#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < CB; j++) accr[i][j] = acci[i][j] = f32x4{0, 0, 0, 0};

    bf16x8 bq[PPT], ah[RB], al[RB];
#pragma unroll
    for (int t = 0; t < PPT; t++) bq[t] = *reinterpret_cast<const bf16x8 *>(bpk + bsrc[t]);
#pragma unroll
    for (int t = 0; t < PPT; t++) *reinterpret_cast<bf16x8 *>(&Bs[bdst[t]]) = bq[t];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        ah[rb] = *reinterpret_cast<const bf16x8 *>(apk + aoff[rb]);
        al[rb] = *reinterpret_cast<const bf16x8 *>(apk + aoff[rb] + 32);
    }
    __syncthreads();

    for (int s = 0; s < nrs; s++) {
        const int cur = s & 1;
        const bool more = s + 1 < nrs;
        bf16x8 nh[RB], nl[RB];
        if (more) {
            const __bf16 *bg = bpk + (s + 1) * bstage;
#pragma unroll
            for (int t = 0; t < PPT; t++) bq[t] = *reinterpret_cast<const bf16x8 *>(bg + bsrc[t]);
            if constexpr (PF) {
                const __bf16 *ag = apk + (s + 1) * astage;
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    nh[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb]);
                    nl[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb] + 32);
                }
            }
        }
        if constexpr (!PF) {
            if (s > 0) {
                const __bf16 *ag = apk + s * astage;
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    ah[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb]);
                    al[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb] + 32);
                }
            }
        }
        const __bf16 *bb = Bs + cur * (TN * FTP_BLD);
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            if constexpr (MID) {
                if (cb == CB / 2 && more) {
#pragma unroll
                    for (int t = 0; t < PPT; t++)
                        *reinterpret_cast<bf16x8 *>(&Bs[(cur ^ 1) * (TN * FTP_BLD) + bdst[t]]) = bq[t];
                }
            }
            const __bf16 *col = bb + (wc * (CB * 16) + cb * 16 + il) * FTP_BLD;
            const bf16x8 b2h = *reinterpret_cast<const bf16x8 *>(col + 8 * c);
            const bf16x8 b2l = *reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c);
            const bf16x8 b1h = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 8 * c1), nmask);
            const bf16x8 b1l = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c1), nmask);
            // small terms first; consecutive MFMAs update different accumulators
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1l, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2l, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
        }
        if (more) {
            if constexpr (!MID) {
#pragma unroll
                for (int t = 0; t < PPT; t++)
                    *reinterpret_cast<bf16x8 *>(&Bs[(cur ^ 1) * (TN * FTP_BLD) + bdst[t]]) = bq[t];
            }
            if constexpr (PF) {
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    ah[rb] = nh[rb];
                    al[rb] = nl[rb];
                }
            }
        }
        __syncthreads();
    }
}

// Pipelined form of the main loop (unrolled by two, parity register sets, no copies): the phase
// pieces of stage s + 2 and the operand fragments of stage s + 1 are loaded at the top of stage s, so
// the only wait in a stage is for loads issued one stage earlier.
template <int RB, int CB, int WR, bool MID>
__device__ __forceinline__ void ft_bf16p_mainloop_pipe(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                       const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                       const int *__restrict__ col_index, int64_t j0, int64_t d0,
                                                       __bf16 *Bs, f32x4 (&accr)[RB][CB], f32x4 (&acci)[RB][CB]) {
    using S = FtpShape<RB, CB, WR>;
    constexpr int TN = S::TN, PPT = S::PPT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave % WR, wc = wave / WR;
    const int c = lane >> 4, il = lane & 15;

    uint32_t bsrc[PPT], bdst[PPT];
#pragma unroll
    for (int t = 0; t < PPT; t++) {
        const int e = tid + 256 * t, jl = e / 8, pc = e % 8;
        const int64_t j = min(j0 + jl, (int64_t)nk - 1);
        bsrc[t] = (uint32_t)(col_index ? col_index[j] : j) * 64u + 8u * pc;
        bdst[t] = jl * FTP_BLD + 8 * pc;
    }
    const int64_t bstage = npts * 64;
    const int c1 = (c + 2) & 3;
    const uint32_t nmask = c >= 2 ? 0x80008000u : 0u;
    uint32_t aoff[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int64_t d = min(d0 + wr * (RB * 16) + rb * 16 + il, ndata - 1);
        aoff[rb] = (uint32_t)d * 64u + 8u * c;
    }
    const int64_t astage = ndata * 64;
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < CB; j++) accr[i][j] = acci[i][j] = f32x4{0, 0, 0, 0};

    auto load_b = [&](int s, bf16x8 (&q)[PPT]) {
        const __bf16 *bg = bpk + s * bstage;
#pragma unroll
        for (int t = 0; t < PPT; t++) q[t] = *reinterpret_cast<const bf16x8 *>(bg + bsrc[t]);
    };
    auto store_b = [&](int buf, const bf16x8 (&q)[PPT]) {
#pragma unroll
        for (int t = 0; t < PPT; t++) *reinterpret_cast<bf16x8 *>(&Bs[buf * (TN * FTP_BLD) + bdst[t]]) = q[t];
    };
    auto load_a = [&](int s, bf16x8 (&h)[RB], bf16x8 (&l)[RB]) {
        const __bf16 *ag = apk + s * astage;
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            h[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb]);
            l[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb] + 32);
        }
    };
    // one stage: qn holds the phase pieces of s + 1 (written to the other buffer), qf receives those of
    // s + 2; (ch, cl) are this stage's operand fragments, (nh, nl) receive those of s + 1
    auto stage = [&](int s, bf16x8 (&qn)[PPT], bf16x8 (&qf)[PPT], bf16x8 (&ch)[RB], bf16x8 (&cl)[RB],
                     bf16x8 (&nh)[RB], bf16x8 (&nl)[RB]) {
        const int cur = s & 1;
        const bool more = s + 1 < nrs;
        if (s + 2 < nrs) load_b(s + 2, qf);
        if (more) load_a(s + 1, nh, nl);
        const __bf16 *bb = Bs + cur * (TN * FTP_BLD);
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            if constexpr (MID) {
                if (cb == CB / 2 && more) store_b(cur ^ 1, qn);
            }
            const __bf16 *col = bb + (wc * (CB * 16) + cb * 16 + il) * FTP_BLD;
            const bf16x8 b2h = *reinterpret_cast<const bf16x8 *>(col + 8 * c);
            const bf16x8 b2l = *reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c);
            const bf16x8 b1h = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 8 * c1), nmask);
            const bf16x8 b1l = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c1), nmask);
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cl[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cl[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ch[rb], b1l, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ch[rb], b2l, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ch[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ch[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
        }
        if constexpr (!MID) {
            if (more) store_b(cur ^ 1, qn);
        }
        __syncthreads();
    };

    bf16x8 qa[PPT], qb[PPT], xah[RB], xal[RB], xbh[RB], xbl[RB];
    load_b(0, qa);
    store_b(0, qa);
    load_a(0, xah, xal);
    if (nrs > 1) load_b(1, qb);
    __syncthreads();
    int s = 0;
    for (; s + 1 < nrs; s += 2) {
        stage(s, qb, qa, xah, xal, xbh, xbl);
        stage(s + 1, qa, qb, xbh, xbl, xah, xal);
    }
    if (s < nrs) stage(s, qb, qa, xah, xal, xbh, xbl);
}

// XCD grouping (XG): workgroups are dispatched round-robin over the 8 XCDs in linear order (x
// fastest). Linear id 8 gx a + 8 b + r is given row tile 8 a + r and column tile b, so the gx column
// tiles of one row tile run on the same XCD and share its L2 copy of the operand rows; the rows past
// the last full group of 8 keep the plain order.
__device__ __forceinline__ void xcd_grouped_tile(int64_t &bx, int64_t &by) {
    const int64_t gx = gridDim.x, gy = gridDim.y;
    const int64_t l = blockIdx.x + gx * blockIdx.y;
    const int64_t full = 8 * gx * (gy / 8);
    if (l < full) {
        const int64_t a = l / (8 * gx), rem = l % (8 * gx);
        by = 8 * a + rem % 8;
        bx = rem / 8;
    } else {
        bx = l % gx;
        by = l / gx;
    }
}

// Eight-wave main loop: 512 threads as WR = 4 (rows) x WC = 2 (columns), tile 192 x 128. The
// operand tile of a stage (192 rows x 64 bf16, contiguous in the stage-major layout) is staged once
// into LDS with coalesced 16-byte loads and read by both column halves, so operand traffic into the
// CU per column is half that of the 4-wave 192 x 64 tile. Operand rows and phase columns use the
// conflict-free 144-byte pitch; both are double-buffered, one barrier per stage.
constexpr int FT8_ROWS = 192, FT8_COLS = 128;

__device__ __forceinline__ void ft8_mainloop(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                             const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                             const int *__restrict__ col_index, int64_t j0, int64_t d0, __bf16 *As,
                                             __bf16 *Bs, f32x4 (&accr)[3][4], f32x4 (&acci)[3][4]) {
    constexpr int RB = 3, CB = 4, TN = FT8_COLS, TM = FT8_ROWS;
    constexpr int APT = TM * 8 / 512, BPT = TN * 8 / 512;  // 16-byte pieces per thread and stage: 3, 2
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 3, wc = wave >> 2;
    const int c = lane >> 4, il = lane & 15;
    const int c1 = (c + 2) & 3;
    const uint32_t nmask = c >= 2 ? 0x80008000u : 0u;
    uint32_t asrc[APT], adst[APT], bsrc[BPT], bdst[BPT];
#pragma unroll
    for (int t = 0; t < APT; t++) {
        const int e = tid + 512 * t, row = e / 8, pc = e % 8;
        asrc[t] = (uint32_t)min(d0 + row, ndata - 1) * 64u + 8u * pc;
        adst[t] = row * FTP_BLD + 8 * pc;
    }
#pragma unroll
    for (int t = 0; t < BPT; t++) {
        const int e = tid + 512 * t, jl = e / 8, pc = e % 8;
        const int64_t j = min(j0 + jl, (int64_t)nk - 1);
        bsrc[t] = (uint32_t)(col_index ? col_index[j] : j) * 64u + 8u * pc;
        bdst[t] = jl * FTP_BLD + 8 * pc;
    }
    const int64_t astage = ndata * 64, bstage = npts * 64;
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < CB; j++) accr[i][j] = acci[i][j] = f32x4{0, 0, 0, 0};
    bf16x8 qa[APT], qb[BPT];
#pragma unroll
    for (int t = 0; t < APT; t++) qa[t] = *reinterpret_cast<const bf16x8 *>(apk + asrc[t]);
#pragma unroll
    for (int t = 0; t < BPT; t++) qb[t] = *reinterpret_cast<const bf16x8 *>(bpk + bsrc[t]);
#pragma unroll
    for (int t = 0; t < APT; t++) *reinterpret_cast<bf16x8 *>(&As[adst[t]]) = qa[t];
#pragma unroll
    for (int t = 0; t < BPT; t++) *reinterpret_cast<bf16x8 *>(&Bs[bdst[t]]) = qb[t];
    __syncthreads();
    for (int s = 0; s < nrs; s++) {
        const int cur = s & 1;
        const bool more = s + 1 < nrs;
        if (more) {
            const __bf16 *ag = apk + (s + 1) * astage;
            const __bf16 *bg = bpk + (s + 1) * bstage;
#pragma unroll
            for (int t = 0; t < APT; t++) qa[t] = *reinterpret_cast<const bf16x8 *>(ag + asrc[t]);
#pragma unroll
            for (int t = 0; t < BPT; t++) qb[t] = *reinterpret_cast<const bf16x8 *>(bg + bsrc[t]);
        }
        const __bf16 *aa = As + cur * (TM * FTP_BLD);
        bf16x8 ah[RB], al[RB];
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const __bf16 *row = aa + (wr * (RB * 16) + rb * 16 + il) * FTP_BLD + 8 * c;
            ah[rb] = *reinterpret_cast<const bf16x8 *>(row);
            al[rb] = *reinterpret_cast<const bf16x8 *>(row + 32);
        }
        const __bf16 *bb = Bs + cur * (TN * FTP_BLD);
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            if (cb == CB / 2 && more) {
#pragma unroll
                for (int t = 0; t < APT; t++)
                    *reinterpret_cast<bf16x8 *>(&As[(cur ^ 1) * (TM * FTP_BLD) + adst[t]]) = qa[t];
#pragma unroll
                for (int t = 0; t < BPT; t++)
                    *reinterpret_cast<bf16x8 *>(&Bs[(cur ^ 1) * (TN * FTP_BLD) + bdst[t]]) = qb[t];
            }
            const __bf16 *col = bb + (wc * (CB * 16) + cb * 16 + il) * FTP_BLD;
            const bf16x8 b2h = *reinterpret_cast<const bf16x8 *>(col + 8 * c);
            const bf16x8 b2l = *reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c);
            const bf16x8 b1h = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 8 * c1), nmask);
            const bf16x8 b1l = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c1), nmask);
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1l, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2l, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
}

// ft8_mainloop with the global loads two stages ahead: the operand and phase pieces of stage s+2 are
// issued at the start of stage s into the register set that held stage s, and written to LDS halfway
// through stage s+1 (about 1.5 stages after issue instead of half a stage). Two register sets; the stage
// loop is unrolled by two so both sets stay in fixed registers. Same products in the same order.
__device__ __forceinline__ void ft8_mainloop_deep(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                  const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                  const int *__restrict__ col_index, int64_t j0, int64_t d0,
                                                  __bf16 *As, __bf16 *Bs, f32x4 (&accr)[3][4], f32x4 (&acci)[3][4]) {
    constexpr int RB = 3, CB = 4, TN = FT8_COLS, TM = FT8_ROWS;
    constexpr int APT = TM * 8 / 512, BPT = TN * 8 / 512;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 3, wc = wave >> 2;
    const int c = lane >> 4, il = lane & 15;
    const int c1 = (c + 2) & 3;
    const uint32_t nmask = c >= 2 ? 0x80008000u : 0u;
    uint32_t asrc[APT], adst[APT], bsrc[BPT], bdst[BPT];
#pragma unroll
    for (int t = 0; t < APT; t++) {
        const int e = tid + 512 * t, row = e / 8, pc = e % 8;
        asrc[t] = (uint32_t)min(d0 + row, ndata - 1) * 64u + 8u * pc;
        adst[t] = row * FTP_BLD + 8 * pc;
    }
#pragma unroll
    for (int t = 0; t < BPT; t++) {
        const int e = tid + 512 * t, jl = e / 8, pc = e % 8;
        const int64_t j = min(j0 + jl, (int64_t)nk - 1);
        bsrc[t] = (uint32_t)(col_index ? col_index[j] : j) * 64u + 8u * pc;
        bdst[t] = jl * FTP_BLD + 8 * pc;
    }
    const int64_t astage = ndata * 64, bstage = npts * 64;
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < CB; j++) accr[i][j] = acci[i][j] = f32x4{0, 0, 0, 0};
    bf16x8 qa[2][APT], qb[2][BPT];
    auto load = [&](int st, bf16x8(&ra)[APT], bf16x8(&rb)[BPT]) {
        const __bf16 *ag = apk + st * astage;
        const __bf16 *bg = bpk + st * bstage;
#pragma unroll
        for (int t = 0; t < APT; t++) ra[t] = *reinterpret_cast<const bf16x8 *>(ag + asrc[t]);
#pragma unroll
        for (int t = 0; t < BPT; t++) rb[t] = *reinterpret_cast<const bf16x8 *>(bg + bsrc[t]);
    };
    auto store = [&](int buf, const bf16x8(&ra)[APT], const bf16x8(&rb)[BPT]) {
#pragma unroll
        for (int t = 0; t < APT; t++) *reinterpret_cast<bf16x8 *>(&As[buf * (TM * FTP_BLD) + adst[t]]) = ra[t];
#pragma unroll
        for (int t = 0; t < BPT; t++) *reinterpret_cast<bf16x8 *>(&Bs[buf * (TN * FTP_BLD) + bdst[t]]) = rb[t];
    };
    load(0, qa[0], qb[0]);
    store(0, qa[0], qb[0]);
    if (nrs > 1) load(1, qa[1], qb[1]);
    __syncthreads();
    // stage s computes from LDS buffer s & 1; register set s & 1 is free at its start
    auto stage = [&](int s, bf16x8(&fa)[APT], bf16x8(&fb)[BPT], const bf16x8(&na)[APT], const bf16x8(&nb)[BPT]) {
        const int cur = s & 1;
        if (s + 2 < nrs) load(s + 2, fa, fb);
        const __bf16 *aa = As + cur * (TM * FTP_BLD);
        bf16x8 ah[RB], al[RB];
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const __bf16 *row = aa + (wr * (RB * 16) + rb * 16 + il) * FTP_BLD + 8 * c;
            ah[rb] = *reinterpret_cast<const bf16x8 *>(row);
            al[rb] = *reinterpret_cast<const bf16x8 *>(row + 32);
        }
        const __bf16 *bb = Bs + cur * (TN * FTP_BLD);
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            if (cb == CB / 2 && s + 1 < nrs) store(cur ^ 1, na, nb);
            const __bf16 *col = bb + (wc * (CB * 16) + cb * 16 + il) * FTP_BLD;
            const bf16x8 b2h = *reinterpret_cast<const bf16x8 *>(col + 8 * c);
            const bf16x8 b2l = *reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c);
            const bf16x8 b1h = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 8 * c1), nmask);
            const bf16x8 b1l = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c1), nmask);
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1l, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2l, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
        }
        __syncthreads();
    };
    for (int s = 0; s < nrs; s += 2) {
        stage(s, qa[0], qb[0], qa[1], qb[1]);
        if (s + 1 < nrs) stage(s + 1, qa[1], qb[1], qa[0], qb[0]);
    }
}

// RR->kR-type contraction with the eight-wave 192 x 128 tile (single unit, XCD-grouped order).
__global__ __launch_bounds__(512, 2) void ft_bf16p8_kernel(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                           const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                           cplx<float> *__restrict__ out, int64_t ldout) {
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * FT8_ROWS * FTP_BLD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * FT8_COLS * FTP_BLD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave & 3, wc = wave >> 2;
    const int c = lane >> 4, il = lane & 15;
    int64_t bx, by;
    xcd_grouped_tile(bx, by);
    const int64_t j0 = bx * FT8_COLS;
    if (j0 >= nk) return;  // uniform for the block
    const int64_t d0 = by * FT8_ROWS;
    f32x4 accr[3][4], acci[3][4];
    ft8_mainloop(ndata, nrs, apk, bpk, npts, nk, nullptr, j0, d0, As, Bs, accr, acci);
#pragma unroll
    for (int rb = 0; rb < 3; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const int64_t j = j0 + wc * 64 + cb * 16 + il;
            if (j >= nk) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t d = d0 + wr * 48 + rb * 16 + 4 * c + i;
                if (d < ndata) out[d + ldout * j] = cplx<float>{accr[rb][cb][i], acci[rb][cb][i]};
            }
        }
}

// RR->kR Fourier over Re fused with the k-side window projection (nw = 16). The operand rows are packed
// j-fastest, d = j + 16 (i + 16 (nu + nm Rp)) (ephip_ft_pack_operand_swapped with nb = 16), so each 16-row
// MFMA block is one complete j group of the row group G = i + 16 (nu + nm Rp). For batch column kk (the
// k-point) and each of its units u (one per in-window band n, in band order, kus[kk] <= u < kus[kk+1]):
// A[u][G] = sum_j g(k)[i, j, nu, Rp] U_k[j, n], a 4-lane shuffle reduction of the accumulators. The
// g(k, Rp) tensor itself is never written.
template <bool DEEP = false>
__global__ __launch_bounds__(512, 2) void ft_bf16p8_proj_kernel(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                                const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                                const int *__restrict__ kus,
                                                                const int64_t *__restrict__ ustate,
                                                                const int *__restrict__ uband,
                                                                const cplx<float> *__restrict__ U, int nbw,
                                                                cplx<float> *__restrict__ A, int64_t lda) {
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * FT8_ROWS * FTP_BLD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * FT8_COLS * FTP_BLD];
    constexpr int NW = 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave & 3, wc = wave >> 2;
    const int c = lane >> 4, il = lane & 15;
    int64_t bx, by;
    xcd_grouped_tile(bx, by);
    const int64_t j0 = bx * FT8_COLS;
    if (j0 >= nk) return;  // uniform for the block
    const int64_t d0 = by * FT8_ROWS;
    f32x4 accr[3][4], acci[3][4];
    if constexpr (DEEP)
        ft8_mainloop_deep(ndata, nrs, apk, bpk, npts, nk, nullptr, j0, d0, As, Bs, accr, acci);
    else
        ft8_mainloop(ndata, nrs, apk, bpk, npts, nk, nullptr, j0, d0, As, Bs, accr, acci);
    const int ku0 = kus[0];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const int64_t kk = j0 + wc * 64 + cb * 16 + il;
        const bool kv = kk < nk;
        const int ub = kv ? kus[kk] - ku0 : 0;
        const int nb = kv ? kus[kk + 1] - kus[kk] : 0;
        const cplx<float> *uk = U + (int64_t)NW * NW * (nb > 0 ? ustate[ub] : 0);
        for (int m = 0; m < nbw; m++) {
            const bool act = m < nb;
            const int band = act ? uband[ub + m] : 0;
            cplx<float> w[4];
#pragma unroll
            for (int r = 0; r < 4; r++) w[r] = act ? uk[4 * c + r + NW * band] : cplx<float>{0.f, 0.f};
#pragma unroll
            for (int rb = 0; rb < 3; rb++) {
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float ar = accr[rb][cb][r], ai = acci[rb][cb][r];
                    sr += ar * w[r].re - ai * w[r].im;
                    si += ar * w[r].im + ai * w[r].re;
                }
                sr += __shfl_xor(sr, 16, 64);
                si += __shfl_xor(si, 16, 64);
                sr += __shfl_xor(sr, 32, 64);
                si += __shfl_xor(si, 32, 64);
                const int64_t G = (d0 + wr * 48 + rb * 16) / 16;
                if (act && c == rb) A[(int64_t)(ub + m) * lda + G] = cplx<float>{sr, si};
            }
        }
    }
}

template <int RB, int CB, int WR, bool PF, bool XG = false, bool MID = false, bool PIPE = false>
__device__ __forceinline__ void ft_bf16p_body(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                              const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                              cplx<float> *__restrict__ out, int64_t ldout, int64_t out_unit,
                                              const int *__restrict__ col_index, int64_t col_stride,
                                              const int *__restrict__ ncol) {
    using S = FtpShape<RB, CB, WR>;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * S::TN * FTP_BLD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave % WR, wc = wave / WR;
    const int c = lane >> 4, il = lane & 15;
    const int u = blockIdx.z;
    if (ncol) nk = ncol[u];
    int64_t bx = blockIdx.x, by = blockIdx.y;
    if constexpr (XG) xcd_grouped_tile(bx, by);
    const int64_t j0 = bx * S::TN;
    if (j0 >= nk) return;  // uniform for the block
    const int64_t d0 = by * S::TM;
    f32x4 accr[RB][CB], acci[RB][CB];
    if constexpr (PIPE)
        ft_bf16p_mainloop_pipe<RB, CB, WR, MID>(ndata, nrs, apk + (int64_t)u * nrs * ndata * 64, bpk, npts, nk,
                                                col_index ? col_index + u * col_stride : nullptr, j0, d0, Bs, accr,
                                                acci);
    else
        ft_bf16p_mainloop<RB, CB, WR, PF, MID>(ndata, nrs, apk + (int64_t)u * nrs * ndata * 64, bpk, npts, nk,
                                               col_index ? col_index + u * col_stride : nullptr, j0, d0, Bs, accr,
                                               acci);
    out += u * out_unit;
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int64_t j = j0 + wc * (CB * 16) + cb * 16 + il;
            if (j >= nk) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t d = d0 + wr * (RB * 16) + rb * 16 + 4 * c + i;
                if (d < ndata) out[d + ldout * j] = cplx<float>{accr[rb][cb][i], acci[rb][cb][i]};
            }
        }
}

// Two resident waves per SIMD (at most 256 registers) for the shapes that fit without spilling; the
// 128 x 128 prefetching shape needs more registers and runs one wave per SIMD.
template <int RB, int CB, int WR, bool PF, bool XG = false, bool MID = false, bool PIPE = false>
__global__ __launch_bounds__(256, 2) void ft_bf16p_kernel(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                          const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                          cplx<float> *__restrict__ out, int64_t ldout,
                                                          int64_t out_unit, const int *__restrict__ col_index,
                                                          int64_t col_stride, const int *__restrict__ ncol) {
    ft_bf16p_body<RB, CB, WR, PF, XG, MID, PIPE>(ndata, nrs, apk, bpk, npts, nk, out, ldout, out_unit, col_index,
                                                 col_stride, ncol);
}

__global__ __launch_bounds__(256) void ft_bf16p_wide_pf_kernel(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                               const __bf16 *__restrict__ bpk, int64_t npts, int nk,
                                                               cplx<float> *__restrict__ out, int64_t ldout,
                                                               int64_t out_unit, const int *__restrict__ col_index,
                                                               int64_t col_stride, const int *__restrict__ ncol) {
    ft_bf16p_body<4, 4, 2, true>(ndata, nrs, apk, bpk, npts, nk, out, ldout, out_unit, col_index, col_stride, ncol);
}

// Main loop of the 192 x 64 unit tile with the phases of a uniform point grid generated on the fly
// instead of gathered from a packed table: point p = (i1, i2, i3) of an n1 x n2 x n3 grid (third index
// fastest, no shift) has e^{2 pi i R.x_p} = prod_a t_a[(R_a i_a) mod n_a], t_a[m] = e^{2 pi i m / n_a},
// an exact integer reduction and two FP32 complex products. Lane l of wave w generates column l's
// phases for R slots 2w, 2w+1, 8+2w, 9+2w of each stage and writes them split into the LDS block.
constexpr int FTG_NMAX = 64;  // largest grid dimension of the on-the-fly phases

__device__ __forceinline__ void ftg_phase(const float2 *tab, const int *__restrict__ R, const int i[3],
                                          const int n[3], float &c, float &s) {
    float2 p[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int m = (R[a] * i[a]) % n[a];
        m += m < 0 ? n[a] : 0;
        p[a] = tab[a * FTG_NMAX + m];
    }
    const float r01 = p[0].x * p[1].x - p[0].y * p[1].y, i01 = p[0].x * p[1].y + p[0].y * p[1].x;
    c = r01 * p[2].x - i01 * p[2].y;
    s = r01 * p[2].y + i01 * p[2].x;
}

template <int RB, int CB, int WR>
__device__ __forceinline__ void ft_bf16p_mainloop_gp(int64_t ndata, int nrs, const __bf16 *__restrict__ apk,
                                                     int nr, const int *__restrict__ irvec, const int n[3],
                                                     const float2 *tab, int nk, const int *__restrict__ col_index,
                                                     int64_t j0, int64_t d0, __bf16 *Bs, f32x4 (&accr)[RB][CB],
                                                     f32x4 (&acci)[RB][CB]) {
    using S = FtpShape<RB, CB, WR>;
    constexpr int TN = S::TN;
    static_assert(TN == 64 && WR == 4, "on-the-fly phases assume 64 columns and 4 waves");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave % WR, wc = wave / WR;
    const int c = lane >> 4, il = lane & 15;
    const int c1 = (c + 2) & 3;
    const uint32_t nmask = c >= 2 ? 0x80008000u : 0u;
    // grid coordinates of this lane's column (columns past the count reuse the last one)
    int gi[3];
    {
        const int p = col_index ? col_index[min(j0 + lane, (int64_t)nk - 1)] : (int)min(j0 + lane, (int64_t)nk - 1);
        gi[0] = p / (n[1] * n[2]);
        gi[1] = (p / n[2]) % n[1];
        gi[2] = p % n[2];
    }
    const int rs[2] = {2 * wave, 8 + 2 * wave};
    auto gen = [&](int s, int buf) {
        __bf16 *col = Bs + buf * (TN * FTP_BLD) + lane * FTP_BLD;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r0 = 16 * s + rs[h];
            float c0 = 0.f, s0 = 0.f, cc1 = 0.f, ss1 = 0.f;
            if (r0 < nr) ftg_phase(tab, irvec + 3 * r0, gi, n, c0, s0);
            if (r0 + 1 < nr) ftg_phase(tab, irvec + 3 * (r0 + 1), gi, n, cc1, ss1);
            __bf16 sh0, sl0, sh1, sl1, ch0, cl0, ch1, cl1;
            split_bf16(s0, sh0, sl0);
            split_bf16(ss1, sh1, sl1);
            split_bf16(c0, ch0, cl0);
            split_bf16(cc1, ch1, cl1);
            const int r = rs[h];
            *reinterpret_cast<bf16x2 *>(col + r) = bf16x2{sh0, sh1};
            *reinterpret_cast<bf16x2 *>(col + 16 + r) = bf16x2{ch0, ch1};
            *reinterpret_cast<bf16x2 *>(col + 32 + r) = bf16x2{sl0, sl1};
            *reinterpret_cast<bf16x2 *>(col + 48 + r) = bf16x2{cl0, cl1};
        }
    };
    uint32_t aoff[RB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int64_t d = min(d0 + wr * (RB * 16) + rb * 16 + il, ndata - 1);
        aoff[rb] = (uint32_t)d * 64u + 8u * c;
    }
    const int64_t astage = ndata * 64;
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < CB; j++) accr[i][j] = acci[i][j] = f32x4{0, 0, 0, 0};

    bf16x8 ah[RB], al[RB];
    gen(0, 0);
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        ah[rb] = *reinterpret_cast<const bf16x8 *>(apk + aoff[rb]);
        al[rb] = *reinterpret_cast<const bf16x8 *>(apk + aoff[rb] + 32);
    }
    __syncthreads();
    for (int s = 0; s < nrs; s++) {
        const int cur = s & 1;
        const bool more = s + 1 < nrs;
        bf16x8 nh[RB], nl[RB];
        if (more) {
            const __bf16 *ag = apk + (s + 1) * astage;
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                nh[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb]);
                nl[rb] = *reinterpret_cast<const bf16x8 *>(ag + aoff[rb] + 32);
            }
            gen(s + 1, cur ^ 1);
        }
        const __bf16 *bb = Bs + cur * (TN * FTP_BLD);
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const __bf16 *col = bb + (wc * (CB * 16) + cb * 16 + il) * FTP_BLD;
            const bf16x8 b2h = *reinterpret_cast<const bf16x8 *>(col + 8 * c);
            const bf16x8 b2l = *reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c);
            const bf16x8 b1h = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 8 * c1), nmask);
            const bf16x8 b1l = bf16_flip(*reinterpret_cast<const bf16x8 *>(col + 32 + 8 * c1), nmask);
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1l, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2l, acci[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                accr[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b1h, accr[rb][cb], 0, 0, 0);
                acci[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rb], b2h, acci[rb][cb], 0, 0, 0);
            }
        }
        if (more) {
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                ah[rb] = nh[rb];
                al[rb] = nl[rb];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// kR->kq Fourier over Rp fused with the k+q-window and phonon rotations and g2, for the e-ph unit
// shape nw = 16, nm = 12 (operand rows d = i + 16 nu): the 192 x 64 tile puts modes 3w..3w+2 of all
// 16 Wannier rows of 64 kept columns in wave w. Per column, t[m][nu] = sum_i conj(U_kq[i, mlo+m])
// g[i + 16 nu] is a 4-lane shuffle reduction of the accumulators, each wave's share of
// ep[m][nu'] = sum_nu t[m][nu] u_ph(q)[nu + 12 nu'] is added into LDS in wave order (deterministic),
// and g2[m][nu'] = |ep|^2 / (2 omega_q,nu') is the only output (no g(k,q) round trip through memory).
// ---------------------------------------------------------------------------------------------
constexpr int FTR_NBW_MAX = 4;

// Epilogue of the fused kR->kq kernels. The wave with row group wr holds modes 3wr..3wr+2 of all 16
// Wannier rows for the columns jw + cb*16 + il of the tile (jw = the wave's column offset). Lane (c, il)
// owns nu' = c, c+4, c+8 of its column; the four row groups' shares of ep are added into EP in row-group
// order (deterministic; waves of different column offsets write disjoint columns), then g2 is written
// for the tile's valid columns, layout [m][nu'] per kept (unit, column), slots m < nbw. Called by every
// thread of the block (it holds barriers); EP holds TN * NBW * 12 complex values.
template <int NBW, int TN>
__device__ __forceinline__ void ftr_epilogue(const f32x4 (&accr)[3][4], const f32x4 (&acci)[3][4], int wr, int jw,
                                             int u, int nk, int64_t j0, const int *__restrict__ col_index,
                                             const int *__restrict__ ikq, int64_t cs,
                                             const cplx<float> *__restrict__ U, const int64_t *__restrict__ bmin,
                                             const int64_t *__restrict__ bmax, const cplx<float> *__restrict__ uph,
                                             const double *__restrict__ omega, int nbw, float *__restrict__ g2,
                                             float *EP) {
    constexpr int RB = 3, CB = 4, NW = 16, NM = 12;
    const int lane = threadIdx.x & 63;
    const int c = lane >> 4, il = lane & 15;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
        const int jj = jw + cb * 16 + il;
        const int64_t j = j0 + jj;
        const bool jv = j < nk;
        const int64_t row = (int64_t)u * cs + (jv ? j : j0);
        const int64_t kq = ikq[row];
        const int iq = col_index[row];
        int nbkq = bmin[kq] > 0 ? (int)(bmax[kq] - bmin[kq] + 1) : 0;
        nbkq = jv ? min(nbkq, nbw) : 0;
        const int mlo = bmin[kq] > 0 ? (int)bmin[kq] - 1 : 0;
        const cplx<float> *uk = U + (int64_t)NW * NW * kq;
        const cplx<float> *up = uph + (int64_t)NM * NM * iq;
        float pr[NBW][3], pi[NBW][3];
#pragma unroll
        for (int m = 0; m < NBW; m++) {
            // w = U_kq[4c + r, mlo + m]
            cplx<float> w[4];
#pragma unroll
            for (int r = 0; r < 4; r++)
                w[r] = m < nbkq ? uk[4 * c + r + NW * (mlo + m)] : cplx<float>{0.f, 0.f};
            float tr[RB], ti[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float ar = accr[rb][cb][r], ai = acci[rb][cb][r];
                    sr += w[r].re * ar + w[r].im * ai;
                    si += w[r].re * ai - w[r].im * ar;
                }
                sr += __shfl_xor(sr, 16, 64);
                si += __shfl_xor(si, 16, 64);
                sr += __shfl_xor(sr, 32, 64);
                si += __shfl_xor(si, 32, 64);
                tr[rb] = sr;
                ti[rb] = si;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int nu2 = c + 4 * k;
                float er = 0.f, ei = 0.f;
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    const cplx<float> p = up[(3 * wr + rb) + NM * nu2];
                    er += tr[rb] * p.re - ti[rb] * p.im;
                    ei += tr[rb] * p.im + ti[rb] * p.re;
                }
                pr[m][k] = er;
                pi[m][k] = ei;
            }
        }
#pragma unroll
        for (int w = 0; w < 4; w++) {
            if (wr == w) {
#pragma unroll
                for (int m = 0; m < NBW; m++)
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const int o = 2 * ((jj * NBW + m) * NM + c + 4 * k);
                        EP[o] = (w == 0 ? 0.f : EP[o]) + pr[m][k];
                        EP[o + 1] = (w == 0 ? 0.f : EP[o + 1]) + pi[m][k];
                    }
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < TN * nbw * NM; e += blockDim.x) {
        const int jj = e / (nbw * NM), rr = e % (nbw * NM), m = rr / NM, nu2 = rr % NM;
        const int64_t j = j0 + jj;
        if (j >= nk) continue;
        const int64_t row = (int64_t)u * cs + j;
        const int o = 2 * ((jj * NBW + m) * NM + nu2);
        const float er = EP[o], ei = EP[o + 1];
        const float w2 = 2.f * static_cast<float>(omega[(int64_t)NM * col_index[row] + nu2]);
        g2[row * nbw * NM + rr] = (er * er + ei * ei) / w2;
    }
}

// Column metadata of the fused epilogue for the lane's four column blocks (k+q state, q, number of k+q
// window bands clipped to nbw, first band), loaded before the main loop so that the epilogue's only
// global reads are the U_kq and u_ph values, which it fetches one column block ahead.
struct FtrCols {
    int kq[4], iq[4], nbkq[4], mlo[4];
};

__device__ __forceinline__ void ftr_columns(FtrCols &m, int jw, int u, int nk, int64_t j0,
                                            const int *__restrict__ col_index, const int *__restrict__ ikq,
                                            int64_t cs, const int64_t *__restrict__ bmin,
                                            const int64_t *__restrict__ bmax, int nbw) {
    const int il = threadIdx.x & 15;
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const int64_t j = j0 + jw + cb * 16 + il;
        const bool jv = j < nk;
        const int64_t row = (int64_t)u * cs + (jv ? j : j0);
        const int kq = ikq[row];
        m.kq[cb] = kq;
        m.iq[cb] = col_index[row];
        const int64_t b0 = bmin[kq], b1 = bmax[kq];
        const int nb = b0 > 0 ? (int)(b1 - b0 + 1) : 0;
        m.nbkq[cb] = jv ? min(nb, nbw) : 0;
        m.mlo[cb] = b0 > 0 ? (int)b0 - 1 : 0;
    }
}

// ftr_epilogue with the column metadata of ftr_columns and the U_kq / u_ph reads of the next column block
// issued before the current block's arithmetic and barriers. Same arithmetic and summation order.
template <int NBW, int TN>
__device__ __forceinline__ void ftr_epilogue_cols(const f32x4 (&accr)[3][4], const f32x4 (&acci)[3][4], int wr,
                                                  int jw, const FtrCols &mc, int u, int nk, int64_t j0,
                                                  const int *__restrict__ col_index, int64_t cs,
                                                  const cplx<float> *__restrict__ U,
                                                  const cplx<float> *__restrict__ uph,
                                                  const double *__restrict__ omega, int nbw, float *__restrict__ g2,
                                                  float *EP) {
    constexpr int RB = 3, CB = 4, NW = 16, NM = 12;
    const int lane = threadIdx.x & 63;
    const int c = lane >> 4, il = lane & 15;
    // two register sets (the next block fetched during this one) for NBW <= 2; one set for wider
    // windows, whose doubled U_kq values would not fit beside the accumulators
    constexpr int NSET = NBW <= 2 ? 2 : 1;
    cplx<float> w[NSET][NBW][4], p[NSET][RB][3];
    auto fetch = [&](int cb, cplx<float> (&wd)[NBW][4], cplx<float> (&pd)[RB][3]) {
        const cplx<float> *uk = U + (int64_t)NW * NW * mc.kq[cb];
        const cplx<float> *up = uph + (int64_t)NM * NM * mc.iq[cb];
#pragma unroll
        for (int m = 0; m < NBW; m++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                wd[m][r] = m < mc.nbkq[cb] ? uk[4 * c + r + NW * (mc.mlo[cb] + m)] : cplx<float>{0.f, 0.f};
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int k = 0; k < 3; k++) pd[rb][k] = up[(3 * wr + rb) + NM * (c + 4 * k)];
    };
    if constexpr (NSET == 2) fetch(0, w[0], p[0]);
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
        const int cur = NSET == 2 ? (cb & 1) : 0;
        if constexpr (NSET == 2) {
            if (cb + 1 < CB) fetch(cb + 1, w[cur ^ 1], p[cur ^ 1]);
        } else {
            fetch(cb, w[0], p[0]);
        }
        const int jj = jw + cb * 16 + il;
        float pr[NBW][3], pi[NBW][3];
#pragma unroll
        for (int m = 0; m < NBW; m++) {
            float tr[RB], ti[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float ar = accr[rb][cb][r], ai = acci[rb][cb][r];
                    sr += w[cur][m][r].re * ar + w[cur][m][r].im * ai;
                    si += w[cur][m][r].re * ai - w[cur][m][r].im * ar;
                }
                sr += __shfl_xor(sr, 16, 64);
                si += __shfl_xor(si, 16, 64);
                sr += __shfl_xor(sr, 32, 64);
                si += __shfl_xor(si, 32, 64);
                tr[rb] = sr;
                ti[rb] = si;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float er = 0.f, ei = 0.f;
#pragma unroll
                for (int rb = 0; rb < RB; rb++) {
                    const cplx<float> q = p[cur][rb][k];
                    er += tr[rb] * q.re - ti[rb] * q.im;
                    ei += tr[rb] * q.im + ti[rb] * q.re;
                }
                pr[m][k] = er;
                pi[m][k] = ei;
            }
        }
#pragma unroll
        for (int wv = 0; wv < 4; wv++) {
            if (wr == wv) {
#pragma unroll
                for (int m = 0; m < NBW; m++)
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const int o = 2 * ((jj * NBW + m) * NM + c + 4 * k);
                        EP[o] = (wv == 0 ? 0.f : EP[o]) + pr[m][k];
                        EP[o + 1] = (wv == 0 ? 0.f : EP[o + 1]) + pi[m][k];
                    }
            }
            __syncthreads();
        }
    }
    for (int e = threadIdx.x; e < TN * nbw * NM; e += blockDim.x) {
        const int jj = e / (nbw * NM), rr = e % (nbw * NM), m = rr / NM, nu2 = rr % NM;
        const int64_t j = j0 + jj;
        if (j >= nk) continue;
        const int64_t row = (int64_t)u * cs + j;
        const int o = 2 * ((jj * NBW + m) * NM + nu2);
        const float er = EP[o], ei = EP[o + 1];
        const float w2 = 2.f * static_cast<float>(omega[(int64_t)NM * col_index[row] + nu2]);
        g2[row * nbw * NM + rr] = (er * er + ei * ei) / w2;
    }
}

template <int NBW, bool MID = false, bool PIPE = false, bool GP = false>
__global__ __launch_bounds__(256, 2) void ft_bf16p_rot_kernel(int nrs, const __bf16 *__restrict__ apk,
                                                              const __bf16 *__restrict__ bpk, int64_t npts,
                                                              const int *__restrict__ ncol,
                                                              const int *__restrict__ col_index,
                                                              const int *__restrict__ ikq, int64_t cs,
                                                              const cplx<float> *__restrict__ U,
                                                              const int64_t *__restrict__ bmin,
                                                              const int64_t *__restrict__ bmax,
                                                              const cplx<float> *__restrict__ uph,
                                                              const double *__restrict__ omega, int nbw,
                                                              float *__restrict__ g2, int nr,
                                                              const int *__restrict__ irvec, int3 qgrid) {
    constexpr int RB = 3, CB = 4, WR = 4, NW = 16, NM = 12, TN = 64;
    constexpr int64_t NDATA = NW * NM;
    // phase staging during the main loop, then the ep tile [jj][m][nu'] (re, im) of the epilogue
    constexpr int BS_ELEMS = 2 * TN * FTP_BLD > 4 * TN * NBW * NM ? 2 * TN * FTP_BLD : 4 * TN * NBW * NM;
    __shared__ __attribute__((aligned(16))) __bf16 Bs[BS_ELEMS];
    const int lane = threadIdx.x & 63, wr = threadIdx.x >> 6;
    const int c = lane >> 4, il = lane & 15;
    const int u = blockIdx.z;
    const int nk = ncol[u];
    const int64_t j0 = (int64_t)blockIdx.x * TN;
    if (j0 >= nk) return;  // uniform for the block
    f32x4 accr[RB][CB], acci[RB][CB];
    if constexpr (GP) {
        __shared__ float2 tab[3 * FTG_NMAX];
        const int n[3] = {qgrid.x, qgrid.y, qgrid.z};
        for (int e = threadIdx.x; e < 3 * FTG_NMAX; e += blockDim.x) {
            const int a = e / FTG_NMAX, m = e % FTG_NMAX;
            float cr = 0.f, si = 0.f;
            if (m < n[a]) cis2pi((double)m / n[a], cr, si);
            tab[e] = make_float2(cr, si);
        }
        __syncthreads();
        ft_bf16p_mainloop_gp<RB, CB, WR>(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, nr, irvec, n, tab, nk,
                                         col_index + u * cs, j0, 0, Bs, accr, acci);
    } else if constexpr (PIPE)
        ft_bf16p_mainloop_pipe<RB, CB, WR, MID>(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, bpk, npts, nk,
                                                col_index + u * cs, j0, 0, Bs, accr, acci);
    else
        ft_bf16p_mainloop<RB, CB, WR, true, MID>(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, bpk, npts, nk,
                                                 col_index + u * cs, j0, 0, Bs, accr, acci);
    // after the main loop's final barrier the staging buffer is free: ep tile [jj][m][nu'] (re, im)
    ftr_epilogue<NBW, TN>(accr, acci, wr, 0, u, nk, j0, col_index, ikq, cs, U, bmin, bmax, uph, omega, nbw, g2,
                          reinterpret_cast<float *>(Bs));
}

// Eight-wave form of the fused kR->kq kernel: 128 kept columns per workgroup as two column halves of
// 64 sharing the unit's operand tile, staged once per stage through LDS (the operand is read into the
// CU once per 128 columns instead of once per 64). The epilogue reuses the operand buffer for ep.
//
// Workgroup order (ug > 0): the units are dealt to the 8 XCDs (unit x + 8v on XCD x, workgroups being
// dispatched round-robin over the XCDs), and each XCD walks its units in groups of ug, column tile by
// column tile, so the tiles running together on one XCD share its L2 for the group's operands and for the
// q range of the phase table they gather. The grid is then (ntiles, 1, 8 ceil(nunits / 8)). ug = 0 keeps
// the plain order (all tiles of a unit in sequence).
template <int NBW, bool PFE = false, bool DEEP = false>
__global__ __launch_bounds__(512, 2) void ft_bf16p8_rot_kernel(int nrs, const __bf16 *__restrict__ apk,
                                                               const __bf16 *__restrict__ bpk, int64_t npts,
                                                               const int *__restrict__ ncol,
                                                               const int *__restrict__ col_index,
                                                               const int *__restrict__ ikq, int64_t cs,
                                                               const cplx<float> *__restrict__ U,
                                                               const int64_t *__restrict__ bmin,
                                                               const int64_t *__restrict__ bmax,
                                                               const cplx<float> *__restrict__ uph,
                                                               const double *__restrict__ omega, int nbw,
                                                               float *__restrict__ g2, int nunits, int ug) {
    constexpr int NW = 16, NM = 12, TN = FT8_COLS;
    constexpr int64_t NDATA = NW * NM;
    static_assert(NDATA == FT8_ROWS, "one e-ph unit per row tile");
    static_assert(2 * FT8_ROWS * FTP_BLD * 2 >= TN * NBW * NM * 2 * 4, "ep tile must fit the operand buffer");
    __shared__ __attribute__((aligned(16))) __bf16 As[2 * FT8_ROWS * FTP_BLD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * FT8_COLS * FTP_BLD];
    const int wave = threadIdx.x >> 6;
    const int wr = wave & 3, wc = wave >> 2;
    int u, tile;
    if (ug > 0) {
        const int64_t T = gridDim.x, V = gridDim.z / 8;
        const int64_t l = blockIdx.x + T * (int64_t)blockIdx.z;
        const int64_t x = l % 8, n = l / 8;  // n-th workgroup of XCD x, n < V T
        const int64_t gi = n / ((int64_t)ug * T), w = n % ((int64_t)ug * T);
        const int64_t g = min((int64_t)ug, V - gi * ug);  // units in this group (the last may be short)
        tile = (int)(w / g);
        u = (int)(x + 8 * (gi * ug + w % g));
        if (u >= nunits) return;  // uniform for the block
    } else {
        u = blockIdx.z;
        tile = blockIdx.x;
    }
    const int nk = ncol[u];
    // column tiles tile, tile + gridDim.x, ... of the unit: the grid may be sized below the longest kept
    // list (the host's bound on count), and a list longer than the bound is still covered
    for (int64_t j0 = (int64_t)tile * TN; j0 < nk; j0 += (int64_t)gridDim.x * TN) {  // uniform for the block
        f32x4 accr[3][4], acci[3][4];
        if constexpr (PFE) {
            FtrCols mc;
            ftr_columns(mc, 64 * wc, u, nk, j0, col_index, ikq, cs, bmin, bmax, nbw);
            if constexpr (DEEP)
                ft8_mainloop_deep(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, bpk, npts, nk, col_index + u * cs,
                                  j0, 0, As, Bs, accr, acci);
            else
                ft8_mainloop(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, bpk, npts, nk, col_index + u * cs, j0,
                             0, As, Bs, accr, acci);
            ftr_epilogue_cols<NBW, TN>(accr, acci, wr, 64 * wc, mc, u, nk, j0, col_index, cs, U, uph, omega, nbw,
                                       g2, reinterpret_cast<float *>(As));
        } else {
            ft8_mainloop(NDATA, nrs, apk + (int64_t)u * nrs * NDATA * 64, bpk, npts, nk, col_index + u * cs, j0, 0,
                         As, Bs, accr, acci);
            ftr_epilogue<NBW, TN>(accr, acci, wr, 64 * wc, u, nk, j0, col_index, ikq, cs, U, bmin, bmax, uph, omega,
                                  nbw, g2, reinterpret_cast<float *>(As));
        }
        __syncthreads();  // the epilogue's reads of the ep tile finish before the next tile's staging writes
    }
}

// One thread per (unit, stage, row): 16 complex values -> 64 bf16 of the packed operand. With swap = nb > 0
// the two fastest row indices (each of size nb) trade places: packed row j + nb (i + nb rest) holds operand
// row i + nb (j + nb rest).
__global__ void ft_pack_operand_kernel(int64_t ndata, int nr, int nrs, const cplx<float> *__restrict__ op,
                                       int64_t ldop, int64_t op_stride, __bf16 *__restrict__ apk, int swap = 0) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int u = blockIdx.y;
    if (e >= ndata * nrs) return;
    const int64_t d = e % ndata;
    const int s = (int)(e / ndata);
    int64_t src = d;
    if (swap > 0) {
        const int64_t j = d % swap, i = (d / swap) % swap, rest = d / ((int64_t)swap * swap);
        src = i + swap * (j + swap * rest);
    }
    const cplx<float> *o = op + u * op_stride + src;
    bf16x8 v[8];
    __bf16 *w = reinterpret_cast<__bf16 *>(v);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int r = 16 * s + k;
        const cplx<float> x = r < nr ? o[ldop * r] : cplx<float>{0.f, 0.f};
        split_bf16(x.re, w[k], w[32 + k]);
        split_bf16(x.im, w[16 + k], w[48 + k]);
    }
    bf16x8 *dst = reinterpret_cast<bf16x8 *>(apk + (((int64_t)u * nrs + s) * ndata + d) * 64);
#pragma unroll
    for (int k = 0; k < 8; k++) dst[k] = v[k];
}

// One thread per (stage, point): 16 phases -> 64 bf16 of the packed phase table.
__global__ void ft_pack_phase_kernel(int nr, const int *__restrict__ irvec, int64_t npts, const double *__restrict__ xk,
                                     int nrs, __bf16 *__restrict__ bpk) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npts * nrs) return;
    const int64_t p = e % npts;
    const int s = (int)(e / npts);
    const double x0 = xk[3 * p], x1 = xk[3 * p + 1], x2 = xk[3 * p + 2];
    bf16x8 v[8];
    __bf16 *w = reinterpret_cast<__bf16 *>(v);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int r = 16 * s + k;
        float cr = 0.f, si = 0.f;
        if (r < nr) cis2pi(irvec[3 * r] * x0 + irvec[3 * r + 1] * x1 + irvec[3 * r + 2] * x2, cr, si);
        split_bf16(si, w[k], w[32 + k]);
        split_bf16(cr, w[16 + k], w[48 + k]);
    }
    bf16x8 *dst = reinterpret_cast<bf16x8 *>(bpk + ((int64_t)s * npts + p) * 64);
#pragma unroll
    for (int k = 0; k < 8; k++) dst[k] = v[k];
}

template <class T>
__global__ void phase_kernel(int nr, const int *__restrict__ irvec, int nk, const double *__restrict__ xk,
                             cplx<T> *__restrict__ phase) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nr * nk) return;
    const int r = (int)(e % nr);
    const int64_t j = e / nr;
    const double t = irvec[3 * r] * xk[3 * j] + irvec[3 * r + 1] * xk[3 * j + 1] + irvec[3 * r + 2] * xk[3 * j + 2];
    T c, s;
    cis2pi(t, c, s);
    phase[e] = cplx<T>{c, s};
}

static int check_ft_args(int64_t ndata, int nr, const void *op, int64_t ldop, const int *irvec, int nk,
                         const double *xk, void *out, int64_t ldout) {
    EPHIP_REQUIRE(ndata >= 0 && nr >= 0 && nk >= 0, "fourier: negative size (ndata=%lld nr=%d nk=%d)",
                  (long long)ndata, nr, nk);
    EPHIP_REQUIRE(ldop >= ndata && ldout >= ndata, "fourier: leading dimension smaller than ndata");
    EPHIP_REQUIRE(ndata == 0 || nk == 0 || (op && irvec && xk && out) || nr == 0,
                  "fourier: null device pointer");
    EPHIP_REQUIRE((ndata + FT_TM - 1) / FT_TM <= 65535, "fourier: ndata=%lld exceeds %lld rows per call",
                  (long long)ndata, (long long)FT_TM * 65535);
    return EPHIP_OK;
}

template <class T>
static int zero_out(int64_t ndata, int nk, void *out, int64_t ldout, hipStream_t st) {
    EPHIP_TRY_HIP(hipMemset2DAsync(out, ldout * sizeof(cplx<T>), 0, ndata * sizeof(cplx<T>), nk, st));
    return EPHIP_OK;
}

}  // namespace ephip

using namespace ephip;

extern "C" int ephip_fourier_c32(int64_t ndata, int32_t nr, const void *op, int64_t ldop, const int32_t *irvec,
                                 int32_t nk, const double *xk, void *out, int64_t ldout, int32_t precision,
                                 void *stream) {
    int st = check_ft_args(ndata, nr, op, ldop, irvec, nk, xk, out, ldout);
    if (st) return st;
    EPHIP_REQUIRE(precision == EPHIP_PREC_FP32 || precision == EPHIP_PREC_BF16X3,
                  "fourier: unknown precision %d", precision);
    if (ndata == 0 || nk == 0) return EPHIP_OK;
    hipStream_t s = as_stream(stream);
    if (nr == 0) return zero_out<float>(ndata, nk, out, ldout, s);
    const unsigned gy = (unsigned)((ndata + FT_TM - 1) / FT_TM);
    auto o = static_cast<const cplx<float> *>(op);
    auto c = static_cast<cplx<float> *>(out);
    if (precision == EPHIP_PREC_FP32) {
        if (nk <= 64) {
            dim3 g((nk + 63) / 64, gy);
            launch_ft<float, 64>(g_ft_variant[1], g, s, ndata, nr, o, ldop, irvec, nk, xk, c, ldout, FtBatch{});
        } else {
            dim3 g((nk + 127) / 128, gy);
            launch_ft<float, 128>(g_ft_variant[0], g, s, ndata, nr, o, ldop, irvec, nk, xk, c, ldout, FtBatch{});
        }
    } else {
        if (nk <= 64) {
            dim3 g((nk + 63) / 64, gy);
            ft_bf16x3_kernel<64><<<g, 256, 0, s>>>(ndata, nr, o, ldop, irvec, nk, xk, c, ldout, FtBatch{});
        } else {
            dim3 g((nk + 127) / 128, gy);
            ft_bf16x3_kernel<128><<<g, 256, 0, s>>>(ndata, nr, o, ldop, irvec, nk, xk, c, ldout, FtBatch{});
        }
    }
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_fourier_c64(int64_t ndata, int32_t nr, const void *op, int64_t ldop, const int32_t *irvec,
                                 int32_t nk, const double *xk, void *out, int64_t ldout, void *stream) {
    int st = check_ft_args(ndata, nr, op, ldop, irvec, nk, xk, out, ldout);
    if (st) return st;
    if (ndata == 0 || nk == 0) return EPHIP_OK;
    hipStream_t s = as_stream(stream);
    if (nr == 0) return zero_out<double>(ndata, nk, out, ldout, s);
    dim3 g((nk + 63) / 64, (unsigned)((ndata + FT_TM - 1) / FT_TM));
    launch_ft<double, 64>(g_ft_variant[2], g, s, ndata, nr, static_cast<const cplx<double> *>(op), ldop, irvec, nk,
                          xk, static_cast<cplx<double> *>(out), ldout, FtBatch{});
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_fourier_phase_c64(int32_t nr, const int32_t *irvec, int32_t nk, const double *xk, void *phase,
                                       void *stream) {
    EPHIP_REQUIRE(nr >= 0 && nk >= 0, "fourier_phase: negative size");
    const int64_t n = (int64_t)nr * nk;
    if (n == 0) return EPHIP_OK;
    EPHIP_REQUIRE(irvec && xk && phase, "fourier_phase: null device pointer");
    phase_kernel<double><<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(
        nr, irvec, nk, xk, static_cast<cplx<double> *>(phase));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_fourier_batched_c32(int64_t ndata, int32_t nr, const void *op, int64_t ldop, int64_t op_stride,
                                         const int32_t *irvec, const double *xk, int32_t nunits, int32_t nk_max,
                                         const int32_t *ncol, const int32_t *col_index, int64_t col_stride, void *out,
                                         int64_t ldout, int64_t out_stride, int32_t precision, void *stream) {
    int st = check_ft_args(ndata, nr, op, ldop, irvec, nk_max, xk, out, ldout);
    if (st) return st;
    EPHIP_REQUIRE(precision == EPHIP_PREC_FP32 || precision == EPHIP_PREC_BF16X3,
                  "fourier_batched: unknown precision %d", precision);
    EPHIP_REQUIRE(nunits >= 0 && nunits <= 65535, "fourier_batched: nunits=%d outside 0..65535", nunits);
    EPHIP_REQUIRE(op_stride >= 0 && out_stride >= 0 && col_stride >= 0, "fourier_batched: negative stride");
    EPHIP_REQUIRE(!col_index || col_stride >= nk_max, "fourier_batched: col_stride < nk_max");
    if (ndata == 0 || nk_max == 0 || nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(nr > 0, "fourier_batched: nr must be positive");
    FtBatch bt;
    bt.op_stride = op_stride;
    bt.out_stride = out_stride;
    bt.col_index = col_index;
    bt.col_stride = col_stride;
    bt.ncol = ncol;
    hipStream_t s = as_stream(stream);
    const unsigned gy = (unsigned)((ndata + FT_TM - 1) / FT_TM);
    auto o = static_cast<const cplx<float> *>(op);
    auto c = static_cast<cplx<float> *>(out);
    const bool small = nk_max <= 64;
    dim3 g(small ? (nk_max + 63) / 64 : (nk_max + 127) / 128, gy, nunits);
    if (precision == EPHIP_PREC_FP32) {
        if (small)
            launch_ft<float, 64>(g_ft_variant[1], g, s, ndata, nr, o, ldop, irvec, nk_max, xk, c, ldout, bt);
        else
            launch_ft<float, 128>(g_ft_variant[0], g, s, ndata, nr, o, ldop, irvec, nk_max, xk, c, ldout, bt);
    } else {
        if (small)
            ft_bf16x3_kernel<64><<<g, 256, 0, s>>>(ndata, nr, o, ldop, irvec, nk_max, xk, c, ldout, bt);
        else
            ft_bf16x3_kernel<128><<<g, 256, 0, s>>>(ndata, nr, o, ldop, irvec, nk_max, xk, c, ldout, bt);
    }
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_ft_set_variant(int32_t kind, int32_t variant) {
    static const int nvar[9] = {3, 3, 3, 8, 13, 2, 2, 2, 2};
    EPHIP_REQUIRE(kind >= 0 && kind < 9, "ft_set_variant: kind %d outside 0..8", kind);
    EPHIP_REQUIRE(variant >= 0 && variant < nvar[kind], "ft_set_variant: variant %d outside the range of kind %d",
                  variant, kind);
    g_ft_variant[kind] = variant;
    return EPHIP_OK;
}

extern "C" int ephip_ft_get_variant(int32_t kind) { return (kind >= 0 && kind < 9) ? g_ft_variant[kind] : -1; }

extern "C" int ephip_ft_pack_operand(int64_t ndata, int32_t nr, const void *op, int64_t ldop, int32_t nunits,
                                     int64_t op_stride, void *apk, void *stream) {
    EPHIP_REQUIRE(ndata >= 0 && nr >= 0 && nunits >= 0 && nunits <= 65535 && ldop >= ndata && op_stride >= 0,
                  "ft_pack_operand: bad sizes (ndata=%lld nr=%d nunits=%d)", (long long)ndata, nr, nunits);
    const int nrs = (nr + 15) / 16;
    if (ndata == 0 || nrs == 0 || nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(op && apk, "ft_pack_operand: null device pointer");
    dim3 g((unsigned)((ndata * nrs + 255) / 256), (unsigned)nunits);
    ft_pack_operand_kernel<<<g, 256, 0, as_stream(stream)>>>(ndata, nr, nrs, static_cast<const cplx<float> *>(op),
                                                             ldop, op_stride, static_cast<__bf16 *>(apk));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_ft_pack_operand_swapped(int64_t ndata, int32_t nr, int32_t nb, const void *op, int64_t ldop,
                                             void *apk, void *stream) {
    EPHIP_REQUIRE(ndata >= 0 && nr >= 0 && ldop >= ndata && nb >= 1,
                  "ft_pack_operand_swapped: bad sizes (ndata=%lld nr=%d nb=%d)", (long long)ndata, nr, nb);
    EPHIP_REQUIRE(ndata % ((int64_t)nb * nb) == 0, "ft_pack_operand_swapped: ndata=%lld not a multiple of nb^2=%d",
                  (long long)ndata, nb * nb);
    const int nrs = (nr + 15) / 16;
    if (ndata == 0 || nrs == 0) return EPHIP_OK;
    EPHIP_REQUIRE(op && apk, "ft_pack_operand_swapped: null device pointer");
    dim3 g((unsigned)((ndata * nrs + 255) / 256), 1);
    ft_pack_operand_kernel<<<g, 256, 0, as_stream(stream)>>>(ndata, nr, nrs, static_cast<const cplx<float> *>(op),
                                                             ldop, 0, static_cast<__bf16 *>(apk), nb);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_ft_pack_phases(int32_t nr, const int32_t *irvec, int64_t npts, const double *xk, void *bpk,
                                    void *stream) {
    EPHIP_REQUIRE(nr >= 0 && npts >= 0, "ft_pack_phases: negative size");
    const int nrs = (nr + 15) / 16;
    if (nrs == 0 || npts == 0) return EPHIP_OK;
    EPHIP_REQUIRE(irvec && xk && bpk, "ft_pack_phases: null device pointer");
    ft_pack_phase_kernel<<<(unsigned)((npts * nrs + 255) / 256), 256, 0, as_stream(stream)>>>(
        nr, irvec, npts, xk, nrs, static_cast<__bf16 *>(bpk));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_fourier_packed_bf16x3(int64_t ndata, int32_t nr, const void *apk, int32_t nunits, const void *bpk,
                                           int64_t npts, int32_t nk_max, const int32_t *ncol, const int32_t *col_index,
                                           int64_t col_stride, void *out, int64_t ldout, int64_t out_stride,
                                           void *stream) {
    EPHIP_REQUIRE(ndata >= 0 && nr >= 0 && nunits >= 0 && nunits <= 65535 && nk_max >= 0 && npts >= 0,
                  "fourier_packed: bad sizes (ndata=%lld nr=%d nunits=%d nk_max=%d)", (long long)ndata, nr, nunits,
                  nk_max);
    EPHIP_REQUIRE(ldout >= ndata && out_stride >= 0, "fourier_packed: leading dimension smaller than ndata");
    EPHIP_REQUIRE(col_index || nk_max <= npts, "fourier_packed: nk_max=%d exceeds the %lld table points", nk_max,
                  (long long)npts);
    EPHIP_REQUIRE(!col_index || col_stride >= nk_max, "fourier_packed: col_stride < nk_max");
    if (ndata == 0 || nk_max == 0 || nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(nr > 0, "fourier_packed: nr must be positive");
    EPHIP_REQUIRE(apk && bpk && out, "fourier_packed: null device pointer");
    const int nrs = (nr + 15) / 16;
    EPHIP_REQUIRE(ndata * 64 < (1LL << 32) && npts * 64 < (1LL << 32),
                  "fourier_packed: operand or table too large for 32-bit offsets");
    // Tile shape: 192 x 64 for 192-row operands (one e-ph unit); for operands of 128 rows or more the
    // tuner's variant of kind 3 (0: 128 x 128, 1: 128 x 128 with A prefetch, 2: 192 x 64 when
    // ndata % 192 == 0, 3: 192 x 64 in XCD-grouped order, one unit only, 4-6: variants of 3 with
    // mid-sequence LDS writes and/or a pipelined main loop, 7: eight waves on 192 x 128 with the
    // operand staged through LDS, one unit without column lists); otherwise 64 x 128.
    int shape;
    const int var = g_ft_variant[3];
    if (ndata % 192 == 0 && (nk_max <= 64 || ndata % 128 != 0))
        shape = 2;
    else if (ndata >= 128)
        shape = ((var >= 2) && ndata % 192 == 0) ? 2 : (var == 1 ? 3 : 1);
    else
        shape = 0;
    const bool grouped = shape == 2 && var >= 3 && nunits == 1;
    const bool mid = grouped && (var == 4 || var == 5);
    const bool pipe = grouped && (var == 5 || var == 6);
    const int tm = shape == 2 ? 192 : (shape == 0 ? 64 : 128), tn = shape == 2 ? 64 : 128;
    EPHIP_REQUIRE((ndata + tm - 1) / tm <= 65535, "fourier_packed: ndata=%lld too large for one call", (long long)ndata);
    if (shape == 2 && var == 7 && nunits == 1 && !col_index && !ncol) {
        dim3 g8((unsigned)((nk_max + FT8_COLS - 1) / FT8_COLS), (unsigned)((ndata + FT8_ROWS - 1) / FT8_ROWS), 1);
        ft_bf16p8_kernel<<<g8, 512, 0, as_stream(stream)>>>(ndata, nrs, static_cast<const __bf16 *>(apk),
                                                            static_cast<const __bf16 *>(bpk), npts, nk_max,
                                                            static_cast<cplx<float> *>(out), ldout);
        EPHIP_LAUNCHED();
        return EPHIP_OK;
    }
    dim3 g((unsigned)((nk_max + tn - 1) / tn), (unsigned)((ndata + tm - 1) / tm), (unsigned)nunits);
    auto a = static_cast<const __bf16 *>(apk);
    auto b = static_cast<const __bf16 *>(bpk);
    auto o = static_cast<cplx<float> *>(out);
    hipStream_t st = as_stream(stream);
    if (pipe && mid)
        ft_bf16p_kernel<3, 4, 4, true, true, true, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout,
                                                                             out_stride, col_index, col_stride, ncol);
    else if (pipe)
        ft_bf16p_kernel<3, 4, 4, true, true, false, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout,
                                                                              out_stride, col_index, col_stride, ncol);
    else if (mid)
        ft_bf16p_kernel<3, 4, 4, true, true, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout,
                                                                       out_stride, col_index, col_stride, ncol);
    else if (grouped)
        ft_bf16p_kernel<3, 4, 4, true, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout, out_stride,
                                                                 col_index, col_stride, ncol);
    else if (shape == 2)
        ft_bf16p_kernel<3, 4, 4, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout, out_stride,
                                                           col_index, col_stride, ncol);
    else if (shape == 3)
        ft_bf16p_wide_pf_kernel<<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout, out_stride, col_index,
                                                    col_stride, ncol);
    else if (shape == 1)
        ft_bf16p_kernel<4, 4, 2, false><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout, out_stride,
                                                            col_index, col_stride, ncol);
    else
        ft_bf16p_kernel<2, 4, 2, true><<<g, 256, 0, st>>>(ndata, nrs, a, b, npts, nk_max, o, ldout, out_stride,
                                                           col_index, col_stride, ncol);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_RR_to_kR_project_packed_bf16x3(int32_t nw, int32_t nm, int32_t nrp, int32_t nre,
                                                       const void *apk, const void *bpk, int64_t npts, int32_t nk,
                                                       const int32_t *kus, const int64_t *ustate,
                                                       const int32_t *uband, const void *U, int32_t nbw, void *A,
                                                       int64_t lda, void *stream) {
    EPHIP_REQUIRE(nw == 16, "eph_RR_to_kR_project: fused projection needs nw = 16 (got %d)", nw);
    EPHIP_REQUIRE(nm >= 1 && nrp >= 1 && nre >= 1 && nk >= 0 && npts >= nk && nbw >= 1,
                  "eph_RR_to_kR_project: bad sizes (nm=%d nrp=%d nre=%d nk=%d npts=%lld nbw=%d)", nm, nrp, nre, nk,
                  (long long)npts, nbw);
    const int64_t ndata = (int64_t)nw * nw * nm * nrp;
    EPHIP_REQUIRE(ndata % FT8_ROWS == 0, "eph_RR_to_kR_project: nw^2 nm nrp = %lld not a multiple of %d",
                  (long long)ndata, FT8_ROWS);
    EPHIP_REQUIRE(lda >= (int64_t)nw * nm * nrp, "eph_RR_to_kR_project: lda smaller than nw nm nrp");
    EPHIP_REQUIRE(ndata * 64 < (1LL << 32) && npts * 64 < (1LL << 32),
                  "eph_RR_to_kR_project: operand or table too large for 32-bit offsets");
    EPHIP_REQUIRE(ndata / FT8_ROWS <= 65535, "eph_RR_to_kR_project: ndata=%lld too large for one call",
                  (long long)ndata);
    if (nk == 0) return EPHIP_OK;
    EPHIP_REQUIRE(apk && bpk && kus && ustate && uband && U && A, "eph_RR_to_kR_project: null device pointer");
    const int nrs = (nre + 15) / 16;
    dim3 g((unsigned)((nk + FT8_COLS - 1) / FT8_COLS), (unsigned)(ndata / FT8_ROWS), 1);
    // tuner kind 5: variant 0 = loads one stage ahead, 1 = two stages ahead
    auto kern = g_ft_variant[5] == 1 ? ft_bf16p8_proj_kernel<true> : ft_bf16p8_proj_kernel<false>;
    kern<<<g, 512, 0, as_stream(stream)>>>(ndata, nrs, static_cast<const __bf16 *>(apk),
                                           static_cast<const __bf16 *>(bpk), npts, nk, kus, ustate, uband,
                                           static_cast<const cplx<float> *>(U), nbw, static_cast<cplx<float> *>(A),
                                           lda);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_kq_rotate_packed_bf16x3(int32_t nw, int32_t nm, int32_t nbw, int32_t nr, const void *apk,
                                                 int32_t nunits, const void *bpk, int64_t npts, int64_t cs,
                                                 int64_t max_count, const int32_t *count, const int32_t *col_index,
                                                 const int32_t *ikq,
                                                 const void *U, const int64_t *bmin, const int64_t *bmax,
                                                 const void *uph, const double *omega, float *g2,
                                                 const int32_t *irvec, const int32_t *qgrid, void *stream) {
    EPHIP_REQUIRE(nw == 16 && nm == 12, "eph_kq_rotate_packed: fused form needs nw = 16 and nm = 12 (got %d, %d)", nw,
                  nm);
    EPHIP_REQUIRE(nbw >= 1 && nbw <= FTR_NBW_MAX, "eph_kq_rotate_packed: nbw=%d outside 1..%d", nbw, FTR_NBW_MAX);
    EPHIP_REQUIRE(nr > 0 && nunits >= 0 && nunits <= 65535 && cs >= 0 && npts > 0, "eph_kq_rotate_packed: bad sizes");
    EPHIP_REQUIRE(npts * 64 < (1LL << 32), "eph_kq_rotate_packed: phase table too large for 32-bit offsets");
    if (nunits == 0 || cs == 0) return EPHIP_OK;
    EPHIP_REQUIRE(apk && bpk && count && col_index && ikq && U && bmin && bmax && uph && omega && g2,
                  "eph_kq_rotate_packed: null device pointer");
    const int nrs = (nr + 15) / 16;
    dim3 g((unsigned)((cs + 63) / 64), 1, (unsigned)nunits);
    auto a = static_cast<const __bf16 *>(apk);
    auto b = static_cast<const __bf16 *>(bpk);
    auto uu = static_cast<const cplx<float> *>(U);
    auto up = static_cast<const cplx<float> *>(uph);
    hipStream_t st = as_stream(stream);
    const int fv = g_ft_variant[4];
    const bool gp = fv == 4;
    const bool mid = fv == 1 || fv == 3, pipe = fv == 2 || fv == 3;
    int3 qg = make_int3(0, 0, 0);
    if (gp) {
        EPHIP_REQUIRE(irvec && qgrid, "eph_kq_rotate_packed: on-the-fly phases need irvec and the q-grid");
        EPHIP_REQUIRE(qgrid[0] >= 1 && qgrid[1] >= 1 && qgrid[2] >= 1 && qgrid[0] <= FTG_NMAX &&
                          qgrid[1] <= FTG_NMAX && qgrid[2] <= FTG_NMAX &&
                          (int64_t)qgrid[0] * qgrid[1] * qgrid[2] == npts,
                      "eph_kq_rotate_packed: q-grid dimensions must be 1..%d with n1 n2 n3 = npts", FTG_NMAX);
        qg = make_int3(qgrid[0], qgrid[1], qgrid[2]);
    }
    if (fv >= 5) {
        // variant 5: plain order; 6, 7, 8: XCD-dealt units walked in groups of 2, 4, 8; 9, 10: as 7, 8 with
        // the epilogue's column metadata loaded before the main loop and U_kq / u_ph fetched ahead
        // 11, 12: as 9, 10 with the main loop's loads two stages ahead
        const int ug = fv == 5 ? 0 : (fv == 6 ? 2 : ((fv == 7 || fv == 9 || fv == 11) ? 4 : 8));
        const bool pfe = fv >= 9, deep = fv >= 11;
        const unsigned gz = ug > 0 ? (unsigned)(8 * ((nunits + 7) / 8)) : (unsigned)nunits;
        const int64_t ncols = max_count > 0 ? min(max_count, cs) : cs;  // grid width; longer lists loop
        dim3 g8((unsigned)((ncols + FT8_COLS - 1) / FT8_COLS), 1, gz);
        // instantiations by window width: 2 (nbw <= 2), 3, 4; the two-stage-ahead variants by 2 and 4
        auto kern = ft_bf16p8_rot_kernel<2>;
        if (deep)
            kern = nbw <= 2 ? ft_bf16p8_rot_kernel<2, true, true> : ft_bf16p8_rot_kernel<4, true, true>;
        else if (pfe)
            kern = nbw <= 2 ? ft_bf16p8_rot_kernel<2, true>
                            : (nbw == 3 ? ft_bf16p8_rot_kernel<3, true> : ft_bf16p8_rot_kernel<4, true>);
        else
            kern = nbw <= 2 ? ft_bf16p8_rot_kernel<2> : (nbw == 3 ? ft_bf16p8_rot_kernel<3> : ft_bf16p8_rot_kernel<4>);
        kern<<<g8, 512, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax, up, omega, nbw, g2, nunits,
                                 ug);
        EPHIP_LAUNCHED();
        return EPHIP_OK;
    }
    if (gp && nbw <= 2) {
        ft_bf16p_rot_kernel<2, false, false, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu,
                                                                      bmin, bmax, up, omega, nbw, g2, nr, irvec, qg);
        EPHIP_LAUNCHED();
        return EPHIP_OK;
    }
    if (gp) {
        ft_bf16p_rot_kernel<4, false, false, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu,
                                                                      bmin, bmax, up, omega, nbw, g2, nr, irvec, qg);
        EPHIP_LAUNCHED();
        return EPHIP_OK;
    }
    if (nbw <= 2 && pipe && mid)
        ft_bf16p_rot_kernel<2, true, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax,
                                                              up, omega, nbw, g2, nr, irvec, qg);
    else if (nbw <= 2 && pipe)
        ft_bf16p_rot_kernel<2, false, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin,
                                                               bmax, up, omega, nbw, g2, nr, irvec, qg);
    else if (nbw > 2 && pipe)
        ft_bf16p_rot_kernel<4, false, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin,
                                                               bmax, up, omega, nbw, g2, nr, irvec, qg);
    else if (nbw <= 2 && mid)
        ft_bf16p_rot_kernel<2, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax, up,
                                                        omega, nbw, g2, nr, irvec, qg);
    else if (nbw <= 2)
        ft_bf16p_rot_kernel<2><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax, up, omega,
                                                  nbw, g2, nr, irvec, qg);
    else if (mid)
        ft_bf16p_rot_kernel<4, true><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax, up,
                                                        omega, nbw, g2, nr, irvec, qg);
    else
        ft_bf16p_rot_kernel<4><<<g, 256, 0, st>>>(nrs, a, b, npts, count, col_index, ikq, cs, uu, bmin, bmax, up, omega,
                                                  nbw, g2, nr, irvec, qg);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
