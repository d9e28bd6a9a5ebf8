// Staged ("GridOpt") Fourier transform of the e-ph tensor over Re for k-points that share coordinates,
// the device form of src/wannier/GridOpt.jl (gridopt_set23! :73-107, gridopt_set3! :110-131,
// gridopt_get3! :133-139) and of the batched k3 step of src/wannier/batched_gridopt_interpolator.jl,
// with the outer-k window projection (src/calculator/run_eph_over_k_and_kq.jl:657-659) fused in.
//
// The Re set is embedded in its bounding box r = (r1min + a, r2min + b, r3min + c), a < n1, b < n2,
// c < n3; rmap[a + n1 (b + n2 c)] is the column of that R in the operand, or -1 (absent: zero); the
// (b, c) face is carried padded to np x np, np = max(n2, n3, 3). Then
//   e^{2 pi i k.R} = e^{2 pi i k1 r1} e^{2 pi i k2 r2} e^{2 pi i k3 r3}
// and g(k) = sum_c e3 sum_b e2 sum_a e1 op[., (a, b, c)]:
//   stage 1 (once per k1 plane): op1[., (b, c)] = sum_a e^{2 pi i k1 r1} op[., (a, b, c)];
//   stage 2 (once per (k1, k2) line): op2[., c] = sum_b e^{2 pi i k2 r2} op1[., (b, c)];
//   stage 3 (per k): g(k) = sum_c e^{2 pi i k3 r3} op2[., c], projected on the k-side window bands.
// All in FP32 with phases from FP64 arguments reduced mod 1. Flops per k fall from 8 nd nr to about
// 8 nd (n3 + n2 n3 / (k per line)) plus stage 1 per plane, which streams the operand once.
#include "common.hpp"

namespace ephip {

// (sr, si) += (ar + i ai)(br + i bi) as four FMAs (the form `sr += ar*br - ai*bi` compiles to six
// instructions: a product, an FMA and an add per component).
__device__ __forceinline__ void cmac(float &sr, float &si, float ar, float ai, float br, float bi) {
    sr = fmaf(ar, br, sr);
    sr = fmaf(-ai, bi, sr);
    si = fmaf(ar, bi, si);
    si = fmaf(ai, br, si);
}

// Symmetric boxes (N = 2h + 1 entries along an axis at r = -h .. h, phases e[h + j] = e^{2 pi i q j / m}):
//   sum_b e[b] B[b] = B[h] + sum_{j=1..h} ( cos_j (B[h+j] + B[h-j]) + i sin_j (B[h+j] - B[h-j]) ),
// with (cos_j, sin_j) = e[h + j]: h real-times-complex pairs (4 FMAs each) instead of N complex
// multiply-adds. kq_fold stores S_j = B[h+j] + B[h-j] at h + j and D_j = B[h+j] - B[h-j] at h - j.
template <int N>
__device__ __forceinline__ void kq_fold(float (&vr)[N], float (&vi)[N]) {
    constexpr int H = N / 2;
#pragma unroll
    for (int j = 1; j <= H; j++) {
        const float pr = vr[H + j], pi = vi[H + j], mr = vr[H - j], mi = vi[H - j];
        vr[H + j] = pr + mr;
        vi[H + j] = pi + mi;
        vr[H - j] = pr - mr;
        vi[H - j] = pi - mi;
    }
}

// sum_b e[b] B[b] from the folded B (kq_fold) and the phases e[H + j] of the positive offsets
template <int N>
__device__ __forceinline__ void kq_folded_sum(float &sr, float &si, const float2 *e, const float (&vr)[N],
                                              const float (&vi)[N]) {
    constexpr int H = N / 2;
    sr = vr[H];
    si = vi[H];
#pragma unroll
    for (int j = 1; j <= H; j++) {
        const float c = e[H + j].x, sn = e[H + j].y;
        sr = fmaf(c, vr[H + j], sr);
        sr = fmaf(-sn, vi[H - j], sr);
        si = fmaf(c, vi[H + j], si);
        si = fmaf(sn, vr[H - j], si);
    }
}

constexpr int GO_N1_MAX = 32;

// Stage 1. One thread per operand row d; op1 rows are written in the j-fastest order
// d' = j + nb (i + nb rest) of d = i + nb (j + nb rest) when nb > 1 (the order the VALU projection
// reduces over), in order d when nb <= 1. The (b, c) face is written padded to np x np (zeros outside
// n2 x n3), so the next stage runs on a fixed square face. P planes (k1 values x1[p], outputs op1[p])
// come from one pass over the operand: each operand element loaded feeds P sums.
constexpr int GO_PLANES_MAX = 8;
struct GoPlanes {
    double x1[GO_PLANES_MAX];
    cplx<float> *op1[GO_PLANES_MAX];
};

template <int P>
__global__ __launch_bounds__(256) void gridopt_stage1_kernel(int64_t nd, int n1, int n2, int n3, int np, int r1min,
                                                             const int *__restrict__ rmap,
                                                             const cplx<float> *__restrict__ op, int64_t ldop,
                                                             GoPlanes pl, int nb) {
    __shared__ float2 ph[P][GO_N1_MAX];  // zero past n1 (the tail of the last block of 8 reads it)
    for (int e = threadIdx.x; e < P * GO_N1_MAX; e += blockDim.x) {
        const int p = e / GO_N1_MAX, a = e % GO_N1_MAX;
        float c = 0.f, s = 0.f;
        if (a < n1) cis2pi(pl.x1[p] * (double)(r1min + a), c, s);
        ph[p][a] = make_float2(c, s);
    }
    __syncthreads();
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    int64_t dd = d;
    if (nb > 1) {
        const int64_t i = d % nb, j = (d / nb) % nb, rest = d / ((int64_t)nb * nb);
        dd = j + nb * (i + nb * rest);
    }
    const cplx<float> *o = op + d;
#pragma unroll 1
    for (int c = 0; c < np; c++)
#pragma unroll 1
        for (int b = 0; b < np; b++) {
            float sr[P], si[P];
#pragma unroll
            for (int p = 0; p < P; p++) sr[p] = si[p] = 0.f;
            if (b < n2 && c < n3) {
                const int bc = b + n2 * c;
                // r1 in blocks of 8, branch-free, so a block's index and operand loads issue together: an
                // absent R (rmap -1) and the block's tail past n1 load column 0 and contribute zero
#pragma unroll 1
                for (int a0 = 0; a0 < n1; a0 += 8) {
#pragma unroll
                    for (int da = 0; da < 8; da++) {
                        const int a = a0 + da;
                        const int ir = a < n1 ? rmap[a + n1 * bc] : -1;  // uniform across the block
                        cplx<float> v = o[ldop * max(ir, 0)];
                        if (ir < 0) v = cplx<float>{0.f, 0.f};
#pragma unroll
                        for (int p = 0; p < P; p++) {
                            const float2 e = ph[p][min(a, GO_N1_MAX - 1)];
                            cmac(sr[p], si[p], e.x, e.y, v.re, v.im);
                        }
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < P; p++) pl.op1[p][dd + nd * (b + np * c)] = cplx<float>{sr[p], si[p]};
        }
}

// Stages 2 and 3 with the window projection, for the k of one k1 plane. One thread per row d'
// (j-fastest, NW consecutive lanes = one j group of G = d' / NW); the thread holds op1[d', (b, c)] in
// registers for the whole launch. For each line l in [l0, l1) (surviving k line_start[l] ..
// line_start[l+1]-1, clipped to the batch [k0, k1)), op2 = sum_b e2 op1, then per k of the line
// g = sum_c e3 op2 and, per window band n of k (units kus[k] .. kus[k+1]-1, relative to kus[k0]):
//   A[u][G] = sum_j g[j + NW G] U_k[j, n]   (an NW-lane shuffle reduction).
// xk holds the surviving k (crystal coordinates, 3 per k, FP64).
constexpr int GO_KLINE_MAX = 64, GO_NBW_MAX = 4;

// Sum over groups of NW = 8 or 16 consecutive lanes with DPP moves inside a 16-lane row (no LDS):
// partner lane i^1 and i^2 (quad permutations), then i <-> 7-i within each 8-lane half (pairs quads
// 0/1 and 2/3), then for NW = 16 i <-> 15-i within the row (pairs the halves). Every lane of the group
// ends with the group's sum.
template <int NW>
__device__ __forceinline__ float dpp_group_sum(float v) {
    auto step = [](float x, int ctrl_sel) {
        int xi = __builtin_bit_cast(int, x), y;
        switch (ctrl_sel) {
            case 0: y = __builtin_amdgcn_update_dpp(0, xi, 0xB1, 0xF, 0xF, false); break;   // quad_perm [1,0,3,2]
            case 1: y = __builtin_amdgcn_update_dpp(0, xi, 0x4E, 0xF, 0xF, false); break;   // quad_perm [2,3,0,1]
            case 2: y = __builtin_amdgcn_update_dpp(0, xi, 0x141, 0xF, 0xF, false); break;  // row_half_mirror
            default: y = __builtin_amdgcn_update_dpp(0, xi, 0x140, 0xF, 0xF, false); break; // row_mirror
        }
        return x + __builtin_bit_cast(float, y);
    };
    v = step(v, 0);
    v = step(v, 1);
    v = step(v, 2);
    if constexpr (NW == 16) v = step(v, 3);
    return v;
}

// The value of x in lane l ^ 8 (row rotation by 8), 7 - l within each half row (half-row mirror),
// l ^ 2 and l ^ 1 (quad permutations), inside each 16-lane row.
constexpr int DPP_ROR8 = 0x128, DPP_HALF_MIRROR = 0x141, DPP_QUAD_2301 = 0x4E, DPP_QUAD_1032 = 0xB1;
template <int CTRL>
__device__ __forceinline__ float dpp_take(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF,
                                                                 false));
}

// Row sums of four values at once, scattered over the 16-lane row: on return lanes 4s .. 4s + 3 of every
// row hold the row sum of the s-th argument. A reduce-scatter: lanes split by bit 3 (partner l ^ 8, row
// rotation by 8) and by bit 2 (partner l ^ 7, half-row mirror) keep one half of their values each, then
// two quad permutations (partners l ^ 2, l ^ 1) finish the sums; {8, 7, 2, 1} span all 16 lanes. Eleven
// VALU operations where four separate row sums take sixteen DPP additions and their moves.
__device__ __forceinline__ float dpp_row_sum4(float a, float b, float c, float d) {
    const int l = threadIdx.x & 15;
    const bool h3 = (l & 8) != 0, h2 = (l & 4) != 0;
    float x0 = h3 ? c : a, y0 = h3 ? a : c, x1 = h3 ? d : b, y1 = h3 ? b : d;
    x0 += dpp_take<DPP_ROR8>(y0);
    x1 += dpp_take<DPP_ROR8>(y1);
    float x = h2 ? x1 : x0;
    const float y = h2 ? x0 : x1;
    x += dpp_take<DPP_HALF_MIRROR>(y);
    x += dpp_take<DPP_QUAD_2301>(x);
    x += dpp_take<DPP_QUAD_1032>(x);
    return x;
}

// dpp_row_sum4 fed so that it needs no lane selects: the row sums of conj(w) g for two columns a, b of
// w, scattered as in dpp_row_sum4 (lanes 4s .. 4s + 3 hold re a, im a, re b, im b for s = 0 .. 3). Each
// lane passes its own column wo (b if lane bit 3 is set, else a) and the other one wx, and the parts of
// g arranged by lane bit 2: (p1, p2) = (gr, gi) or (gi, -gr), the part the lane keeps (real or
// imaginary), and (q1, q2) the other; then w.x p1 + w.y p2 is that part of conj(w) g.
__device__ __forceinline__ float dpp_row_sum4_pairs(float2 wo, float2 wx, float p1, float p2, float q1, float q2) {
    float x0 = wo.x * p1 + wo.y * p2;  // own column, kept part
    float x1 = wo.x * q1 + wo.y * q2;  // own column, other part
    const float y0 = wx.x * p1 + wx.y * p2, y1 = wx.x * q1 + wx.y * q2;
    x0 += dpp_take<DPP_ROR8>(y0);  // partner l ^ 8 holds the other column as its own, same parts
    x1 += dpp_take<DPP_ROR8>(y1);
    float x = x0 + dpp_take<DPP_HALF_MIRROR>(x1);  // partner l ^ 7 keeps the other part
    x += dpp_take<DPP_QUAD_2301>(x);
    x += dpp_take<DPP_QUAD_1032>(x);
    return x;
}

// The same for one column: lanes with bit 3 clear end with the row sum of re(conj(w) g), lanes with
// bit 3 set with im; (p1, p2) is the lane's part of g ((gr, gi) or (gi, -gr) by bit 3), (q1, q2) the other.
__device__ __forceinline__ float dpp_row_sum2(float2 w, float p1, float p2, float q1, float q2) {
    float x = w.x * p1 + w.y * p2;
    const float y = w.x * q1 + w.y * q2;
    x += dpp_take<DPP_ROR8>(y);  // partner l ^ 8 keeps the other part
    x += dpp_take<DPP_HALF_MIRROR>(x);
    x += dpp_take<DPP_QUAD_2301>(x);
    x += dpp_take<DPP_QUAD_1032>(x);
    return x;
}

// The same with the parts split by lane bit 2, as in dpp_row_sum4_pairs: lanes with bit 2 clear end with
// the row sum of re(conj(w) g), lanes with bit 2 set with im.
__device__ __forceinline__ float dpp_row_sum2_bit2(float2 w, float p1, float p2, float q1, float q2) {
    float x = w.x * p1 + w.y * p2;
    const float y = w.x * q1 + w.y * q2;
    x += dpp_take<DPP_HALF_MIRROR>(y);  // partner 7 - l in the half row keeps the other part
    x += dpp_take<DPP_QUAD_2301>(x);
    x += dpp_take<DPP_QUAD_1032>(x);
    x += dpp_take<DPP_ROR8>(x);
    return x;
}

template <int N, int NW>
__global__ __launch_bounds__(256) void gridopt_kR_project_kernel(
    int64_t nd, int r2min, int r3min, const cplx<float> *__restrict__ op1,
    const double *__restrict__ xk, int64_t k0, int64_t k1, int l0, int l1, const int *__restrict__ line_start,
    const int *__restrict__ kus, const int64_t *__restrict__ ustate, const int *__restrict__ uband,
    const cplx<float> *__restrict__ U, int nbw, cplx<float> *__restrict__ A, int64_t lda) {
    // per line: stage-2 phases, stage-3 phases of its k, and each k's first unit, unit count and U_k
    // columns of its window bands (read once per line instead of by every thread per k)
    constexpr int N2 = N, N3 = N;
    __shared__ float2 ph2[N2];
    __shared__ float2 ph3[GO_KLINE_MAX * N3];
    __shared__ int kub[GO_KLINE_MAX], knb[GO_KLINE_MAX];
    __shared__ float2 uc[GO_KLINE_MAX * GO_NBW_MAX * NW];
    const int t = threadIdx.x;
    const int64_t dp = (int64_t)blockIdx.x * blockDim.x + t;
    const bool dv = dp < nd;
    const int j = t % NW;
    const int64_t G = dp / NW;
    float vr[N3][N2], vi[N3][N2];
#pragma unroll
    for (int c = 0; c < N3; c++)
#pragma unroll
        for (int b = 0; b < N2; b++) {
            const cplx<float> v = dv ? op1[dp + nd * (b + N2 * c)] : cplx<float>{0.f, 0.f};
            vr[c][b] = v.re;
            vi[c][b] = v.im;
        }
    const int ku0 = kus[k0];
    for (int l = l0; l < l1; l++) {
        const int64_t kb = max((int64_t)line_start[l], k0), ke = min((int64_t)line_start[l + 1], k1);
        const int nkl = (int)(ke - kb);  // uniform; at most GO_KLINE_MAX (host check)
        if (nkl <= 0) continue;
        __syncthreads();  // the previous line's tables are no longer read
        if (t < N2) {
            float c, s;
            cis2pi(xk[3 * kb + 1] * (double)(r2min + t), c, s);
            ph2[t] = make_float2(c, s);
        }
        for (int e = t; e < nkl * N3; e += blockDim.x) {
            const int kk = e / N3, c3 = e % N3;
            float c, s;
            cis2pi(xk[3 * (kb + kk) + 2] * (double)(r3min + c3), c, s);
            ph3[kk * N3 + c3] = make_float2(c, s);
        }
        for (int kk = t; kk < nkl; kk += blockDim.x) {
            kub[kk] = kus[kb + kk] - ku0;
            knb[kk] = min(kus[kb + kk + 1] - kus[kb + kk], nbw);
        }
        __syncthreads();
        for (int e = t; e < nkl * nbw * NW; e += blockDim.x) {
            const int kk = e / (nbw * NW), m = (e / NW) % nbw, jj = e % NW;
            float2 w = make_float2(0.f, 0.f);
            if (m < knb[kk]) {
                const int ub = kub[kk];
                const cplx<float> v = U[(int64_t)NW * NW * ustate[ub] + jj + NW * uband[ub + m]];
                w = make_float2(v.re, v.im);
            }
            uc[(kk * GO_NBW_MAX + m) * NW + jj] = w;
        }
        float o2r[N3], o2i[N3];
#pragma unroll
        for (int c = 0; c < N3; c++) {
            float sr = 0.f, si = 0.f;
#pragma unroll
            for (int b = 0; b < N2; b++) {
                const float2 p = ph2[b];
                cmac(sr, si, p.x, p.y, vr[c][b], vi[c][b]);
            }
            o2r[c] = sr;
            o2i[c] = si;
        }
        __syncthreads();  // U columns staged
        // each lane keeps one k's first unit and band count (nkl <= 64 = wave size); the loop reads them
        // with readlane, so no LDS round trip sits in front of a k's loads
        const int lane = t & 63;
        const int kubv = lane < nkl ? kub[lane] : 0, knbv = lane < nkl ? knb[lane] : 0;
        for (int kk = 0; kk < nkl; kk++) {
            float2 w[GO_NBW_MAX];  // all slots read up front (in bounds; slots past the k's bands unused)
#pragma unroll
            for (int m = 0; m < GO_NBW_MAX; m++) w[m] = uc[(kk * GO_NBW_MAX + m) * NW + j];
            float gr = 0.f, gi = 0.f;
#pragma unroll
            for (int c = 0; c < N3; c++) {
                const float2 p = ph3[kk * N3 + c];
                cmac(gr, gi, p.x, p.y, o2r[c], o2i[c]);
            }
            const int ub = __builtin_amdgcn_readlane(kubv, kk), nbk = __builtin_amdgcn_readlane(knbv, kk);
#pragma unroll
            for (int m = 0; m < GO_NBW_MAX; m++) {
                if (m < nbk) {  // uniform across the block
                    const float sr = dpp_group_sum<NW>(gr * w[m].x - gi * w[m].y);
                    const float si = dpp_group_sum<NW>(gr * w[m].y + gi * w[m].x);
                    if (dv && j == 0) A[(int64_t)(ub + m) * lda + G] = cplx<float>{sr, si};
                }
            }
        }
    }
}

// Per-batch tables of the MFMA form: the U column of every unit of the batch (ucol[u nw + j] =
// U[ustate[u]][j, uband[u]]), the stage-3 phases of every k of the batch (ph3[(k - k0) np + c] =
// e^{2 pi i k3 (r3min + c)}) and the stage-2 phases of every line (ph2[(l - l0) np + b] =
// e^{2 pi i k2 (r2min + b)}, k2 from the line's first k in the batch). One thread per entry.
__global__ __launch_bounds__(256) void gridopt_kR_tables_kernel(int nw, int np, int r2min, int r3min,
                                                                const double *__restrict__ xk, int64_t k0, int64_t k1,
                                                                int l0, int l1, const int *__restrict__ line_start,
                                                                int nunits, const int64_t *__restrict__ ustate,
                                                                const int *__restrict__ uband,
                                                                const cplx<float> *__restrict__ U,
                                                                cplx<float> *__restrict__ ucol,
                                                                cplx<float> *__restrict__ ph2,
                                                                cplx<float> *__restrict__ ph3) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n_u = (int64_t)nunits * nw, n_3 = (k1 - k0) * np, n_2 = (int64_t)(l1 - l0) * np;
    if (e < n_u) {
        const int64_t u = e / nw, j = e % nw;
        ucol[e] = U[(int64_t)nw * nw * ustate[u] + j + (int64_t)nw * uband[u]];
    } else if (e < n_u + n_3) {
        const int64_t f = e - n_u, k = k0 + f / np;
        const int c = (int)(f % np);
        float cr, ci;
        cis2pi(xk[3 * k + 2] * (double)(r3min + c), cr, ci);
        ph3[f] = cplx<float>{cr, ci};
    } else if (e < n_u + n_3 + n_2) {
        const int64_t f = e - n_u - n_3;
        const int l = l0 + (int)(f / np), b = (int)(f % np);
        const int64_t k = max((int64_t)line_start[l], k0);
        float cr, ci;
        cis2pi(xk[3 * k + 1] * (double)(r2min + b), cr, ci);
        ph2[f] = cplx<float>{cr, ci};
    }
}

// Stages 2 and 3 with the window projection on the matrix cores (nw = 16). Along a line the projected
// operators are one complex product with a k-independent right factor:
//   A[(k, n)][G] = sum_{j, c} W[(k, n)][(j, c)] op2[(j, c)][G],   W = U_k[j, n] e^{2 pi i k3 (r3min + c)}.
// A workgroup holds one block of 16 G (i = lane % 16) and all 16 j (four per wave, j = thread / 16): op1
// in its natural row order d = i + 16 (j + 16 rest), so each lane's op2[c] is its own B operand of the
// 16x16x4 FP32 MFMA (B[j][i]) and W is built in the A operand (A[pair][j]). The pairs (k, n) of a line
// (its units, consecutive) are taken 16 at a time; per c each wave runs three MFMAs (a complex product
// from three real products, three independent accumulators) over its four j, and the four waves'
// partial tiles are summed through LDS. The line's phases and U columns come from the batch tables
// (gridopt_kR_tables_kernel) into LDS buffers that alternate between lines, so one barrier per line
// separates a line's staging from the previous line's reads; partial tiles alternate likewise, one
// barrier per tile.
// SYM: the Re box spans -h .. h along axis 2 (N = 2h + 1): the stage-2 sums in the folded form (kq_fold).
template <int N, bool SYM = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N <= 7 ? 3 : 2))) void
gridopt_kR_project_mfma_kernel(int64_t nd, const cplx<float> *__restrict__ op1, int64_t k0, int64_t k1, int l0,
                               int l1, int lb0, const int *__restrict__ line_start, int max_pairs,
                               const int *__restrict__ kus, const cplx<float> *__restrict__ ucol,
                               const cplx<float> *__restrict__ ph2g, const cplx<float> *__restrict__ ph3g, int nbw,
                               cplx<float> *__restrict__ A, int64_t lda) {
    constexpr int NW = 16, NP = GO_KLINE_MAX * GO_NBW_MAX;
    __shared__ float2 ph2[2][N];
    __shared__ float2 ph3[2][GO_KLINE_MAX * N];
    extern __shared__ float2 uc[];    // two buffers of max_pairs x 16: U_k[j, n] of the line's pairs
    __shared__ int pk[2][NP];         // pair -> its k on the line
    __shared__ f32x4 red[2][2][4][64];  // the waves' partial tiles (real, imaginary), two buffers
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int j = t >> 4;  // = 4 wave + lane / 16
    const int64_t dp = (int64_t)blockIdx.x * blockDim.x + t;
    const bool dv = dp < nd;
    const int64_t gcol = (int64_t)blockIdx.x * NW + (t & 15);  // G = i + 16 rest
    float vr[N][N], vi[N][N];
#pragma unroll
    for (int c = 0; c < N; c++)
#pragma unroll
        for (int b = 0; b < N; b++) {
            const cplx<float> v = dv ? op1[dp + nd * (b + N * c)] : cplx<float>{0.f, 0.f};
            vr[c][b] = v.re;
            vi[c][b] = v.im;
        }
    if constexpr (SYM) {
#pragma unroll
        for (int c = 0; c < N; c++) kq_fold<N>(vr[c], vi[c]);  // over b, for the stage-2 sums
    }
    const int ku0 = kus[k0];
    int lb = 0;  // line buffer
    for (int l = l0; l < l1; l++) {
        const int64_t kb = max((int64_t)line_start[l], k0), ke = min((int64_t)line_start[l + 1], k1);
        const int nkl = (int)(ke - kb);  // uniform; at most GO_KLINE_MAX (host check)
        if (nkl <= 0) continue;
        // the line's units, consecutive (at most nbw per k, so at most max_pairs)
        const int u_first = kus[kb] - ku0, np_l = min(kus[ke] - kus[kb], max_pairs);
        float2 *ucl = uc + (size_t)lb * max_pairs * NW;
        // the line's table loads for the first round of each table issued together (one memory round trip
        // per line for up to blockDim q3 phases and 2 blockDim U entries), then written to LDS; the unfolded
        // N = 7 form, at its register bound (12 to 20 bytes of spills so), loads them table by table
        const int n3 = nkl * N, nuc = np_l * NW, bd = blockDim.x;
        int kus0 = 0, kus1 = 0;
        int e3 = t, eu = t;  // where the remaining rounds of the q3 phases and U entries start
        if constexpr (SYM || N != 7) {
            cplx<float> a2{0.f, 0.f}, a3{0.f, 0.f}, au0{0.f, 0.f}, au1{0.f, 0.f};
            if (t < N) a2 = ph2g[(l - lb0) * N + t];
            if (t < n3) a3 = ph3g[(kb - k0) * N + t];
            if (t < nuc) au0 = ucol[(int64_t)u_first * NW + t];
            if (t + bd < nuc) au1 = ucol[(int64_t)u_first * NW + t + bd];
            if (t < 64 && t < nkl) {
                kus0 = kus[kb + t];
                kus1 = kus[kb + t + 1];
            }
            if (t < N) ph2[lb][t] = make_float2(a2.re, a2.im);
            if (t < n3) ph3[lb][t] = make_float2(a3.re, a3.im);
            if (t < nuc) ucl[t] = make_float2(au0.re, au0.im);
            if (t + bd < nuc) ucl[t + bd] = make_float2(au1.re, au1.im);
            e3 = t + bd;
            eu = t + 2 * bd;
        } else {
            if (t < N) {
                const cplx<float> v = ph2g[(l - lb0) * N + t];
                ph2[lb][t] = make_float2(v.re, v.im);
            }
        }
        for (int e = e3; e < n3; e += bd) {
            const cplx<float> v = ph3g[(kb - k0) * N + e];
            ph3[lb][e] = make_float2(v.re, v.im);
        }
        for (int e = eu; e < nuc; e += bd) {
            const cplx<float> v = ucol[(int64_t)u_first * NW + e];
            ucl[e] = make_float2(v.re, v.im);
        }
        if constexpr (!(SYM || N != 7)) {
            if (t < 64 && t < nkl) {
                kus0 = kus[kb + t];
                kus1 = kus[kb + t + 1];
            }
        }
        if (t < 64) {  // pair -> k: exclusive prefix of the band counts (<= 4) over the line's k
            const int nbk = t < nkl ? min(kus1 - kus0, nbw) : 0;
            const unsigned long long below = (1ull << t) - 1ull;
            int start = 0;
#pragma unroll
            for (int bit = 0; bit < 3; bit++) start += __popcll(__ballot((nbk >> bit) & 1) & below) << bit;
            for (int m = 0; m < nbk; m++) pk[lb][start + m] = t;
        }
        __syncthreads();  // this line's tables staged (the other buffers were last read before it)
        float o2r[N], o2i[N], o2s[N];  // op2 and its real + imaginary part (three-product complex MFMA)
#pragma unroll
        for (int c = 0; c < N; c++) {
            float sr = 0.f, si = 0.f;
            if constexpr (SYM) {
                kq_folded_sum<N>(sr, si, ph2[lb], vr[c], vi[c]);
            } else {
#pragma unroll
                for (int b = 0; b < N; b++) {
                    const float2 p = ph2[lb][b];
                    cmac(sr, si, p.x, p.y, vr[c][b], vi[c][b]);
                }
            }
            o2r[c] = sr;
            o2i[c] = si;
            o2s[c] = sr + si;
        }
        for (int p0 = 0; p0 < np_l; p0 += 16) {
            // A operand: row pair p0 + lane % 16 (zero past the line's pairs), column j
            const int p = p0 + (lane & 15);
            int kk = 0;
            float ur = 0.f, ui = 0.f;
            if (p < np_l) {
                kk = pk[lb][p];
                const float2 w = ucl[p * NW + j];
                ur = w.x;
                ui = w.y;
            }
            // W op2 from three real products (Gauss):
            //   re = wr o2r - wi o2i,  im = (wr + wi)(o2r + o2i) - wr o2r - wi o2i
            f32x4 prr = {0.f, 0.f, 0.f, 0.f}, pii = prr, pss = prr;
#pragma unroll
            for (int c = 0; c < N; c++) {
                const float2 e = ph3[lb][kk * N + c];
                const float wr = ur * e.x - ui * e.y, wi = ur * e.y + ui * e.x;
                prr = __builtin_amdgcn_mfma_f32_16x16x4f32(wr, o2r[c], prr, 0, 0, 0);
                pii = __builtin_amdgcn_mfma_f32_16x16x4f32(wi, o2i[c], pii, 0, 0, 0);
                pss = __builtin_amdgcn_mfma_f32_16x16x4f32(wr + wi, o2s[c], pss, 0, 0, 0);
            }
            // this wave's four j: lane holds pair rows 4 (lane / 16) + r, column i = lane % 16
            const int buf = (p0 >> 4) & 1;
            red[buf][0][wave][lane] = prr - pii;
            red[buf][1][wave][lane] = pss - prr - pii;
            __syncthreads();
            {
                const int row = t >> 4, src = (t & 15) + 16 * (row >> 2), r = row & 3;
                const float *rf = reinterpret_cast<const float *>(&red[buf][0][0][0]);
                float sr = 0.f, si = 0.f;
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    sr += rf[(w * 64 + src) * 4 + r];
                    si += rf[((4 + w) * 64 + src) * 4 + r];
                }
                if (dv && p0 + row < np_l) A[(int64_t)(u_first + p0 + row) * lda + gcol] = cplx<float>{sr, si};
            }
        }
        lb ^= 1;
    }
}

}  // namespace ephip

using namespace ephip;

extern "C" int ephip_gridopt_stage1_planes(int64_t nd, int32_t n1, int32_t n2, int32_t n3, int32_t np,
                                           int32_t r1min, const int32_t *rmap, const void *op, int64_t ldop,
                                           int32_t nplanes, const double *x1, int32_t nb, void *const *op1,
                                           void *stream) {
    EPHIP_REQUIRE(nd >= 0 && n1 >= 1 && n2 >= 1 && n3 >= 1 && n1 <= GO_N1_MAX && ldop >= nd && np >= n2 && np >= n3,
                  "gridopt_stage1: bad sizes (nd=%lld n1=%d n2=%d n3=%d np=%d, n1 <= %d)", (long long)nd, n1, n2, n3,
                  np, GO_N1_MAX);
    EPHIP_REQUIRE(nplanes >= 1 && nplanes <= GO_PLANES_MAX, "gridopt_stage1: %d planes (1..%d)", nplanes,
                  GO_PLANES_MAX);
    EPHIP_REQUIRE(nb <= 1 || nd % ((int64_t)nb * nb) == 0, "gridopt_stage1: nd not a multiple of nb^2");
    if (nd == 0) return EPHIP_OK;
    EPHIP_REQUIRE(rmap && op && x1 && op1, "gridopt_stage1: null pointer");
    GoPlanes pl{};
    for (int p = 0; p < nplanes; p++) {
        EPHIP_REQUIRE(op1[p], "gridopt_stage1: null output for plane %d", p);
        pl.x1[p] = x1[p];
        pl.op1[p] = static_cast<cplx<float> *>(op1[p]);
    }
    using K = decltype(&gridopt_stage1_kernel<1>);
    static const K kern[GO_PLANES_MAX] = {gridopt_stage1_kernel<1>, gridopt_stage1_kernel<2>, gridopt_stage1_kernel<3>,
                                          gridopt_stage1_kernel<4>, gridopt_stage1_kernel<5>, gridopt_stage1_kernel<6>,
                                          gridopt_stage1_kernel<7>, gridopt_stage1_kernel<8>};
    kern[nplanes - 1]<<<(unsigned)((nd + 255) / 256), 256, 0, as_stream(stream)>>>(
        nd, n1, n2, n3, np, r1min, rmap, static_cast<const cplx<float> *>(op), ldop, pl, nb);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gridopt_stage1(int64_t nd, int32_t n1, int32_t n2, int32_t n3, int32_t np, int32_t r1min,
                                    const int32_t *rmap, const void *op, int64_t ldop, double x1, int32_t nb,
                                    void *op1, void *stream) {
    return ephip_gridopt_stage1_planes(nd, n1, n2, n3, np, r1min, rmap, op, ldop, 1, &x1, nb, &op1, stream);
}

extern "C" int ephip_gridopt_kR_project(int32_t nw, int32_t nm, int32_t nrp, int32_t np, int32_t r2min,
                                        int32_t r3min, const void *op1, const double *xk, int64_t k0, int64_t k1,
                                        int32_t l0, int32_t l1, const int32_t *line_start, int32_t max_line_k,
                                        const int32_t *kus, const int64_t *ustate, const int32_t *uband,
                                        const void *U, int32_t nbw, void *A, int64_t lda, void *stream) {
    EPHIP_REQUIRE(nw == 8 || nw == 16, "gridopt_kR_project: nw must be 8 or 16 (got %d)", nw);
    EPHIP_REQUIRE(nm >= 1 && nrp >= 1 && nbw >= 1 && nbw <= GO_NBW_MAX && k0 >= 0 && k1 >= k0 && l1 >= l0 &&
                      l0 >= 0,
                  "gridopt_kR_project: bad sizes (nbw=%d, at most %d window bands)", nbw, GO_NBW_MAX);
    EPHIP_REQUIRE(np >= 3 && np <= 8, "gridopt_kR_project: padded box face np x np with 3 <= np <= 8 (got %d)", np);
    EPHIP_REQUIRE(max_line_k >= 0 && max_line_k <= GO_KLINE_MAX, "gridopt_kR_project: %d k on a line > %d",
                  max_line_k, GO_KLINE_MAX);
    EPHIP_REQUIRE(lda >= (int64_t)nw * nm * nrp, "gridopt_kR_project: lda smaller than nw nm nrp");
    const int64_t nd = (int64_t)nw * nw * nm * nrp;
    if (k1 == k0 || l1 == l0) return EPHIP_OK;
    EPHIP_REQUIRE(op1 && xk && line_start && kus && ustate && uband && U && A,
                  "gridopt_kR_project: null device pointer");
    const unsigned g = (unsigned)((nd + 255) / 256);
    auto o1 = static_cast<const cplx<float> *>(op1);
    auto uu = static_cast<const cplx<float> *>(U);
    auto aa = static_cast<cplx<float> *>(A);
    hipStream_t st = as_stream(stream);
    // one instantiation per padded face size (registers hold the thread's np x np stage-1 values)
    using K = decltype(&gridopt_kR_project_kernel<3, 16>);
    static const K k16[6] = {gridopt_kR_project_kernel<3, 16>, gridopt_kR_project_kernel<4, 16>,
                             gridopt_kR_project_kernel<5, 16>, gridopt_kR_project_kernel<6, 16>,
                             gridopt_kR_project_kernel<7, 16>, gridopt_kR_project_kernel<8, 16>};
    static const K k8[6] = {gridopt_kR_project_kernel<3, 8>, gridopt_kR_project_kernel<4, 8>,
                            gridopt_kR_project_kernel<5, 8>, gridopt_kR_project_kernel<6, 8>,
                            gridopt_kR_project_kernel<7, 8>, gridopt_kR_project_kernel<8, 8>};
    const K kern = (nw == 16 ? k16 : k8)[np - 3];
    kern<<<g, 256, 0, st>>>(nd, r2min, r3min, o1, xk, k0, k1, l0, l1, line_start, kus, ustate, uband, uu, nbw, aa,
                            lda);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gridopt_kR_tables(int32_t nw, int32_t np, int32_t r2min, int32_t r3min, const double *xk,
                                      int64_t k0, int64_t k1, int32_t l0, int32_t l1, const int32_t *line_start,
                                      int32_t nunits, const int64_t *ustate, const int32_t *uband, const void *U,
                                      void *ucol, void *ph2, void *ph3, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && np >= 1 && k1 >= k0 && k0 >= 0 && l1 >= l0 && l0 >= 0 && nunits >= 0,
                  "gridopt_kR_tables: bad sizes");
    const int64_t n = (int64_t)nunits * nw + (k1 - k0) * np + (int64_t)(l1 - l0) * np;
    if (n == 0) return EPHIP_OK;
    EPHIP_REQUIRE(xk && line_start && ustate && uband && U && ucol && ph2 && ph3, "gridopt_kR_tables: null pointer");
    gridopt_kR_tables_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(
        nw, np, r2min, r3min, xk, k0, k1, l0, l1, line_start, nunits, ustate, uband,
        static_cast<const cplx<float> *>(U), static_cast<cplx<float> *>(ucol), static_cast<cplx<float> *>(ph2),
        static_cast<cplx<float> *>(ph3));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gridopt_kR_project_mfma(int32_t nw, int32_t nm, int32_t nrp, int32_t np, int32_t r2min,
                                             const void *op1, int64_t k0, int64_t k1, int32_t l0, int32_t l1, int32_t lb0,
                                             const int32_t *line_start, int32_t max_line_k,
                                             int32_t max_line_units, const int32_t *kus, const void *ucol,
                                             const void *ph2, const void *ph3, int32_t nbw, void *A, int64_t lda,
                                             void *stream) {
    EPHIP_REQUIRE(nw == 16, "gridopt_kR_project_mfma: nw must be 16 (got %d)", nw);
    EPHIP_REQUIRE(nm >= 1 && nrp >= 1 && nbw >= 1 && nbw <= GO_NBW_MAX && k0 >= 0 && k1 >= k0 && l1 >= l0 &&
                      l0 >= lb0 && lb0 >= 0,
                  "gridopt_kR_project_mfma: bad sizes (nbw=%d, at most %d window bands)", nbw, GO_NBW_MAX);
    EPHIP_REQUIRE(np >= 3 && np <= 8, "gridopt_kR_project_mfma: padded box face np x np with 3 <= np <= 8 (got %d)",
                  np);
    EPHIP_REQUIRE(max_line_k >= 0 && max_line_k <= GO_KLINE_MAX, "gridopt_kR_project_mfma: %d k on a line > %d",
                  max_line_k, GO_KLINE_MAX);
    EPHIP_REQUIRE(lda >= (int64_t)nw * nm * nrp, "gridopt_kR_project_mfma: lda smaller than nw nm nrp");
    const int64_t nd = (int64_t)nw * nw * nm * nrp;
    if (k1 == k0 || l1 == l0) return EPHIP_OK;
    EPHIP_REQUIRE(op1 && line_start && kus && ucol && ph2 && ph3 && A, "gridopt_kR_project_mfma: null device pointer");
    using K = decltype(&gridopt_kR_project_mfma_kernel<3>);
    static const K kern[6] = {gridopt_kR_project_mfma_kernel<3>, gridopt_kR_project_mfma_kernel<4>,
                              gridopt_kR_project_mfma_kernel<5>, gridopt_kR_project_mfma_kernel<6>,
                              gridopt_kR_project_mfma_kernel<7>, gridopt_kR_project_mfma_kernel<8>};
    static const K kern_sym[3] = {gridopt_kR_project_mfma_kernel<3, true>, gridopt_kR_project_mfma_kernel<5, true>,
                                  gridopt_kR_project_mfma_kernel<7, true>};
    // folded stage-2 sums when the Re box spans -h .. h along axis 2 (odd np, tuner kind 8 = 0)
    const bool sym = (np & 1) && np <= 7 && r2min == -(np / 2) && ephip_ft_get_variant(8) == 0;
    EPHIP_REQUIRE(max_line_units >= 0 && max_line_units <= max_line_k * nbw,
                  "gridopt_kR_project_mfma: max_line_units %d > max_line_k * nbw", max_line_units);
    const int max_pairs = max_line_units > 0 ? max_line_units : max(max_line_k, 1) * nbw;
    const size_t lds = 2 * (size_t)max_pairs * nw * sizeof(float2);  // U columns of a line's pairs, two buffers
    (sym ? kern_sym[(np - 3) / 2] : kern[np - 3])<<<(unsigned)(nd / 256), 256, lds, as_stream(stream)>>>(
        nd, static_cast<const cplx<float> *>(op1), k0, k1, l0, l1, lb0, line_start, max_pairs, kus,
        static_cast<const cplx<float> *>(ucol), static_cast<const cplx<float> *>(ph2),
        static_cast<const cplx<float> *>(ph3), nbw, static_cast<cplx<float> *>(A), lda);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

// ---------------------------------------------------------------------------------------------
// Staged kR->kq over Rp for the q of a uniform grid (m1, m2, m3), fused with the k+q-window and phonon
// rotations and g2: the GridOpt factorisation applied to the unit operators A_u[G, Rp]
// (G = i + nw nu) of the outer-k loop, replacing the per-(unit, q) Fourier sum of
// src/wannier_to_bloch_batched.jl:245-273 and the rotations :442-455.
// ---------------------------------------------------------------------------------------------
namespace ephip {

// Per unit, the first kept-list position of every q line (q1, q2): line_off[u (nlines + 1) + l] =
// lower_bound(col_index[u, 0:count[u]], l m3), and count[u] at l = nlines. One workgroup per unit.
__global__ __launch_bounds__(256) void kq_line_offsets_kernel(int nunits, const int *__restrict__ count,
                                                              const int *__restrict__ col_index, int64_t cs,
                                                              int m3, int nlines, int *__restrict__ line_off) {
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int n = count[u];
    const int *col = col_index + (int64_t)u * cs;
    for (int l = threadIdx.x; l <= nlines; l += blockDim.x) {
        const int v = l * m3;
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (col[mid] < v)
                lo = mid + 1;
            else
                hi = mid;
        }
        line_off[(int64_t)u * (nlines + 1) + l] = lo;
    }
}

constexpr int GQ_M3_MAX = 64, GQ_SLAB_MAX = 1024;

// One workgroup per (q1, unit), one thread per row G < nw nm (threads past it only help stage the
// tables). Rp box: (p1min + a, p2min + b, p3min + c), a < n1, b, c < N; rmap[b + 8 (a + n1 c)] = Rp
// index or -1 (b fastest, rows padded to 8, so the N indices of one (a, c) come in one scalar load).
// Stage 1 (registers): B1[b][c] = sum_a e^{2 pi i (q1/m1)(p1min + a)} A[u][G + ldr Rp]. The kept q of
// the (u, q1) slab are taken in chunks of whole q lines with at most GQ_CHUNK entries (a line with more
// kept q is split into pieces of GQ_CHUNK); per chunk, per
// kept q (in line order; B2[c] = sum_b e2 B1[b][c] is recomputed when the line changes): g = sum_c e3 B2[c];
// t[m, nu] = sum_i conj(U_kq[i, mlo + m]) g[i + 16 nu] (DPP row reduction over i); ep[m, nu'] = sum_nu
// t[m, nu] uph(q)[nu, nu'], g2 = |ep|^2 / (2 omega(q)[nu']) at row u cs + (kept position), slots m < nbw
// (zero past the k+q window), the layout of ephip_eph_kq_rotate_packed_bf16x3.
constexpr int GQ_CHUNK = 32, GQ_M2_MAX = 64;
// U_kq loads per thread issued together when a chunk's window columns are staged (measured at 4 and 6: 6 at
// window width 2, 2.99 against 3.13 ms per cfg2 batch; 4 at width 3, 7.88 against 8.07 ms at cfg4)
template <int NBW>
constexpr int gq_stage() { return NBW == 2 ? 6 : 4; }

// Stage 1 of the staged kR->kq kernels for the row `row` of a unit operator (abase = the unit's first
// element): B1[c][b] = sum_a e1[a] abase[row + ldr Rp(a, b, c)], a < n1, b, c < N.
// Dense boxes (every Rp of the N^3 box present, Rp = s1 a + s2 b + s3 c, n1 = N): the operator is
// addressed arithmetically (uniform offset + the thread's row), fully unrolled, so the loads are not
// chained behind index-map reads and can be issued ahead of their use. Otherwise the index map rmap
// (b + 8 (a + n1 c) -> Rp or -1) is read per (a, c), one face row of loads in flight.
template <int N, bool DENSE>
__device__ __forceinline__ void kq_stage1(float (&br)[N][N], float (&bi)[N][N], const cplx<float> *__restrict__ abase,
                                          int row, int64_t ldr, int n1, const int *__restrict__ rmap,
                                          const float2 *ph1, int s1, int s2, int s3) {
#pragma unroll
    for (int c = 0; c < N; c++)
#pragma unroll
        for (int b = 0; b < N; b++) br[c][b] = bi[c][b] = 0.f;
    if constexpr (DENSE) {
#pragma unroll
        for (int c = 0; c < N; c++)
#pragma unroll
            for (int a = 0; a < N; a++) {
                const float2 p = ph1[a];
#pragma unroll
                for (int b = 0; b < N; b++) {
                    const cplx<float> *col = abase + ldr * (int64_t)(s1 * a + s2 * b + s3 * c);  // uniform
                    const cplx<float> v = col[row];
                    cmac(br[c][b], bi[c][b], p.x, p.y, v.re, v.im);
                }
            }
    } else {
        const cplx<float> *a_u = abase + row;
#pragma unroll
        for (int c = 0; c < N; c++)
            for (int a = 0; a < n1; a++) {
                const float2 p = ph1[a];
                cplx<float> v[N];
#pragma unroll
                for (int b = 0; b < N; b++) {  // branch-free (absent R: column 0 loaded, zero used)
                    const int ir = rmap[b + 8 * (a + n1 * c)];  // uniform; the N entries are contiguous
                    v[b] = a_u[ldr * max(ir, 0)];
                    if (ir < 0) v[b] = cplx<float>{0.f, 0.f};
                }
#pragma unroll
                for (int b = 0; b < N; b++) cmac(br[c][b], bi[c][b], p.x, p.y, v[b].re, v[b].im);
            }
    }
}

// Stage 1 for every q1 of a batch's units at once, ahead of the kR->kq kernel (cube boxes, n1 = N):
//   b1[((u m1 + q1) N^2 + b + N c) nd2 + G] = sum_a e^{2 pi i q1 (p1min + a) / m1} A_u[G, Rp(a, b, c)]
// (zero for absent Rp). One thread per (unit, (b, c), G), G fastest: the N operator values of its column
// are loaded once and feed all m1 sums (folded over a when r1 spans -h .. h); the kR->kq workgroup of
// (unit, q1) then reads its 75 KB slice instead of forming stage 1 from the unit's 527 KB operator.
constexpr int GQ_PRE_M1_MAX = 64;
template <int N, bool DENSE, bool SYM>
__global__ __launch_bounds__(256) void gridopt_kq_stage1_all_kernel(int nunits, int nd2, int m1, int p1min,
                                                                    const int *__restrict__ rmap, int s1, int s2,
                                                                    int s3, const cplx<float> *__restrict__ A,
                                                                    int64_t lda, int64_t ldr, float2 *__restrict__ b1) {
    __shared__ float2 e1[GQ_PRE_M1_MAX * N];
    for (int e = threadIdx.x; e < m1 * N; e += blockDim.x) {
        float c, s;
        cis2pi((double)(e / N) / m1 * (double)(p1min + e % N), c, s);
        e1[e] = make_float2(c, s);
    }
    __syncthreads();
    constexpr int NN = N * N;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nunits * NN * nd2) return;
    const int g = (int)(e % nd2), bc = (int)((e / nd2) % NN);
    const int64_t u = e / ((int64_t)nd2 * NN);
    const int b = bc % N, c = bc / N;
    float vr[N], vi[N];
    const cplx<float> *au = A + u * lda + g;
#pragma unroll
    for (int a = 0; a < N; a++) {
        const int ir = DENSE ? s1 * a + s2 * b + s3 * c : rmap[b + 8 * (a + N * c)];
        const cplx<float> v = au[ldr * max(ir, 0)];
        vr[a] = ir >= 0 ? v.re : 0.f;
        vi[a] = ir >= 0 ? v.im : 0.f;
    }
    if constexpr (SYM) kq_fold<N>(vr, vi);
    float2 *out = b1 + (u * m1 * NN + bc) * nd2 + g;
    for (int q1 = 0; q1 < m1; q1++) {
        float sr = 0.f, si = 0.f;
        if constexpr (SYM) {
            kq_folded_sum<N>(sr, si, e1 + q1 * N, vr, vi);
        } else {
#pragma unroll
            for (int a = 0; a < N; a++) cmac(sr, si, e1[q1 * N + a].x, e1[q1 * N + a].y, vr[a], vi[a]);
        }
        out[(int64_t)q1 * NN * nd2] = make_float2(sr, si);
    }
}

// The kR->kq workgroup's stage-1 sums from the pre-pass output (b1 slice of (unit, q1), rows G).
template <int N>
__device__ __forceinline__ void kq_stage1_load(float (&br)[N][N], float (&bi)[N][N], const float2 *__restrict__ b1u,
                                               int row, int nd2) {
#pragma unroll
    for (int c = 0; c < N; c++)
#pragma unroll
        for (int b = 0; b < N; b++) {
            const float2 v = b1u[(int64_t)(b + N * c) * nd2 + row];
            br[c][b] = v.x;
            bi[c][b] = v.y;
        }
}

// Workgroup order of the staged kR->kq kernels (grid m1 x nunits). Workgroups are dispatched round-robin
// over the 8 XCDs in linear order (x fastest), so in the plain order (q1 = blockIdx.x, unit = blockIdx.y)
// the m1 q1-workgroups of one unit spread over every XCD and each XCD's L2 fetches the unit's operator
// rows. The grouped order gives XCD x one contiguous run of (unit, q1) pairs, unit-major, so one unit's
// workgroups share one L2: linear id L = x + 8 n maps to run position n of XCD x, whose run starts after
// the runs of XCDs 0 .. x-1 (T / 8 pairs each, one more for the first T % 8 XCDs).
__device__ __forceinline__ void kq_block(int grouped, int &q1, int &u) {
    if (!grouped) {
        q1 = blockIdx.x;
        u = blockIdx.y;
        return;
    }
    const int64_t T = (int64_t)gridDim.x * gridDim.y;
    const int64_t L = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const int64_t x = L & 7, n = L >> 3, base = T >> 3, rem = T & 7;
    const int64_t id = x * base + (x < rem ? x : rem) + n;
    q1 = (int)(id % gridDim.x);
    u = (int)(id / gridDim.x);
}

// SYM: the Rp box spans -h .. h along axes 2 and 3 (N = 2h + 1): stages 2 and 3 in the folded form.
// SEG: the slab's kept q may exceed GQ_SLAB_MAX (n2 n3 > GQ_SLAB_MAX): taken in segments of whole q lines
template <int N, int NBW, int NMC = 0, bool DENSE = false, bool SYM = false, bool SEG = false>  // NMC: compile-time mode count (12), or 0
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N <= 7 ? 3 : 1))) void gridopt_kq_rotate_kernel(
    int nm, int n1, int p1min, int p2min, int p3min, const int *__restrict__ rmap, int s1, int s2, int s3, int m1,
    int m2, int m3,
    const cplx<float> *__restrict__ A, int64_t lda, int64_t ldr, const int *__restrict__ line_off,
    const int *__restrict__ col_index, const int *__restrict__ ikq, int64_t cs, const cplx<float> *__restrict__ U,
    const int64_t *__restrict__ bmin, const int64_t *__restrict__ bmax, const cplx<float> *__restrict__ uph,
    const double *__restrict__ omega, int nbw, float *__restrict__ g2, float2 *__restrict__ epo, int grouped,
    const float2 *__restrict__ b1) {
    constexpr int NW = 16;
    // q1, all q2, all q3 phases; the q2/q3 rows padded to an even count and read as 16-byte pairs
    constexpr int NE = (N + 1) & ~1;
    __shared__ float2 ph1[32];
    __shared__ float4 ph2[GQ_M2_MAX * NE / 2], ph3[GQ_M3_MAX * NE / 2];
    __shared__ int slo[GQ_M2_MAX + 1];  // kept-list offsets of the slab's lines
    // metadata of every kept q of the (u, q1) slab: k+q state, q, and the packed word
    // (window bands | first band << 8 | q3 << 16 | q2 << 22), q2, q3 < 64
    __shared__ int sk_kq[GQ_SLAB_MAX], sk_iq[GQ_SLAB_MAX], sk_nb[GQ_SLAB_MAX];
    // an even column count is summed in pairs fed in lane order (dpp_row_sum4_pairs); an odd one with
    // dpp_row_sum4 (padding it to pairs costs more in staging than the selects it saves)
    constexpr bool PAIRS = NBW % 2 == 0;
    // on the symmetric box the lane parts are arranged once per workgroup instead of per column, by bit 2
    // of i for the pair sums and by bit 3 for the one-column sums of odd widths
    constexpr bool SWAP = SYM;
    constexpr int SWAP_BIT = PAIRS ? 4 : 8;
    // width 2 on the symmetric box: a kept q whose k+q window holds one band (most) takes the one-column
    // sum, its U column staged in both slots of the pair
    constexpr bool ONE2 = SWAP && NBW == 2;
    // chunk size: GQ_CHUNK kept q, fewer for wide windows so that four workgroups still fit a CU's LDS
    // (cfg4, window width 3: 12.6 instead of 15.7 ms per 1152-k batch)
    constexpr int CHUNK = NBW >= 4 ? GQ_CHUNK / 2 : NBW == 3 ? GQ_CHUNK * 3 / 4 : GQ_CHUNK;
    __shared__ float2 ucol[CHUNK * NBW * NW];
    // t[kk][m][nu], nu < nm <= 16. Every lane of a row ends its row sum holding one of the values the row
    // writes, and the lanes holding one value store it together (the same word from 4 or 8 lanes), so the
    // column loop stores with no branch and no lane select; the rows of t >= 16 nm (a partly filled last
    // wave, nu < 16 as blocks hold at most 256 threads) store into slots nu >= nm, which nothing reads.
    __shared__ __attribute__((aligned(16))) float2 tq[CHUNK * NBW * NW];
    int q1, u;
    kq_block(grouped, q1, u);
    const int t = threadIdx.x;
    const int nd2 = NW * nm;
    const bool rv = t < nd2;
    const int i = t % NW, nu = t / NW;
    const int nlines = m1 * m2;
    const int *lo_u = line_off + (int64_t)u * (nlines + 1);
    const int64_t row0 = (int64_t)u * cs;
    for (int l = t; l <= m2; l += blockDim.x) slo[l] = lo_u[q1 * m2 + l];  // m2 + 1 may exceed the block
    if (t < n1) {
        float c, s;
        cis2pi((double)q1 / m1 * (double)(p1min + t), c, s);
        ph1[t] = make_float2(c, s);
    }
    for (int e = t; e < m2 * NE; e += blockDim.x) {
        float c = 0.f, s = 0.f;
        if (e % NE < N) cis2pi((double)(e / NE) / m2 * (double)(p2min + e % NE), c, s);
        reinterpret_cast<float2 *>(ph2)[e] = make_float2(c, s);
    }
    for (int e = t; e < m3 * NE; e += blockDim.x) {
        float c = 0.f, s = 0.f;
        if (e % NE < N) cis2pi((double)(e / NE) / m3 * (double)(p3min + e % NE), c, s);
        reinterpret_cast<float2 *>(ph3)[e] = make_float2(c, s);
    }
    __syncthreads();
    // The slab's kept q in segments of whole q lines with at most GQ_SLAB_MAX entries (a line holds at
    // most m3 <= 64): lines [L0, L1) of the segment, its tables in LDS from entry slo[L0] on.
    auto segment_end = [&](int L0) {  // uniform, kept in a scalar register
        if constexpr (!SEG) return m2;  // the whole slab (at most GQ_SLAB_MAX entries)
        int L1 = L0 + 1;
        while (L1 < m2 && slo[L1 + 1] - slo[L0] <= GQ_SLAB_MAX) L1++;
        return __builtin_amdgcn_readfirstlane(L1);
    };
    auto fill = [&](int L0, int L1) {
        const int sb = slo[L0], n = slo[L1] - sb;
        for (int kk = t; kk < n; kk += blockDim.x) {
            const int64_t row = row0 + sb + kk;
            const int kq = ikq[row];
            const int iq = col_index[row];
            sk_kq[kk] = kq;
            sk_iq[kk] = iq;
            const int64_t b0 = bmin[kq], b1 = bmax[kq];
            const int nb = b0 > 0 ? (int)(b1 - b0 + 1) : 0;
            sk_nb[kk] = min(nb, nbw) | ((b0 > 0 ? (int)b0 - 1 : 0) << 8) | ((iq % m3) << 16) |
                        (((iq / m3) % m2) << 22);
        }
    };
    int L0 = 0, L1 = segment_end(0);
    fill(L0, L1);
    // stage 1
    float br[N][N], bi[N][N];
    if (b1)  // precomputed by the stage-1 pre-pass
        kq_stage1_load<N>(br, bi, b1 + ((int64_t)u * m1 + q1) * N * N * nd2, rv ? t : 0, nd2);
    else
        kq_stage1<N, DENSE>(br, bi, A + (int64_t)u * lda, rv ? t : 0, ldr, n1, rmap, ph1, s1, s2, s3);
    if constexpr (SWAP) {
        // rows with bit SWAP_BIT of i set keep i conj(B1) mirrored over b and c. The q2 and q3 phases of
        // the symmetric box satisfy e[N - 1 - x] = conj(e[x]), so their stage-2/3 sums give i conj(g) =
        // (g_i, g_r): the parts of g in the order the pair and one-column sums take them for these lanes
        // (with conj(U) staged for their rows), and the per-column lane selects drop out of the loop.
        if (i & SWAP_BIT) {
            float tr[N][N], ti[N][N];
#pragma unroll
            for (int c = 0; c < N; c++)
#pragma unroll
                for (int b = 0; b < N; b++) {
                    tr[c][b] = bi[N - 1 - c][N - 1 - b];
                    ti[c][b] = br[N - 1 - c][N - 1 - b];
                }
#pragma unroll
            for (int c = 0; c < N; c++)
#pragma unroll
                for (int b = 0; b < N; b++) {
                    br[c][b] = tr[c][b];
                    bi[c][b] = ti[c][b];
                }
        }
    }
    if constexpr (SYM) {
#pragma unroll
        for (int c = 0; c < N; c++) kq_fold<N>(br[c], bi[c]);  // over b, for the stage-2 sums
        // over c, for the stage-3 sums: stage 2 acts on each c alone, so its sums come out folded
        constexpr int H = N / 2;
#pragma unroll
        for (int b = 0; b < N; b++)
#pragma unroll
            for (int j = 1; j <= H; j++) {
                const float pr = br[H + j][b], pi = bi[H + j][b], mr = br[H - j][b], mi = bi[H - j][b];
                br[H + j][b] = pr + mr;
                bi[H + j][b] = pi + mi;
                br[H - j][b] = pr - mr;
                bi[H - j][b] = pi - mi;
            }
    }
    const int lane = t & 63;
    for (;;) {  // segments (uniform)
        // the segment's first slab entry and entry count (uniform, in scalar registers)
        const int sb = __builtin_amdgcn_readfirstlane(slo[L0]), nslab = __builtin_amdgcn_readfirstlane(slo[L1] - sb);
        int e0 = 0, lc = L0;  // chunk start in the segment's tables and the line that holds it
        while (e0 < nslab) {  // uniform
            while (slo[lc + 1] - sb <= e0) lc++;
            // the chunk: whole lines up to CHUNK entries, or the next CHUNK entries of a longer line
            const int e1 = min(e0 + CHUNK, nslab);
            int cut = e1;
            for (int l = lc + 1; l <= L1 && slo[l] - sb <= e1; l++) cut = slo[l] - sb;
            const int ks = e0, nk = cut - e0;  // 0 < nk <= CHUNK
            e0 = cut;
            __syncthreads();  // segment tables written; the previous chunk's tables no longer read
            // STAGE loads in flight per thread: every load of a round is issued (its index clamped to a
            // written table entry, U[0] where no band is read) before the first is waited for
            constexpr int STAGE = gq_stage<NBW>();
            const int nstage = nk * NBW * NW;
            for (int e0 = t; e0 < nstage; e0 += STAGE * blockDim.x) {
                cplx<float> v[STAGE];
                bool ok[STAGE];
#pragma unroll
                for (int r = 0; r < STAGE; r++) {
                    const int e = e0 + r * blockDim.x;
                    const int kk = min(e / (NBW * NW), nk - 1), m = (e / NW) % NBW, ii = e % NW;
                    const int nbm = sk_nb[ks + kk], nbk = nbm & 0xff, mlo = (nbm >> 8) & 0xff;
                    const int ms = (ONE2 && nbk == 1) ? 0 : m;
                    ok[r] = e < nstage && ms < nbk;
                    v[r] = U[ok[r] ? (int64_t)NW * NW * sk_kq[ks + kk] + ii + NW * (mlo + ms) : 0];
                }
#pragma unroll
                for (int r = 0; r < STAGE; r++) {
                    const int e = e0 + r * blockDim.x;
                    if (e >= nstage) break;
                    const int kk = e / (NBW * NW), m = (e / NW) % NBW, ii = e % NW;
                    const float2 w = ok[r] ? make_float2(v[r].re, (SWAP && (ii & SWAP_BIT)) ? -v[r].im : v[r].im)
                                           : make_float2(0.f, 0.f);
                    if constexpr (PAIRS) {  // per pair, the reading lane's own column first (bit 3 of i picks it)
                        ucol[(kk * NBW + (m & ~1) + ((m & 1) ^ ((ii >> 3) & 1))) * NW + ii] = w;
                    } else {
                        ucol[e] = w;
                    }
                }
            }
            __syncthreads();  // U columns staged
            // lane kk keeps entry kk's metadata word (nk <= CHUNK < 64); the loop takes it with readlane
            const int meta = lane < nk ? sk_nb[ks + lane] : 0;
            float b2r[N], b2i[N];
            int cur = -1;
            // the columns of each pair in lane order (staged so): lanes with bit 3 set take the pair's second
            // column as their own (see dpp_row_sum4_pairs); read before the stage-2/3 sums that hide the latency
            const int h3 = (i >> 3) & 1;
            for (int kk = 0; kk < nk; kk++) {
                float2 wo[(NBW + 1) / 2], wx[(NBW + 1) / 2], w[NBW];
                if constexpr (PAIRS) {
#pragma unroll
                    for (int h = 0; h < NBW / 2; h++) {
                        wo[h] = ucol[(kk * NBW + 2 * h) * NW + i];
                        wx[h] = ucol[(kk * NBW + 2 * h + 1) * NW + i];
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < NBW; m++) w[m] = ucol[(kk * NBW + m) * NW + i];
                }
                const int w23 = __builtin_amdgcn_readlane(meta, kk);
                const int q2 = (w23 >> 22) & 63, q3 = (w23 >> 16) & 63;
                if (__builtin_expect(q2 != cur, 0)) {  // uniform; the next line (rare) out of the loop's path
                    cur = q2;
                    float2 e2[NE];
#pragma unroll
                    for (int h = 0; h < NE / 2; h++) {
                        const float4 v = ph2[q2 * (NE / 2) + h];
                        e2[2 * h] = make_float2(v.x, v.y);
                        e2[2 * h + 1] = make_float2(v.z, v.w);
                    }
#pragma unroll
                    for (int c = 0; c < N; c++) {
                        float sr = 0.f, si = 0.f;
                        if constexpr (SYM) {
                            kq_folded_sum<N>(sr, si, e2, br[c], bi[c]);
                        } else {
#pragma unroll
                            for (int b = 0; b < N; b++) cmac(sr, si, e2[b].x, e2[b].y, br[c][b], bi[c][b]);
                        }
                        b2r[c] = sr;  // folded over c on symmetric boxes (the stage-1 sums are)
                        b2i[c] = si;
                    }
                }
                float2 e3[NE];
#pragma unroll
                for (int h = 0; h < NE / 2; h++) {
                    const float4 v = ph3[q3 * (NE / 2) + h];
                    e3[2 * h] = make_float2(v.x, v.y);
                    e3[2 * h + 1] = make_float2(v.z, v.w);
                }
                float gr = 0.f, gi = 0.f;
                if constexpr (SYM) {
                    kq_folded_sum<N>(gr, gi, e3, b2r, b2i);
                } else {
#pragma unroll
                    for (int c = 0; c < N; c++) cmac(gr, gi, e3[c].x, e3[c].y, b2r[c], b2i[c]);
                }
                // conj(w) g summed over the 16 lanes of the row. Even widths: the column pairs, lanes 4s .. 4s + 3
                // of the row write value s of each pair (re, im of its first column, re, im of its second). Odd
                // widths: one column when the k+q window holds one band (uniform; most kept q), lanes 0 .. 7 and
                // 8 .. 15 of the row writing its re and im, else dpp_row_sum4 over the columns. Width 2 on the
                // symmetric box: one column when the window holds one band, the lanes with bit 2 of i clear
                // writing re and those with it set im
                float *tqf = reinterpret_cast<float *>(tq);
                if constexpr (ONE2) {
                    if ((w23 & 0xff) <= 1) {  // uniform
                        const float x = dpp_row_sum2_bit2(wo[0], gr, gi, gi, -gr);
                        tqf[(kk * NBW * NW + nu) * 2 + ((i >> 2) & 1)] = x;
                        continue;
                    }
                }
                if constexpr (!PAIRS) {
                    if ((w23 & 0xff) <= 1) {
                        // the parts of g by bit 3 as in the pair sums (bit 2), held so with SWAP
                        const float p1 = SWAP ? gr : h3 ? gi : gr, p2 = SWAP ? gi : h3 ? -gr : gi;
                        const float q1 = SWAP ? gi : h3 ? gr : gi, q2 = SWAP ? -gr : h3 ? gi : -gr;
                        const float x = dpp_row_sum2(w[0], p1, p2, q1, q2);
                        tqf[(kk * NBW * NW + nu) * 2 + h3] = x;
                        continue;
                    }
                }
                if constexpr (PAIRS) {
                    // the parts of g by bit 2 of i: (p1, p2) = (gr, gi) or (gi, -gr), (q1, q2) the other; with
                    // SWAP the lanes already hold (p1, p2) and their conj(U) makes (q1, q2) = (p2, -p1)
                    const bool h2 = (i & 4) != 0;
                    const float gp1 = SWAP ? gr : h2 ? gi : gr, gp2 = SWAP ? gi : h2 ? -gr : gi;
                    const float gq1 = SWAP ? gi : h2 ? gr : gi, gq2 = SWAP ? -gr : h2 ? gi : -gr;
#pragma unroll
                    for (int h = 0; h < NBW / 2; h++) {
                        const float x = dpp_row_sum4_pairs(wo[h], wx[h], gp1, gp2, gq1, gq2);
                        const int vi = 4 * h + (i >> 2);  // value index: column vi / 2, real (0) or imaginary (1)
                        tqf[((kk * NBW + (vi >> 1)) * NW + nu) * 2 + (vi & 1)] = x;
                    }
                } else {
                    float v[2 * NBW];
                    // g and U as they are (SWAP rows with bit 3 of i hold (g_i, g_r) and conj(U))
                    const float tr = SWAP && h3 ? gi : gr, ti = SWAP && h3 ? gr : gi;
#pragma unroll
                    for (int m = 0; m < NBW; m++) {
                        const float wy = SWAP && h3 ? -w[m].y : w[m].y;
                        v[2 * m] = w[m].x * tr + wy * ti;
                        v[2 * m + 1] = w[m].x * ti - wy * tr;
                    }
#pragma unroll
                    for (int g0 = 0; g0 < 2 * NBW; g0 += 4) {
                        const float x = dpp_row_sum4(v[g0], v[g0 + 1], g0 + 2 < 2 * NBW ? v[g0 + 2] : 0.f,
                                                     g0 + 3 < 2 * NBW ? v[g0 + 3] : 0.f);
                        const int vi = g0 + (i >> 2);
                        if (rv && (i & 3) == 0 && vi < 2 * NBW)
                            tqf[((kk * NBW + (vi >> 1)) * NW + nu) * 2 + (vi & 1)] = x;
                    }
                }
            }
            __syncthreads();  // t of every kept q of the chunk
            if constexpr (NMC > 0) {
                // entries (slot m, kk, mode) with the slot outermost: past the k+q window (most entries of the
                // slots m >= 1) a whole wave skips the rotation and writes zeros; compile-time index arithmetic
                for (int m = 0; m < nbw; m++) {  // uniform
                    for (int e = t; e < nk * NMC; e += blockDim.x) {
                        const int kk = e / NMC, nu2 = e % NMC;
                        const int64_t oi = (row0 + sb + ks + kk) * nbw * NMC + m * NMC + nu2;
                        if (m >= (sk_nb[ks + kk] & 0xff)) {  // past the k+q window: zero
                            g2[oi] = 0.f;
                            if (epo) epo[oi] = make_float2(0.f, 0.f);
                            continue;
                        }
                        const int iq = sk_iq[ks + kk];
                        const cplx<float> *up = uph + (int64_t)NMC * NMC * iq + NMC * nu2;
                        float er = 0.f, ei = 0.f;
                        if constexpr (NMC % 2 == 0) {  // u_ph column and t row as 16-byte loads (even offsets)
                            const float4 *up4 = reinterpret_cast<const float4 *>(up);
                            const float4 *tq4 = reinterpret_cast<const float4 *>(tq + (kk * NBW + m) * NW);
#pragma unroll
                            for (int h = 0; h < NMC / 2; h++) {
                                const float4 p = up4[h], tv = tq4[h];
                                cmac(er, ei, tv.x, tv.y, p.x, p.y);
                                cmac(er, ei, tv.z, tv.w, p.z, p.w);
                            }
                        } else {
#pragma unroll
                            for (int n = 0; n < NMC; n++) {  // constant offsets from one base address
                                const float2 tv = tq[(kk * NBW + m) * NW + n];
                                const cplx<float> p = up[n];
                                cmac(er, ei, tv.x, tv.y, p.re, p.im);
                            }
                        }
                        const float w2 = 2.f * static_cast<float>(omega[(int64_t)NMC * iq + nu2]);
                        g2[oi] = (er * er + ei * ei) / w2;
                        if (epo) epo[oi] = make_float2(er, ei);
                    }
                }
            } else {
                for (int e = t; e < nk * nbw * nm; e += blockDim.x) {
                    const int kk = e / (nbw * nm), rr = e % (nbw * nm), m = rr / nm, nu2 = rr % nm;
                    if constexpr (!PAIRS || ONE2) {  // past the k+q window: zero (its t may not have been formed)
                        if (m >= (sk_nb[ks + kk] & 0xff)) {
                            g2[(row0 + sb + ks + kk) * nbw * nm + rr] = 0.f;
                            if (epo) epo[(row0 + sb + ks + kk) * nbw * nm + rr] = make_float2(0.f, 0.f);
                            continue;
                        }
                    }
                    const int iq = sk_iq[ks + kk];
                    const cplx<float> *up = uph + (int64_t)nm * nm * iq + nm * nu2;
                    float er = 0.f, ei = 0.f;
#pragma unroll
                    for (int n = 0; n < NW; n++) {  // nm <= NW: fixed trip count, loads issued together
                        const int nn = min(n, nm - 1);
                        const float on = n < nm ? 1.f : 0.f;
                        const float2 tv = tq[(kk * NBW + m) * NW + nn];
                        const cplx<float> p = up[nn];
                        cmac(er, ei, on * tv.x, on * tv.y, p.re, p.im);
                    }
                    const float w2 = 2.f * static_cast<float>(omega[(int64_t)nm * iq + nu2]);
                    g2[(row0 + sb + ks + kk) * nbw * nm + rr] = (er * er + ei * ei) / w2;
                    if (epo) epo[(row0 + sb + ks + kk) * nbw * nm + rr] = make_float2(er, ei);
                }
            }
        }
        if (!SEG || L1 >= m2) break;
        L0 = L1;
        L1 = segment_end(L0);
        __syncthreads();  // the last chunk's reads of the segment tables done
        fill(L0, L1);
    }
}

// The same staged kR->kq with stage 3 and the k+q-window rotation as one contraction on FP32 MFMAs.
// Per q line (q1, q2) and per column (kept q, k+q window band m) of the line:
//   t[nu][(q, m)] = sum_{i, c} B2[i + 16 nu][c] W[(i, c)][(q, m)],  W = conj(U_kq[i, mlo + m]) e3[q3][c],
// a (16 x 16 N) by (16 N x columns) product: nu on the MFMA rows (nm <= 16), the columns in tiles of 16
// (only the k+q window's bands, packed per line), K = (i, c) split over the four waves by i (wave w
// takes i = 4w .. 4w + 3), W built in the B operand from U_kq and the q3 phases, the complex product
// from three real MFMAs (Gauss). Stages 1 and 2 stay on the vector ALUs with one row G = i + 16 nu per
// thread (threads t < 16 nm, coalesced loads); the stage-2 sums B2 of a line pass through LDS into the
// MFMA operand layout (lane: nu = lane % 16, i = 4 wave + lane / 16). The four waves' partial tiles are
// summed through LDS into t, and the phonon rotation and g2 run on up to GQM_TQ collected columns at a
// time. The U_kq columns of a line's first GQM_UC columns are gathered into LDS by the wave without
// stage-2 rows (nm <= 12) while the others form the stage-2 sums. LDS arrays read across the 16 lanes of
// an MFMA row use odd pitches (b2s: nm + 1 or 17 per i, tq and ucol: 17 per column), so those reads and
// the writes that fill them are free of bank conflicts. Output: the layout of gridopt_kq_rotate_kernel
// (zeros past the k+q window).
constexpr int GQM_TQ = 48, GQM_UC = 32;

// SYM: the Rp box spans -h .. h along axis 2 (N = 2h + 1): stage 2 in the folded form.
template <int N, int NMC = 0, bool DENSE = false, bool SYM = false>  // NMC: compile-time mode count (12), or 0
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N <= 7 ? 3 : 2))) void
gridopt_kq_rotate_mfma_kernel(int nm_rt, int n1, int p1min, int p2min, int p3min, const int *__restrict__ rmap,
                              int s1, int s2, int s3, int m1, int m2, int m3, const cplx<float> *__restrict__ A, int64_t lda, int64_t ldr,
                              const int *__restrict__ line_off, const int *__restrict__ col_index,
                              const int *__restrict__ ikq, int64_t cs, const cplx<float> *__restrict__ U,
                              const int64_t *__restrict__ bmin, const int64_t *__restrict__ bmax,
                              const cplx<float> *__restrict__ uph, const double *__restrict__ omega, int nbw,
                              float *__restrict__ g2, float2 *__restrict__ epo, int grouped,
                              const float2 *__restrict__ b1) {
    constexpr int NW = 16;
    constexpr int NE = (N + 1) & ~1;
    constexpr int BP = NMC > 0 ? NMC + 1 : 17;  // b2s pitch per i (odd)
    const int nm = NMC > 0 ? NMC : nm_rt;
    // dynamic: q2 and q3 phase rows (m2, m3 rows of NE / 2 float4), the slab's k+q states and metadata words
    extern __shared__ float4 gqm_dyn[];
    float4 *ph2 = gqm_dyn;
    float4 *ph3 = ph2 + m2 * (NE / 2);
    int *sk_kq = reinterpret_cast<int *>(ph3 + m3 * (NE / 2));
    int *sk_nb = sk_kq + m2 * m3;
    __shared__ float2 ph1[GO_N1_MAX];
    __shared__ int slo[GQ_M2_MAX + 1];
    __shared__ float2 b2s[N][NW * BP];       // stage-2 sums of the current line, [c][nu + BP i]
    __shared__ f32x4 red[2][2][4][64];       // the waves' partial tiles (real, imaginary), two buffers
    __shared__ int colk[2][GQ_M3_MAX * 4];   // the line's columns: entry | band slot << 12 | q3 << 16
    __shared__ int ncols[2];
    __shared__ float2 tq[GQM_TQ * 17];       // t of the collected columns, [nu + 17 column]
    __shared__ int tqid[GQM_TQ];
    __shared__ float2 ucol[GQM_UC * 17];     // U_kq[i, band] of the line's first columns, [i + 17 column]
    int q1, u;
    kq_block(grouped, q1, u);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nd2 = NW * nm;
    const bool rv = t < nd2;
    const int nlines = m1 * m2;
    const int *lo_u = line_off + (int64_t)u * (nlines + 1);
    const int64_t row0 = (int64_t)u * cs;
    for (int l = t; l <= m2; l += blockDim.x) slo[l] = lo_u[q1 * m2 + l];
    if (t < n1) {
        float c, s;
        cis2pi((double)q1 / m1 * (double)(p1min + t), c, s);
        ph1[t] = make_float2(c, s);
    }
    for (int e = t; e < m2 * NE; e += blockDim.x) {
        float c = 0.f, s = 0.f;
        if (e % NE < N) cis2pi((double)(e / NE) / m2 * (double)(p2min + e % NE), c, s);
        reinterpret_cast<float2 *>(ph2)[e] = make_float2(c, s);
    }
    for (int e = t; e < m3 * NE; e += blockDim.x) {
        float c = 0.f, s = 0.f;
        if (e % NE < N) cis2pi((double)(e / NE) / m3 * (double)(p3min + e % NE), c, s);
        reinterpret_cast<float2 *>(ph3)[e] = make_float2(c, s);
    }
    __syncthreads();
    const int slab0 = slo[0], nslab = slo[m2] - slab0;  // <= m2 m3
    for (int kk = t; kk < nslab; kk += blockDim.x) {
        const int64_t row = row0 + slab0 + kk;
        const int kq = ikq[row];
        const int iq = col_index[row];
        sk_kq[kk] = kq;
        const int64_t b0 = bmin[kq], b1 = bmax[kq];
        const int nb = b0 > 0 ? min((int)(b1 - b0 + 1), nbw) : 0;
        sk_nb[kk] = nb | ((b0 > 0 ? (int)b0 - 1 : 0) << 8) | ((iq % m3) << 16) | (((iq / m3) % m2) << 22);
        float *o = g2 + row * nbw * nm;
        for (int e = nb * nm; e < nbw * nm; e++) o[e] = 0.f;  // slots past the k+q window
        if (epo)
            for (int e = nb * nm; e < nbw * nm; e++) epo[row * nbw * nm + e] = make_float2(0.f, 0.f);
    }
    __syncthreads();  // the slab tables, read by every wave from the first line on
    // stage 1 (rows G = t)
    float br[N][N], bi[N][N];
    if (b1)  // precomputed by the stage-1 pre-pass
        kq_stage1_load<N>(br, bi, b1 + ((int64_t)u * m1 + q1) * N * N * nd2, rv ? t : 0, nd2);
    else
        kq_stage1<N, DENSE>(br, bi, A + (int64_t)u * lda, rv ? t : 0, ldr, n1, rmap, ph1, s1, s2, s3);
    if constexpr (SYM) {
#pragma unroll
        for (int c = 0; c < N; c++) kq_fold<N>(br[c], bi[c]);  // over b, for the stage-2 sums
    }
    // the phonon rotation and g2 of the collected columns: ep[nu'] = sum_nu t[nu] uph(q)[nu, nu']
    int tqn = 0;
    auto flush = [&]() {
        __syncthreads();  // t of the collected columns written
        for (int e = t; e < tqn * nm; e += blockDim.x) {
            const int col = e / nm, nu2 = e - col * nm;
            const int cid = tqid[col], kk = cid & 0xfff, m = (cid >> 12) & 15;
            const int iq = (q1 * m2 + ((sk_nb[kk] >> 22) & 63)) * m3 + ((cid >> 16) & 63);
            const cplx<float> *up = uph + (int64_t)nm * nm * iq + nm * nu2;
            float er = 0.f, ei = 0.f;
            if constexpr (NMC > 0 && NMC % 2 == 0) {  // the u_ph column as 16-byte loads (nm^2 iq + nm nu2 even)
                const float4 *up4 = reinterpret_cast<const float4 *>(up);
#pragma unroll
                for (int h = 0; h < NMC / 2; h++) {
                    const float4 p = up4[h];
                    const float2 t0 = tq[2 * h + 17 * col], t1 = tq[2 * h + 1 + 17 * col];
                    cmac(er, ei, t0.x, t0.y, p.x, p.y);
                    cmac(er, ei, t1.x, t1.y, p.z, p.w);
                }
            } else if constexpr (NMC > 0) {
#pragma unroll
                for (int n = 0; n < NMC; n++) {
                    const float2 tv = tq[n + 17 * col];
                    const cplx<float> p = up[n];
                    cmac(er, ei, tv.x, tv.y, p.re, p.im);
                }
            } else {
#pragma unroll
                for (int n = 0; n < NW; n++) {
                    const int nn = min(n, nm - 1);
                    const float on = n < nm ? 1.f : 0.f;
                    const float2 tv = tq[nn + 17 * col];
                    const cplx<float> p = up[nn];
                    cmac(er, ei, on * tv.x, on * tv.y, p.re, p.im);
                }
            }
            const float w2 = 2.f * static_cast<float>(omega[(int64_t)nm * iq + nu2]);
            g2[(row0 + slab0 + kk) * nbw * nm + m * nm + nu2] = (er * er + ei * ei) / w2;
            if (epo) epo[(row0 + slab0 + kk) * nbw * nm + m * nm + nu2] = make_float2(er, ei);
        }
        __syncthreads();  // the collected columns consumed
        tqn = 0;
    };
    // MFMA operand coordinates of this lane: row nu = lane % 16, K index i = 4 wave + lane / 16
    const int anu = lane & 15, ai_ = 4 * wave + (lane >> 4);
    const bool aok = anu < nm;
    int tb = 0, lb = 0;
    for (int q2 = 0; q2 < m2; q2++) {  // uniform
        const int eb = slo[q2] - slab0, ee = slo[q2 + 1] - slab0;
        if (ee == eb) continue;
        if (rv) {  // stage 2 of this line for row G = t
            float2 e2[NE];
#pragma unroll
            for (int h = 0; h < NE / 2; h++) {
                const float4 v = ph2[q2 * (NE / 2) + h];
                e2[2 * h] = make_float2(v.x, v.y);
                e2[2 * h + 1] = make_float2(v.z, v.w);
            }
#pragma unroll
            for (int c = 0; c < N; c++) {
                float sr = 0.f, si = 0.f;
                if constexpr (SYM) {
                    kq_folded_sum<N>(sr, si, e2, br[c], bi[c]);
                } else {
#pragma unroll
                    for (int b = 0; b < N; b++) cmac(sr, si, e2[b].x, e2[b].y, br[c][b], bi[c][b]);
                }
                b2s[c][(t >> 4) + BP * (t & 15)] = make_float2(sr, si);
            }
        }
        if (wave == 3) {  // the line's columns (entry, band slot): exclusive prefix of the band counts (<= 4)
            const int ne = ee - eb;  // <= m3 <= 64
            const int meta = lane < ne ? sk_nb[eb + lane] : 0;
            const int nb = meta & 0xff;
            const unsigned long long below = (1ull << lane) - 1ull;
            int start = 0;
#pragma unroll
            for (int bit = 0; bit < 3; bit++) start += __popcll(__ballot((nb >> bit) & 1) & below) << bit;
            for (int m = 0; m < nb; m++) colk[lb][start + m] = (eb + lane) | (m << 12) | (meta & (63 << 16));
            if (lane == 63) ncols[lb] = start + nb;
            // U_kq columns of the first GQM_UC columns: lane = (column % 4, i), 16 columns per pass
            const int nc = min(__shfl(start + nb, 63, 64), GQM_UC);
            wave_lds_sync();  // this wave's column table written
            for (int j = lane >> 4; j < nc; j += 4) {
                const int cid = colk[lb][j], kk = cid & 0xfff, m = (cid >> 12) & 15;
                const int mlo = (sk_nb[kk] >> 8) & 0xff;
                const cplx<float> v = U[(int64_t)NW * NW * sk_kq[kk] + (lane & 15) + NW * (mlo + m)];
                ucol[(lane & 15) + 17 * j] = make_float2(v.re, v.im);
            }
        }
        __syncthreads();  // B2 and the column table of this line
        const int ncol = ncols[lb];
        if (ncol > 0) {  // uniform
            for (int col0 = 0; col0 < ncol; col0 += 16) {  // uniform
                if (tqn + 16 > GQM_TQ) flush();
                const int j = col0 + (lane & 15);
                int cid = 0, q3 = 0;
                float ur = 0.f, ui = 0.f;
                if (j < ncol) {
                    cid = colk[lb][j];
                    q3 = (cid >> 16) & 63;
                    if (j < GQM_UC) {
                        const float2 v = ucol[ai_ + 17 * j];
                        ur = v.x;
                        ui = v.y;
                    } else {
                        const int kk = cid & 0xfff, m = (cid >> 12) & 15;
                        const cplx<float> v = U[(int64_t)NW * NW * sk_kq[kk] + ai_ + NW * (((sk_nb[kk] >> 8) & 0xff) + m)];
                        ur = v.re;
                        ui = v.im;
                    }
                }
                // t = sum B2 conj(U) e3 from three real products (Gauss):
                //   re = b_r w_r - b_i w_i,  im = (b_r + b_i)(w_r + w_i) - b_r w_r - b_i w_i
                f32x4 prr = {0.f, 0.f, 0.f, 0.f}, pii = prr, pss = prr;
#pragma unroll
                for (int h = 0; h < NE / 2; h++) {
                    const float4 ev = ph3[q3 * (NE / 2) + h];
#pragma unroll
                    for (int s = 0; s < 2; s++) {
                        const int c = 2 * h + s;
                        if (c < N) {
                            const float ex = s ? ev.z : ev.x, ey = s ? ev.w : ev.y;
                            const float wr = ur * ex + ui * ey, wi = ur * ey - ui * ex;
                            // A operand: this lane's B2 value, read per tile (the barrier below orders it)
                            const float2 av = aok ? b2s[c][anu + BP * ai_] : make_float2(0.f, 0.f);
                            prr = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, wr, prr, 0, 0, 0);
                            pii = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, wi, pii, 0, 0, 0);
                            pss = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x + av.y, wr + wi, pss, 0, 0, 0);
                        }
                    }
                }
                // this wave's partial: lane holds rows nu = 4 (lane / 16) + r, column lane % 16
                red[tb][0][wave][lane] = prr - pii;
                red[tb][1][wave][lane] = pss - prr - pii;
                __syncthreads();
                {
                    const int nu = t >> 4, col = t & 15, src = col + 16 * (nu >> 2), r = nu & 3;
                    const float *rf = reinterpret_cast<const float *>(&red[tb][0][0][0]);
                    float sr = 0.f, si = 0.f;
#pragma unroll
                    for (int w = 0; w < 4; w++) {
                        sr += rf[(w * 64 + src) * 4 + r];
                        si += rf[((4 + w) * 64 + src) * 4 + r];
                    }
                    if (col0 + col < ncol) {
                        if (nu < nm) tq[nu + 17 * (tqn + col)] = make_float2(sr, si);
                        if (nu == 0) tqid[tqn + col] = cid;  // lanes 0..15 of wave 0 built column col0 + col
                    }
                }
                tqn += min(16, ncol - col0);
                tb ^= 1;
            }
        }
        lb ^= 1;
    }
    flush();
}

}  // namespace ephip

extern "C" int ephip_kq_line_offsets(int32_t nunits, const int32_t *count, const int32_t *col_index, int64_t cs,
                                     int32_t m3, int32_t nlines, int32_t *line_off, void *stream) {
    EPHIP_REQUIRE(nunits >= 0 && cs >= 0 && m3 >= 1 && nlines >= 1, "kq_line_offsets: bad sizes");
    if (nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(count && col_index && line_off, "kq_line_offsets: null device pointer");
    kq_line_offsets_kernel<<<nunits, 256, 0, as_stream(stream)>>>(nunits, count, col_index, cs, m3, nlines, line_off);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_gridopt_kq_rotate(int32_t nw, int32_t nm, int32_t nbw, int32_t nunits, int32_t n1, int32_t np,
                                       int32_t p1min, int32_t p2min, int32_t p3min, const int32_t *rmap,
                                       const int32_t *qgrid, const void *A, int64_t lda, int64_t ldr,
                                       const int32_t *line_off, const int32_t *col_index, const int32_t *ikq,
                                       int64_t cs, const void *U, const int64_t *bmin, const int64_t *bmax,
                                       const void *uph, const double *omega, const int32_t *dense_strides,
                                       float *g2, void *ep, void *b1, void *stream) {
    EPHIP_REQUIRE(nw == 16 && nm >= 1 && nm <= 16, "gridopt_kq_rotate: nw must be 16 and nm <= 16 (got %d, %d)", nw,
                  nm);
    EPHIP_REQUIRE(nbw >= 1 && nbw <= 4, "gridopt_kq_rotate: nbw=%d outside 1..4", nbw);
    EPHIP_REQUIRE(n1 >= 1 && n1 <= 32 && np >= 3 && np <= 8, "gridopt_kq_rotate: Rp box n1=%d (<= 32), np=%d (3..8)",
                  n1, np);
    EPHIP_REQUIRE(qgrid && qgrid[0] >= 1 && qgrid[1] >= 1 && qgrid[2] >= 1 && qgrid[2] <= GQ_M3_MAX &&
                      qgrid[1] <= GQ_M2_MAX && qgrid[0] <= 65535,
                  "gridopt_kq_rotate: q grid must be n1 n2 n3 with n1 <= 65535 and n2, n3 <= %d", GQ_M3_MAX);
    EPHIP_REQUIRE(nunits >= 0 && nunits <= 65535 && ldr >= (int64_t)nw * nm && lda >= ldr && cs >= 0,
                  "gridopt_kq_rotate: bad sizes");
    if (nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(rmap && A && line_off && col_index && ikq && U && bmin && bmax && uph && omega && g2,
                  "gridopt_kq_rotate: null device pointer");
    dim3 g((unsigned)qgrid[0], (unsigned)nunits);
    // dense box: every Rp of the np^3 box present at index s1 a + s2 b + s3 c (the caller's check)
    const bool dense = dense_strides != nullptr && n1 == np;
    const int s1 = dense ? dense_strides[0] : 0, s2 = dense ? dense_strides[1] : 0, s3 = dense ? dense_strides[2] : 0;
    if (dense)
        EPHIP_REQUIRE(s1 >= 0 && s2 >= 0 && s3 >= 0 && (int64_t)(s1 + s2 + s3) * (np - 1) < ((int64_t)1 << 31),
                      "gridopt_kq_rotate: bad dense-box strides");
    const int grouped = ephip_ft_get_variant(7) == 1;  // XCD-grouped workgroup order
    if (b1) {  // stage 1 of every (unit, q1) by the pre-pass into b1 (cube boxes, n1 = np, m1 <= 64)
        EPHIP_REQUIRE(n1 == np && qgrid[0] <= GQ_PRE_M1_MAX && (np & 1),
                      "gridopt_kq_rotate: the stage-1 pre-pass needs a cube box with odd np (n1 = np = %d) and "
                      "m1 <= %d (got m1 = %d)", np, GQ_PRE_M1_MAX, qgrid[0]);
        const bool sym1 = p1min == -(np / 2);
        using KS = decltype(&gridopt_kq_stage1_all_kernel<3, false, false>);
#define GQP_ROW(d, y) {gridopt_kq_stage1_all_kernel<3, d, y>, gridopt_kq_stage1_all_kernel<5, d, y>, \
                       gridopt_kq_stage1_all_kernel<7, d, y>}
        static const KS ptable[2][2][3] = {{GQP_ROW(false, false), GQP_ROW(false, true)},
                                           {GQP_ROW(true, false), GQP_ROW(true, true)}};
#undef GQP_ROW
        EPHIP_REQUIRE(np <= 7, "gridopt_kq_rotate: the stage-1 pre-pass covers np <= 7 (got %d)", np);
        const int64_t nthr = (int64_t)nunits * np * np * nw * nm;
        ptable[dense ? 1 : 0][sym1 ? 1 : 0][(np - 3) / 2]<<<(unsigned)((nthr + 255) / 256), 256, 0,
                                                             as_stream(stream)>>>(
            nunits, nw * nm, qgrid[0], p1min, rmap, s1, s2, s3, static_cast<const cplx<float> *>(A), lda, ldr,
            static_cast<float2 *>(b1));
        EPHIP_LAUNCHED();
    }
    // folded stage-2/3 sums when the Rp box spans -h .. h along axes 2 and 3 (odd np, tuner kind 8 = 0)
    const bool sym = (np & 1) && p2min == -(np / 2) && p3min == -(np / 2) && ephip_ft_get_variant(8) == 0;
    // stage 3 and the k+q-window rotation on MFMAs (its slab tables in LDS at once: n2 n3 <= GQ_SLAB_MAX;
    // the VALU form takes larger slabs in segments)
    if (ephip_ft_get_variant(6) == 1 && qgrid[1] * qgrid[2] <= GQ_SLAB_MAX) {
        using KM = decltype(&gridopt_kq_rotate_mfma_kernel<3>);
#define GQM_ROW(c, d) {gridopt_kq_rotate_mfma_kernel<3, c, d>, gridopt_kq_rotate_mfma_kernel<4, c, d>, \
                       gridopt_kq_rotate_mfma_kernel<5, c, d>, gridopt_kq_rotate_mfma_kernel<6, c, d>, \
                       gridopt_kq_rotate_mfma_kernel<7, c, d>, gridopt_kq_rotate_mfma_kernel<8, c, d>}
        static const KM mtable[2][2][6] = {{GQM_ROW(0, false), GQM_ROW(12, false)},
                                           {GQM_ROW(0, true), GQM_ROW(12, true)}};
#undef GQM_ROW
#define GQMS_ROW(c, d) {gridopt_kq_rotate_mfma_kernel<3, c, d, true>, gridopt_kq_rotate_mfma_kernel<5, c, d, true>, \
                        gridopt_kq_rotate_mfma_kernel<7, c, d, true>}
        static const KM mstable[2][2][3] = {{GQMS_ROW(0, false), GQMS_ROW(12, false)},
                                            {GQMS_ROW(0, true), GQMS_ROW(12, true)}};
#undef GQMS_ROW
        const int ne = (np + 1) & ~1;
        const size_t lds = (size_t)(qgrid[1] + qgrid[2]) * (ne / 2) * 16 + 2 * (size_t)qgrid[1] * qgrid[2] * 4;
        const KM km = (sym && np <= 7) ? mstable[dense ? 1 : 0][nm == 12 ? 1 : 0][(np - 3) / 2]
                                       : mtable[dense ? 1 : 0][nm == 12 ? 1 : 0][np - 3];
        km<<<g, 256, lds, as_stream(stream)>>>(
            nm, n1, p1min, p2min, p3min, rmap, s1, s2, s3, qgrid[0], qgrid[1], qgrid[2],
            static_cast<const cplx<float> *>(A), lda, ldr, line_off, col_index, ikq, cs,
            static_cast<const cplx<float> *>(U), bmin, bmax, static_cast<const cplx<float> *>(uph), omega, nbw, g2,
            static_cast<float2 *>(ep), grouped, static_cast<const float2 *>(b1));
        EPHIP_LAUNCHED();
        return EPHIP_OK;
    }
    const int nd2 = nw * nm;
    const unsigned threads = (unsigned)(((nd2 + 63) / 64) * 64);
    using K = decltype(&gridopt_kq_rotate_kernel<3, 2>);
#define GQ_ROW(nb, c, d, g) {gridopt_kq_rotate_kernel<3, nb, c, d, false, g>, gridopt_kq_rotate_kernel<4, nb, c, d, false, g>, \
                             gridopt_kq_rotate_kernel<5, nb, c, d, false, g>, gridopt_kq_rotate_kernel<6, nb, c, d, false, g>, \
                             gridopt_kq_rotate_kernel<7, nb, c, d, false, g>, gridopt_kq_rotate_kernel<8, nb, c, d, false, g>}
#define GQ_TAB(g) {{{GQ_ROW(2, 0, false, g), GQ_ROW(3, 0, false, g), GQ_ROW(4, 0, false, g)}, \
                    {GQ_ROW(2, 12, false, g), GQ_ROW(3, 12, false, g), GQ_ROW(4, 12, false, g)}}, \
                   {{GQ_ROW(2, 0, true, g), GQ_ROW(3, 0, true, g), GQ_ROW(4, 0, true, g)}, \
                    {GQ_ROW(2, 12, true, g), GQ_ROW(3, 12, true, g), GQ_ROW(4, 12, true, g)}}}
    static const K table[2][2][2][3][6] = {GQ_TAB(false), GQ_TAB(true)};  // [seg][dense][nm 12][nbw][np]
#undef GQ_TAB
#undef GQ_ROW
#define GQS_ROW(nb, c, d, g) {gridopt_kq_rotate_kernel<3, nb, c, d, true, g>, gridopt_kq_rotate_kernel<5, nb, c, d, true, g>, \
                              gridopt_kq_rotate_kernel<7, nb, c, d, true, g>}
#define GQS_TAB(g) {{{GQS_ROW(2, 0, false, g), GQS_ROW(3, 0, false, g), GQS_ROW(4, 0, false, g)}, \
                     {GQS_ROW(2, 12, false, g), GQS_ROW(3, 12, false, g), GQS_ROW(4, 12, false, g)}}, \
                    {{GQS_ROW(2, 0, true, g), GQS_ROW(3, 0, true, g), GQS_ROW(4, 0, true, g)}, \
                     {GQS_ROW(2, 12, true, g), GQS_ROW(3, 12, true, g), GQS_ROW(4, 12, true, g)}}}
    static const K stable[2][2][2][3][3] = {GQS_TAB(false), GQS_TAB(true)};
#undef GQS_TAB
#undef GQS_ROW
    const int seg = qgrid[1] * qgrid[2] > GQ_SLAB_MAX;  // the slab in segments of whole q lines
    const K kern = (sym && np <= 7)
                       ? stable[seg][dense ? 1 : 0][nm == 12 ? 1 : 0][nbw <= 2 ? 0 : nbw - 2][(np - 3) / 2]
                       : table[seg][dense ? 1 : 0][nm == 12 ? 1 : 0][nbw <= 2 ? 0 : nbw - 2][np - 3];
    kern<<<g, threads, 0, as_stream(stream)>>>(nm, n1, p1min, p2min, p3min, rmap, s1, s2, s3, qgrid[0], qgrid[1],
                                                qgrid[2], static_cast<const cplx<float> *>(A), lda, ldr, line_off,
                                                col_index, ikq, cs, static_cast<const cplx<float> *>(U), bmin, bmax,
                                                static_cast<const cplx<float> *>(uph), omega, nbw, g2,
                                                static_cast<float2 *>(ep), grouped, static_cast<const float2 *>(b1));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
