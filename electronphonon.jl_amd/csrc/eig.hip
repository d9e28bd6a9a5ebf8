// Batched Hermitian eigensolve for small matrices (n <= 32), one wavefront per matrix.
//
// Replaces the reference's eigen_batched / eigvals_batched (src/wannier_to_bloch_batched.jl:30-65,
// CUDA method ext/ElectronPhononCUDAExt.jl:80-115): ascending eigenvalues, orthonormal eigenvectors,
// no degeneracy gauge fixing (the reference's batched path applies none either). The per-k path's
// degenerate-band gauge fixing (solve_eigen_el!, src/diagonalize.jl:47-85) is a separate kernel at the end
// of this file, applied to the eigenpairs afterwards.
//
// Method (FP32, all in LDS + registers of one wave):
//   1. Householder reduction of the Hermitian matrix to tridiagonal form, one reflector per column;
//      the reflector vectors are kept in the zeroed part of the matrix for the back-transform.
//   2. Diagonal phase scaling turns the complex tridiagonal into a real symmetric one.
//   3. Implicit QL with Wilkinson-type shifts on (d, e) (the published EISPACK imtql2 scheme): the
//      scalar recurrence is uniform across the wave (lane i holds d_i, e_i; reads go through
//      readlane), the rotations are applied to the rows of Z in parallel (lane = row).
//   4. Eigenvectors V = H_0 ... H_{n-3} D Z, one column per lane.
// A second kernel refines in FP64 against the FP64 matrix: a full Rayleigh-Ritz step in the basis of
// the FP32 eigenvectors (FP64 orthonormalisation, projection, parallel Jacobi), which gives FP64-grade
// eigenvalues and orthonormal vectors also for exactly and nearly degenerate bands.
#include "common.hpp"

#include <mutex>

namespace ephip {

struct c32 {
    float re, im;
};
struct c64 {
    double re, im;
};

__device__ __forceinline__ float rdl(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int rdl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <class TIn>
__device__ __forceinline__ c32 load_c32(const TIn *p, int64_t i) {
    return c32{static_cast<float>(p[i].re), static_cast<float>(p[i].im)};
}

// LDS carve-up per wave (n <= 32): A (n*n c32) | Y (n*n c32) | Z (n*n f32) | v, w (n c32 each) | ec (n c32)
__host__ __device__ constexpr int heev_lds_floats(int n) { return 2 * n * n + 2 * n * n + n * n + 2 * n + 2 * n + 2 * n; }

template <class TIn>
__global__ __launch_bounds__(256) void heev_c32_kernel(int n, int64_t nb, const TIn *__restrict__ H,
                                                       c32 *__restrict__ U, float *__restrict__ E,
                                                       int *__restrict__ info) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (b >= nb) return;  // whole wave exits together; no workgroup barriers below
    float *base = lds + (int64_t)wave * heev_lds_floats(n);
    c32 *A = reinterpret_cast<c32 *>(base);
    c32 *Y = A + n * n;
    float *Z = reinterpret_cast<float *>(Y + n * n);
    c32 *Vh = reinterpret_cast<c32 *>(Z + n * n);
    c32 *Wh = Vh + n;
    c32 *EC = Wh + n;
    const int nn = n * n;
    const TIn *Hb = H + b * nn;

    // 1. Hermitian matrix from the upper triangle ('U' convention)
    for (int e = lane; e < nn; e += 64) {
        const int i = e % n, j = e / n;
        c32 v;
        if (i <= j) {
            v = load_c32(Hb, i + (int64_t)n * j);
        } else {
            v = load_c32(Hb, j + (int64_t)n * i);
            v.im = -v.im;
        }
        if (i == j) v.im = 0.f;
        A[e] = v;
    }
    wave_lds_sync();

    for (int j = 0; j + 2 < n; j++) {
        const int m = n - j - 1;
        c32 x{0.f, 0.f};
        if (lane < m) x = A[(j + 1 + lane) + n * j];
        const float xn2 = wave_sum(x.re * x.re + x.im * x.im);
        if (xn2 == 0.f) {
            if (lane == 0) EC[j] = c32{0.f, 0.f};
            wave_lds_sync();
            continue;
        }
        const float xn = sqrtf(xn2);
        const c32 x0 = A[(j + 1) + n * j];
        const float ax0 = hypotf(x0.re, x0.im);
        const float p0r = ax0 > 0.f ? x0.re / ax0 : 1.f, p0i = ax0 > 0.f ? x0.im / ax0 : 0.f;
        const c32 alpha{-p0r * xn, -p0i * xn};
        c32 v = x;
        if (lane == 0) {
            v.re -= alpha.re;
            v.im -= alpha.im;
        }
        const float vn = sqrtf(wave_sum(v.re * v.re + v.im * v.im));
        v.re /= vn;
        v.im /= vn;
        if (lane < m) {
            Vh[lane] = v;
            A[(j + 1 + lane) + n * j] = v;
        }
        wave_lds_sync();
        // p = A22 v
        c32 p{0.f, 0.f};
        if (lane < m) {
            for (int k = 0; k < m; k++) {
                const c32 a = A[(j + 1 + lane) + n * (j + 1 + k)];
                const c32 vk = Vh[k];
                p.re += a.re * vk.re - a.im * vk.im;
                p.im += a.re * vk.im + a.im * vk.re;
            }
        }
        const float K = wave_sum(v.re * p.re + v.im * p.im);  // Re(v^H p), zero on idle lanes
        c32 w{p.re - K * v.re, p.im - K * v.im};
        if (lane < m) Wh[lane] = w;
        wave_lds_sync();
        // A22 -= 2 (v w^H + w v^H)
        for (int e = lane; e < m * m; e += 64) {
            const int r = e % m, c = e / m;
            const c32 vr = Vh[r], wr = Wh[r], vc = Vh[c], wc = Wh[c];
            c32 &a = A[(j + 1 + r) + n * (j + 1 + c)];
            a.re -= 2.f * ((vr.re * wc.re + vr.im * wc.im) + (wr.re * vc.re + wr.im * vc.im));
            a.im -= 2.f * ((vr.im * wc.re - vr.re * wc.im) + (wr.im * vc.re - wr.re * vc.im));
        }
        if (lane == 0) EC[j] = alpha;
        wave_lds_sync();
    }
    if (n >= 2 && lane == 0) EC[n - 2] = A[(n - 1) + n * (n - 2)];
    wave_lds_sync();

    // 2. real tridiagonal: lane i holds d_i, e_i and the phase tau_i
    float d = 0.f, ee = 0.f, tr = 1.f, ti = 0.f;
    if (lane < n) d = A[lane + n * lane].re;
    {
        float cr = 1.f, ci = 0.f;
        for (int jj = 0; jj + 1 < n; jj++) {
            const c32 ec = EC[jj];
            const float ae = hypotf(ec.re, ec.im);
            if (lane == jj) ee = ae;
            if (ae > 0.f) {
                const float er = ec.re / ae, ei = ec.im / ae;
                const float nr = cr * er - ci * ei, ni = cr * ei + ci * er;
                cr = nr;
                ci = ni;
            }
            if (lane == jj + 1) {
                tr = cr;
                ti = ci;
            }
        }
    }
    const bool want_v = (U != nullptr);
    if (want_v && lane < n)
        for (int c = 0; c < n; c++) Z[lane + n * c] = (lane == c) ? 1.f : 0.f;

    // 3. implicit QL
    int fail = 0;
    const float eps = 1.1920929e-07f;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        while (true) {
            int m = l;
            for (; m < n - 1; m++) {
                const float dd = fabsf(rdl(d, m)) + fabsf(rdl(d, m + 1));
                if (fabsf(rdl(ee, m)) <= eps * dd) break;
            }
            if (m == l) break;
            if (iter++ == 60) {
                fail = 1;
                break;
            }
            const float dlv = rdl(d, l), el = rdl(ee, l);
            float g = (rdl(d, l + 1) - dlv) / (2.f * el);
            float r = hypotf(g, 1.f);
            g = rdl(d, m) - dlv + el / (g + copysignf(r, g));
            float s = 1.f, c = 1.f, p = 0.f;
            bool deflated = false;
            for (int i = m - 1; i >= l; i--) {
                const float ei = rdl(ee, i);
                const float f = s * ei, bb = c * ei;
                r = hypotf(f, g);
                if (lane == i + 1) ee = r;
                if (r == 0.f) {
                    if (lane == i + 1) d -= p;
                    if (lane == m) ee = 0.f;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = rdl(d, i + 1) - p;
                r = (rdl(d, i) - g) * s + 2.f * c * bb;
                p = s * r;
                if (lane == i + 1) d = g + p;
                g = c * r - bb;
                if (want_v && lane < n) {
                    const float zf = Z[lane + n * (i + 1)], zi = Z[lane + n * i];
                    Z[lane + n * (i + 1)] = s * zi + c * zf;
                    Z[lane + n * i] = c * zi - s * zf;
                }
            }
            if (deflated) continue;
            if (lane == l) {
                d -= p;
                ee = g;
            }
            if (lane == m) ee = 0.f;
        }
        if (fail) break;
    }

    // ascending order: lane i's eigenvalue goes to position rank_i (stable for ties)
    int rank = 0;
    for (int k = 0; k < n; k++) {
        const float dk = rdl(d, k);
        rank += (dk < d || (dk == d && k < lane)) ? 1 : 0;
    }
    if (lane < n) E[b * n + rank] = d;
    if (info && lane == 0) info[b] = fail;
    if (!want_v) return;

    // 4. eigenvectors: Y = D Z (columns permuted into ascending order), then the reflectors
    wave_lds_sync();
    if (lane < n) {
        for (int c = 0; c < n; c++) {
            const int rc = rdl(rank, c);
            const float z = Z[lane + n * c];
            Y[lane + n * rc] = c32{tr * z, ti * z};
        }
    }
    wave_lds_sync();
    if (lane < n) {
        for (int j = n - 3; j >= 0; j--) {
            const int m = n - j - 1;
            float sr = 0.f, si = 0.f;
            for (int i = 0; i < m; i++) {
                const c32 vi = A[(j + 1 + i) + n * j];
                const c32 y = Y[(j + 1 + i) + n * lane];
                sr += vi.re * y.re + vi.im * y.im;
                si += vi.re * y.im - vi.im * y.re;
            }
            for (int i = 0; i < m; i++) {
                const c32 vi = A[(j + 1 + i) + n * j];
                c32 &y = Y[(j + 1 + i) + n * lane];
                y.re -= 2.f * (vi.re * sr - vi.im * si);
                y.im -= 2.f * (vi.re * si + vi.im * sr);
            }
        }
    }
    wave_lds_sync();
    c32 *Ub = U + b * nn;
    for (int e = lane; e < nn; e += 64) Ub[e] = Y[e];
}

// ---------------------------------------------------------------------------------------------
// FP64 refinement, one wave per matrix, in the basis of the FP32 eigenvectors:
//   1. Q = V orthonormalised in FP64 by two Loewdin steps Q <- Q (3/2 I - 1/2 Q^H Q)
//      (V is orthonormal to ~1e-6, two steps reach ~1e-23);
//   2. M = Q^H H Q is unitarily similar to H and nearly diagonal (off-diagonal ~ FP32 residual);
//   3. cyclic Jacobi sweeps on M in FP64, the n/2 disjoint rotations of a round-robin round applied
//      in parallel (columns: lane = row; rows: lane = column), accumulated into Q. Convergence is
//      quadratic from a nearly diagonal start, and exact or close eigenvalue pairs need no special
//      handling (a degenerate pair gets a well-defined 45-degree rotation).
// Eigenvalues are diag(M) sorted ascending, eigenvectors the matching columns of Q (rounded to FP32).
// LDS per wave: H | Q | W | M (n columns of c64 each, column pitch ld = n | 1) + rotation table and
// ranks. The odd pitch spreads the 16-byte column-strided accesses (lane = column) over the banks; with
// ld = n = 16 every lane of such an access fell on the same banks.
// ---------------------------------------------------------------------------------------------
constexpr int REFINE_EXTRA_DOUBLES = 176;
__host__ __device__ constexpr int refine_ld(int n) { return n | 1; }
__host__ __device__ constexpr int refine_lds_doubles(int n) { return 8 * n * refine_ld(n) + REFINE_EXTRA_DOUBLES; }

__device__ __forceinline__ double rdl64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 cmulc(c64 a, c64 b) {  // conj(a) * b
    return c64{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return c64{a.re + b.re, a.im + b.im}; }

struct JacobiRot {
    int p, q;
    c64 gpp, gpq, gqp, gqq;  // G = diag(1, e^{-i phi}) [[cs, sn], [-sn, cs]] on (p, q)
};

// n x n matrices in LDS with column pitch ld (element (r, j) at r + ld*j); lanes over entries.
// out = alpha * A B + beta * C
__device__ void lds_gemm_nn(int n, int ld, const c64 *A, const c64 *B, c64 *out, double alpha, double beta,
                            const c64 *C, int lane) {
    for (int e = lane; e < n * n; e += 64) {
        const int r = e % n, j = e / n;
        c64 s{0.0, 0.0};
        for (int k = 0; k < n; k++) s = cadd(s, cmul(A[r + ld * k], B[k + ld * j]));
        c64 v{alpha * s.re, alpha * s.im};
        if (beta != 0.0) v = c64{v.re + beta * C[r + ld * j].re, v.im + beta * C[r + ld * j].im};
        out[r + ld * j] = v;
    }
}

// out = A^H B
__device__ void lds_gemm_cn(int n, int ld, const c64 *A, const c64 *B, c64 *out, int lane) {
    for (int e = lane; e < n * n; e += 64) {
        const int i = e % n, j = e / n;
        c64 s{0.0, 0.0};
        for (int r = 0; r < n; r++) s = cadd(s, cmulc(A[r + ld * i], B[r + ld * j]));
        out[i + ld * j] = s;
    }
}

// Parallel cyclic Jacobi on the n x n Hermitian M (LDS, column pitch ldm, n <= 32): the n/2 disjoint
// rotations of a round-robin round are applied in parallel (columns: lane = row; rows: lane = column) and
// accumulated into the n columns of Q (qrows x n, column pitch ldq, qrows <= 64). Stops when the
// strict upper triangle's Frobenius norm (the part the rotations remove; rounding leaves M Hermitian only
// to a few ulps, and unitary rotations cannot remove that anti-Hermitian residue) is <= tol/sqrt(2) x the
// Frobenius norm, or after 12 sweeps. A pivot below 1e-20 of its diagonal scale is left alone: rotating it
// would only shrink it further, towards the subnormal range where its phase can no longer be normalised
// and the rotation stops being unitary. rot: 16 entries of LDS.
__device__ void wave_jacobi(int n, c64 *M, int ldm, c64 *Q, int ldq, int qrows, JacobiRot *rot, int lane,
                            double tol) {
    const int nn = n * n;
    const int N = n + (n & 1);
    const int npairs = N / 2;
    for (int sweep = 0; sweep < 12; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int e = lane; e < nn; e += 64) {
            const c64 me = M[e % n + ldm * (e / n)];
            const double a2 = me.re * me.re + me.im * me.im;
            tot += a2;
            if (e % n < e / n) off += 2.0 * a2;
        }
        off = wave_sum(off);
        tot = wave_sum(tot);
        if (off <= tol * tol * tot) break;
        for (int rd = 0; rd < N - 1; rd++) {
            if (lane < npairs) {
                int a, c;
                if (lane == 0) {
                    a = N - 1;
                    c = rd;
                } else {
                    a = (rd + lane) % (N - 1);
                    c = (rd - lane + (N - 1)) % (N - 1);
                }
                JacobiRot g;
                g.p = min(a, c);
                g.q = max(a, c);
                g.gpp = c64{1.0, 0.0};
                g.gpq = c64{0.0, 0.0};
                g.gqp = c64{0.0, 0.0};
                g.gqq = c64{1.0, 0.0};
                if (g.q < n) {
                    const c64 mpq = M[g.p + ldm * g.q];
                    const double amp = hypot(mpq.re, mpq.im);
                    const double scale = fabs(M[g.p + ldm * g.p].re) + fabs(M[g.q + ldm * g.q].re);
                    if (amp > 1e-20 * scale && amp > 1e-280) {
                        const double er = mpq.re / amp, ei = mpq.im / amp;
                        const double tau = (M[g.q + ldm * g.q].re - M[g.p + ldm * g.p].re) / (2.0 * amp);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                        g.gpp = c64{cs, 0.0};
                        g.gpq = c64{sn, 0.0};
                        g.gqp = c64{-sn * er, sn * ei};
                        g.gqq = c64{cs * er, -cs * ei};
                    }
                } else {
                    g.q = g.p;  // dummy partner of an odd n: identity on a single index
                }
                rot[lane] = g;
            }
            wave_lds_sync();
            if (lane < n || lane < qrows) {  // M <- M G, Q <- Q G (lane = row)
                for (int k = 0; k < npairs; k++) {
                    const JacobiRot g = rot[k];
                    if (g.p == g.q) continue;
                    if (lane < n) {
                        const c64 mp = M[lane + ldm * g.p], mq = M[lane + ldm * g.q];
                        M[lane + ldm * g.p] = cadd(cmul(mp, g.gpp), cmul(mq, g.gqp));
                        M[lane + ldm * g.q] = cadd(cmul(mp, g.gpq), cmul(mq, g.gqq));
                    }
                    if (lane < qrows) {
                        const c64 qp = Q[lane + ldq * g.p], qq = Q[lane + ldq * g.q];
                        Q[lane + ldq * g.p] = cadd(cmul(qp, g.gpp), cmul(qq, g.gqp));
                        Q[lane + ldq * g.q] = cadd(cmul(qp, g.gpq), cmul(qq, g.gqq));
                    }
                }
            }
            wave_lds_sync();
            if (lane < n) {  // M <- G^H M (lane = column)
                for (int k = 0; k < npairs; k++) {
                    const JacobiRot g = rot[k];
                    if (g.p == g.q) continue;
                    const c64 mp = M[g.p + ldm * lane], mq = M[g.q + ldm * lane];
                    M[g.p + ldm * lane] = cadd(cmulc(g.gpp, mp), cmulc(g.gqp, mq));
                    M[g.q + ldm * lane] = cadd(cmulc(g.gpq, mp), cmulc(g.gqq, mq));
                }
            }
            wave_lds_sync();
        }
    }
}

__global__ __launch_bounds__(256) void heev_refine_kernel(int n, int64_t nb, const c64 *__restrict__ H,
                                                          c32 *__restrict__ U, double *__restrict__ E64,
                                                          double tol) {
    extern __shared__ double lds64[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (b >= nb) return;
    const int nn = n * n, ld = refine_ld(n), nl = n * ld;
    c64 *Hs = reinterpret_cast<c64 *>(lds64 + (int64_t)wave * refine_lds_doubles(n));
    c64 *Q = Hs + nl;
    c64 *W = Q + nl;
    c64 *M = W + nl;
    JacobiRot *rot = reinterpret_cast<JacobiRot *>(M + nl);  // 16 x 72 bytes
    int *rank_s = reinterpret_cast<int *>(rot + 16);           // n ints
    for (int e = lane; e < nn; e += 64) {
        const int i = e % n, j = e / n;
        c64 h = (i <= j) ? H[b * nn + i + (int64_t)n * j] : H[b * nn + j + (int64_t)n * i];
        if (i > j) h.im = -h.im;
        if (i == j) h.im = 0.0;
        Hs[i + ld * j] = h;
        const c32 u = U[b * nn + e];
        Q[i + ld * j] = c64{u.re, u.im};
    }
    wave_lds_sync();
    // 1. two Loewdin steps: S = Q^H Q (in W), M = 3/2 Q - 1/2 Q S, Q <- M
    for (int it = 0; it < 2; it++) {
        lds_gemm_cn(n, ld, Q, Q, W, lane);
        wave_lds_sync();
        lds_gemm_nn(n, ld, Q, W, M, -0.5, 1.5, Q, lane);
        wave_lds_sync();
        for (int e = lane; e < nn; e += 64) Q[e % n + ld * (e / n)] = M[e % n + ld * (e / n)];
        wave_lds_sync();
    }
    // 2. M = Q^H (H Q)
    lds_gemm_nn(n, ld, Hs, Q, W, 1.0, 0.0, nullptr, lane);
    wave_lds_sync();
    lds_gemm_cn(n, ld, Q, W, M, lane);
    wave_lds_sync();
    // 3. parallel cyclic Jacobi
    wave_jacobi(n, M, ld, Q, ld, n, rot, lane, tol);
    // 4. ascending order
    double d = 0.0;
    if (lane < n) d = M[lane + ld * lane].re;
    int rank = 0;
    for (int k = 0; k < n; k++) {
        const double dk = rdl64(d, k);
        rank += (dk < d || (dk == d && k < lane)) ? 1 : 0;
    }
    if (lane < n) {
        E64[b * n + rank] = d;
        rank_s[lane] = rank;
    }
    wave_lds_sync();
    for (int e = lane; e < nn; e += 64) {
        const int r = e % n, j = e / n;
        const c64 q = Q[r + ld * j];
        U[b * nn + r + (int64_t)n * rank_s[j]] = c32{static_cast<float>(q.re), static_cast<float>(q.im)};
    }
}


// ---------------------------------------------------------------------------------------------
// Degenerate-band gauge fixing of the reference's per-k eigensolve (solve_eigen_el!,
// src/diagonalize.jl:47-85), one wave per matrix, after the eigenpairs are known. Bands are grouped as
// the reference groups them: a group starts at band a and runs while |E_i - E_a| <= cutoff. In each group
// of m > 1 bands the fixed Hermitian perturbation P (src/diagonalize.jl:11-18; the matrix of EPW's
// hamwan2bloch) is projected, M = U_g^H P U_g, M is diagonalised by the FP64 Jacobi and U_g <- U_g V with
// the columns in ascending order of M's eigenvalues. The gauge inside each group is then fixed by H and P
// alone (up to one phase per column), whatever gauge the solver returned. Matrices without a pair closer
// than the cutoff are not touched. LDS (one wave per workgroup): U_g | W = P U_g | M (n*n c64 each),
// rotation table, ranks.
// ---------------------------------------------------------------------------------------------
__host__ __device__ constexpr int gauge_lds_bytes(int n) { return 3 * n * n * 16 + 16 * (int)sizeof(JacobiRot) + 4 * n; }

__device__ __forceinline__ c64 degeneracy_perturbation(int r, int c) {  // P[r, c], 0-based row r, column c
    return c64{0.25644832 + 0.005 * (double)(r + c + 2), 0.0125 * (double)(r - c)};
}

__device__ void gauge_fix_group(int n, int a, int m, c32 *__restrict__ Ub, c64 *Ug, c64 *W, c64 *M,
                                JacobiRot *rot, int *rank_s, int lane) {
    for (int t = lane; t < n * m; t += 64) {
        const c32 u = Ub[t % n + n * (a + t / n)];
        Ug[t] = c64{u.re, u.im};
    }
    wave_lds_sync();
    for (int t = lane; t < n * m; t += 64) {  // W = P U_g
        const int r = t % n, c = t / n;
        c64 acc{0.0, 0.0};
        for (int k = 0; k < n; k++) acc = cadd(acc, cmul(degeneracy_perturbation(r, k), Ug[k + n * c]));
        W[t] = acc;
    }
    wave_lds_sync();
    for (int t = lane; t < m * m; t += 64) {  // M = U_g^H W, upper triangle, mirrored: exactly Hermitian
        const int i = t % m, j = t / m;
        if (i > j) continue;
        c64 acc{0.0, 0.0};
        for (int r = 0; r < n; r++) acc = cadd(acc, cmulc(Ug[r + n * i], W[r + n * j]));
        if (i == j) acc.im = 0.0;
        M[t] = acc;
        M[j + m * i] = c64{acc.re, -acc.im};
    }
    wave_lds_sync();
    wave_jacobi(m, M, m, Ug, n, n, rot, lane, 1e-15);
    const double d = lane < m ? M[lane + m * lane].re : 0.0;
    int rank = 0;
    for (int k = 0; k < m; k++) {
        const double dk = rdl64(d, k);
        rank += (dk < d || (dk == d && k < lane)) ? 1 : 0;
    }
    if (lane < m) rank_s[lane] = rank;
    wave_lds_sync();
    for (int t = lane; t < n * m; t += 64) {
        const int r = t % n, c = t / n;
        Ub[r + n * (a + rank_s[c])] = c32{static_cast<float>(Ug[t].re), static_cast<float>(Ug[t].im)};
    }
    wave_lds_sync();
}

__global__ __launch_bounds__(64) void degenerate_gauge_kernel(int n, int64_t nb, const double *__restrict__ E64,
                                                              c32 *__restrict__ U, double cutoff) {
    extern __shared__ double lds64[];
    const int lane = threadIdx.x;
    const int64_t b = blockIdx.x;
    const double e = lane < n ? E64[b * n + lane] : 0.0;
    bool any = false;  // is_degenerate: some neighbouring pair closer than the cutoff
    for (int i = 0; i + 1 < n && !any; i++) any = fabs(rdl64(e, i + 1) - rdl64(e, i)) < cutoff;
    if (!any) return;
    c64 *Ug = reinterpret_cast<c64 *>(lds64);
    c64 *W = Ug + n * n;
    c64 *M = W + n * n;
    JacobiRot *rot = reinterpret_cast<JacobiRot *>(M + n * n);
    int *rank_s = reinterpret_cast<int *>(rot + 16);
    c32 *Ub = U + b * n * n;
    double first = rdl64(e, 0);
    int a = 0;
    for (int i = 0; i + 1 < n; i++) {  // wave-uniform walk over the ascending eigenvalues
        const double next = rdl64(e, i + 1);
        int last;
        if (fabs(first - next) > cutoff)
            last = i;
        else if (i == n - 2)
            last = i + 1;
        else
            continue;
        if (last > a) gauge_fix_group(n, a, last - a + 1, Ub, Ug, W, M, rot, rank_s, lane);
        first = next;
        a = i + 1;
    }
}

}  // namespace ephip

using namespace ephip;

// Waves (one matrix each) per workgroup: up to 4, keeping the workgroup's LDS within 64 KiB.
static int waves_per_block(size_t lds_per_wave) {
    int w = (int)((64 * 1024) / lds_per_wave);
    return w < 1 ? 1 : (w > 4 ? 4 : w);
}

static int heev_launch_check(int n, int64_t nb) {
    EPHIP_REQUIRE(n >= 1 && n <= 32, "heev: n=%d outside 1..32 (one wavefront per matrix)", n);
    EPHIP_REQUIRE(nb >= 0, "heev: negative batch");
    EPHIP_REQUIRE((nb + 3) / 4 <= 0x7fffffffLL, "heev: batch too large");
    return EPHIP_OK;
}

extern "C" int ephip_heev_c32(int32_t n, int64_t nb, const void *H, int32_t h_is_c64, void *U, float *E,
                              int32_t *info, void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(H && E, "heev: null device pointer");
    const int wpb = waves_per_block(sizeof(float) * heev_lds_floats(n));
    const size_t lds = wpb * sizeof(float) * heev_lds_floats(n);
    const unsigned grid = (unsigned)((nb + wpb - 1) / wpb);
    hipStream_t s = as_stream(stream);
    if (h_is_c64)
        heev_c32_kernel<c64><<<grid, 64 * wpb, lds, s>>>(n, nb, static_cast<const c64 *>(H), static_cast<c32 *>(U), E, info);
    else
        heev_c32_kernel<c32><<<grid, 64 * wpb, lds, s>>>(n, nb, static_cast<const c32 *>(H), static_cast<c32 *>(U), E, info);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_heev_refine(int32_t n, int64_t nb, const void *H64, void *U, const float *E32, double *E64,
                                 double tol, void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(H64 && U && E64, "heev_refine: null device pointer");
    EPHIP_REQUIRE(tol >= 0.0, "heev_refine: negative tolerance");
    (void)E32;
    const int wpb = waves_per_block(sizeof(double) * refine_lds_doubles(n));
    const size_t lds = wpb * sizeof(double) * refine_lds_doubles(n);
    if (lds > 64 * 1024) {  // n = 32: one matrix's FP64 working set (67 KiB) in a one-wave block; gfx950 has
                            // 160 KiB of LDS per workgroup
        EPHIP_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(heev_refine_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    heev_refine_kernel<<<(unsigned)((nb + wpb - 1) / wpb), 64 * wpb, lds, as_stream(stream)>>>(
        n, nb, static_cast<const c64 *>(H64), static_cast<c32 *>(U), E64, tol > 0.0 ? tol : 1e-15);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_degenerate_gauge_fix(int32_t n, int64_t nb, const double *E64, void *U, double cutoff,
                                          void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    EPHIP_REQUIRE(cutoff >= 0.0, "degenerate_gauge_fix: negative cutoff");
    EPHIP_REQUIRE(nb <= 0x7fffffffLL, "degenerate_gauge_fix: batch too large");
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(E64 && U, "degenerate_gauge_fix: null device pointer");
    degenerate_gauge_kernel<<<(unsigned)nb, 64, gauge_lds_bytes(n), as_stream(stream)>>>(
        n, nb, E64, static_cast<c32 *>(U), cutoff);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

// eigen_batched / eigvals_batched with the FP64 accuracy of the reference's zheev: the FP32 wave solve
// followed by the FP64 refinement, as one entry each. Per matrix, eigvals_batched keeps the FP32
// eigenvectors (8 n^2 bytes) and eigenvalues (4 n bytes) in the caller's workspace: the refinement needs
// the vectors even when only values are returned.
static int64_t eigvals_work_bytes_per_matrix(int n) { return 8LL * n * n + 4LL * n; }

static int eigen_refined_run(int32_t n, int64_t nb, const void *H64, double *E64, void *U, float *E32, int32_t *info,
                             void *stream) {
    int st = ephip_heev_c32(n, nb, H64, 1, U, E32, info, stream);
    if (st) return st;
    return ephip_heev_refine(n, nb, H64, U, E32, E64, 0.0, stream);
}

extern "C" int ephip_eigen_batched(int32_t n, int64_t nb, const void *H64, double *E64, void *U, float *work,
                                   int32_t *info, void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(H64 && E64 && U && work, "eigen_batched: null device pointer");
    return eigen_refined_run(n, nb, H64, E64, U, work, info, stream);
}

extern "C" int ephip_eigvals_batched(int32_t n, int64_t nb, const void *H64, double *E64, void *work, int32_t *info,
                                     void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(H64 && E64 && work, "eigvals_batched: null device pointer");
    char *w = static_cast<char *>(work);
    return eigen_refined_run(n, nb, H64, E64, w, reinterpret_cast<float *>(w + 8LL * n * n * nb), info, stream);
}

// ---------------------------------------------------------------------------------------------
// Allocation-free eigenvalue entry: the workspace of ephip_eigvals_batched (FP32 eigenvectors, then FP32
// eigenvalues) lives in a file-scope device array instead of a caller buffer. Each GPU the library runs
// on holds its own copy until the process ends; its address on a device is looked up once, with that
// device current, and kept. Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------------------------
namespace ephip {

constexpr size_t EIG_SCRATCH_BYTES = size_t(16) << 20;  // 16 MiB per GPU
constexpr int EIG_SCRATCH_DEVICES = 64;

static __device__ __align__(256) char scratch[EIG_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[EIG_SCRATCH_DEVICES];

// Device address of this GPU's scratch array (the host-side name is not a device address).
static int scratch_address(void **out) {
    int dev = -1;
    EPHIP_TRY_HIP(hipGetDevice(&dev));
    EPHIP_REQUIRE(dev >= 0 && dev < EIG_SCRATCH_DEVICES, "eigvals_batched_noalloc: device %d outside 0..%d", dev,
                  EIG_SCRATCH_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        EPHIP_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        EPHIP_REQUIRE(p != nullptr, "eigvals_batched_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return EPHIP_OK;
}

}  // namespace ephip

extern "C" int ephip_eigvals_batched_noalloc(int32_t n, int64_t nb, const void *H64, double *E64, int32_t *info,
                                             void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    const int64_t per = eigvals_work_bytes_per_matrix(n);
    const int64_t cap = (int64_t)(EIG_SCRATCH_BYTES / (size_t)per);
    EPHIP_REQUIRE(nb <= cap,
                  "eigvals_batched_noalloc: %lld matrices of n=%d need %lld bytes of workspace; the %zu-byte scratch "
                  "holds %lld (call in chunks, or use ephip_eigvals_batched with a caller workspace)",
                  (long long)nb, n, (long long)(nb * per), EIG_SCRATCH_BYTES, (long long)cap);
    EPHIP_REQUIRE(H64 && E64, "eigvals_batched_noalloc: null device pointer");
    void *w = nullptr;
    st = scratch_address(&w);
    if (st) return st;
    char *wc = static_cast<char *>(w);
    // every entry of the workspace is written by the FP32 solve before the refinement reads it: no clearing
    return eigen_refined_run(n, nb, H64, E64, wc, reinterpret_cast<float *>(wc + 8LL * n * n * nb), info, stream);
}
