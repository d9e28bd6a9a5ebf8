// Batched Hermitian eigensolve for small matrices (n <= 32), one wavefront per matrix.
//
// Replaces the reference's eigen_batched / eigvals_batched (src/wannier_to_bloch_batched.jl:30-65,
// CUDA method ext/ElectronPhononCUDAExt.jl:80-115): ascending eigenvalues, orthonormal eigenvectors,
// no degeneracy gauge fixing (the reference's batched path applies none either).
//
// Method (FP32, all in LDS + registers of one wave):
//   1. Householder reduction of the Hermitian matrix to tridiagonal form, one reflector per column;
//      the reflector vectors are kept in the zeroed part of the matrix for the back-transform.
//   2. Diagonal phase scaling turns the complex tridiagonal into a real symmetric one.
//   3. Implicit QL with Wilkinson-type shifts on (d, e): the scalar recurrence is uniform across the
//      wave (lane i holds d_i, e_i; reads go through readlane), the rotations are applied to the rows
//      of Z in parallel (lane = row).
//   4. Eigenvectors V = H_0 ... H_{n-3} D Z, one column per lane.
// A second kernel refines in FP64 against the FP64 matrix: Rayleigh quotients for every vector and a
// Rayleigh-Ritz step (orthonormalise + project + Jacobi) for clusters whose eigenvalue gaps are below
// a tolerance, which fixes near-degenerate eigenvalues and re-orthogonalises their vectors.
#include "common.hpp"

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
// FP64 refinement. One wave per matrix. LDS: H (n*n c64) | V (n*n c64) | W (n*n c64) | M (n*n c64).
// ---------------------------------------------------------------------------------------------
__host__ __device__ constexpr int refine_lds_doubles(int n) { return 8 * n * n + 2 * n; }

__device__ __forceinline__ double rdl64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Cyclic complex Jacobi on the s x s Hermitian M (column-major, ld s) run by one lane; X accumulates
// the rotations (s x s). Eigenvalues end on the diagonal of M.
__device__ void jacobi_herm(int s, c64 *M, c64 *X) {
    for (int i = 0; i < s * s; i++) X[i] = c64{(i % s == i / s) ? 1.0 : 0.0, 0.0};
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int c = 0; c < s; c++)
            for (int r = 0; r < s; r++) {
                const double a2 = M[r + s * c].re * M[r + s * c].re + M[r + s * c].im * M[r + s * c].im;
                tot += a2;
                if (r != c) off += a2;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < s - 1; p++)
            for (int q = p + 1; q < s; q++) {
                const c64 mpq = M[p + s * q];
                const double amp = sqrt(mpq.re * mpq.re + mpq.im * mpq.im);
                if (amp == 0.0) continue;
                const double er = mpq.re / amp, ei = mpq.im / amp;  // e^{i phi}
                const double tau = (M[q + s * q].re - M[p + s * p].re) / (2.0 * amp);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                // G = diag(1, e^{-i phi}) [[cs, sn], [-sn, cs]]
                const c64 gpp{cs, 0.0}, gpq{sn, 0.0};
                const c64 gqp{-sn * er, sn * ei}, gqq{cs * er, -cs * ei};
                auto mul = [](c64 a, c64 b) { return c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; };
                auto cmul = [](c64 a, c64 b) { return c64{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; };
                for (int r = 0; r < s; r++) {  // M <- M G, X <- X G
                    const c64 mp = M[r + s * p], mq = M[r + s * q];
                    const c64 a = mul(mp, gpp), bq = mul(mq, gqp), c2 = mul(mp, gpq), d2 = mul(mq, gqq);
                    M[r + s * p] = c64{a.re + bq.re, a.im + bq.im};
                    M[r + s * q] = c64{c2.re + d2.re, c2.im + d2.im};
                    const c64 xp = X[r + s * p], xq = X[r + s * q];
                    const c64 xa = mul(xp, gpp), xb = mul(xq, gqp), xc = mul(xp, gpq), xd = mul(xq, gqq);
                    X[r + s * p] = c64{xa.re + xb.re, xa.im + xb.im};
                    X[r + s * q] = c64{xc.re + xd.re, xc.im + xd.im};
                }
                for (int c = 0; c < s; c++) {  // M <- G^H M
                    const c64 mp = M[p + s * c], mq = M[q + s * c];
                    const c64 a = cmul(gpp, mp), bq = cmul(gqp, mq), c2 = cmul(gpq, mp), d2 = cmul(gqq, mq);
                    M[p + s * c] = c64{a.re + bq.re, a.im + bq.im};
                    M[q + s * c] = c64{c2.re + d2.re, c2.im + d2.im};
                }
                M[p + s * q] = c64{0.0, 0.0};
                M[q + s * p] = c64{0.0, 0.0};
            }
    }
}

__global__ __launch_bounds__(256) void heev_refine_kernel(int n, int64_t nb, const c64 *__restrict__ H,
                                                          c32 *__restrict__ U, const float *__restrict__ E32,
                                                          double *__restrict__ E64, double tol) {
    extern __shared__ double lds64[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (b >= nb) return;
    const int nn = n * n;
    c64 *Hs = reinterpret_cast<c64 *>(lds64 + (int64_t)wave * refine_lds_doubles(n));
    c64 *V = Hs + nn;
    c64 *M = V + nn;
    c64 *X = M + nn;
    double *lam_s = reinterpret_cast<double *>(X + nn);
    for (int e = lane; e < nn; e += 64) {
        const int i = e % n, j = e / n;
        c64 h = (i <= j) ? H[b * nn + i + (int64_t)n * j] : H[b * nn + j + (int64_t)n * i];
        if (i > j) h.im = -h.im;
        if (i == j) h.im = 0.0;
        Hs[e] = h;
        const c32 u = U[b * nn + e];
        V[e] = c64{u.re, u.im};
    }
    wave_lds_sync();
    // Rayleigh quotient of column `lane`
    double lam = 0.0;
    if (lane < n) {
        double num = 0.0, den = 0.0;
        for (int r = 0; r < n; r++) {
            double wr = 0.0, wi = 0.0;
            for (int k = 0; k < n; k++) {
                const c64 h = Hs[r + n * k], u = V[k + n * lane];
                wr += h.re * u.re - h.im * u.im;
                wi += h.re * u.im + h.im * u.re;
            }
            const c64 ur = V[r + n * lane];
            num += ur.re * wr + ur.im * wi;
            den += ur.re * ur.re + ur.im * ur.im;
        }
        lam = num / den;
    }
    // clusters of consecutive eigenvalues closer than tol (in the FP32 ascending order)
    int start = 0;
    for (int c = 1; c <= n; c++) {
        const bool brk = (c == n) || (rdl64(lam, c) - rdl64(lam, c - 1) > tol);
        if (!brk) continue;
        const int s = c - start;
        if (s >= 2) {
            // modified Gram-Schmidt on V[:, start:c] (lanes over rows)
            for (int a = start; a < c; a++) {
                c64 va{0.0, 0.0};
                if (lane < n) va = V[lane + n * a];
                for (int a2 = start; a2 < a; a2++) {
                    c64 vb{0.0, 0.0};
                    if (lane < n) vb = V[lane + n * a2];
                    const double pr = wave_sum(vb.re * va.re + vb.im * va.im);
                    const double pi = wave_sum(vb.re * va.im - vb.im * va.re);
                    va.re -= pr * vb.re - pi * vb.im;
                    va.im -= pr * vb.im + pi * vb.re;
                }
                const double nrm = sqrt(wave_sum(va.re * va.re + va.im * va.im));
                if (lane < n) V[lane + n * a] = c64{va.re / nrm, va.im / nrm};
                wave_lds_sync();
            }
            // M = V_c^H H V_c (s x s): lane handles one entry at a time
            for (int e = lane; e < s * s; e += 64) {
                const int p = start + e % s, q = start + e / s;
                double mr = 0.0, mi = 0.0;
                for (int r = 0; r < n; r++) {
                    double wr = 0.0, wi = 0.0;
                    for (int k = 0; k < n; k++) {
                        const c64 h = Hs[r + n * k], u = V[k + n * q];
                        wr += h.re * u.re - h.im * u.im;
                        wi += h.re * u.im + h.im * u.re;
                    }
                    const c64 up = V[r + n * p];
                    mr += up.re * wr + up.im * wi;
                    mi += up.re * wi - up.im * wr;
                }
                M[e] = c64{mr, mi};
            }
            wave_lds_sync();
            if (lane == 0) {
                jacobi_herm(s, M, X);
                // ascending order inside the cluster
                for (int i = 0; i < s; i++) lam_s[i] = M[i + s * i].re;
                for (int i = 0; i < s - 1; i++) {
                    int km = i;
                    for (int k = i + 1; k < s; k++)
                        if (lam_s[k] < lam_s[km]) km = k;
                    if (km != i) {
                        const double t = lam_s[i];
                        lam_s[i] = lam_s[km];
                        lam_s[km] = t;
                        for (int r = 0; r < s; r++) {
                            const c64 tx = X[r + s * i];
                            X[r + s * i] = X[r + s * km];
                            X[r + s * km] = tx;
                        }
                    }
                }
            }
            wave_lds_sync();
            // V_c <- V_c X ; lanes over rows
            if (lane < n) {
                c64 row[32];
                for (int a = 0; a < s; a++) row[a] = V[lane + n * (start + a)];
                for (int a = 0; a < s; a++) {
                    double vr = 0.0, vi = 0.0;
                    for (int a2 = 0; a2 < s; a2++) {
                        const c64 x = X[a2 + s * a];
                        vr += row[a2].re * x.re - row[a2].im * x.im;
                        vi += row[a2].re * x.im + row[a2].im * x.re;
                    }
                    V[lane + n * (start + a)] = c64{vr, vi};
                }
            }
            if (lane >= start && lane < c) lam = lam_s[lane - start];
            wave_lds_sync();
        }
        start = c;
    }
    if (lane < n) E64[b * n + lane] = lam;
    for (int e = lane; e < nn; e += 64) U[b * nn + e] = c32{static_cast<float>(V[e].re), static_cast<float>(V[e].im)};
    (void)E32;
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
                                 double cluster_tol, void *stream) {
    int st = heev_launch_check(n, nb);
    if (st) return st;
    if (nb == 0) return EPHIP_OK;
    EPHIP_REQUIRE(n <= 31, "heev_refine: n=%d > 31 (the FP64 working set of one matrix exceeds 64 KiB of LDS)", n);
    EPHIP_REQUIRE(H64 && U && E64, "heev_refine: null device pointer");
    EPHIP_REQUIRE(cluster_tol >= 0.0, "heev_refine: negative cluster tolerance");
    const int wpb = waves_per_block(sizeof(double) * refine_lds_doubles(n));
    const size_t lds = wpb * sizeof(double) * refine_lds_doubles(n);
    heev_refine_kernel<<<(unsigned)((nb + wpb - 1) / wpb), 64 * wpb, lds, as_stream(stream)>>>(
        n, nb, static_cast<const c64 *>(H64), static_cast<c32 *>(U), E32, E64, cluster_tol);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
