// Energy-window filter, phonon post-processing, k-point index maps and the e-ph rotations.
// These are bandwidth-bound: one thread (or one wave) per point, coalesced along the fastest index.
#include <algorithm>

#include "common.hpp"

namespace ephip {

struct c32 {
    float re, im;
};
struct c64 {
    double re, im;
};

// inside_window + _filter_kpoints (src/filter.jl:19, :40-100): E is (nw, nk) ascending per k.
// Band indices are 1-based; a k with no in-window band gets keep = 0 and bmin = bmax = 0.
__global__ void window_filter_kernel(int nw, int64_t nk, const double *__restrict__ E, double wmin, double wmax,
                                     uint8_t *__restrict__ keep, int64_t *__restrict__ bmin,
                                     int64_t *__restrict__ bmax, int64_t *__restrict__ nbelow) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double *e = E + (int64_t)nw * k;
    int first = nw + 1, last = 0;
    for (int i = 0; i < nw; i++)
        if (e[i] >= wmin) {
            first = i + 1;
            break;
        }
    for (int i = nw - 1; i >= 0; i--)
        if (e[i] <= wmax) {
            last = i + 1;
            break;
        }
    const bool in = first <= last;
    keep[k] = in ? 1 : 0;
    bmin[k] = in ? first : 0;
    bmax[k] = in ? last : 0;
    if (nbelow) nbelow[k] = first - 1;
}

// D(q)[i, j] /= sqrt(m_i); /= sqrt(m_j)  (src/compute_states.jl:295-297, src/wannier_to_bloch.jl:199-202)
__global__ void ph_mass_scale_kernel(int nm, int64_t nq, c64 *__restrict__ D, const double *__restrict__ mass) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nm * nm * nq) return;
    const int i = (int)(e % nm), j = (int)((e / nm) % nm);
    c64 v = D[e];
    const double si = sqrt(mass[i]), sj = sqrt(mass[j]);
    v.re = v.re / si / sj;
    v.im = v.im / si / sj;
    D[e] = v;
}

// omega = sign(w2) sqrt(|w2|); u[i, :] /= sqrt(m_i)  (src/diagonalize.jl:118-126, src/compute_states.jl:299-300)
__global__ void ph_finalize_kernel(int nm, int64_t nq, const double *__restrict__ w2, double *__restrict__ omega,
                                   c32 *__restrict__ U, const double *__restrict__ mass) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (int64_t)nm * nq) {
        const double w = w2[e];
        omega[e] = w >= 0 ? sqrt(w) : -sqrt(-w);
    }
    if (U && e < (int64_t)nm * nm * nq) {
        const int i = (int)(e % nm);
        const float s = static_cast<float>(sqrt(mass[i]));
        c32 u = U[e];
        u.re /= s;
        u.im /= s;
        U[e] = u;
    }
}

// _grid_coords_reduced + _hash_xk + dense-table _ik_from_hash (src/common/kpoints.jl:614-656):
// c_d = mod(round(Int, (x_d + dx_d - s_d) n_d), n_d) (ties to even), hash = (c1 n2 + c2) n3 + c3,
// idx = table[hash] - 1 (table holds 1-based indices, 0 = not a point) or hash itself without a table.
__global__ void kq_index_kernel(int64_t nk, const double *__restrict__ xk, double dx0, double dx1, double dx2,
                                int n1, int n2, int n3, double s0, double s1, double s2,
                                const int64_t *__restrict__ table, int64_t *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    auto red = [](double x, double dx, double s, int n) {
        int64_t c = (int64_t)rint((x + dx - s) * n) % n;
        return c < 0 ? c + n : c;
    };
    const int64_t c1 = red(xk[3 * k], dx0, s0, n1), c2 = red(xk[3 * k + 1], dx1, s1, n2),
                  c3 = red(xk[3 * k + 2], dx2, s2, n3);
    const int64_t h = (c1 * n2 + c2) * n3 + c3;
    out[k] = table ? table[h] - 1 : h;
}

// RR->Rq phonon-mode rotation (src/wannier_to_bloch.jl:292-297) for a batch of q:
// out[b, nu', Re, q] = sum_nu g[b, nu, Re, q] uph[nu, nu', q]. One thread per (b, Re, q).
__global__ void rotate_modes_kernel(int64_t nbas, int nm, int nr_el, int nq, const c32 *__restrict__ g, int64_t ldg,
                                    const c32 *__restrict__ uph, c32 *__restrict__ out, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_q = nbas * nr_el;
    if (e >= per_q * nq) return;
    const int64_t q = e / per_q, rem = e % per_q;
    const int64_t b = rem % nbas, re = rem / nbas;
    const c32 *gi = g + ldg * q + b + nbas * nm * re;
    c32 *go = out + ldo * q + b + nbas * nm * re;
    const c32 *u = uph + (int64_t)nm * nm * q;
    c32 in[48];
    for (int nu = 0; nu < nm; nu++) in[nu] = gi[nbas * nu];
    for (int nu2 = 0; nu2 < nm; nu2++) {
        float sr = 0.f, si = 0.f;
        for (int nu = 0; nu < nm; nu++) {
            const c32 a = in[nu], w = u[nu + nm * nu2];
            sr += a.re * w.re - a.im * w.im;
            si += a.re * w.im + a.im * w.re;
        }
        go[nbas * nu2] = c32{sr, si};
    }
}

// eph_apply_rotations_rqkq! + g2 fold (src/wannier_to_bloch_batched.jl:360-377, :451-453;
// CUDA kernel ext/ElectronPhononCUDAExt.jl:218-239) restricted to the energy window, as the outer-q
// loop's column masks make it (src/calculator/run_eph_over_q_and_k.jl:555-585):
//   ep[m, n, nu] = sum_ij conj(ukq[i, m]) g[i, j, nu] uk[j, n],  m, n in the in-window bands,
//   g2 = |ep|^2 / (2 omega_nu).
// One wave per batch column t; output slots [m - mlo, n - nlo, nu] in an (nbw, nbw, nm) block,
// zero outside the window.
__global__ __launch_bounds__(256) void rotate_rqkq_g2_kernel(
    int nw, int nm, int64_t ncol, const c32 *__restrict__ g, int64_t ldg, const int64_t *__restrict__ ik,
    const int64_t *__restrict__ ikq, const c32 *__restrict__ U, const int64_t *__restrict__ bmin,
    const int64_t *__restrict__ bmax, const double *__restrict__ omega, int nbw, float *__restrict__ g2,
    c32 *__restrict__ ep_out) {
    extern __shared__ c32 tsh[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t t = (int64_t)blockIdx.x * 4 + wave;
    if (t >= ncol) return;
    c32 *T1 = tsh + (int64_t)wave * nw * nbw * nm;  // T1[i, n', nu]
    const int64_t k = ik[t], kq = ikq[t];
    const int nlo = (int)bmin[k] - 1, nbk = (int)(bmax[k] - bmin[k] + 1) * (bmin[k] > 0);
    const int mlo = (int)bmin[kq] - 1, nbkq = (int)(bmax[kq] - bmin[kq] + 1) * (bmin[kq] > 0);
    // Slots cover nbw bands; a wider window (a caller error) is truncated, never written past LDS.
    const int nbk_s = min(nbk, nbw), nbkq_s = min(nbkq, nbw);
    const c32 *gt = g + ldg * t;
    const c32 *uk = U + (int64_t)nw * nw * k;
    const c32 *ukq = U + (int64_t)nw * nw * kq;
    if (nbk_s > 0 && nbkq_s > 0) {
        for (int p = lane; p < nw * nm; p += 64) {
            const int i = p % nw, nu = p / nw;
            for (int n = 0; n < nbk_s; n++) {
                float sr = 0.f, si = 0.f;
                for (int j = 0; j < nw; j++) {
                    const c32 a = gt[i + (int64_t)nw * (j + (int64_t)nw * nu)];
                    const c32 u = uk[j + (int64_t)nw * (nlo + n)];
                    sr += a.re * u.re - a.im * u.im;
                    si += a.re * u.im + a.im * u.re;
                }
                T1[i + nw * (n + nbw * nu)] = c32{sr, si};
            }
        }
    }
    wave_lds_sync();
    const int nslot = nbw * nbw * nm;
    for (int p = lane; p < nslot; p += 64) {
        const int ms = p % nbw, ns = (p / nbw) % nbw, nu = p / (nbw * nbw);
        float er = 0.f, ei = 0.f;
        if (ms < nbkq_s && ns < nbk_s) {
            for (int i = 0; i < nw; i++) {
                const c32 u = ukq[i + (int64_t)nw * (mlo + ms)];
                const c32 a = T1[i + nw * (ns + nbw * nu)];
                er += u.re * a.re + u.im * a.im;
                ei += u.re * a.im - u.im * a.re;
            }
        }
        const float w = static_cast<float>(omega[nu]);
        if (g2) g2[t * nslot + p] = (er * er + ei * ei) / (2.f * w);
        if (ep_out) ep_out[t * nslot + p] = c32{er, ei};
    }
}

}  // namespace ephip

using namespace ephip;

static inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int ephip_window_filter(int32_t nw, int64_t nk, const double *E, double wmin, double wmax, uint8_t *keep,
                                   int64_t *bmin, int64_t *bmax, int64_t *nbelow, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nk >= 0, "window_filter: bad sizes nw=%d nk=%lld", nw, (long long)nk);
    if (nk == 0) return EPHIP_OK;
    EPHIP_REQUIRE(E && keep && bmin && bmax, "window_filter: null device pointer");
    window_filter_kernel<<<blocks_for(nk), 256, 0, as_stream(stream)>>>(nw, nk, E, wmin, wmax, keep, bmin, bmax,
                                                                         nbelow);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_ph_mass_scale(int32_t nm, int64_t nq, void *D, const double *mass, void *stream) {
    EPHIP_REQUIRE(nm >= 1 && nq >= 0, "ph_mass_scale: bad sizes");
    if (nq == 0) return EPHIP_OK;
    EPHIP_REQUIRE(D && mass, "ph_mass_scale: null device pointer");
    ph_mass_scale_kernel<<<blocks_for((int64_t)nm * nm * nq), 256, 0, as_stream(stream)>>>(
        nm, nq, static_cast<c64 *>(D), mass);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_ph_finalize(int32_t nm, int64_t nq, const double *w2, double *omega, void *U,
                                 const double *mass, void *stream) {
    EPHIP_REQUIRE(nm >= 1 && nq >= 0, "ph_finalize: bad sizes");
    if (nq == 0) return EPHIP_OK;
    EPHIP_REQUIRE(w2 && omega && mass, "ph_finalize: null device pointer");
    ph_finalize_kernel<<<blocks_for((int64_t)nm * nm * nq), 256, 0, as_stream(stream)>>>(
        nm, nq, w2, omega, static_cast<c32 *>(U), mass);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_kq_index(int64_t nk, const double *xk, const double *dx, const int32_t *ngrid,
                              const double *shift, const int64_t *table, int64_t *out, void *stream) {
    EPHIP_REQUIRE(nk >= 0, "kq_index: negative size");
    EPHIP_REQUIRE(ngrid && ngrid[0] > 0 && ngrid[1] > 0 && ngrid[2] > 0, "kq_index: bad grid");
    if (nk == 0) return EPHIP_OK;
    EPHIP_REQUIRE(xk && out, "kq_index: null device pointer");
    const double z[3] = {0.0, 0.0, 0.0};
    const double *d = dx ? dx : z, *s = shift ? shift : z;
    kq_index_kernel<<<blocks_for(nk), 256, 0, as_stream(stream)>>>(nk, xk, d[0], d[1], d[2], ngrid[0], ngrid[1],
                                                                   ngrid[2], s[0], s[1], s[2], table, out);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_rotate_modes(int64_t nbasis, int32_t nm, int32_t nr_el, int32_t nq, const void *g,
                                      int64_t ldg, const void *uph, void *out, int64_t ldout, void *stream) {
    EPHIP_REQUIRE(nbasis >= 1 && nm >= 1 && nm <= 48 && nr_el >= 1 && nq >= 0,
                  "eph_rotate_modes: bad sizes (nm must be <= 48)");
    EPHIP_REQUIRE(ldg >= nbasis * nm * nr_el && ldout >= nbasis * nm * nr_el, "eph_rotate_modes: bad ld");
    if (nq == 0) return EPHIP_OK;
    EPHIP_REQUIRE(g && uph && out && g != out, "eph_rotate_modes: null or aliased device pointer");
    rotate_modes_kernel<<<blocks_for(nbasis * nr_el * nq), 256, 0, as_stream(stream)>>>(
        nbasis, nm, nr_el, nq, static_cast<const c32 *>(g), ldg, static_cast<const c32 *>(uph),
        static_cast<c32 *>(out), ldout);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_rotate_rqkq_g2(int32_t nw, int32_t nm, int64_t ncol, const void *g, int64_t ldg,
                                        const int64_t *ik, const int64_t *ikq, const void *U, const int64_t *bmin,
                                        const int64_t *bmax, const double *omega, int32_t nbw, float *g2,
                                        void *ep, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nw <= 64 && nm >= 1 && ncol >= 0,
                  "eph_rotate_rqkq_g2: bad sizes nw=%d nm=%d ncol=%lld", nw, nm, (long long)ncol);
    EPHIP_REQUIRE(nbw >= 1 && nbw <= nw, "eph_rotate_rqkq_g2: nbw=%d must be in 1..nw=%d", nbw, nw);
    EPHIP_REQUIRE(ldg >= (int64_t)nw * nw * nm, "eph_rotate_rqkq_g2: ldg too small");
    const size_t lds = 4 * sizeof(c32) * (size_t)nw * nbw * nm;
    EPHIP_REQUIRE(lds <= 64 * 1024, "eph_rotate_rqkq_g2: nw*nbw*nm too large for LDS");
    if (ncol == 0) return EPHIP_OK;
    EPHIP_REQUIRE(g && ik && ikq && U && bmin && bmax && omega && (g2 || ep),
                  "eph_rotate_rqkq_g2: null device pointer");
    rotate_rqkq_g2_kernel<<<blocks_for(ncol, 4), 256, lds, as_stream(stream)>>>(
        nw, nm, ncol, static_cast<const c32 *>(g), ldg, ik, ikq, static_cast<const c32 *>(U), bmin, bmax, omega,
        nbw, g2, static_cast<c32 *>(ep));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

// ---------------------------------------------------------------------------------------------
// Generic device primitives of the backend method table: strided-batched complex GEMM (alpha 1,
// beta 0) and the window scatter of device-resident calculators.
// ---------------------------------------------------------------------------------------------
namespace ephip {

// C_b[i, j] = sum_p op(A_b)[i, p] op(B_b)[p, j], one thread per (b, i, j); op is the matrix ('N'), its
// transpose ('T') or its adjoint ('C'). Column-major operands; accumulation in the element type.
template <typename V, typename S>
__global__ void batched_gemm_kernel(int ta, int tb, int m, int n, int k, int64_t nbatch, const V *__restrict__ A,
                                    int64_t lda, int64_t sa, const V *__restrict__ B, int64_t ldb, int64_t sb,
                                    V *__restrict__ C, int64_t ldc, int64_t sc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t mn = (int64_t)m * n;
    if (e >= mn * nbatch) return;
    const int64_t b = e / mn;
    const int r = (int)(e % mn), i = r % m, j = r / m;
    const V *a = A + sa * b;
    const V *bb = B + sb * b;
    S re = 0, im = 0;
    for (int p = 0; p < k; p++) {
        V x = ta == 'N' ? a[i + lda * p] : a[p + lda * i];
        V y = tb == 'N' ? bb[p + ldb * j] : bb[j + ldb * p];
        if (ta == 'C') x.y = -x.y;
        if (tb == 'C') y.y = -y.y;
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
    }
    V out;
    out.x = re;
    out.y = im;
    C[sc * b + i + ldc * j] = out;
}

// For every (m, n, nu, iq) of g2vals (nbkq, nbk, nm, nq) column-major: i = imap_i[n], f = imap_f[m + ld_f
// (ikqs[iq] - 1)] (1-based states, 0 = outside the window; imap_f has nk_f columns); when both are positive,
// slot nu + nm (i - i0 - 1) + nm ni_stride (f - 1) of g2_out / w_out takes g2vals and w[nu + nm iq]. k+q
// indices outside 1..nk_f and slots outside [0, n_out) are not written; each such entry adds 1 to
// *dropped (when given), so a wrong i0 / ni_stride surfaces instead of losing data silently.
template <typename T>
__global__ void window_scatter_kernel(T *__restrict__ g2_out, T *__restrict__ w_out, const T *__restrict__ g2vals,
                                      const int64_t *__restrict__ imap_i, const int64_t *__restrict__ imap_f,
                                      int64_t ld_f, int64_t nk_f, const int64_t *__restrict__ ikqs, const T *__restrict__ w,
                                      int nbkq, int nbk, int nm, int64_t nq, int64_t ni_stride, int64_t i0,
                                      int64_t n_out, unsigned int *__restrict__ dropped) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbkq * nbk * nm * nq) return;
    const int m = (int)(e % nbkq);
    const int n = (int)((e / nbkq) % nbk);
    const int nu = (int)((e / ((int64_t)nbkq * nbk)) % nm);
    const int64_t iq = e / ((int64_t)nbkq * nbk * nm);
    const int64_t i = imap_i[n];
    const int64_t kq = ikqs[iq];
    if (kq < 1 || kq > nk_f) {
        if (dropped) atomicAdd(dropped, 1u);
        return;
    }
    const int64_t f = imap_f[m + ld_f * (kq - 1)];
    if (i <= 0 || f <= 0) return;
    const int64_t lin = nu + (int64_t)nm * (i - i0 - 1) + (int64_t)nm * ni_stride * (f - 1);
    if (lin < 0 || lin >= n_out) {
        if (dropped) atomicAdd(dropped, 1u);
        return;
    }
    g2_out[lin] = g2vals[e];
    w_out[lin] = w[nu + (int64_t)nm * iq];
}

// eph_apply_rotations! (ext/ElectronPhononCUDAExt.jl:177-201): one thread per (m, n, nu', q) of ep
// (nbkq, nbk, nm, nq), dense column-major: ep = sum_nu (sum_iw conj(ukq[iw, m, q]) g[iw, n, nu, q]) uph[nu, nu', q],
// g2 = |ep|^2 / (2 w[nu', q]) when g2 is given. g is read through its four strides (elements), so strided
// views are taken as they are; ukq (nw, nbkq, nq), uph (nm, nm, nq) and w (nm, nq) dense.
template <typename V, typename R>
__global__ void apply_rotations_kernel(int nw, int nbkq, int nbk, int nm, int64_t nq, V *__restrict__ ep,
                                       const V *__restrict__ g, int64_t s0, int64_t s1, int64_t s2, int64_t s3,
                                       const V *__restrict__ ukq, const V *__restrict__ uph, R *__restrict__ g2,
                                       const R *__restrict__ w) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbkq * nbk * nm * nq) return;
    const int m = (int)(e % nbkq);
    const int n = (int)((e / nbkq) % nbk);
    const int nu2 = (int)((e / ((int64_t)nbkq * nbk)) % nm);
    const int64_t q = e / ((int64_t)nbkq * nbk * nm);
    const V *uk = ukq + (int64_t)nw * nbkq * q + (int64_t)nw * m;
    const V *up = uph + (int64_t)nm * nm * q + (int64_t)nm * nu2;
    const V *gq = g + s1 * n + s3 * q;
    R er = 0, ei = 0;
    for (int nu = 0; nu < nm; nu++) {
        R tr = 0, ti = 0;
        for (int iw = 0; iw < nw; iw++) {
            const V a = uk[iw], b = gq[s0 * iw + s2 * nu];
            tr += a.x * b.x + a.y * b.y;
            ti += a.x * b.y - a.y * b.x;
        }
        const V p = up[nu];
        er += tr * p.x - ti * p.y;
        ei += tr * p.y + ti * p.x;
    }
    V out;
    out.x = er;
    out.y = ei;
    ep[e] = out;
    if (g2) g2[e] = (er * er + ei * ei) / (2 * w[nu2 + (int64_t)nm * q]);
}

// eph_apply_rotations_rqkq! (ext/ElectronPhononCUDAExt.jl:241-258, generic src/wannier_to_bloch_batched.jl:
// 360-377): one thread per (m, n, nu, k) of ep (nbkq, nbk, nm, nk), dense column-major:
// ep = sum_jw (sum_iw conj(ukq[iw, m, k]) g[iw, jw, nu, k]) uk[jw, n, k]; g (nw, nw, nm, nk), uk (nw, nbk, nk),
// ukq (nw, nbkq, nk) dense.
template <typename V, typename R>
__global__ void apply_rotations_rqkq_kernel(int nw, int nbkq, int nbk, int nm, int64_t nk, V *__restrict__ ep,
                                            const V *__restrict__ g, const V *__restrict__ uk,
                                            const V *__restrict__ ukq) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbkq * nbk * nm * nk) return;
    const int m = (int)(e % nbkq);
    const int n = (int)((e / nbkq) % nbk);
    const int nu = (int)((e / ((int64_t)nbkq * nbk)) % nm);
    const int64_t k = e / ((int64_t)nbkq * nbk * nm);
    const V *ul = ukq + (int64_t)nw * nbkq * k + (int64_t)nw * m;
    const V *ur = uk + (int64_t)nw * nbk * k + (int64_t)nw * n;
    const V *gk = g + (int64_t)nw * nw * (nu + (int64_t)nm * k);
    R er = 0, ei = 0;
    for (int jw = 0; jw < nw; jw++) {
        R tr = 0, ti = 0;
        for (int iw = 0; iw < nw; iw++) {
            const V a = ul[iw], b = gk[iw + (int64_t)nw * jw];
            tr += a.x * b.x + a.y * b.y;
            ti += a.x * b.y - a.y * b.x;
        }
        const V p = ur[jw];
        er += tr * p.x - ti * p.y;
        ei += tr * p.y + ti * p.x;
    }
    V out;
    out.x = er;
    out.y = ei;
    ep[e] = out;
}

}  // namespace ephip

static bool valid_trans(int32_t t) { return t == 'N' || t == 'T' || t == 'C'; }

extern "C" int ephip_batched_gemm(int32_t fp64, int32_t transA, int32_t transB, int32_t m, int32_t n, int32_t k,
                                  int64_t nbatch, const void *A, int64_t lda, int64_t strideA, const void *B,
                                  int64_t ldb, int64_t strideB, void *C, int64_t ldc, int64_t strideC,
                                  void *stream) {
    EPHIP_REQUIRE(valid_trans(transA) && valid_trans(transB), "batched_gemm: trans must be 'N', 'T' or 'C'");
    EPHIP_REQUIRE(m >= 0 && n >= 0 && k >= 0 && nbatch >= 0, "batched_gemm: negative size");
    EPHIP_REQUIRE(lda >= std::max(1, transA == 'N' ? m : k) && ldb >= std::max(1, transB == 'N' ? k : n) &&
                      ldc >= std::max(1, m),
                  "batched_gemm: leading dimension too small");
    const int64_t mn = (int64_t)m * n;
    if (mn == 0 || nbatch == 0) return EPHIP_OK;
    EPHIP_REQUIRE(strideC >= ldc * n || nbatch == 1, "batched_gemm: batches of C overlap");
    EPHIP_REQUIRE(C && (k == 0 || (A && B)), "batched_gemm: null device pointer");  // k = 0: C = 0
    EPHIP_REQUIRE(mn * nbatch <= 0x7fffffffLL * 256, "batched_gemm: too many entries");
    if (fp64) {
        batched_gemm_kernel<double2, double><<<blocks_for(mn * nbatch), 256, 0, as_stream(stream)>>>(
            transA, transB, m, n, k, nbatch, static_cast<const double2 *>(A), lda, strideA,
            static_cast<const double2 *>(B), ldb, strideB, static_cast<double2 *>(C), ldc, strideC);
    } else {
        batched_gemm_kernel<float2, float><<<blocks_for(mn * nbatch), 256, 0, as_stream(stream)>>>(
            transA, transB, m, n, k, nbatch, static_cast<const float2 *>(A), lda, strideA,
            static_cast<const float2 *>(B), ldb, strideB, static_cast<float2 *>(C), ldc, strideC);
    }
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_window_scatter(int32_t fp64, void *g2_out, void *w_out, int64_t n_out, const void *g2vals,
                                        const int64_t *imap_i, const int64_t *imap_f, int64_t ld_f, int64_t nk_f,
                                        const int64_t *ikqs, const void *w, int32_t nbkq, int32_t nbk, int32_t nm,
                                        int64_t nq, int64_t ni_stride, int64_t i0, uint32_t *dropped,
                                        void *stream) {
    EPHIP_REQUIRE(nbkq >= 0 && nbk >= 0 && nm >= 0 && nq >= 0 && n_out >= 0 && ni_stride >= 0 && i0 >= 0,
                  "eph_window_scatter: negative size");
    EPHIP_REQUIRE(nk_f >= 0, "eph_window_scatter: negative nk_f");
    EPHIP_REQUIRE(ld_f >= std::max(1, nbkq), "eph_window_scatter: ld_f below nbkq");
    const int64_t ne = (int64_t)nbkq * nbk * nm * nq;
    if (ne == 0) return EPHIP_OK;
    EPHIP_REQUIRE(g2_out && w_out && g2vals && imap_i && imap_f && ikqs && w, "eph_window_scatter: null device pointer");
    EPHIP_REQUIRE(ne <= 0x7fffffffLL * 256, "eph_window_scatter: too many entries");
    if (fp64)
        window_scatter_kernel<double><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            static_cast<double *>(g2_out), static_cast<double *>(w_out), static_cast<const double *>(g2vals), imap_i,
            imap_f, ld_f, nk_f, ikqs, static_cast<const double *>(w), nbkq, nbk, nm, nq, ni_stride, i0, n_out,
            dropped);
    else
        window_scatter_kernel<float><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            static_cast<float *>(g2_out), static_cast<float *>(w_out), static_cast<const float *>(g2vals), imap_i,
            imap_f, ld_f, nk_f, ikqs, static_cast<const float *>(w), nbkq, nbk, nm, nq, ni_stride, i0, n_out,
            dropped);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_apply_rotations(int32_t fp64, int32_t nw, int32_t nbkq, int32_t nbk, int32_t nm, int64_t nq,
                                         void *ep, const void *g, int64_t gs0, int64_t gs1, int64_t gs2, int64_t gs3,
                                         const void *ukqs, const void *uphs, void *g2_out, const void *wq,
                                         void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nbkq >= 0 && nbk >= 0 && nm >= 0 && nq >= 0, "eph_apply_rotations: bad sizes");
    EPHIP_REQUIRE(nbkq <= nw && nbk <= nw, "eph_apply_rotations: band counts above nw");
    const int64_t ne = (int64_t)nbkq * nbk * nm * nq;
    if (ne == 0) return EPHIP_OK;
    EPHIP_REQUIRE(ep && g && ukqs && uphs && (!g2_out || wq), "eph_apply_rotations: null device pointer");
    EPHIP_REQUIRE(ne <= 0x7fffffffLL * 256, "eph_apply_rotations: too many entries");
    if (fp64)
        apply_rotations_kernel<double2, double><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            nw, nbkq, nbk, nm, nq, static_cast<double2 *>(ep), static_cast<const double2 *>(g), gs0, gs1, gs2, gs3,
            static_cast<const double2 *>(ukqs), static_cast<const double2 *>(uphs), static_cast<double *>(g2_out),
            static_cast<const double *>(wq));
    else
        apply_rotations_kernel<float2, float><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            nw, nbkq, nbk, nm, nq, static_cast<float2 *>(ep), static_cast<const float2 *>(g), gs0, gs1, gs2, gs3,
            static_cast<const float2 *>(ukqs), static_cast<const float2 *>(uphs), static_cast<float *>(g2_out),
            static_cast<const float *>(wq));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_apply_rotations_rqkq(int32_t fp64, int32_t nw, int32_t nbkq, int32_t nbk, int32_t nm,
                                              int64_t nk, void *ep, const void *g, const void *uks, const void *ukqs,
                                              void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nbkq >= 0 && nbk >= 0 && nm >= 0 && nk >= 0, "eph_apply_rotations_rqkq: bad sizes");
    EPHIP_REQUIRE(nbkq <= nw && nbk <= nw, "eph_apply_rotations_rqkq: band counts above nw");
    const int64_t ne = (int64_t)nbkq * nbk * nm * nk;
    if (ne == 0) return EPHIP_OK;
    EPHIP_REQUIRE(ep && g && uks && ukqs, "eph_apply_rotations_rqkq: null device pointer");
    EPHIP_REQUIRE(ne <= 0x7fffffffLL * 256, "eph_apply_rotations_rqkq: too many entries");
    if (fp64)
        apply_rotations_rqkq_kernel<double2, double><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            nw, nbkq, nbk, nm, nk, static_cast<double2 *>(ep), static_cast<const double2 *>(g),
            static_cast<const double2 *>(uks), static_cast<const double2 *>(ukqs));
    else
        apply_rotations_rqkq_kernel<float2, float><<<blocks_for(ne), 256, 0, as_stream(stream)>>>(
            nw, nbkq, nbk, nm, nk, static_cast<float2 *>(ep), static_cast<const float2 *>(g),
            static_cast<const float2 *>(uks), static_cast<const float2 *>(ukqs));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
