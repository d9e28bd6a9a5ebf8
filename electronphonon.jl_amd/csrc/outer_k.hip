// Outer-k e-ph path with the k-side window projection: the reference's list-batched RR->kR and
// kR->kq drivers (src/wannier_to_bloch_batched.jl:178-273) and rotations (:442-455, CUDA
// ext/ElectronPhononCUDAExt.jl:146-201), with the per-k window projection of the outer-k loop
// (src/calculator/run_eph_over_k_and_kq.jl:657-659). A "unit" is one (k, in-window band n) pair;
// each unit owns a projected operator g(k)[i, n, nu, Rp] and a compacted list of q whose k+q has an
// in-window band (the other pairs are exactly zero, as the reference's k+q window leaves them).
#include "common.hpp"

namespace ephip {

struct c32 {
    float re, im;
};

// "ph" layout (b, Re) x Rp  ->  "el" layout (b, Rp) x Re, b in [0, B) with B = nw*nw*nm
// (src/model_new.jl:551-565: the two outer-momentum layouts of the same tensor).
__global__ void ph_to_el_kernel(int64_t B, int nre, int nrp, const c32 *__restrict__ x, c32 *__restrict__ y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * nre * nrp) return;
    const int64_t b = e % B, rest = e / B;
    const int64_t rp = rest % nrp, re = rest / nrp;
    y[e] = x[b + B * (re + (int64_t)nre * rp)];
}

// get_eph_RR_to_kR! projection (src/wannier_to_bloch.jl:381-388 per k; batched :193-197):
// A_u[(i + nw*nu) + nw*nm*Rp] = sum_j G_col(u)[i + nw*(j + nw*(nu + nm*Rp))] * U_state(u)[j, band(u)].
__global__ void project_kR_kernel(int nw, int nm, int nrp, int nunits, const c32 *__restrict__ G, int64_t ldg,
                                  const int *__restrict__ ucol, const int64_t *__restrict__ ustate,
                                  const int *__restrict__ uband, const c32 *__restrict__ U, c32 *__restrict__ A) {
    const int64_t per = (int64_t)nw * nm * nrp;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * nunits) return;
    const int u = (int)(e / per);
    const int64_t r = e % per;
    const int i = (int)(r % nw);
    const int64_t nurp = r / nw;  // nu + nm * Rp
    const c32 *g = G + ldg * ucol[u] + i + (int64_t)nw * nw * nurp;
    const c32 *uk = U + (int64_t)nw * nw * ustate[u] + (int64_t)nw * uband[u];
    float sr = 0.f, si = 0.f;
    for (int j = 0; j < nw; j++) {
        const c32 a = g[(int64_t)nw * j], w = uk[j];
        sr += a.re * w.re - a.im * w.im;
        si += a.re * w.im + a.im * w.re;
    }
    A[e] = c32{sr, si};
}

// Per unit: the q of the list whose k+q (on the k grid) has an in-window band, in q order.
// col_index[u*cs + c] = q, ikq[u*cs + c] = state index of k+q, count[u] = number kept.
// One workgroup of 256 threads per unit; positions from wave ballots + a 4-wave prefix. A pass takes
// KQC_CHUNKS chunks of 256 q (their loads in flight together) and one barrier; the wave counts alternate
// between two LDS slots, so that barrier separates their writes from their reads (a slot is rewritten two
// passes later, after the next barrier). Grid coordinates in int32 (|rint(x n)| < 2^31 for the reduced
// coordinates of a k-grid the size check admits).
constexpr int KQC_CHUNKS = 4;
__global__ __launch_bounds__(256) void kq_compact_kernel(int nunits, const int64_t *__restrict__ ustate,
                                                        const double *__restrict__ xk, int nq,
                                                        const double *__restrict__ xq, int n1, int n2, int n3,
                                                        const int64_t *__restrict__ bmin, int64_t cs,
                                                        int *__restrict__ col_index, int *__restrict__ ikq,
                                                        int *__restrict__ count) {
    __shared__ int wave_cnt[2][KQC_CHUNKS][4];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t s = ustate[u];
    const double k0 = xk[3 * s], k1 = xk[3 * s + 1], k2 = xk[3 * s + 2];
    auto red = [](double x, int n) {
        const int c = (int)rint(x * n) % n;
        return c < 0 ? c + n : c;
    };
    const unsigned long long below = (1ull << lane) - 1ull;
    int base = 0, slot = 0;
    for (int c0 = 0; c0 < nq; c0 += 256 * KQC_CHUNKS) {
        bool keep[KQC_CHUNKS];
        int h[KQC_CHUNKS];
        unsigned long long ball[KQC_CHUNKS];
#pragma unroll
        for (int q = 0; q < KQC_CHUNKS; q++) {
            const int c = c0 + 256 * q + tid;
            keep[q] = false;
            h[q] = 0;
            if (c < nq) {
                h[q] = (red(k0 + xq[3 * c], n1) * n2 + red(k1 + xq[3 * c + 1], n2)) * n3 +
                       red(k2 + xq[3 * c + 2], n3);
                keep[q] = bmin[h[q]] > 0;
            }
        }
#pragma unroll
        for (int q = 0; q < KQC_CHUNKS; q++) {
            ball[q] = __ballot(keep[q]);
            if (lane == 0) wave_cnt[slot][q][wave] = __popcll(ball[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KQC_CHUNKS; q++) {
            int off = base;
            for (int w = 0; w < wave; w++) off += wave_cnt[slot][q][w];
            if (keep[q]) {
                const int64_t o = u * cs + off + __popcll(ball[q] & below);
                col_index[o] = c0 + 256 * q + tid;
                ikq[o] = h[q];
            }
            base += wave_cnt[slot][q][0] + wave_cnt[slot][q][1] + wave_cnt[slot][q][2] + wave_cnt[slot][q][3];
        }
        slot ^= 1;
    }
    if (tid == 0) count[u] = base;
}

// kR->kq rotations + g2 fold for one (unit, kept q) per wave (eph_apply_rotations!,
// src/wannier_to_bloch_batched.jl:442-455; src/wannier_to_bloch.jl:428-440):
//   t[m, nu]   = sum_i conj(U_kq[i, mlo + m]) Gq[i + nw*nu]
//   ep[m, nu'] = sum_nu t[m, nu] u_ph(q)[nu, nu'],   g2 = |ep|^2 / (2 omega(q)[nu'])
// Output row (u*cs + c) holds nbw*nm values [m*nm + nu'], zero for m >= window width at k+q.
__global__ __launch_bounds__(256) void rotate_kq_g2_kernel(int nw, int nm, int nbw, int nunits, int64_t cs,
                                                          const int *__restrict__ count,
                                                          const int *__restrict__ col_index,
                                                          const int *__restrict__ ikq, const c32 *__restrict__ Gq,
                                                          int64_t ldg, int64_t gstride, const c32 *__restrict__ U,
                                                          const int64_t *__restrict__ bmin,
                                                          const int64_t *__restrict__ bmax,
                                                          const c32 *__restrict__ uph,
                                                          const double *__restrict__ omega, float *__restrict__ g2,
                                                          c32 *__restrict__ ep_out) {
    extern __shared__ c32 tsh[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int u = blockIdx.y;
    const int64_t c = (int64_t)blockIdx.x * 4 + wave;
    if (u >= nunits || c >= count[u]) return;
    c32 *T = tsh + wave * nbw * nm;
    const int64_t row = (int64_t)u * cs + c;
    const int iq = col_index[row];
    const int64_t kq = ikq[row];
    const int mlo = (int)bmin[kq] - 1;
    int nbkq = bmin[kq] > 0 ? (int)(bmax[kq] - bmin[kq] + 1) : 0;  // no window band: all slots zero
    nbkq = nbkq < nbw ? nbkq : nbw;
    const c32 *g = Gq + gstride * u + ldg * c;
    const c32 *uk = U + (int64_t)nw * nw * kq;
    for (int p = lane; p < nbw * nm; p += 64) {
        const int m = p % nbw, nu = p / nbw;
        float sr = 0.f, si = 0.f;
        if (m < nbkq) {
            for (int i = 0; i < nw; i++) {
                const c32 w = uk[i + (int64_t)nw * (mlo + m)], a = g[i + (int64_t)nw * nu];
                sr += w.re * a.re + w.im * a.im;
                si += w.re * a.im - w.im * a.re;
            }
        }
        T[m + nbw * nu] = c32{sr, si};
    }
    wave_lds_sync();
    const c32 *up = uph + (int64_t)nm * nm * iq;
    const double *om = omega + (int64_t)nm * iq;
    for (int p = lane; p < nbw * nm; p += 64) {
        const int nu2 = p % nm, m = p / nm;
        float er = 0.f, ei = 0.f;
        if (m < nbkq) {
            for (int nu = 0; nu < nm; nu++) {
                const c32 t = T[m + nbw * nu], w = up[nu + nm * nu2];
                er += t.re * w.re - t.im * w.im;
                ei += t.re * w.im + t.im * w.re;
            }
        }
        const int64_t o = row * nbw * nm + p;
        if (g2) g2[o] = (er * er + ei * ei) / (2.f * static_cast<float>(om[nu2]));
        if (ep_out) ep_out[o] = c32{er, ei};
    }
}

// Same contraction, 16 kept columns of one unit per 256-thread workgroup with every operand staged in
// LDS: the columns' Fourier outputs (contiguous in memory, one coalesced sweep), their k+q window
// eigenvector columns, phonon eigenvectors and frequencies. Then t = U_kq^H g and ep = t u_ph from LDS.
// The per-mode blocks of g and u_ph are stored with one complex of padding (pitch nw + 1, nm + 1) so
// that lanes reading different modes hit different LDS banks.
constexpr int RKQ_COLS = 16;

__host__ __device__ inline size_t rkq_lds_bytes(int nw, int nm, int nbw) {
    return sizeof(c32) * (size_t)RKQ_COLS *
               ((size_t)(nw + 1) * nm + (size_t)nm * (nm + 1) + (size_t)nbw * nw + (size_t)nbw * nm) +
           sizeof(float) * RKQ_COLS * nm + sizeof(int) * RKQ_COLS * 2;
}

__global__ __launch_bounds__(256) void rotate_kq_g2_lds_kernel(int nw, int nm, int nbw, int nunits, int64_t cs,
                                                              const int *__restrict__ count,
                                                              const int *__restrict__ col_index,
                                                              const int *__restrict__ ikq, const c32 *__restrict__ Gq,
                                                              int64_t ldg, int64_t gstride, const c32 *__restrict__ U,
                                                              const int64_t *__restrict__ bmin,
                                                              const int64_t *__restrict__ bmax,
                                                              const c32 *__restrict__ uph,
                                                              const double *__restrict__ omega,
                                                              float *__restrict__ g2, c32 *__restrict__ ep_out) {
    extern __shared__ __attribute__((aligned(16))) unsigned char rkq_lds[];
    const int u = blockIdx.y;
    const int64_t c0 = (int64_t)blockIdx.x * RKQ_COLS;
    if (u >= nunits) return;
    const int cnt = count[u];
    if (c0 >= cnt) return;  // uniform for the block
    const int ncol = (int)min<int64_t>(RKQ_COLS, cnt - c0);
    const int tid = threadIdx.x;
    const int nd = nw * nm;
    const int ndp = (nw + 1) * nm, npp = (nm + 1) * nm;     // padded per-column sizes
    c32 *Gs = reinterpret_cast<c32 *>(rkq_lds);            // [c][nu][i], pitch nw + 1
    c32 *Ps = Gs + RKQ_COLS * ndp;                          // [c][nu'][nu], pitch nm + 1
    c32 *Us = Ps + RKQ_COLS * npp;                          // [c][m][i]
    c32 *Ts = Us + RKQ_COLS * nbw * nw;                     // [c][m][nu]
    float *Ws = reinterpret_cast<float *>(Ts + RKQ_COLS * nbw * nm);  // [c][nu] omega
    int *Ms = reinterpret_cast<int *>(Ws + RKQ_COLS * nm);             // [c] mlo, [c] nbkq
    const int64_t row0 = (int64_t)u * cs + c0;
    if (tid < RKQ_COLS) {
        int mlo = 0, nb = 0;
        if (tid < ncol) {
            const int64_t kq = ikq[row0 + tid];
            if (bmin[kq] > 0) {
                mlo = (int)bmin[kq] - 1;
                nb = (int)(bmax[kq] - bmin[kq] + 1);
                nb = nb < nbw ? nb : nbw;
            }
        }
        Ms[tid] = mlo;
        Ms[RKQ_COLS + tid] = nb;
    }
    // Fourier outputs of the ncol columns are contiguous: ldg == nd is not assumed, copy column-wise
    const c32 *gsrc = Gq + gstride * u + ldg * c0;
    for (int e = tid; e < ncol * nd; e += 256) {
        const int c = e / nd, d = e % nd;
        Gs[c * ndp + (nw + 1) * (d / nw) + d % nw] = gsrc[ldg * c + d];
    }
    for (int e = tid; e < ncol * nm * nm; e += 256) {
        const int c = e / (nm * nm), r = e % (nm * nm);
        Ps[c * npp + (nm + 1) * (r / nm) + r % nm] = uph[(int64_t)nm * nm * col_index[row0 + c] + r];
    }
    for (int e = tid; e < ncol * nm; e += 256) {
        const int c = e / nm, r = e % nm;
        Ws[e] = static_cast<float>(omega[(int64_t)nm * col_index[row0 + c] + r]);
    }
    __syncthreads();
    for (int e = tid; e < ncol * nbw * nw; e += 256) {
        const int c = e / (nbw * nw), r = e % (nbw * nw), m = r / nw, i = r % nw;
        c32 v{0.f, 0.f};
        if (m < Ms[RKQ_COLS + c]) v = U[(int64_t)nw * nw * ikq[row0 + c] + (int64_t)nw * (Ms[c] + m) + i];
        Us[e] = v;
    }
    __syncthreads();
    // t[c][m][nu] = sum_i conj(U[c][m][i]) g[c][i + nw nu]
    for (int e = tid; e < ncol * nbw * nm; e += 256) {
        const int c = e / (nbw * nm), r = e % (nbw * nm), m = r / nm, nu = r % nm;
        float sr = 0.f, si = 0.f;
        const c32 *uu = Us + (c * nbw + m) * nw;
        const c32 *gg = Gs + c * ndp + (nw + 1) * nu;
        for (int i = 0; i < nw; i++) {
            const c32 w = uu[i], a = gg[i];
            sr += w.re * a.re + w.im * a.im;
            si += w.re * a.im - w.im * a.re;
        }
        Ts[e] = c32{sr, si};
    }
    __syncthreads();
    // ep[c][m][nu'] = sum_nu t[c][m][nu] uph[c][nu + nm nu'], g2 = |ep|^2 / (2 omega)
    for (int e = tid; e < ncol * nbw * nm; e += 256) {
        const int c = e / (nbw * nm), r = e % (nbw * nm), m = r / nm, nu2 = r % nm;
        float er = 0.f, ei = 0.f;
        if (m < Ms[RKQ_COLS + c]) {
            const c32 *tt = Ts + (c * nbw + m) * nm;
            const c32 *pp = Ps + c * npp + (nm + 1) * nu2;
            for (int nu = 0; nu < nm; nu++) {
                const c32 t = tt[nu], w = pp[nu];
                er += t.re * w.re - t.im * w.im;
                ei += t.re * w.im + t.im * w.re;
            }
        }
        const int64_t o = (row0 + c) * nbw * nm + r;
        if (g2) g2[o] = (er * er + ei * ei) / (2.f * Ws[c * nm + nu2]);
        if (ep_out) ep_out[o] = c32{er, ei};
    }
}

// One k of an outer-k payload in the layout of EPDataQBatched (src/EPData.jl:66-73, built by
// src/calculator/run_eph_over_k_and_kq.jl:843-876): eps / g2s (nw, nbandk, nm, nq) over the full k+q band
// range and the k-side window [ibandk_offset + 1, ibandk_offset + nbandk], for the kept columns
// j0 .. j0 + nq - 1 of the k; zero outside both windows. The k's units (one per in-window band, band0 + 1 ..
// band0 + nunits) are consecutive payload rows of stride cs columns; unit row columns hold (slot, nu')
// for k+q band bmin - 1 + slot. wqs (nm, nq) = omega of each column's q; ikqs (nq) = state_pos of each
// column's k+q state (1-based position in the filtered state list).
__global__ void units_to_qbatched_kernel(int nw, int nm, int nbw, int nbandk, int64_t cs,
                                         const int *__restrict__ col_index, const int *__restrict__ ikq, int j0, int nq,
                                         int nunits, int band0, int ibandk_offset, const float2 *__restrict__ ep,
                                         const float *__restrict__ g2, const int64_t *__restrict__ bmin,
                                         const int64_t *__restrict__ bmax, const double *__restrict__ omega,
                                         const int64_t *__restrict__ state_pos, float2 *__restrict__ eps,
                                         float *__restrict__ g2s, double *__restrict__ wqs, int64_t *__restrict__ ikqs) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nw * nbandk * nm * nq) return;
    const int m = (int)(e % nw);
    const int n = (int)((e / nw) % nbandk);
    const int nu2 = (int)((e / ((int64_t)nw * nbandk)) % nm);
    const int j = (int)(e / ((int64_t)nw * nbandk * nm));
    const int kq = ikq[j0 + j];
    const int iq = col_index[j0 + j];
    const int64_t b0 = bmin[kq];
    const int nbq = b0 > 0 ? min((int)(bmax[kq] - b0 + 1), nbw) : 0;
    const int slot = m - (int)(b0 - 1), unit = ibandk_offset + n - band0;
    float2 v = make_float2(0.f, 0.f);
    float w = 0.f;
    if (b0 > 0 && slot >= 0 && slot < nbq && unit >= 0 && unit < nunits) {
        const int64_t src = ((int64_t)unit * cs + j0 + j) * nbw * nm + (int64_t)slot * nm + nu2;
        if (ep) v = ep[src];
        w = g2[src];
    }
    if (eps) eps[e] = v;
    g2s[e] = w;
    if (m == 0 && n == 0) {
        wqs[nu2 + (int64_t)nm * j] = omega[(int64_t)nm * iq + nu2];
        if (nu2 == 0) ikqs[j] = state_pos[kq];
    }
}

}  // namespace ephip

using namespace ephip;

static inline unsigned nblocks(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int ephip_eph_ph_to_el(int64_t nbasis, int32_t nre, int32_t nrp, const void *ph, void *el, void *stream) {
    EPHIP_REQUIRE(nbasis >= 1 && nre >= 1 && nrp >= 1, "eph_ph_to_el: bad sizes");
    EPHIP_REQUIRE(ph && el && ph != el, "eph_ph_to_el: null or aliased device pointer");
    const int64_t n = nbasis * nre * nrp;
    EPHIP_REQUIRE(nblocks(n) < 0x7fffffffu, "eph_ph_to_el: tensor too large");
    ph_to_el_kernel<<<nblocks(n), 256, 0, as_stream(stream)>>>(nbasis, nre, nrp, static_cast<const c32 *>(ph),
                                                              static_cast<c32 *>(el));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_project_kR(int32_t nw, int32_t nm, int32_t nrp, int32_t nunits, const void *G, int64_t ldg,
                                    const int32_t *ucol, const int64_t *ustate, const int32_t *uband, const void *U,
                                    void *A, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nm >= 1 && nrp >= 1 && nunits >= 0, "eph_project_kR: bad sizes");
    EPHIP_REQUIRE(ldg >= (int64_t)nw * nw * nm * nrp, "eph_project_kR: ldg too small");
    if (nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(G && ucol && ustate && uband && U && A, "eph_project_kR: null device pointer");
    const int64_t n = (int64_t)nw * nm * nrp * nunits;
    project_kR_kernel<<<nblocks(n), 256, 0, as_stream(stream)>>>(nw, nm, nrp, nunits, static_cast<const c32 *>(G), ldg,
                                                                ucol, ustate, uband, static_cast<const c32 *>(U),
                                                                static_cast<c32 *>(A));
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_kq_compact(int32_t nunits, const int64_t *ustate, const double *xk, int32_t nq, const double *xq,
                                const int32_t *ngrid, const int64_t *bmin, int64_t cs, int32_t *col_index, int32_t *ikq,
                                int32_t *count, void *stream) {
    EPHIP_REQUIRE(nunits >= 0 && nunits <= 0x7fffffff && nq >= 0, "kq_compact: bad sizes");
    EPHIP_REQUIRE(ngrid && ngrid[0] > 0 && ngrid[1] > 0 && ngrid[2] > 0, "kq_compact: bad grid");
    EPHIP_REQUIRE((int64_t)ngrid[0] * ngrid[1] * ngrid[2] < 0x7fffffffLL, "kq_compact: grid too large for int32 indices");
    EPHIP_REQUIRE(cs >= nq, "kq_compact: column stride %lld < nq %d", (long long)cs, nq);
    if (nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(ustate && xk && xq && bmin && col_index && ikq && count, "kq_compact: null device pointer");
    kq_compact_kernel<<<nunits, 256, 0, as_stream(stream)>>>(nunits, ustate, xk, nq, xq, ngrid[0], ngrid[1], ngrid[2],
                                                           bmin, cs, col_index, ikq, count);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_eph_rotate_kq_g2(int32_t nw, int32_t nm, int32_t nbw, int32_t nunits, int64_t cs,
                                      const int32_t *count, const int32_t *col_index, const int32_t *ikq,
                                      const void *Gq, int64_t ldg, int64_t gstride, const void *U,
                                      const int64_t *bmin, const int64_t *bmax, const void *uph, const double *omega,
                                      float *g2, void *ep, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nm >= 1 && nunits >= 0 && cs >= 0, "eph_rotate_kq_g2: bad sizes");
    EPHIP_REQUIRE(nbw >= 1 && nbw <= nw, "eph_rotate_kq_g2: nbw=%d must be in 1..nw=%d", nbw, nw);
    EPHIP_REQUIRE(ldg >= (int64_t)nw * nm && gstride >= ldg * cs, "eph_rotate_kq_g2: bad ldg/gstride");
    const size_t lds = 4 * sizeof(c32) * (size_t)nbw * nm;
    EPHIP_REQUIRE(lds <= 64 * 1024, "eph_rotate_kq_g2: nbw*nm too large");
    EPHIP_REQUIRE(nunits <= 65535, "eph_rotate_kq_g2: nunits=%d > 65535", nunits);
    if (nunits == 0 || cs == 0) return EPHIP_OK;
    EPHIP_REQUIRE(count && col_index && ikq && Gq && U && bmin && bmax && uph && omega && (g2 || ep),
                  "eph_rotate_kq_g2: null device pointer");
    const size_t lds2 = rkq_lds_bytes(nw, nm, nbw);
    if (lds2 <= 64 * 1024) {
        dim3 grid((unsigned)((cs + RKQ_COLS - 1) / RKQ_COLS), (unsigned)nunits);
        rotate_kq_g2_lds_kernel<<<grid, 256, lds2, as_stream(stream)>>>(
            nw, nm, nbw, nunits, cs, count, col_index, ikq, static_cast<const c32 *>(Gq), ldg, gstride,
            static_cast<const c32 *>(U), bmin, bmax, static_cast<const c32 *>(uph), omega, g2, static_cast<c32 *>(ep));
    } else {  // large nw*nm: one wave per column
        dim3 grid((unsigned)((cs + 3) / 4), (unsigned)nunits);
        rotate_kq_g2_kernel<<<grid, 256, lds, as_stream(stream)>>>(nw, nm, nbw, nunits, cs, count, col_index, ikq,
                                                                  static_cast<const c32 *>(Gq), ldg, gstride,
                                                                  static_cast<const c32 *>(U), bmin, bmax,
                                                                  static_cast<const c32 *>(uph), omega, g2,
                                                                  static_cast<c32 *>(ep));
    }
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_units_to_qbatched(int32_t nw, int32_t nm, int32_t nbw, int32_t nbandk, int64_t cs,
                                       const int32_t *col_index, const int32_t *ikq, int32_t j0, int32_t nq,
                                       int32_t nunits, int32_t band0, int32_t ibandk_offset, const void *ep,
                                       const float *g2, const int64_t *bmin, const int64_t *bmax, const double *omega,
                                       const int64_t *state_pos, void *eps, float *g2s, double *wqs, int64_t *ikqs,
                                       void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nm >= 1 && nbw >= 1 && nbandk >= 1 && nbandk <= nw && cs >= 0 && j0 >= 0 && nq >= 0 &&
                      nunits >= 1 && nunits <= nbandk && band0 >= 0 && ibandk_offset >= 0 &&
                      ibandk_offset + nbandk <= nw && band0 >= ibandk_offset &&
                      band0 + nunits <= ibandk_offset + nbandk && (int64_t)j0 + nq <= cs,
                  "units_to_qbatched: bad sizes or window");
    if (nq == 0) return EPHIP_OK;
    EPHIP_REQUIRE(col_index && ikq && g2 && bmin && bmax && omega && state_pos && g2s && wqs && ikqs,
                  "units_to_qbatched: null device pointer");
    const int64_t ne = (int64_t)nw * nbandk * nm * nq;
    units_to_qbatched_kernel<<<nblocks(ne), 256, 0, as_stream(stream)>>>(
        nw, nm, nbw, nbandk, cs, col_index, ikq, j0, nq, nunits, band0, ibandk_offset,
        static_cast<const float2 *>(ep), g2, bmin, bmax, omega, state_pos, static_cast<float2 *>(eps), g2s, wqs, ikqs);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
