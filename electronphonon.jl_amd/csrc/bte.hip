// Boltzmann transport on the outer-k payload (FP64): scattering increments, the scattering-out rates
// and scattering-in matrix, band velocities, and the sparse product of the iterative solver.
//
// bte_scattering_increments restates src/boltzmann/bte_scattering_core.jl:33-84 (occupation
// conventions 1..6, vanishing-delta guards) with the smeared delta of src/common/smearing.jl;
// the accumulation follows the device kernel ext/ElectronPhononCUDAExt.jl:323-372 /
// bte_window_accumulate! (src/boltzmann/boltzmann_calculator.jl:336-372): atomic S_out, plain S_in.
// S_in is stored in the outer-k loop's own sparse layout: row (unit u, kept column c, k+q band slot m).
#include "common.hpp"

namespace ephip {

__device__ __forceinline__ double occ_fd(double e, double T) { return 1.0 / (exp(e / T) + 1.0); }
__device__ __forceinline__ double occ_fd_deriv(double e, double T) {
    return -1.0 / (2.0 + exp(e / T) + exp(-e / T)) / T;
}
__device__ __forceinline__ double occ_boson(double e, double T) { return e == 0.0 ? 0.0 : 1.0 / expm1(e / T); }
__device__ __forceinline__ double smear(int method, double eta, double x) {
    if (method == 1) return (1.0 / sqrt(M_PI)) / eta * exp(-(x / eta) * (x / eta));
    return (eta / M_PI) / (x * x + eta * eta);
}

__device__ void bte_increments(int method, double ek, double ekq, double wq, double g2, double wtq, double mu,
                               double T, int smearing, double eta, double &so, double &si) {
    const double d1 = smear(smearing, eta, ek - ekq + wq);
    const double d2 = smear(smearing, eta, ek - ekq - wq);
    so = 0.0;
    si = 0.0;
    if (d1 == 0.0 && d2 == 0.0) return;
    const double nq = occ_boson(wq, T), fk = occ_fd(ek - mu, T), fkq = occ_fd(ekq - mu, T);
    double f1o, f2o, f1i, f2i;
    if (method == 1) {
        f1o = nq + fkq; f2o = nq + 1 - fkq;
        f1i = nq + fk;  f2i = nq + 1 - fk;
    } else if (method == 2) {
        const double preo = fkq * (1 - fkq) / (fk * (1 - fk));
        f1o = preo * (nq + 1 - fk); f2o = preo * (nq + fk);
        const double prei = fk * (1 - fk) / (fkq * (1 - fkq));
        f1i = prei * (nq + 1 - fkq); f2i = prei * (nq + fkq);
    } else if (method == 3) {
        const double preo = (1 - fkq) / (1 - fk);
        f1o = preo * nq; f2o = preo * (nq + 1);
        const double prei = (1 - fk) / (1 - fkq);
        f1i = prei * nq; f2i = prei * (nq + 1);
    } else if (method == 4) {
        const double preo = fkq / fk;
        f1o = preo * (nq + 1); f2o = preo * nq;
        const double prei = fk / fkq;
        f1i = prei * (nq + 1); f2i = prei * nq;
    } else if (method == 5) {
        const double dk = occ_fd_deriv(ek - mu, T), dkq = occ_fd_deriv(ekq - mu, T);
        f1o = sqrt(nq * (nq + 1) * dkq / dk); f2o = f1o;
        f1i = sqrt(nq * (nq + 1) * dk / dkq); f2i = f1i;
    } else {
        const double no = occ_boson(ekq - ek, T);
        f1o = no + fkq; f2o = -(no + fkq);
        const double ni = occ_boson(ek - ekq, T);
        f1i = ni + fk; f2i = -(ni + fk);
    }
    const double pref = 2.0 * M_PI * wtq * g2;
    so = ((d1 == 0.0 ? 0.0 : d1 * f1o) + (d2 == 0.0 ? 0.0 : d2 * f2o)) * pref;
    si = ((d2 == 0.0 ? 0.0 : d2 * f1i) + (d1 == 0.0 ? 0.0 : d1 * f2i)) * pref;
}

// in: (n x 8) rows {ek, ekq, wq, g2, wtq, mu, T, eta}; out: (n x 2) {s_out, s_in}
__global__ void bte_increments_kernel(int64_t n, int method, int smearing, const double *__restrict__ in,
                                      double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double *r = in + 8 * e;
    double so, si;
    bte_increments(method, r[0], r[1], r[2], r[3], r[4], r[5], r[6], smearing, r[7], so, si);
    out[2 * e] = so;
    out[2 * e + 1] = si;
}

// One thread per (unit u, kept column c, k+q slot m). i = unit_state_index[u] (initial state);
// f = state_offset[ikq] + m (final state); energies e_state[i], e_state[f]; modes below w_cut skipped.
// S_out[i + nstates*iT] += sum_nu s_out (atomic). S_in is compacted: the unit's kept columns are
// rows rowptr[u] .. rowptr[u] + count[u] - 1, S_in[(row*nbw + m)*nT + iT] = sum_nu s_in, and
// row_code[row] = first final state | (number of k+q window slots << 27) of the row.
__global__ void bte_accumulate_kernel(int nunits, int64_t cs, int nbw, int nm, const int *__restrict__ count,
                                      const int *__restrict__ col_index, const int *__restrict__ ikq,
                                      const float *__restrict__ g2, const int64_t *__restrict__ unit_state_index,
                                      const int64_t *__restrict__ state_offset, const int64_t *__restrict__ bmin,
                                      const int64_t *__restrict__ bmax, const double *__restrict__ e_state,
                                      const double *__restrict__ omega, double wtq, int nT,
                                      const double *__restrict__ mus, const double *__restrict__ Ts,
                                      const double *__restrict__ etas, int method, int smearing, double w_cut,
                                      int64_t nstates, const int64_t *__restrict__ rowptr, double *__restrict__ S_out,
                                      double *__restrict__ S_in, uint32_t *__restrict__ row_code) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int u = blockIdx.y;
    if (u >= nunits || e >= cs * nbw) return;
    const int64_t c = e / nbw;
    const int m = (int)(e % nbw);
    if (c >= count[u]) return;
    const int64_t row = (int64_t)u * cs + c;      // payload row
    const int64_t orow = rowptr[u] + c;           // compacted output row
    const int64_t kq = ikq[row];
    const int nbkq = bmin[kq] > 0 ? min((int)(bmax[kq] - bmin[kq] + 1), nbw) : 0;
    if (m == 0) row_code[orow] = (uint32_t)state_offset[kq] | ((uint32_t)nbkq << 27);
    double *sin_row = S_in + (orow * nbw + m) * nT;
    if (m >= nbkq) {
        for (int t = 0; t < nT; t++) sin_row[t] = 0.0;
        return;
    }
    const int64_t i = unit_state_index[u];
    const int64_t f = state_offset[kq] + m;
    const double ek = e_state[i], ekq = e_state[f];
    const int iq = col_index[row];
    const float *gr = g2 + row * nbw * nm + (int64_t)m * nm;
    for (int t = 0; t < nT; t++) {
        double so = 0.0, si = 0.0;
        for (int nu = 0; nu < nm; nu++) {
            const double w = omega[(int64_t)nm * iq + nu];
            if (w < w_cut) continue;
            double a, b;
            bte_increments(method, ek, ekq, w, (double)gr[nu], wtq, mus[t], Ts[t], smearing, etas[t], a, b);
            so += a;
            si += b;
        }
        atomicAdd(&S_out[i + nstates * t], so);
        sin_row[t] = si;
    }
}

// bte_window_accumulate! with the reference's arguments (ext/ElectronPhononCUDAExt.jl:323-372, generic
// src/boltzmann/boltzmann_calculator.jl:336-372): one thread per (m, n, j) of a q batch. i = imap_i[n]
// (outer state, 1-based, 0 = outside the window), f = imap_f[m + ld_f (ikqs[j] - 1)] (inner state); per
// temperature the increments of every mode with w >= w_cut are summed; S_out[(i - 1) + ld_so iT] gets an
// atomic add, S_in[(i - i0 - 1) + ld_si1 ((f - 1) + ld_si2 iT)] the plain write (each (i, f) hit once).
// g2vals (nbkq, nbk, nm, nq) and w (nm, nq) in R (float or double); everything else FP64 / int64.
template <typename R>
__global__ void bte_window_accumulate_kernel(double *__restrict__ S_out, int64_t ld_so, double *__restrict__ S_in,
                                             int64_t ld_si1, int64_t ld_si2, const R *__restrict__ g2vals,
                                             const R *__restrict__ wq, const int64_t *__restrict__ imap_i,
                                             const int64_t *__restrict__ imap_f, int64_t ld_f, int64_t nk_f,
                                             const int64_t *__restrict__ ikqs, const double *__restrict__ e_i,
                                             const double *__restrict__ e_f, const double *__restrict__ wf, int nT,
                                             const double *__restrict__ mus, const double *__restrict__ Ts,
                                             const double *__restrict__ etas, int method, int smearing, double w_cut,
                                             int nbkq, int nbk, int nm, int64_t nq, int64_t i0,
                                             uint32_t *__restrict__ dropped) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbkq * nbk * nq) return;
    const int m = (int)(e % nbkq);
    const int n = (int)((e / nbkq) % nbk);
    const int64_t j = e / ((int64_t)nbkq * nbk);
    const int64_t i = imap_i[n];
    if (i <= 0) return;
    const int64_t kq = ikqs[j];
    if (kq < 1 || kq > nk_f || i - i0 < 1 || i - i0 > ld_si1) {
        if (dropped) atomicAdd(dropped, 1u);
        return;
    }
    const int64_t f = imap_f[m + ld_f * (kq - 1)];
    if (f <= 0) return;
    if (f > ld_si2) {
        if (dropped) atomicAdd(dropped, 1u);
        return;
    }
    const double ek = e_i[i - 1], ekq = e_f[f - 1], wtq = wf[f - 1];
    const R *gr = g2vals + m + (int64_t)nbkq * (n + (int64_t)nbk * nm * j);
    const int64_t nbb = (int64_t)nbkq * nbk;
    for (int t = 0; t < nT; t++) {
        double so = 0.0, si = 0.0;
        for (int nu = 0; nu < nm; nu++) {
            const double w = (double)wq[nu + (int64_t)nm * j];
            if (w < w_cut) continue;
            double a, b;
            bte_increments(method, ek, ekq, w, (double)gr[nbb * nu], wtq, mus[t], Ts[t], smearing, etas[t], a, b);
            so += a;
            si += b;
        }
        atomicAdd(&S_out[(i - 1) + ld_so * t], so);
        S_in[(i - i0 - 1) + ld_si1 * ((f - 1) + ld_si2 * t)] = si;
    }
}

// Diagonal band velocities: v[s][a] = Re(u^H dH_a u) for state s = (k = state_k[s], band = state_band[s]);
// dH: per state-k column, 3 operators of nw*nw (c64), column index state_col[s].
__global__ void band_velocity_kernel(int nw, int64_t nstates, const int64_t *__restrict__ state_col,
                                     const int *__restrict__ state_band, const int64_t *__restrict__ state_k,
                                     const double2 *__restrict__ dH, int64_t ld_dh,
                                     const float2 *__restrict__ U, double *__restrict__ v) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nstates * 3) return;
    const int64_t s = e / 3;
    const int a = (int)(e % 3);
    const double2 *M = dH + ld_dh * state_col[s] + (int64_t)a * nw * nw;
    const float2 *u = U + (int64_t)nw * nw * state_k[s] + (int64_t)nw * state_band[s];
    double acc = 0.0;
    for (int j = 0; j < nw; j++) {
        double wr = 0.0, wi = 0.0;
        for (int i = 0; i < nw; i++) {
            const double2 mij = M[i + (int64_t)nw * j];
            const float2 ui = u[i];
            // conj(u_i) M_ij
            wr += ui.x * mij.x + ui.y * mij.y;
            wi += ui.x * mij.y - ui.y * mij.x;
        }
        const float2 uj = u[j];
        acc += wr * uj.x - wi * uj.y;
    }
    v[3 * s + a] = acc;
}

// Full rotated operator of one (direction a, k) per workgroup, FP64 (U promoted from complex64):
// out[m + nw n] = (U_k^H M_a(k) U_k)[m, n], plus the Berry-connection term i (e_m - e_n) rbar[m, n] when the
// energies and rbar are given. U_k, M_a(k) and T = M_a U_k sit in LDS (nw <= 32: 3 x 16 KB); each thread
// forms whole entries of T, then of out, in the same order for every k.
constexpr int VEL_NW_MAX = 32;
__global__ void __launch_bounds__(256) el_velocity_full_kernel(int nw, const double2 *__restrict__ M, int64_t ld_m,
                                                               const float2 *__restrict__ U,
                                                               const double *__restrict__ E,
                                                               const double2 *__restrict__ rbar, int64_t ld_r,
                                                               double2 *__restrict__ out, int64_t ld_out) {
    __shared__ double2 us[VEL_NW_MAX * VEL_NW_MAX], ms[VEL_NW_MAX * VEL_NW_MAX], ts[VEL_NW_MAX * VEL_NW_MAX];
    const int a = (int)(blockIdx.x % 3);
    const int64_t k = blockIdx.x / 3;
    const int n2 = nw * nw;
    const double2 *mk = M + ld_m * k + (int64_t)a * n2;
    const float2 *uk = U + (int64_t)n2 * k;
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {
        const float2 v = uk[e];
        us[e] = make_double2(v.x, v.y);
        ms[e] = mk[e];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {  // T[i, n] = sum_j M[i, j] U[j, n]
        const int i = e % nw, n = e / nw;
        double re = 0.0, im = 0.0;
        for (int j = 0; j < nw; j++) {
            const double2 x = ms[i + nw * j], y = us[j + nw * n];
            re += x.x * y.x - x.y * y.y;
            im += x.x * y.y + x.y * y.x;
        }
        ts[e] = make_double2(re, im);
    }
    __syncthreads();
    double2 *ok = out + ld_out * k + (int64_t)a * n2;
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {  // out[m, n] = sum_i conj(U[i, m]) T[i, n]
        const int m = e % nw, n = e / nw;
        double re = 0.0, im = 0.0;
        for (int i = 0; i < nw; i++) {
            const double2 x = us[i + nw * m], y = ts[i + nw * n];
            re += x.x * y.x + x.y * y.y;
            im += x.x * y.y - x.y * y.x;
        }
        if (rbar) {  // + i (e_m - e_n) rbar[m, n]
            const double de = E[(int64_t)nw * k + m] - E[(int64_t)nw * k + n];
            const double2 r = rbar[ld_r * k + (int64_t)a * n2 + e];
            re -= de * r.y;
            im += de * r.x;
        }
        ok[e] = make_double2(re, im);
    }
}

// y[i, :] = sum over the unit's rows and k+q slots S_in[row, m, iT] x[f, :] (3 components), one
// wavefront per unit (i = unit_state_index[u]); lanes take consecutive rows, f = first state of the
// row's code + m for m below the row's slot count.
__global__ void __launch_bounds__(256) bte_spmv_kernel(int nunits, int nbw, int nT, int iT,
                                                      const int64_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ row_code,
                                                      const int64_t *__restrict__ unit_state_index,
                                                      const double *__restrict__ S_in,
                                                      const double *__restrict__ x, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int u = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (u >= nunits) return;  // whole wavefront exits together
    const int64_t r0 = rowptr[u], r1 = rowptr[u + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t row = r0 + lane; row < r1; row += 64) {
        const uint32_t code = row_code[row];
        const int f0 = (int)(code & 0x7ffffffu), nb = (int)(code >> 27);
        const double *sr = S_in + row * nbw * nT + iT;
        for (int m = 0; m < nb; m++) {
            const double sv = sr[(int64_t)m * nT];
            const double *xf = x + 3 * (int64_t)(f0 + m);
            a0 += sv * xf[0];
            a1 += sv * xf[1];
            a2 += sv * xf[2];
        }
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane == 0) {
        double *yi = y + 3 * unit_state_index[u];
        yi[0] = a0;
        yi[1] = a1;
        yi[2] = a2;
    }
}

}  // namespace ephip

using namespace ephip;

static inline unsigned nbk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int ephip_bte_increments(int64_t n, int32_t method, int32_t smearing, const double *in, double *out,
                                    void *stream) {
    EPHIP_REQUIRE(method >= 1 && method <= 6, "bte_increments: method %d outside 1..6", method);
    EPHIP_REQUIRE(smearing == 1 || smearing == 2, "bte_increments: smearing must be 1 (Gaussian) or 2 (Lorentzian)");
    if (n <= 0) return EPHIP_OK;
    EPHIP_REQUIRE(in && out, "bte_increments: null device pointer");
    bte_increments_kernel<<<nbk(n), 256, 0, as_stream(stream)>>>(n, method, smearing, in, out);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_bte_accumulate_units(int32_t nunits, int64_t cs, int32_t nbw, int32_t nm, const int32_t *count,
                                          const int32_t *col_index, const int32_t *ikq, const float *g2,
                                          const int64_t *unit_state_index, const int64_t *state_offset,
                                          const int64_t *bmin, const int64_t *bmax, const double *e_state,
                                          const double *omega, double wtq, int32_t nT, const double *mus,
                                          const double *Ts, const double *etas, int32_t method, int32_t smearing,
                                          double w_cut, int64_t nstates, const int64_t *rowptr, double *S_out,
                                          double *S_in, uint32_t *row_code, void *stream) {
    EPHIP_REQUIRE(nunits >= 0 && nunits <= 65535 && cs >= 0 && nbw >= 1 && nbw <= 31 && nm >= 1 && nT >= 1,
                  "bte_accumulate_units: bad sizes");
    EPHIP_REQUIRE(nstates < (1LL << 27), "bte_accumulate_units: %lld states exceed the row code range",
                  (long long)nstates);
    EPHIP_REQUIRE(method >= 1 && method <= 6 && (smearing == 1 || smearing == 2),
                  "bte_accumulate_units: bad method/smearing");
    if (nunits == 0 || cs == 0) return EPHIP_OK;
    EPHIP_REQUIRE(count && col_index && ikq && g2 && unit_state_index && state_offset && bmin && bmax && e_state &&
                      omega && mus && Ts && etas && rowptr && S_out && S_in && row_code,
                  "bte_accumulate_units: null device pointer");
    dim3 g(nbk(cs * nbw), (unsigned)nunits);
    bte_accumulate_kernel<<<g, 256, 0, as_stream(stream)>>>(nunits, cs, nbw, nm, count, col_index, ikq, g2,
                                                            unit_state_index, state_offset, bmin, bmax, e_state, omega,
                                                            wtq, nT, mus, Ts, etas, method, smearing, w_cut, nstates,
                                                            rowptr, S_out, S_in, row_code);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_band_velocity(int32_t nw, int64_t nstates, const int64_t *state_col, const int32_t *state_band,
                                   const int64_t *state_k, const void *dH, int64_t ld_dh, const void *U, double *v,
                                   void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nstates >= 0 && ld_dh >= 3LL * nw * nw, "band_velocity: bad sizes");
    if (nstates == 0) return EPHIP_OK;
    EPHIP_REQUIRE(state_col && state_band && state_k && dH && U && v, "band_velocity: null device pointer");
    band_velocity_kernel<<<nbk(nstates * 3), 256, 0, as_stream(stream)>>>(
        nw, nstates, state_col, state_band, state_k, static_cast<const double2 *>(dH), ld_dh,
        static_cast<const float2 *>(U), v);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_el_velocity_direct_batched(int32_t nw, int64_t nk, const void *M, int64_t ld_m, const void *U,
                                                const double *E, const void *rbar, int64_t ld_r, void *out,
                                                int64_t ld_out, void *stream) {
    EPHIP_REQUIRE(nw >= 1 && nw <= VEL_NW_MAX && nk >= 0 && nk <= 0x7fffffffLL / 3,
                  "el_velocity_direct_batched: nw=%d (1..%d), nk=%lld", nw, VEL_NW_MAX, (long long)nk);
    EPHIP_REQUIRE(ld_m >= 3LL * nw * nw && ld_out >= 3LL * nw * nw, "el_velocity_direct_batched: ld below 3 nw^2");
    EPHIP_REQUIRE((E == nullptr) == (rbar == nullptr), "el_velocity_direct_batched: E and rbar come together");
    EPHIP_REQUIRE(rbar == nullptr || ld_r >= 3LL * nw * nw, "el_velocity_direct_batched: ld_r below 3 nw^2");
    if (nk == 0) return EPHIP_OK;
    EPHIP_REQUIRE(M && U && out, "el_velocity_direct_batched: null device pointer");
    el_velocity_full_kernel<<<(unsigned)(3 * nk), 256, 0, as_stream(stream)>>>(
        nw, static_cast<const double2 *>(M), ld_m, static_cast<const float2 *>(U), E,
        static_cast<const double2 *>(rbar), ld_r, static_cast<double2 *>(out), ld_out);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_bte_spmv(int32_t nunits, int32_t nbw, int32_t nT, int32_t iT, const int64_t *rowptr,
                              const uint32_t *row_code, const int64_t *unit_state_index, const double *S_in,
                              const double *x, double *y, void *stream) {
    EPHIP_REQUIRE(nunits >= 0 && nbw >= 1 && nT >= 1 && iT >= 0 && iT < nT, "bte_spmv: bad sizes");
    if (nunits == 0) return EPHIP_OK;
    EPHIP_REQUIRE(rowptr && row_code && unit_state_index && S_in && x && y, "bte_spmv: null device pointer");
    bte_spmv_kernel<<<(unsigned)((nunits + 3) / 4), 256, 0, as_stream(stream)>>>(nunits, nbw, nT, iT, rowptr, row_code,
                                                                                unit_state_index, S_in, x, y);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}

extern "C" int ephip_bte_window_accumulate(int32_t fp64, double *S_out, int64_t ld_so, double *S_in, int64_t ld_si1,
                                           int64_t ld_si2, const void *g2vals, const void *wq, const int64_t *imap_i,
                                           const int64_t *imap_f, int64_t ld_f, int64_t nk_f, const int64_t *ikqs,
                                           const double *e_i, const double *e_f, const double *wf, int32_t nT,
                                           const double *mus, const double *Ts, const double *etas, int32_t method,
                                           int32_t smearing, double w_cut, int32_t nbandkq, int32_t nbandk, int32_t nm,
                                           int64_t nq, int64_t i0, uint32_t *dropped, void *stream) {
    EPHIP_REQUIRE(nbandkq >= 0 && nbandk >= 0 && nm >= 0 && nq >= 0 && nT >= 1 && i0 >= 0 && nk_f >= 0,
                  "bte_window_accumulate: bad sizes");
    EPHIP_REQUIRE(ld_f >= (nbandkq > 0 ? nbandkq : 1) && ld_si1 >= 1 && ld_si2 >= 1 && ld_so >= 1,
                  "bte_window_accumulate: bad leading dimensions");
    EPHIP_REQUIRE(method >= 1 && method <= 6 && (smearing == 1 || smearing == 2),
                  "bte_window_accumulate: bad method/smearing");
    const int64_t ne = (int64_t)nbandkq * nbandk * nq;
    if (ne == 0) return EPHIP_OK;
    EPHIP_REQUIRE(S_out && S_in && g2vals && wq && imap_i && imap_f && ikqs && e_i && e_f && wf && mus && Ts && etas,
                  "bte_window_accumulate: null device pointer");
    EPHIP_REQUIRE(ne <= 0x7fffffffLL * 256, "bte_window_accumulate: too many entries");
    if (fp64)
        bte_window_accumulate_kernel<double><<<nbk(ne), 256, 0, as_stream(stream)>>>(
            S_out, ld_so, S_in, ld_si1, ld_si2, static_cast<const double *>(g2vals), static_cast<const double *>(wq),
            imap_i, imap_f, ld_f, nk_f, ikqs, e_i, e_f, wf, nT, mus, Ts, etas, method, smearing, w_cut, nbandkq, nbandk, nm, nq,
            i0, dropped);
    else
        bte_window_accumulate_kernel<float><<<nbk(ne), 256, 0, as_stream(stream)>>>(
            S_out, ld_so, S_in, ld_si1, ld_si2, static_cast<const float *>(g2vals), static_cast<const float *>(wq),
            imap_i, imap_f, ld_f, nk_f, ikqs, e_i, e_f, wf, nT, mus, Ts, etas, method, smearing, w_cut, nbandkq, nbandk, nm, nq,
            i0, dropped);
    EPHIP_LAUNCHED();
    return EPHIP_OK;
}
