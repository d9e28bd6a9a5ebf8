/*
 * ephip.h -- C-ABI of libephip, the MI355X (gfx950) Wannier->Bloch engine.
 *
 * Drop-in boundary for ElectronPhonon.jl's device method table (src/common/gpu_utils.jl:1-78 and the
 * CUDA extension ext/ElectronPhononCUDAExt.jl). Each entry names the reference interface it replaces
 * (file:line in the ElectronPhonon.jl tree). A binding (Julia ccall, Python ctypes) passes plain
 * device pointers and sizes; no runtime types cross the boundary.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer unless marked (host). Arrays are column-major with the
 *     reference's index order; complex values are interleaved (re, im) pairs: "c32" = 2 x float,
 *     "c64" = 2 x double. Leading dimensions are in elements.
 *   - Indices are 0-based, except window band indices (bmin/bmax), which are 1-based like the
 *     reference's band ranges, with 0 meaning "no band in the window".
 *   - O(k) = sum_R exp(+2 pi i k.R) O(R), k in crystal coordinates (3 doubles per point), R integer
 *     (3 int32 per vector) -- src/wannier/WannierInterpolator.jl:127,135.
 *   - `stream` is a hipStream_t (NULL = default stream). All calls are asynchronous on `stream`
 *     unless stated otherwise; errors in argument checks are reported synchronously.
 *   - Return value: EPHIP_OK or an error code; ephip_last_error() returns the message of the last
 *     failing call on the calling thread.
 */
#ifndef EPHIP_H
#define EPHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPHIP_OK 0
#define EPHIP_ERR_ARG 1
#define EPHIP_ERR_HIP 2
#define EPHIP_ERR_NOCONV 3

/* Precision of a Fourier contraction (inputs and outputs are c32 in both). */
#define EPHIP_PREC_FP32 0   /* FP32 MFMA, FP32 accumulation */
#define EPHIP_PREC_BF16X3 1 /* split-BF16 hi+lo operands, three BF16 MFMA products, FP32 accumulation */

/* Message of the last failing call on this thread ("" if none). Host-side, synchronous. */
const char *ephip_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int ephip_version(void);

/* free_bytes(::GPUBackend) -- src/common/gpu_utils.jl:51, ext/ElectronPhononCUDAExt.jl:36.
 * (host) free/total device memory in bytes of the current device. Synchronous. */
int ephip_free_bytes(int64_t *free_bytes, int64_t *total_bytes);

/* synchronize(::GPUBackend) -- src/common/gpu_utils.jl:60, ext/ElectronPhononCUDAExt.jl:37.
 * Blocks until `stream` is drained, and reports any asynchronous kernel fault. */
int ephip_synchronize(void *stream);

/* fourier_phase! -- src/wannier/batched_fourier_core.jl:24-30.
 * phase[r + nr*j] (c64) = exp(2 pi i R_r . x_j) for irvec (3 x nr int32) and xk (3 x nk double). */
int ephip_fourier_phase_c64(int32_t nr, const int32_t *irvec, int32_t nk, const double *xk, void *phase,
                            void *stream);

/* fourier_batched! / get_fourier_batched! -- src/wannier/batched_fourier_core.jl:117-139,
 * src/wannier/batched_interpolator.jl:200-214 (the ZGEMM at :136 of batched_fourier_core.jl).
 * out[d + ldout*j] = sum_r op[d + ldop*r] * exp(2 pi i R_r . x_j), d < ndata, j < nk.
 * op: c32 (ldop x nr), out: c32 (ldout x nk). The phase tile is generated on chip from the FP64
 * argument R.x reduced mod 1; precision = EPHIP_PREC_FP32 or EPHIP_PREC_BF16X3. */
int ephip_fourier_c32(int64_t ndata, int32_t nr, const void *op, int64_t ldop, const int32_t *irvec,
                      int32_t nk, const double *xk, void *out, int64_t ldout, int32_t precision,
                      void *stream);

/* Same contraction in FP64 (c64 in and out), FP64 MFMA. Used where the reference's FP64 result is
 * needed on the device (band energies for the window filter, dynamical matrices). */
int ephip_fourier_c64(int64_t ndata, int32_t nr, const void *op, int64_t ldop, const int32_t *irvec,
                      int32_t nk, const double *xk, void *out, int64_t ldout, void *stream);

/* eigvals_batched / eigen_batched -- src/wannier_to_bloch_batched.jl:30-65,
 * ext/ElectronPhononCUDAExt.jl:80-115. Hermitian eigensolve of nb matrices H (n x n x nb, upper
 * triangle read), n <= 32, one wavefront per matrix, FP32 arithmetic. H is c64 if h_is_c64 else c32
 * (not modified). E: float (n x nb) ascending. U: c32 (n x n x nb) eigenvectors, or NULL for values
 * only (U may alias a c32 H). info: int32 per matrix (0 = converged) or NULL. */
int ephip_heev_c32(int32_t n, int64_t nb, const void *H, int32_t h_is_c64, void *U, float *E,
                   int32_t *info, void *stream);

/* FP64 refinement of ephip_heev_c32 results against the FP64 matrices H64 (c64, n x n x nb): the FP32
 * eigenvectors U are orthonormalised in FP64, H64 is projected onto them (M = Q^H H Q, unitarily similar
 * to H and nearly diagonal) and M is diagonalised by parallel cyclic Jacobi sweeps in FP64 until the
 * off-diagonal Frobenius norm is <= tol x the Frobenius norm (tol = 0: 1e-15). U (c32) is overwritten
 * with the refined vectors; E64: double (n x nb) ascending. n <= 32. E32 is not read. Gives the FP64
 * accuracy of the reference's zheev eigenvalues (src/common/AllocatedLAPACK.jl:44-108), also for
 * degenerate bands. */
int ephip_heev_refine(int32_t n, int64_t nb, const void *H64, void *U, const float *E32, double *E64,
                      double tol, void *stream);

/* Degenerate-band gauge fixing of the per-k eigensolve solve_eigen_el! (src/diagonalize.jl:11-18, 47-85),
 * in place on eigenpairs already computed (E64 double (n x nb) ascending, U c32 (n x n x nb)): bands are
 * grouped as the reference groups them (a group starting at band a runs while |E_i - E_a| <= cutoff); in
 * each group of m > 1 bands, U_g <- U_g V where V diagonalises U_g^H P U_g (P: the reference's fixed
 * perturbation matrix) with ascending eigenvalues. Matrices with no neighbouring pair closer than cutoff
 * are untouched. cutoff: the reference's electron_degen_cutoff is 0.01 meV (src/common/constants.jl:6),
 * 1e-5 in eV. n <= 32. */
int ephip_degenerate_gauge_fix(int32_t n, int64_t nb, const double *E64, void *U, double cutoff, void *stream);

/* eigen_batched(Hk) -- src/wannier_to_bloch_batched.jl:43-65 (CUDA method ext/ElectronPhononCUDAExt.jl:101-115),
 * at the FP64 accuracy of the reference's zheev path: ephip_heev_c32 then ephip_heev_refine. H64: c64
 * (n x n x nb), upper triangle read, not modified. E64: double (n x nb) ascending; U: c32 (n x n x nb)
 * eigenvectors; work: float (n x nb) device workspace (the FP32 eigenvalues); info: int32 per matrix
 * (0 = the FP32 solve converged) or NULL. n <= 32. */
int ephip_eigen_batched(int32_t n, int64_t nb, const void *H64, double *E64, void *U, float *work, int32_t *info,
                        void *stream);

/* eigvals_batched(Hk) -- src/wannier_to_bloch_batched.jl:30-41 (CUDA method ext/ElectronPhononCUDAExt.jl:80-93):
 * eigenvalues only, E64 double (n x nb) ascending, FP64-refined as in ephip_eigen_batched. work: device
 * workspace of nb (8 n^2 + 4 n) bytes, 8-byte aligned (the FP32 eigenvectors the refinement needs, then the
 * FP32 eigenvalues). info as in ephip_eigen_batched. n <= 32. */
int ephip_eigvals_batched(int32_t n, int64_t nb, const void *H64, double *E64, void *work, int32_t *info,
                          void *stream);

/* eigvals_batched without a workspace argument and without any allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same results as ephip_eigvals_batched; its workspace is a
 * file-scope device array of 16 MiB per GPU (held until the process ends), so nb <= floor(16 MiB /
 * (8 n^2 + 4 n)) (7943 matrices at n = 16); larger nb returns EPHIP_ERR_ARG. Calls that overlap on one GPU
 * share the array and race: serialise them (one stream, or events). n <= 32. */
int ephip_eigvals_batched_noalloc(int32_t n, int64_t nb, const void *H64, double *E64, int32_t *info, void *stream);

/* _filter_kpoints / inside_window -- src/filter.jl:19, :40-100 (device branch :46-72).
 * E: double (nw x nk) ascending. keep[k] = any band in [wmin, wmax]; bmin/bmax: 1-based
 * searchsortedfirst(E, wmin) / searchsortedlast(E, wmax), 0 when keep = 0; nbelow[k] = bmin - 1
 * counted as searchsortedfirst - 1 (may be NULL). */
int ephip_window_filter(int32_t nw, int64_t nk, const double *E, double wmin, double wmax, uint8_t *keep,
                        int64_t *bmin, int64_t *bmax, int64_t *nbelow, void *stream);

/* Phonon dynamical-matrix mass scaling -- src/compute_states.jl:295-297 (device phonon states),
 * src/wannier_to_bloch.jl:199-202: D (c64, nm x nm x nq) /= sqrt(m_i) sqrt(m_j); mass: double (nm). */
int ephip_ph_mass_scale(int32_t nm, int64_t nq, void *D, const double *mass, void *stream);

/* Phonon frequencies and eigenvector mass factor -- src/diagonalize.jl:118-126,
 * src/compute_states.jl:298-300: omega = sign(w2) sqrt(|w2|) (double, nm x nq); U (c32, may be NULL)
 * rows divided by sqrt(m_i). */
int ephip_ph_finalize(int32_t nm, int64_t nq, const double *w2, double *omega, void *U, const double *mass,
                      void *stream);

/* k-point index map -- _grid_coords_reduced / _hash_xk / _ik_from_hash / xk_to_ik
 * (src/common/kpoints.jl:614-656) as used by _fill_iqs! (src/calculator/run_eph_over_k_and_kq.jl:565-582).
 * For each point x_k (xk, 3 x nk double) computes the grid hash of x_k + dx on the (host) ngrid[3]
 * grid shifted by (host) shift[3] (NULL = 0), then out[k] = table[hash] - 1 with table a dense
 * int64 map hash -> 1-based index (0 = absent, so out = -1), or out[k] = hash if table is NULL.
 * dx: (host) 3 doubles or NULL. */
int ephip_kq_index(int64_t nk, const double *xk, const double *dx, const int32_t *ngrid, const double *shift,
                   const int64_t *table, int64_t *out, void *stream);

/* Phonon-mode rotation of get_eph_RR_to_Rq! -- src/wannier_to_bloch.jl:292-297, batched over nq:
 * out[b + nb*(nu' + nm*Re) + ldout*q] = sum_nu g[b + nb*(nu + nm*Re) + ldg*q] * uph[nu + nm*nu' + nm*nm*q]
 * with nb = nbasis (= nw*nw), all c32; nm <= 48. g and out must not alias. */
int ephip_eph_rotate_modes(int64_t nbasis, int32_t nm, int32_t nr_el, int32_t nq, const void *g, int64_t ldg,
                           const void *uph, void *out, int64_t ldout, void *stream);

/* eph_apply_rotations_rqkq! + g2 fold -- src/wannier_to_bloch_batched.jl:360-377 and :451-453,
 * ext/ElectronPhononCUDAExt.jl:214-258, with the outer-q loop's window masks
 * (src/calculator/run_eph_over_q_and_k.jl:555-585). For batch column t < ncol:
 *   g column t: c32 (nw, nw, nm) at g + ldg*t, index legend g[i, j, nu] (i: k+q side, j: k side);
 *   k = ik[t], k+q = ikq[t] index the states: U c32 (nw, nw, nstates), bmin/bmax int64 (nstates);
 *   ep[m, n, nu] = sum_ij conj(U_kq[i, m]) g[i, j, nu] U_k[j, n] for m in the k+q window, n in the k
 *   window; g2 = |ep|^2 / (2 omega[nu]) (omega: double nm, no cutoff, as the reference).
 * Output block of column t: slot [ms + nbw*(ns + nbw*nu)], ms = m - (bmin[kq]-1), ns = n - (bmin[k]-1),
 * zero outside the windows; g2 (float) and/or ep (c32) may be NULL. nbw must be >= every window width. */
int ephip_eph_rotate_rqkq_g2(int32_t nw, int32_t nm, int64_t ncol, const void *g, int64_t ldg, const int64_t *ik,
                             const int64_t *ikq, const void *U, const int64_t *bmin, const int64_t *bmax,
                             const double *omega, int32_t nbw, float *g2, void *ep, void *stream);

/* Batched fourier_batched!: nunits independent contractions in one launch (the list-batched Fourier of
 * get_eph_kR_to_kq_batched!, src/wannier_to_bloch_batched.jl:245-273, over many k at once). Unit u uses
 * op + u*op_stride (c32, ldop x nr) and writes out + u*out_stride (c32, ldout x ncol[u]); it has
 * ncol[u] columns (device int32, or nk_max for every unit if NULL); column j is the point
 * xk[col_index[u*col_stride + j]] (device int32, or xk[j] if NULL). nunits <= 65535. */
int ephip_fourier_batched_c32(int64_t ndata, int32_t nr, const void *op, int64_t ldop, int64_t op_stride,
                              const int32_t *irvec, const double *xk, int32_t nunits, int32_t nk_max,
                              const int32_t *ncol, const int32_t *col_index, int64_t col_stride, void *out,
                              int64_t ldout, int64_t out_stride, int32_t precision, void *stream);

/* The two outer-momentum layouts of the e-ph tensor (src/model_new.jl:551-565): "ph" (b, Re) x Rp to
 * "el" (b, Rp) x Re, b < nbasis (= nw*nw*nmodes), c32. ph and el must not alias. */
int ephip_eph_ph_to_el(int64_t nbasis, int32_t nre, int32_t nrp, const void *ph, void *el, void *stream);

/* k-side projection of get_eph_RR_to_kR! (src/wannier_to_bloch.jl:381-388; batched
 * src/wannier_to_bloch_batched.jl:190-197) onto single in-window bands, with the outer-k window
 * projection of src/calculator/run_eph_over_k_and_kq.jl:657-659. G: Fourier output of the "el" tensor
 * for a k batch, column t = g(k_t)[i + nw*(j + nw*(nu + nm*Rp))] (c32, ldg per column). Unit u:
 * A[u*nw*nm*nrp + (i + nw*nu) + nw*nm*Rp] = sum_j G[ldg*ucol[u] + ...j...] * U[ustate[u]][j, uband[u]]
 * with U c32 (nw, nw, nstates) eigenvectors and uband 0-based. */
int ephip_eph_project_kR(int32_t nw, int32_t nm, int32_t nrp, int32_t nunits, const void *G, int64_t ldg,
                         const int32_t *ucol, const int64_t *ustate, const int32_t *uband, const void *U, void *A,
                         void *stream);

/* k+q lookup and window skip for the outer-k loop (_fill_iqs!, src/calculator/run_eph_over_k_and_kq.jl:
 * 565-582, with the hash of src/common/kpoints.jl:614-651 on the unshifted k grid ngrid[3] (host)).
 * For unit u at k = xk[ustate[u]] and every q = xq[c], c < nq: k+q's grid index h; keeps the q with
 * bmin[h] > 0 (an in-window band at k+q), in q order: col_index[u*cs + i] = c, ikq[u*cs + i] = h,
 * count[u] = number kept. cs >= nq. Pairs dropped here are exactly zero in the reference. */
int ephip_kq_compact(int32_t nunits, const int64_t *ustate, const double *xk, int32_t nq, const double *xq,
                     const int32_t *ngrid, const int64_t *bmin, int64_t cs, int32_t *col_index, int32_t *ikq,
                     int32_t *count, void *stream);

/* kR->kq rotations + g2 fold (eph_apply_rotations!, src/wannier_to_bloch_batched.jl:442-455;
 * src/wannier_to_bloch.jl:428-440; CUDA ext/ElectronPhononCUDAExt.jl:146-201) on the k+q window:
 * for unit u and kept column c < count[u], with Gq column = Gq + u*gstride + c*ldg holding
 * g[i + nw*nu] (c32), q = col_index[u*cs + c], k+q state = ikq[u*cs + c]:
 *   ep[m, nu'] = sum_i sum_nu conj(U_kq[i, bmin-1+m]) g[i + nw*nu] uph(q)[nu, nu'],
 *   g2 = |ep|^2 / (2 omega(q)[nu'])   (uph c32 (nm, nm, nq), omega double (nm, nq)).
 * Output row u*cs + c holds nbw*nm floats [m*nm + nu'] (zero for m beyond the k+q window); g2 and/or
 * ep (c32, same layout) may be NULL. nunits <= 65535. */
int ephip_eph_rotate_kq_g2(int32_t nw, int32_t nm, int32_t nbw, int32_t nunits, int64_t cs, const int32_t *count,
                           const int32_t *col_index, const int32_t *ikq, const void *Gq, int64_t ldg,
                           int64_t gstride, const void *U, const int64_t *bmin, const int64_t *bmax, const void *uph,
                           const double *omega, float *g2, void *ep, void *stream);

/* Pre-packed split-BF16 Fourier (no counterpart in the reference; the split-BF16 form of the ZGEMM of
 * src/wannier/batched_fourier_core.jl:117-139 with both operands split once):
 *  - ephip_ft_pack_operand: nunits complex64 operands (ndata x nr, leading dimension ldop, unit u at
 *    op + u*op_stride) -> apk, bf16, nunits*nrs*ndata*64 values, nrs = ceil(nr/16): for stage s and row
 *    d, 64 values at ((u*nrs + s)*ndata + d)*64 = re_hi[16] im_hi[16] re_lo[16] im_lo[16] of R 16s..16s+15
 *    (zero past nr).
 *  - ephip_ft_pack_phases: exp(2 pi i R.x_p) for npts points -> bpk, bf16, nrs*npts*64 values: for stage s
 *    and point p, 64 values at (s*npts + p)*64 = hi{im[16], re[16]} lo{im[16], re[16]}.
 *  - ephip_fourier_packed_bf16x3: out[d + ldout*j] (+ u*out_stride) = sum_r op_u[d, r] e^{2 pi i R_r.x_p},
 *    p = col_index[u*col_stride + j] (or j, then nk_max <= npts), j < ncol[u] (or nk_max); FP32
 *    accumulation of hi*hi + hi*lo + lo*hi on 16x16x32 bf16 MFMAs. */
int ephip_ft_pack_operand(int64_t ndata, int32_t nr, const void *op, int64_t ldop, int32_t nunits, int64_t op_stride,
                          void *apk, void *stream);
int ephip_ft_pack_phases(int32_t nr, const int32_t *irvec, int64_t npts, const double *xk, void *bpk, void *stream);
/* As ephip_ft_pack_operand for one operand, with the two fastest row indices (each of size nb) swapped:
 * packed row j + nb*(i + nb*rest) holds operand row i + nb*(j + nb*rest); ndata a multiple of nb^2. */
int ephip_ft_pack_operand_swapped(int64_t ndata, int32_t nr, int32_t nb, const void *op, int64_t ldop, void *apk,
                                  void *stream);
int ephip_fourier_packed_bf16x3(int64_t ndata, int32_t nr, const void *apk, int32_t nunits, const void *bpk,
                                int64_t npts, int32_t nk_max, const int32_t *ncol, const int32_t *col_index,
                                int64_t col_stride, void *out, int64_t ldout, int64_t out_stride, void *stream);

/* RR->kR Fourier over Re fused with the k-side window projection (get_eph_RR_to_kR! of
 * src/wannier_to_bloch_batched.jl:178-197 followed by the outer-k projection of
 * src/calculator/run_eph_over_k_and_kq.jl:657-659, replacing ephip_fourier_packed_bf16x3 +
 * ephip_eph_project_kR) for nw = 16, split-BF16: apk = the "el" tensor (rows i + nw*(j + nw*(nu + nm*Rp)),
 * nr = nre) packed by ephip_ft_pack_operand_swapped with nb = 16; bpk = packed phases of the batch's nk
 * k-points (npts >= nk). kus: int32, nk+1 entries, units of batch column t are kus[t]-kus[0] ..
 * kus[t+1]-kus[0]-1 (one per in-window band, at most nbw per column), with ustate (int64) and 0-based
 * uband (int32) indexed from the batch's first unit. Writes A[u*lda + (i + nw*nu) + nw*nm*Rp] =
 * sum_j g(k_t)[i, j, nu, Rp] U[ustate[u]][j, uband[u]] as ephip_eph_project_kR does; g(k, Rp) is not
 * stored. nw^2*nm*nrp must be a multiple of 192. */
int ephip_eph_RR_to_kR_project_packed_bf16x3(int32_t nw, int32_t nm, int32_t nrp, int32_t nre, const void *apk,
                                             const void *bpk, int64_t npts, int32_t nk, const int32_t *kus,
                                             const int64_t *ustate, const int32_t *uband, const void *U, int32_t nbw,
                                             void *A, int64_t lda, void *stream);

/* Staged (GridOpt) RR->kR over Re with the k-side window projection fused: the device form of
 * src/wannier/GridOpt.jl (gridopt_set23! :73-107, gridopt_set3! :110-131, gridopt_get3! :133-139) and
 * of the batched k3 step of src/wannier/batched_gridopt_interpolator.jl, feeding the outer-k projection
 * of src/calculator/run_eph_over_k_and_kq.jl:657-659. FP32 throughout, phases from FP64 arguments.
 * The Re list is embedded in its bounding box (r1min + a, r2min + b, r3min + c), a < n1 <= 32,
 * b < n2 <= 8, c < n3 <= 8; rmap (int32, n1*n2*n3) gives the operand column of each box point, or -1.
 *  - ephip_gridopt_stage1: op1[(b + np*c)*nd + d'] = sum_a e^{2 pi i x1 (r1min + a)} op[d + ldop*rmap(a,b,c)]
 *    for the nd rows of op (c64, the "el" layout rows (i, j, nu, Rp)), written j-fastest
 *    (d' = j + nb*(i + nb*rest) for d = i + nb*(j + nb*rest)) when nb > 1; the face is padded to
 *    np x np (np >= n2, n3; zeros outside the box).
 *  - ephip_gridopt_kR_project: for the surviving k k0 <= k < k1 (xk: their crystal coordinates, FP64,
 *    3 per k; op1 padded to np x np, 3 <= np <= 8) on the lines l0 <= l < l1 (line_start: first surviving k of each (k1, k2) line, all in
 *    the k1 plane whose op1 is given; at most max_line_k <= 64 k per line): op2 = sum_b e^{2 pi i k2 r2}
 *    op1, g(k) = sum_c e^{2 pi i k3 r3} op2, and for each unit u of k (kus, ustate, uband as in
 *    ephip_eph_RR_to_kR_project_packed_bf16x3) A[u*lda + (i + nw*nu) + nw*nm*Rp] =
 *    sum_j g(k)[i, j, nu, Rp] U[ustate[u]][j, uband[u]]. nw = 8 or 16, nbw <= 4, nb = nw in stage 1. */
int ephip_gridopt_stage1(int64_t nd, int32_t n1, int32_t n2, int32_t n3, int32_t np, int32_t r1min,
                         const int32_t *rmap, const void *op, int64_t ldop, double x1, int32_t nb, void *op1,
                         void *stream);
/* ephip_gridopt_stage1 for nplanes <= 8 k1 planes in one pass over op (each operand element feeds every
 * plane): x1 (host array, nplanes) holds their k1 coordinates and op1 (host array of nplanes device
 * pointers) their outputs, each laid out as for ephip_gridopt_stage1. */
int ephip_gridopt_stage1_planes(int64_t nd, int32_t n1, int32_t n2, int32_t n3, int32_t np, int32_t r1min,
                                const int32_t *rmap, const void *op, int64_t ldop, int32_t nplanes, const double *x1,
                                int32_t nb, void *const *op1, void *stream);
int ephip_gridopt_kR_project(int32_t nw, int32_t nm, int32_t nrp, int32_t np, int32_t r2min, int32_t r3min,
                             const void *op1, const double *xk, int64_t k0, int64_t k1, int32_t l0, int32_t l1,
                             const int32_t *line_start, int32_t max_line_k, const int32_t *kus,
                             const int64_t *ustate, const int32_t *uband, const void *U, int32_t nbw, void *A,
                             int64_t lda, void *stream);
/* Stages 2/3 and the projection on the matrix cores (16x16x4 FP32 MFMAs), nw = 16: along a line,
 * A[(k, n)][G] = sum_{j, c} U_k[j, n] e^{2 pi i k3 (r3min + c)} op2[(j, c)][G]. op1 in the natural row order
 * d = i + nw*(j + nw*rest) (ephip_gridopt_stage1 with nb = 1). The per-batch tables come from
 * ephip_gridopt_kR_tables for the batch's k k0 <= k < k1, its lines lb0 <= l < lb1 and its units:
 * ucol (complex64, units x nw): the U column of each unit (unit u relative to kus[k0]: U[ustate[u]][:, uband[u]]);
 * ph3 (complex64, (k1 - k0) x np): e^{2 pi i k3 (r3min + c)}; ph2 (complex64, (lb1 - lb0) x np):
 * e^{2 pi i k2 (r2min + b)} of each line. ephip_gridopt_kR_project_mfma then covers the lines
 * lb0 <= l0 <= l < l1 of one k1 plane (as ephip_gridopt_kR_project; its units are consecutive, at most
 * nbw per k). max_line_units (<= max_line_k*nbw; 0 means max_line_k*nbw) bounds the units of any line and
 * sizes the kernel's LDS. r2min: the box's first r2 (the tables' r2min); when the box spans -h .. h along
 * axis 2 (odd np) the stage-2 sums run in the folded form (pairs of opposite offsets; tuner kind 8).
 * Output as ephip_gridopt_kR_project. */
int ephip_gridopt_kR_tables(int32_t nw, int32_t np, int32_t r2min, int32_t r3min, const double *xk, int64_t k0,
                            int64_t k1, int32_t lb0, int32_t lb1, const int32_t *line_start, int32_t nunits,
                            const int64_t *ustate, const int32_t *uband, const void *U, void *ucol, void *ph2,
                            void *ph3, void *stream);
int ephip_gridopt_kR_project_mfma(int32_t nw, int32_t nm, int32_t nrp, int32_t np, int32_t r2min, const void *op1,
                                  int64_t k0,
                                  int64_t k1, int32_t l0, int32_t l1, int32_t lb0, const int32_t *line_start,
                                  int32_t max_line_k, int32_t max_line_units, const int32_t *kus, const void *ucol,
                                  const void *ph2, const void *ph3, int32_t nbw, void *A, int64_t lda,
                                  void *stream);

/* Staged (GridOpt) kR->kq over Rp for the q of a uniform grid qgrid = (m1, m2, m3) (q index
 * (q1*m2 + q2)*m3 + q3 at q = (q1/m1, q2/m2, q3/m3)), fused with the k+q-window and phonon rotations and
 * g2: the same output as ephip_eph_kq_rotate_packed_bf16x3 (the get_eph_kR_to_kq_batched! +
 * eph_apply_rotations! + g2 sequence of src/wannier_to_bloch_batched.jl:245-273 and :442-455), computed
 * in FP32 by per-axis partial sums (src/wannier/GridOpt.jl) over the unit operators A (c64, unit u at
 * A + u*lda, element (G, Rp) at G + ldr*Rp, G = i + nw*nu). The Rp list is embedded in its bounding box
 * (p1min + a, p2min + b, p3min + c), a < n1 <= 32, b, c < np (3..8); rmap (int32, 8*n1*np, index
 * b + 8*(a + n1*c), rows padded to 8) gives the Rp index or -1. line_off (int32, nunits x (m1*m2 + 1)) holds each unit's
 * first kept-list position per q line, from ephip_kq_line_offsets (lower bounds in the sorted kept
 * lists of ephip_kq_compact). nw = 16, nm <= 16, nbw <= 4, m1 <= 65535, m2, m3 <= 64 (the kept q of a (unit, q1)
 * slab are taken in segments of whole q lines of at most 1024 entries; the MFMA form, tuner kind 6 = 1,
 * runs for m2*m3 <= 1024 and the vector form otherwise). dense_strides: host
 * int32[3] (s1, s2, s3) when the box is dense (n1 = np and every Rp present, Rp index = s1 a + s2 b + s3 c;
 * the operator is then addressed without rmap), or NULL. ep: NULL, or complex64 in the layout of g2 (one
 * complex per g2 entry): the rotated matrix elements themselves (zero past the k+q window). b1: NULL (each
 * (unit, q1) workgroup forms its stage-1 sums from the unit operator), or a complex64 workspace of
 * nunits*m1*np*np*nw*nm elements: a pre-pass then forms stage 1 for every (unit, q1) at once (one read of
 * each unit operator) and the workgroups read their slices; cube boxes only (n1 = np, odd np <= 7),
 * m1 <= 64. */
int ephip_kq_line_offsets(int32_t nunits, const int32_t *count, const int32_t *col_index, int64_t cs, int32_t m3,
                          int32_t nlines, int32_t *line_off, void *stream);
int ephip_gridopt_kq_rotate(int32_t nw, int32_t nm, int32_t nbw, int32_t nunits, int32_t n1, int32_t np,
                            int32_t p1min, int32_t p2min, int32_t p3min, const int32_t *rmap, const int32_t *qgrid,
                            const void *A, int64_t lda, int64_t ldr, const int32_t *line_off,
                            const int32_t *col_index, const int32_t *ikq, int64_t cs, const void *U,
                            const int64_t *bmin, const int64_t *bmax, const void *uph, const double *omega,
                            const int32_t *dense_strides, float *g2, void *ep, void *b1, void *stream);

/* kR->kq Fourier over Rp fused with the k+q-window and phonon rotations and g2 (the
 * get_eph_kR_to_kq_batched! + eph_apply_rotations! + g2 sequence of src/wannier_to_bloch_batched.jl:
 * 245-273 and :442-455 in one kernel) for nw = 16, nm = 12, nbw <= 4, split-BF16: apk holds the
 * nunits packed unit operands (rows i + 16 nu, nr = nr_ep), bpk the packed q-grid phase table of npts
 * points; count/col_index/ikq (row stride cs) are the kept-q lists of ephip_kq_compact; U, bmin, bmax,
 * uph, omega as in ephip_eph_rotate_kq_g2. Writes g2 (nbw x nm per kept (unit, q), [m, nu]) only.
 * irvec (int32 nr x 3, the Rp list) and qgrid (3 ints, n1 n2 n3 = npts, third index fastest, no shift)
 * are read by the tuner variant that generates the phases on the fly (kind 4, variant 4); they may be
 * NULL otherwise. max_count (<= 0: cs) sizes the grid of the eight-wave variants (kind 4, 5..12), whose
 * workgroups loop over the column tiles of their unit, so it is a hint (an upper bound on count[u] gives
 * no idle workgroups) and a longer list is still processed in full. */
int ephip_eph_kq_rotate_packed_bf16x3(int32_t nw, int32_t nm, int32_t nbw, int32_t nr, const void *apk,
                                      int32_t nunits, const void *bpk, int64_t npts, int64_t cs, int64_t max_count,
                                      const int32_t *count, const int32_t *col_index, const int32_t *ikq,
                                      const void *U, const int64_t *bmin, const int64_t *bmax, const void *uph,
                                      const double *omega, float *g2, const int32_t *irvec, const int32_t *qgrid,
                                      void *stream);

/* Kernel variant of the Fourier contractions, chosen by the tuner (ephip/tune.py) from timings on the
 * live shapes. kind: 0 = FP32 with nk > 64, 1 = FP32 with nk <= 64, 2 = FP64; variant: 0 = single-
 * buffered LDS stages of 16 R, 1 = double-buffered stages of 16 R, 2 = double-buffered stages of 8 R.
 * kind 3 = pre-packed split-BF16 with ndata >= 128 (except 192-row operands, always 192 x 64);
 * variant: 0 = 128 x 128 tiles, 1 = 128 x 128 with the next stage's operand prefetched into registers,
 * 2 = 192 x 64 tiles (ndata a multiple of 192), 3 = 192 x 64 tiles in XCD-grouped workgroup order (the
 * column tiles of a row tile on one XCD, sharing its L2; single-unit calls), 4 = as 3 with the next
 * stage's phase pieces written to LDS halfway through the MFMA sequence, 5 = as 4 with the pipelined
 * main loop (phase pieces two stages ahead, operand fragments one stage ahead, no register copies),
 * 6 = as 3 pipelined, 7 = eight-wave workgroups on 192 x 128 tiles with the operand tile staged
 * through LDS once for both column halves (single-unit calls without column lists). kind 4 = the fused kR->kq rotation kernel; variant 0 = phase pieces written
 * after the MFMA sequence, 1 = halfway through it, 2 = pipelined, 3 = pipelined and halfway, 4 = phases
 * of the uniform q-grid generated on the fly from per-axis tables (no phase-table reads), 5 = eight-
 * wave workgroups on 128 kept columns with the unit's operand staged through LDS once for both halves,
 * 6, 7, 8 = as 5 with the units dealt to the XCDs and walked in groups of 2, 4, 8, column tile by tile,
 * 9, 10 = as 7, 8 with the epilogue's column metadata read before the main loop and the U_kq / u_ph
 * values of the next column block fetched ahead, 11, 12 = as 9, 10 with the main loop's global loads
 * issued two stages ahead. kind 5 = the fused RR->kR + projection kernel; variant 0 = loads one stage
 * ahead, 1 = two stages ahead. kind 6 = the staged kR->kq kernel (ephip_gridopt_kq_rotate); variant 0 =
 * stage 3 and the k+q-window rotation on the vector ALUs, 1 = as one contraction on FP32 MFMAs. kind 7 =
 * the staged kR->kq kernel's workgroup order; variant 0 = q1 fastest, unit by unit, 1 = XCD-grouped (each
 * XCD takes one contiguous run of (unit, q1) pairs, so a unit's workgroups share one L2). kind 8 = the staged
 * kR->kq kernel's stage-2/3 sums; variant 0 = the folded form (pairs of opposite offsets, cos and sin
 * factors) when the Rp box spans -h .. h along axes 2 and 3 (odd np), the general form otherwise;
 * 1 = always the general form.
 * Process-wide setting (no counterpart in the reference, whose tuning constants are hard-coded:
 * src/wannier_to_bloch_batched.jl:397-406). Host-side, synchronous. */
int ephip_ft_set_variant(int32_t kind, int32_t variant);
int ephip_ft_get_variant(int32_t kind); /* returns the variant, or -1 for a bad kind */

/* ---- Band-energy surrogate (no counterpart in the reference: the engine's own learned pre-filter,
 * used ahead of the exact window filter of src/filter.jl:40-100). All arrays row-major FP32 unless
 * noted. ---- */

/* Periodic Fourier features of k: f[n, j] = cos(2 pi m_j.k_n), f[n, nf2 + j] = sin(...);
 * xk double (3 x npts), mvec int32 (3 x nf2), f float (2 nf2 x npts). */
int ephip_mlp_features(int64_t npts, const double *xk, int32_t nf2, const int32_t *mvec, float *f, void *stream);

/* C (M x N) = op(A) op(B) (+ bias[n]) (+ beta C), row-major with leading dimensions; ta/tb transpose. */
int ephip_gemm_f32(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float *A, int64_t lda,
                   const float *B, int64_t ldb, float *C, int64_t ldc, const float *bias, float beta, void *stream);

/* Split-K form of ephip_gemm_f32 (no bias/beta): slice z of nsplit sums its share of K into
 * Cpart + z*M*ldc; reduce the nsplit partials with ephip_colsum(nsplit, M*ldc, ...). */
int ephip_gemm_f32_splitk(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const float *A, int64_t lda,
                          const float *B, int64_t ldb, float *Cpart, int64_t ldc, int32_t nsplit, void *stream);

/* a = silu(z); gz = ga * silu'(z). */
int ephip_silu_fwd(int64_t n, const float *z, float *a, void *stream);
int ephip_silu_bwd(int64_t n, const float *z, const float *ga, float *gz, void *stream);

/* gy = 2 (y - t) scale; partial[b] (double) = sum of (y - t)^2 over block b of 256 elements. */
int ephip_mse_grad(int64_t n, const float *y, const float *t, float scale, float *gy, double *partial, void *stream);

/* out[c] = sum over rows of x[r*ldx + c]. */
int ephip_colsum(int32_t rows, int32_t cols, const float *x, int64_t ldx, float *out, void *stream);

/* One Adam step (bias-corrected, step counts from 1) on n parameters. */
int ephip_adam_step(int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1, float beta2,
                    float eps, int32_t step, void *stream);

/* y (bf16) = x (float) rounded to nearest-even. */
int ephip_f32_to_bf16(int64_t n, const float *x, void *y, void *stream);

/* Fused BF16 inference of the MLP (features -> W1,b1 -> SiLU -> W2,b2 -> SiLU -> W3,b3) on bf16 MFMAs
 * with FP32 accumulation, and the pruning test: keep[p] = 1 if any predicted band lies in
 * [wmin - margin, wmax + margin], else 0. W1 (hidden x 2nf2), W2 (hidden x hidden), W3 (nout x hidden)
 * bf16 row-major; biases float; y_out (nout x npts float) may be NULL. 2 nf2 and hidden multiples of
 * 32 up to 128, nout a multiple of 16 up to 128. */
int ephip_mlp_bf16_prune(int64_t npts, const double *xk, int32_t nf2, const int32_t *mvec, int32_t hidden,
                         int32_t nout, const void *W1, const float *b1, const void *W2, const float *b2,
                         const void *W3, const float *b3, float wmin, float wmax, float margin, float *y_out,
                         uint8_t *keep, void *stream);

/* ---- Boltzmann transport on the outer-k payload (FP64) ---- */

/* Scattering increments for a list of n (electron k, electron k+q, phonon mode) triples: in is n rows
 * {ek, ekq, wq, g2, wtq, mu, T, eta} (double), out is n rows {s_out, s_in}. method 1..6 selects the
 * occupation convention and smearing 1 = Gaussian, 2 = Lorentzian delta. Replaces the per-triple
 * bte_scattering_increments (src/boltzmann/bte_scattering_core.jl:33-84). */
int ephip_bte_increments(int64_t n, int32_t method, int32_t smearing, const double *in, double *out, void *stream);

/* Scattering-out rates and compacted scattering-in rows of one outer-k batch (the device work of
 * bte_window_accumulate!, src/boltzmann/boltzmann_calculator.jl:336-365). Unit u (initial state
 * unit_state_index[u]) has count[u] kept q (payload arrays col_index, ikq, g2 of ephip_eph_rotate_kq_g2,
 * row stride cs, g2 rows of nbw x nm); final state f = state_offset[ikq] + m for k+q window slot m.
 * For each temperature t: S_out[i + nstates*t] += sum over kept q, m, modes with omega >= w_cut
 * (atomic), S_in[((rowptr[u] + c)*nbw + m)*nT + t] = sum over modes; row r = rowptr[u] + c records
 * row_code[r] = state_offset[ikq] | (number of k+q window slots << 27) (nstates < 2^27, nbw <= 31).
 * e_state: per-state energies; omega (nm x nq); wtq: q weight; mus/Ts/etas: nT per-temperature values. */
int ephip_bte_accumulate_units(int32_t nunits, int64_t cs, int32_t nbw, int32_t nm, const int32_t *count,
                               const int32_t *col_index, const int32_t *ikq, const float *g2,
                               const int64_t *unit_state_index, const int64_t *state_offset, const int64_t *bmin,
                               const int64_t *bmax, const double *e_state, const double *omega, double wtq,
                               int32_t nT, const double *mus, const double *Ts, const double *etas, int32_t method,
                               int32_t smearing, double w_cut, int64_t nstates, const int64_t *rowptr, double *S_out,
                               double *S_in, uint32_t *row_code, void *stream);

/* Band-diagonal velocities v[3*s + a] = Re(u^H dH_a u) of nstates states (k block state_k[s], band
 * state_band[s] of the nw x nw complex64 eigenvector blocks U); dH: complex128 operator columns of
 * ld_dh >= 3 nw^2 entries (direction slowest), column state_col[s]. Replaces the per-k
 * get_el_velocity_diag_berry_connection! (src/wannier_to_bloch.jl:70-81). */
int ephip_band_velocity(int32_t nw, int64_t nstates, const int64_t *state_col, const int32_t *state_band,
                        const int64_t *state_k, const void *dH, int64_t ld_dh, const void *U, double *v,
                        void *stream);

/* C_b = op(A_b) op(B_b) for b < nbatch (alpha 1, beta 0), complex128 (fp64 = 1) or complex64 (fp64 = 0),
 * column-major m x n result, inner size k; op selected by transA / transB = 'N' (the matrix), 'T'
 * (transpose) or 'C' (adjoint); batch b of X starts at X + b strideX (elements). Replaces the backend
 * method batched_gemm! (src/common/gpu_utils.jl:60-78; CUDA method ext/ElectronPhononCUDAExt.jl:116-126). */
int ephip_batched_gemm(int32_t fp64, int32_t transA, int32_t transB, int32_t m, int32_t n, int32_t k, int64_t nbatch,
                       const void *A, int64_t lda, int64_t strideA, const void *B, int64_t ldb, int64_t strideB,
                       void *C, int64_t ldc, int64_t strideC, void *stream);

/* Window scatter of a device-resident calculator: for every (m, n, nu, iq) of g2vals (nbkq, nbk, nm, nq,
 * column-major; float64 when fp64 = 1, else float32, like w, g2_out and w_out), i = imap_i[n] and
 * f = imap_f[m + ld_f (ikqs[iq] - 1)] (int64, 1-based, 0 = outside the window; imap_f has nk_f columns);
 * when both are positive, slot nu + nm (i - i0 - 1) + nm ni_stride (f - 1) of g2_out and w_out (n_out
 * entries each) takes g2vals and w[nu + nm iq]. Entries with a k+q index outside 1..nk_f or a slot
 * outside [0, n_out) are not written; each adds 1 to the device counter *dropped (uint32, may be NULL;
 * the caller zeroes it and reads it back), so a wrong i0 / ni_stride is reported, not lost. The
 * reference has no such entries (its slots are in range by construction).
 * Replaces eph_window_scatter! (src/calculator/calculator_utils.jl:6-52; CUDA method
 * ext/ElectronPhononCUDAExt.jl:280-313). */
int ephip_eph_window_scatter(int32_t fp64, void *g2_out, void *w_out, int64_t n_out, const void *g2vals,
                             const int64_t *imap_i, const int64_t *imap_f, int64_t ld_f, int64_t nk_f,
                             const int64_t *ikqs, const void *w, int32_t nbkq, int32_t nbk, int32_t nm, int64_t nq,
                             int64_t ni_stride, int64_t i0, uint32_t *dropped, void *stream);

/* eph_apply_rotations!(ep, g, ukqs, u_phs, tmp; g2_out, wq) with the reference's arguments
 * (src/wannier_to_bloch_batched.jl:442-455; CUDA method ext/ElectronPhononCUDAExt.jl:177-201):
 * ep[m, n, nu', q] = sum_nu (sum_iw conj(ukqs[iw, m, q]) g[iw, n, nu, q]) u_phs[nu, nu', q], and when g2_out is
 * given g2_out[m, n, nu', q] = |ep|^2 / (2 wq[nu', q]). ep (nbkq, nbk, nm, nq) and g2_out dense column-major;
 * g (nw, nbk, nm, nq) read through its element strides gs0..gs3 (a strided device view is taken as it is);
 * ukqs (nw, nbkq, nq), u_phs (nm, nm, nq), wq (nm, nq) dense. Complex128 / float64 when fp64 = 1, else
 * complex64 / float32. The reference's tmp is not needed. */
int ephip_eph_apply_rotations(int32_t fp64, int32_t nw, int32_t nbkq, int32_t nbk, int32_t nm, int64_t nq, void *ep,
                              const void *g, int64_t gs0, int64_t gs1, int64_t gs2, int64_t gs3, const void *ukqs,
                              const void *uphs, void *g2_out, const void *wq, void *stream);

/* eph_apply_rotations_rqkq!(ep, g, uks, ukqs, tmp, uk_rep) with the reference's arguments
 * (src/wannier_to_bloch_batched.jl:360-377; CUDA method ext/ElectronPhononCUDAExt.jl:241-258):
 * ep[m, n, nu, k] = sum_{iw, jw} conj(ukqs[iw, m, k]) g[iw, jw, nu, k] uks[jw, n, k]; ep (nbkq, nbk, nm, nk),
 * g (nw, nw, nm, nk), uks (nw, nbk, nk), ukqs (nw, nbkq, nk), all dense column-major, complex128 when
 * fp64 = 1 else complex64. The reference's tmp and uk_rep scratch are not needed. */
int ephip_eph_apply_rotations_rqkq(int32_t fp64, int32_t nw, int32_t nbkq, int32_t nbk, int32_t nm, int64_t nk,
                                   void *ep, const void *g, const void *uks, const void *ukqs, void *stream);

/* bte_window_accumulate! with the reference's arguments (src/boltzmann/boltzmann_calculator.jl:336-372; CUDA
 * method ext/ElectronPhononCUDAExt.jl:323-372): for each (m, n, j) of a q batch, i = imap_i[n] and
 * f = imap_f[m + ld_f (ikqs[j] - 1)] (int64, 1-based, 0 = outside the window); per temperature iT (nT of them:
 * mus, Ts, etas) the increments of src/boltzmann/bte_scattering_core.jl:33-84 (method 1..6, smearing 1 =
 * Gaussian, 2 = Lorentzian) of the modes with wq >= w_cut are summed, S_out[(i - 1) + ld_so iT] += sum_out
 * (atomic) and S_in[(i - i0 - 1) + ld_si1 ((f - 1) + ld_si2 iT)] = sum_in (the tile of outer rows i0 + 1 ..
 * i0 + ld_si1). g2vals (nbkq, nbk, nm, nq) and wq (nm, nq) float64 when fp64 = 1 else float32; e_i, e_f, wf
 * (per-final-state weight), S_out, S_in float64. k+q indices outside 1..nk_f and rows outside the tile are not
 * written and add 1 to *dropped (uint32, may be NULL). */
int ephip_bte_window_accumulate(int32_t fp64, double *S_out, int64_t ld_so, double *S_in, int64_t ld_si1,
                                int64_t ld_si2, const void *g2vals, const void *wq, const int64_t *imap_i,
                                const int64_t *imap_f, int64_t ld_f, int64_t nk_f, const int64_t *ikqs,
                                const double *e_i, const double *e_f, const double *wf, int32_t nT, const double *mus,
                                const double *Ts, const double *etas, int32_t method, int32_t smearing, double w_cut,
                                int32_t nbandkq, int32_t nbandk, int32_t nm, int64_t nq, int64_t i0, uint32_t *dropped,
                                void *stream);

/* One k of an outer-k batch payload in the layout of EPDataQBatched (src/EPData.jl:66-73, as the batched
 * outer-k loop builds it, src/calculator/run_eph_over_k_and_kq.jl:843-876): eps (complex64, may be NULL) and
 * g2s (float) (nw, nbandk, nm, nq) column-major over the full k+q band range m and the k-side window
 * n = ibandk_offset + 1 .. ibandk_offset + nbandk, for the kept columns j0 .. j0 + nq - 1 of the k; zero
 * outside the k and k+q windows. col_index / ikq: the kept q-grid indices and k+q states of the k's first
 * unit row; the k's nunits units (bands band0 + 1 .. band0 + nunits, 0-based band0) are consecutive rows of
 * cs columns of ep (complex64, may be NULL) / g2 (the staged kernels' (slot, nu') per column). wqs (nm, nq)
 * float64 = omega (nm x nq_grid) at each column's q; ikqs (nq) int64 = state_pos[k+q state] (1-based
 * position in the filtered state list). */
int ephip_units_to_qbatched(int32_t nw, int32_t nm, int32_t nbw, int32_t nbandk, int64_t cs, const int32_t *col_index,
                            const int32_t *ikq, int32_t j0, int32_t nq, int32_t nunits, int32_t band0,
                            int32_t ibandk_offset, const void *ep, const float *g2, const int64_t *bmin,
                            const int64_t *bmax, const double *omega, const int64_t *state_pos, void *eps, float *g2s,
                            double *wqs, int64_t *ikqs, void *stream);

/* Full rotated operator of nk k-points: out[k ld_out + a nw^2 + m + nw n] = (U_k^H M_a(k) U_k)[m, n] in FP64
 * for the three directions a, with M (complex128, column k of ld_m >= 3 nw^2 entries, direction slowest)
 * the Fourier of a 3 nw^2-data operator (i R H(R) for the velocity, the position operator for rbar) and U
 * the nw x nw complex64 eigenvector blocks of the same k; nw <= 32. With E (nw per k) and rbar (same layout
 * as out, ld_r) both given, the Berry-connection term i (e_m - e_n) rbar[m, n] is added; both NULL gives
 * the plain rotation. Replaces get_el_velocity_direct_batched (src/wannier_to_bloch_batched.jl:121-148)
 * and the Berry-connection update of _compute_electron_states_device! (src/compute_states.jl:149-168). */
int ephip_el_velocity_direct_batched(int32_t nw, int64_t nk, const void *M, int64_t ld_m, const void *U,
                                     const double *E, const void *rbar, int64_t ld_r, void *out, int64_t ld_out,
                                     void *stream);

/* Scattering-in product of the iterative BTE for one temperature: y[3*i + a] = sum over the unit's
 * rows rowptr[u]..rowptr[u+1]-1 and slots m < (row_code[row] >> 27) of S_in[(row*nbw + m)*nT + iT]
 * x[3*f + a], f = (row_code[row] & (2^27 - 1)) + m, i = unit_state_index[u] (the
 * scat_mat * map_i_to_f * dF product of _solve_bte,
 * src/boltzmann/electron_lbte.jl:116-150). */
int ephip_bte_spmv(int32_t nunits, int32_t nbw, int32_t nT, int32_t iT, const int64_t *rowptr,
                   const uint32_t *row_code, const int64_t *unit_state_index, const double *S_in, const double *x,
                   double *y, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EPHIP_H */
