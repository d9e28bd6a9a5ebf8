"""Oracle-side checks of device results on full workload shapes (test infrastructure only).

Used by tests/ and by bench.py's oracle check (after the timed region, never measured): the FP64
filter of a whole fine grid, and the gauge-invariant |g|^2 of sampled (k, q) pairs of an outer-k
batch, both computed by the FP64 C oracle (``oracle.py``) and compared with what the device produced.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def full_grid_states(m, xk, chunk: int = 65536):
    """FP64 band energies (nk, nw) of every k of ``xk`` (oracle Fourier + oracle Hermitian eigensolve),
    in chunks so the FP64 H(k) stack stays small."""
    E = np.empty((len(xk), m.nw))
    for s in range(0, len(xk), chunk):
        e = min(len(xk), s + chunk)
        E[s:e] = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, xk[s:e]), vectors=False)
    return E


def filter_compare(E_o, keep_d, bmin_d, bmax_d, E_d, wmin: float, wmax: float) -> dict:
    """The oracle's window filter (src/filter.jl:40-100) of the FP64 energies ``E_o`` against the device's
    keep / band_min / band_max (integer arrays, compared bit for bit), plus the energy error."""
    keep, bmin, bmax, _ = O.window_filter(E_o, wmin, wmax)
    err = np.abs(E_d - E_o)
    # distance of the nearest band to a window edge: how far the filter is from a tie
    edge = float(np.min(np.minimum(np.abs(E_o - wmin), np.abs(E_o - wmax))))
    return {
        "nk": int(len(keep)),
        "K_surv_oracle": int(keep.sum()),
        "K_surv_device": int(np.asarray(keep_d).astype(bool).sum()),
        "keep_equal": bool(np.array_equal(np.asarray(keep_d).astype(np.uint8), keep)),
        "band_min_equal": bool(np.array_equal(np.asarray(bmin_d, dtype=np.int64), bmin)),
        "band_max_equal": bool(np.array_equal(np.asarray(bmax_d, dtype=np.int64), bmax)),
        "max_abs_dE_eV": float(err.max()),
        "rms_dE_eV": float(np.sqrt((err ** 2).mean())),
        "min_edge_distance_eV": edge,
    }


def sample_units_g2(payload, qsel, nk_sample: int, torch, seed: int = 0):
    """Device g2 of the outer-k payload ``payload`` (an EPDataUnits of the batch just run) for every unit
    of ``nk_sample`` k of the batch (seeded choice) at the q of ``qsel``. Returns
    {q: {k: {band (0-based): (nbw, nm) float32}}}; a (k, q) is present when q is in the unit's kept list."""
    state = payload.unit_state.cpu().numpy()
    band = payload.unit_band.cpu().numpy()
    cnt = payload.count.cpu().numpy()
    kset = np.unique(state)
    rng = np.random.default_rng(seed)
    pick = set(rng.choice(kset, min(nk_sample, len(kset)), replace=False).tolist())
    units = np.flatnonzero(np.isin(state, list(pick)))
    dev = payload.col_index.device
    ci = payload.col_index.index_select(0, torch.from_numpy(units).to(dev)).cpu().numpy()
    nbw = payload.nbw
    nm = payload.g2.shape[-1] // nbw
    out = {int(q): {} for q in qsel}
    for r, u in enumerate(units):
        row = ci[r, :cnt[u]]
        for q in qsel:
            hit = np.flatnonzero(row == q)
            if len(hit):
                g = payload.g2[int(u), int(hit[0])].cpu().numpy().reshape(nbw, nm)
                out[int(q)].setdefault(int(state[u]), {})[int(band[u])] = g
    return out


def g2_parity(m, epmat_ph, kgrid, qgrid, got, bmin, bmax, nbw: int, cutoff: float = 1e-5,
              omega_min: float = 1e-5) -> dict:
    """Norm-wise relative error of the device |g|^2 (``got`` as returned by ``sample_units_g2``) against
    the FP64 oracle, gauge invariant: oracle eigenpairs of every k and k+q involved, RR->Rq->kq in FP64,
    both sides averaged over degenerate bands (src/EPState.jl:205-238) at k and k+q; modes with
    |omega| < omega_min left out (consumers skip them). ``epmat_ph`` is the "ph"-layout tensor as the
    device holds it (complex64), promoted to FP64."""
    nw, nm = m.nw, m.nmodes
    xk = O.kpoints_grid(kgrid)
    xq = O.kpoints_grid(qgrid)
    ks = sorted({k for q in got for k in got[q]})
    need = sorted(set(ks) | {int(O.hash_xk(xk[k] + xq[q], kgrid)[0]) for q in got for k in got[q]})
    E_o, U_o = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, xk[need]))
    pos = {k: i for i, k in enumerate(need)}
    ep128 = np.asarray(epmat_ph, dtype=np.complex128)
    num = den = 0.0
    maxrel = 0.0
    dmax = rmax = 0.0
    npairs = 0
    for q, per_k in got.items():
        if not per_k:
            continue
        om, uph = O.ph_eigen(O.fourier(m.dyn_r, m.irvec_d, xq[q:q + 1])[0], m.mass)
        eRpq = O.eph_RR_to_Rq(nw, nm, ep128, m.irvec_el, m.irvec_ep, xq[q], uph)
        ok = np.abs(om) >= omega_min
        for k, bands in per_k.items():
            kq = int(O.hash_xk(xk[k] + xq[q], kgrid)[0])
            Uk = U_o[pos[k]].reshape(nw, nw, order="F")[:, bmin[k] - 1:bmax[k]]
            Ukq = U_o[pos[kq]].reshape(nw, nw, order="F")[:, bmin[kq] - 1:bmax[kq]]
            nbkq = Ukq.shape[1]
            if set(bands) != set(range(bmin[k] - 1, bmax[k])):
                raise AssertionError(f"units of k={k} at q={q}: bands {sorted(bands)}")
            ep = O.eph_Rq_to_kq(nw, nm, eRpq, m.irvec_el, xk[k], Uk, Ukq)   # [nu, n, m]
            ref = O.g2(ep, om)
            gd = np.stack([bands[n][:nbkq, :].T for n in range(bmin[k] - 1, bmax[k])], axis=1).astype(np.float64)
            ek, ekq = E_o[pos[k], bmin[k] - 1:bmax[k]], E_o[pos[kq], bmin[kq] - 1:bmax[kq]]
            r = O.g2_degenerate_average(ref, ekq, ek, cutoff)[ok]
            gg = O.g2_degenerate_average(gd, ekq, ek, cutoff)[ok]
            num += float(((gg - r) ** 2).sum())
            den += float((r ** 2).sum())
            if np.abs(r).max() > 0:
                maxrel = max(maxrel, float(np.abs(gg - r).max() / np.abs(r).max()))
            dmax = max(dmax, float(np.abs(gg - r).max(initial=0.0)))
            rmax = max(rmax, float(np.abs(r).max(initial=0.0)))
            npairs += 1
    # rel_err: RMS (norm-wise) relative error over all sampled entries; max_rel_err: max|d| / max|ref| over
    # all of them; max_block_rel_err: the worst (k, q) block's max|d| / max|ref|
    return {"pairs": npairs, "rel_err": float(np.sqrt(num / den)) if den > 0 else float("nan"),
            "max_rel_err": dmax / rmax if rmax > 0 else float("nan"), "max_block_rel_err": maxrel}
