"""FP64 transport oracle (test infrastructure only: imported by tests/, smoke() and the benchmark's
CPU leg, never by the product path).

Restates, on the host, the reference's uniform-grid transport pipeline without symmetry:
* electron states with the per-k eigensolve's degenerate-band gauge fixing (src/diagonalize.jl:47-85);
* band velocities Re(u^H dH/dk u), dH/dk = sum_R i R_cart H(R) e^{2 pi i k.R}
  (src/wannier/WannierObject.jl:78-87, src/wannier_to_bloch.jl:70-81);
* the scattering accumulation bte_window_accumulate! (src/boltzmann/boltzmann_calculator.jl:336-365)
  into dense S_out / S_in, with g2 = |g|^2 / (2 omega) from the oracle's e-ph pipeline;
* SERTA occupations and conductivity and the fixed-point linearised BTE
  (src/boltzmann/electron_lbte.jl:43-50, :116-190), plus a direct linear solve of the same system;
* the metal chemical potential by bisection (src/boltzmann/electron_serta.jl:19-52).

Pinning (tests/test_bte_oracle.py): velocities against central finite differences of the oracle
eigenvalues (Hellmann-Feynman); the dense accumulation against an independent per-element
composition of the golden-pinned bte_increments; the fixed-point solution against the direct
linear solve; the chemical potential against a closed form.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def occ(x):
    return 0.5 * (1.0 - np.tanh(0.5 * x))


def dfermi(e, T):
    """-df/de of the Fermi-Dirac occupation."""
    return 1.0 / (2.0 + np.exp(e / T) + np.exp(-e / T)) / T


def velocity_operator(irvec, ham_r, lattice=None):
    lat = np.eye(3) if lattice is None else np.asarray(lattice, dtype=np.float64)
    rc = irvec.astype(np.float64) @ lat.T
    return np.concatenate([1j * rc[:, a:a + 1] * ham_r for a in range(3)], axis=1)


# Degeneracy threshold of the per-k eigensolve's gauge fixing: 0.01 meV (src/common/constants.jl:6), in eV.
DEGEN_CUTOFF = 1e-5


def states(m, kgrid, window, degen_cutoff=DEGEN_CUTOFF):
    """Electron states of the grid (src/compute_states.jl, CPU path): eigenpairs with the degenerate-band
    gauge fixing of get_el_eigen! / solve_eigen_el! (src/diagonalize.jl:47-85), window filter, state
    lists. degen_cutoff None: no gauge fixing."""
    xk = O.kpoints_grid(kgrid)
    E, U = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, xk), degen_cutoff=degen_cutoff)
    keep, bmin, bmax, nbelow = O.window_filter(E, window[0], window[1])
    nb = np.where(bmin > 0, bmax - bmin + 1, 0).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(nb)]).astype(np.int64)
    ks = np.repeat(np.arange(len(xk)), nb)
    ib = np.arange(off[-1]) - off[ks] + bmin[ks] - 1
    return dict(xk=xk, E=E, U=U, keep=keep, bmin=bmin, bmax=bmax, nbelow=nbelow, off=off, ks=ks, ib=ib,
                es=E[ks, ib], weight=1.0 / len(xk))


def velocities(m, s, lattice=None):
    """(n, 3) band-diagonal velocities of the in-window states of ``states``."""
    nw = m.nw
    op = velocity_operator(m.irvec_h, m.ham_r, lattice)
    kl = np.unique(s["ks"])
    dH = O.fourier(op, m.irvec_h, s["xk"][kl])         # (nkl, 3 nw^2)
    pos = {k: i for i, k in enumerate(kl)}
    v = np.zeros((len(s["ks"]), 3))
    for st, (k, b) in enumerate(zip(s["ks"], s["ib"])):
        u = s["U"][k].reshape(nw, nw, order="F")[:, b]
        for a in range(3):
            M = dH[pos[k], a * nw * nw:(a + 1) * nw * nw].reshape(nw, nw, order="F")
            v[st, a] = np.real(np.conj(u) @ M @ u)
    return v


def scattering(m, s, qgrid, kgrid, Tlist, mulist, etas, method=5, smearing=1, w_cut=1e-5):
    """Dense S_out (nT, n) and S_in (nT, n, n) from the oracle e-ph pipeline over every q of the grid."""
    nw, nm = m.nw, m.nmodes
    xq = O.kpoints_grid(qgrid)
    Dq = O.fourier(m.dyn_r, m.irvec_d, xq)
    ns = len(s["ks"])
    nT = len(Tlist)
    S_out = np.zeros((nT, ns))
    S_in = np.zeros((nT, ns, ns))
    ik = np.flatnonzero(s["keep"]).astype(np.int64)
    nbw = int(max(1, (np.where(s["bmin"] > 0, s["bmax"] - s["bmin"] + 1, 0)).max()))
    Ucm = np.ascontiguousarray(s["U"].reshape(len(s["xk"]), nw * nw))
    for iq in range(len(xq)):
        om, uph = O.ph_eigen(Dq[iq], m.mass)
        eRpq = O.eph_RR_to_Rq(nw, nm, m.epmat_ph, m.irvec_el, m.irvec_ep, xq[iq], uph)
        ikq = O.hash_xk(s["xk"][ik] + xq[iq], kgrid)
        ep = O.outer_q_block(nw, nm, eRpq, m.irvec_el, s["xk"], Ucm, s["bmin"], s["bmax"], ik, ikq, nbw)
        g2 = np.abs(ep) ** 2 / (2.0 * np.where(om[None, :, None, None] != 0, om[None, :, None, None], 1.0))
        O.bte_accumulate_q(nm, nbw, ik, ikq, g2, s["bmin"], s["bmax"], s["off"], s["es"], om, 1.0 / len(xq),
                           mulist, Tlist, etas, method, smearing, w_cut, S_out, S_in)
    return S_out, S_in


def conductivity(df, v, w, spin, volume):
    return (df.T @ v) * (w * spin / volume)


def solve_bte(S_out, S_in, es, v, w, Tlist, mulist, spin=2, volume=1.0, max_iter=100, rtol=1e-10, mixing=1.0):
    """SERTA and fixed-point BTE conductivities (nT, 3, 3) and the iteration counts."""
    nT = len(Tlist)
    ss, sb, its = np.zeros((nT, 3, 3)), np.zeros((nT, 3, 3)), []
    for t in range(nT):
        rhs = (dfermi(es - mulist[t], Tlist[t]) / S_out[t])[:, None] * v
        ss[t] = conductivity(rhs, v, w, spin, volume)
        df, obs, it = rhs.copy(), ss[t], 0
        for it in range(1, max_iter + 1):
            new = (S_in[t] @ df) / S_out[t][:, None] + rhs
            df = new * mixing + df * (1 - mixing)
            old, obs = obs, conductivity(df, v, w, spin, volume)
            if np.linalg.norm(obs - old) / np.linalg.norm(obs) < rtol:
                break
        sb[t] = obs
        its.append(it)
    return ss, sb, its


def solve_bte_direct(S_out, S_in, es, v, w, Tlist, mulist, spin=2, volume=1.0):
    """Direct solution of dF = (S_in dF) / S_out + rhs: (diag(S_out) - S_in) dF = S_out rhs."""
    nT = len(Tlist)
    sb = np.zeros((nT, 3, 3))
    for t in range(nT):
        rhs = (dfermi(es - mulist[t], Tlist[t]) / S_out[t])[:, None] * v
        A = np.diag(S_out[t]) - S_in[t]
        df = np.linalg.solve(A, S_out[t][:, None] * rhs)
        sb[t] = conductivity(df, v, w, spin, volume)
    return sb


def find_mu(target, T, es, w):
    lo, hi = es.min() - 50 * T, es.max() + 50 * T
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if w * occ((es - mid) / T).sum() > target:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)
