"""CPU pins of the transport oracle (oracle/bte_ref.py) and of the host-side transport logic.

* e-ph vertex hermiticity: with the symmetrised tensor, |g_mn,nu(k, q)|^2 = |g_nm,nu(k+q, -q)|^2
  through the oracle's RR->Rq->kq pipeline (a physics identity; the raw tensor violates it, and a
  transposed operand fails it);
* velocities against central finite differences of the oracle eigenvalues (Hellmann-Feynman:
  v = dE/dk_cart = (1/2 pi) dE/dk_crys for the a = 1 cubic cell);
* the dense scattering accumulation against a per-element composition of the golden-pinned
  bte_increments;
* the fixed-point BTE iteration against a direct linear solve of the same system;
* the metal chemical potential against a closed form, and the host bisection against the oracle's.
"""
import numpy as np
import pytest

import bte_ref as B
import oracle as O
from ephip import model as M


def _gbloch(m, k, q):
    nw, nm = m.nw, m.nmodes
    Dq = O.fourier(m.dyn_r, m.irvec_d, q[None])[0]
    om, uph = O.ph_eigen(Dq, m.mass)
    eRpq = O.eph_RR_to_Rq(nw, nm, m.epmat_ph, m.irvec_el, m.irvec_ep, q, uph)
    eye = np.eye(nw, dtype=np.complex128)
    return np.abs(O.eph_Rq_to_kq(nw, nm, eRpq, m.irvec_el, k, eye, eye)) ** 2, om   # [nu, n, m]


@pytest.fixture(scope="module")
def small_models():
    kw = dict(nw=4, natoms=1, box_hd=3, box_g=3, seed=77)
    return M.make_model(**kw), M.make_model(hermitian_g=True, **kw)


def test_hermitian_vertex(small_models):
    raw, herm = small_models
    rng = np.random.default_rng(5)
    for _ in range(3):
        k, q = rng.random(3), rng.random(3)
        a, oa = _gbloch(herm, k, q)
        b, ob = _gbloch(herm, k + q, -q)
        assert np.allclose(oa, ob, rtol=0, atol=1e-14)
        assert np.abs(a - np.swapaxes(b, 1, 2)).max() <= 1e-12 * a.max()
        assert np.abs(a - b).max() > 1e-2 * a.max()           # the transpose matters
        c, _ = _gbloch(raw, k, q)
        d, _ = _gbloch(raw, k + q, -q)
        assert np.abs(c - np.swapaxes(d, 1, 2)).max() > 1e-2 * c.max()


def test_time_reversal_model():
    m = M.make_model(nw=4, natoms=1, box_hd=3, box_g=3, seed=78, time_reversal=True)
    k = np.array([[0.13, 0.27, 0.41]])
    e1, _ = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, k))
    e2, _ = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, -k))
    assert np.abs(e1 - e2).max() < 1e-12


def test_velocity_finite_difference():
    m, g = M.make_config_model("cfg1", with_g=False)
    s = B.states(m, (4, 4, 4), (-1.0, 1.0))
    v = B.velocities(m, s)
    h = 1e-5
    err = 0.0
    for st in range(0, len(s["ks"]), 7):
        k, b = s["ks"][st], s["ib"][st]
        for a in range(3):
            d = np.zeros(3)
            d[a] = h
            ep = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, (s["xk"][k] + d)[None]), vectors=False)
            em = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, (s["xk"][k] - d)[None]), vectors=False)
            fd = (ep[0, b] - em[0, b]) / (2 * h) / (2 * np.pi)
            err = max(err, abs(fd - v[st, a]))
    assert err <= 1e-6 * np.abs(v).max(), err


def test_accumulate_composition():
    """The C accumulation loop against a per-element composition of bte_increments."""
    rng = np.random.default_rng(11)
    nm, nbw = 3, 3
    nk = 5
    bmin = np.array([1, 0, 2, 1, 3], np.int64)
    bmax = np.array([3, 0, 3, 1, 4], np.int64)
    nb = np.where(bmin > 0, bmax - bmin + 1, 0)
    off = np.concatenate([[0], np.cumsum(nb)]).astype(np.int64)
    ns = int(off[-1])
    es = rng.uniform(-0.2, 0.2, ns)
    ik = np.array([0, 2, 3, 4, 1], np.int64)
    ikq = np.array([2, 4, 1, 0, 3], np.int64)
    g2 = rng.random((len(ik), nm, nbw, nbw))
    om = np.array([0.0, 0.03, 0.05])
    mus, Ts, etas = np.array([0.01, -0.02]), np.array([0.025, 0.05]), np.array([0.05, 0.08])
    for method in range(1, 7):
        for smearing in (1, 2):
            So = np.zeros((2, ns))
            Si = np.zeros((2, ns, ns))
            O.bte_accumulate_q(nm, nbw, ik, ikq, g2, bmin, bmax, off, es, om, 0.25, mus, Ts, etas, method,
                               smearing, 1e-5, So, Si)
            Ro = np.zeros((2, ns))
            Ri = np.zeros((2, ns, ns))
            for t in range(len(ik)):
                a, b = ik[t], ikq[t]
                for n in range(nb[a]):
                    for mm in range(nb[b]):
                        i, f = off[a] + n, off[b] + mm
                        for it in range(2):
                            for nu in (1, 2):   # mode 0 is below the cutoff
                                so, si = O.bte_increments(method, es[i], es[f], om[nu], g2[t, nu, n, mm], 0.25,
                                                          mus[it], Ts[it], smearing, etas[it])
                                Ro[it, i] += so
                                Ri[it, i, f] += si
            assert np.allclose(So, Ro, rtol=1e-13, atol=0), (method, smearing)
            assert np.allclose(Si, Ri, rtol=1e-13, atol=0), (method, smearing)


def test_fixed_point_matches_direct_solve():
    rng = np.random.default_rng(3)
    n = 40
    S_out = rng.uniform(1.0, 2.0, (1, n))
    S_in = rng.random((1, n, n)) * 0.6 * S_out[0][:, None] / n   # spectral radius < 0.6
    es = rng.uniform(-0.1, 0.1, n)
    v = rng.standard_normal((n, 3))
    T, mu = np.array([0.03]), np.array([0.0])
    ss, sb, its = B.solve_bte(S_out, S_in, es, v, 1.0 / n, T, mu, max_iter=200, rtol=1e-14)
    sd = B.solve_bte_direct(S_out, S_in, es, v, 1.0 / n, T, mu)
    assert its[0] < 200
    assert np.abs(sb - sd).max() <= 1e-12 * np.abs(sd).max()
    assert np.abs(sb - ss).max() > 1e-3 * np.abs(ss).max()   # the scattering-in term is not negligible


def test_chemical_potential():
    from ephip import bte
    T, w, e0 = 0.03, 0.5, 0.12
    target = 0.2   # one level of weight w: target = w f((e0 - mu)/T)
    exact = e0 - T * np.log(w / target - 1.0)
    assert abs(B.find_mu(target, T, np.array([e0]), w) - exact) < 1e-12
    assert abs(bte.find_chemical_potential(target, T, np.array([e0]), w) - exact) < 1e-12
    rng = np.random.default_rng(9)
    es = rng.uniform(-0.5, 0.5, 300)
    assert abs(bte.find_chemical_potential(37.0 / 300, 0.05, es, 1.0 / 300) - B.find_mu(37.0 / 300, 0.05, es, 1.0 / 300)) < 1e-11
    with pytest.raises(ValueError):
        bte.find_chemical_potential(2.0, 0.05, es, 1.0 / 300)


def test_band_states_host():
    """band_states (torch index bookkeeping, here on CPU tensors) orders states like the oracle."""
    torch = pytest.importorskip("torch")
    from ephip import bte, core
    m, g = M.make_config_model("cfg1", with_g=False)
    s = B.states(m, (4, 4, 4), (-1.0, 1.0))

    class El:
        pass
    el = El()
    el.E = torch.from_numpy(s["E"])
    el.filt = core.FilterResult(torch.from_numpy(s["keep"]), torch.from_numpy(s["bmin"]),
                                torch.from_numpy(s["bmax"]), torch.from_numpy(s["nbelow"]))
    bs = bte.band_states(el)
    assert bs.n == len(s["ks"])
    assert np.array_equal(bs.ks.numpy(), s["ks"])
    assert np.array_equal(bs.ibands.numpy(), s["ib"])
    assert np.array_equal(bs.state_offset.numpy(), s["off"])
    assert np.array_equal(bs.es.numpy(), s["es"])
    assert bs.nstates_base == pytest.approx(s["nbelow"].sum() / len(s["xk"]))


def test_calculator_argument_checks():
    from ephip import bte
    p = bte.ElectronTransportParams(Tlist=[0.025], nlist=[0.01], volume=1.0)
    with pytest.raises(ValueError):
        bte.BoltzmannCalculator(p, occupation_method=7)
    with pytest.raises(ValueError):
        bte.BoltzmannCalculator(p, smearing=("Tetrahedron", 0.01))
    with pytest.raises(ValueError):
        bte.BoltzmannCalculator(p, smearing=("Gaussian", 0.0))
    with pytest.raises(ValueError):
        bte.ElectronTransportParams(Tlist=[0.025, 0.03], nlist=[0.01], volume=1.0)


def test_conductivity_drude_closed_form():
    """bte_ref.conductivity on a parabolic band with a constant scattering rate (S_in = 0, so BTE = SERTA)
    against the Drude law sigma = n tau / m* (e = hbar = 1), n = spin / V * w sum f the carrier density:
    by parts, sum (-df/de) v_x^2 = (1 / m*) sum f for E = |k|^2 / (2 m*) once f vanishes at the zone edge,
    which a uniform grid sums to spectral accuracy. Independent of the formula's factors: a missing spin,
    volume or weight factor, a transposed index or a sign flip of -df/de fails it; the off-diagonal
    entries vanish by symmetry."""
    import bte_ref as B
    n1 = 48
    k = (np.arange(n1) - n1 // 2) / n1
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    K = 2 * np.pi * np.stack([kx.ravel(), ky.ravel(), kz.ravel()], 1)   # Cartesian, cubic cell a = 1
    for mstar, spin, vol in ((1.0, 2, 1.7), (0.7, 1, 1.0)):
        E = (K ** 2).sum(1) / (2 * mstar)
        v = K / mstar
        mu, T, tau = 1.2, 0.3, 3.0
        w = 1.0 / len(E)
        S_out = np.full((1, len(E)), 1.0 / tau)
        rhs = (B.dfermi(E - mu, T) / S_out[0])[:, None] * v
        sig = B.conductivity(rhs, v, w, spin, vol)
        n = spin / vol * w * B.occ((E - mu) / T).sum()
        assert abs(sig[0, 0] / (n * tau / mstar) - 1) < 1e-5
        assert np.allclose(np.diag(sig), sig[0, 0], rtol=1e-12)
        assert np.abs(sig - np.diag(np.diag(sig))).max() < 1e-12 * sig[0, 0]
