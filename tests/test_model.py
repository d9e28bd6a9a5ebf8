"""Properties of the synthetic model generator (host side)."""
import numpy as np

import oracle as O
from ephip import model as M


def test_rbox_sorted_by_reversed_R():
    r = M.rbox(5)
    assert r.shape == (125, 3)
    key = [tuple(x[::-1]) for x in r]
    assert key == sorted(key)
    assert np.array_equal(r[::-1], -r)  # -R at the mirrored index


def test_hamiltonian_hermitian_blocks_and_onsite():
    m, _ = M.make_config_model("cfg1", with_g=False)
    nr = len(m.irvec_h)
    for i in range(nr):
        a = m.ham_r[i].reshape(m.nw, m.nw, order="F")
        b = m.ham_r[nr - 1 - i].reshape(m.nw, m.nw, order="F")
        assert np.array_equal(b, a.conj().T)
    h0 = m.ham_r[nr // 2].reshape(m.nw, m.nw, order="F")
    d = np.diag(h0)
    assert np.all(d.imag == 0) and np.all(np.abs(d.real) <= 2.0)


def test_dynamical_matrix_psd_and_acoustic_sum_rule():
    m, _ = M.make_config_model("cfg1", with_g=False)
    nm = m.nmodes
    rng = np.random.default_rng(0)
    xq = np.vstack([np.zeros(3), rng.random((20, 3))])
    Dq = O.fourier(m.dyn_r, m.irvec_d, xq)
    for j in range(len(xq)):
        d = Dq[j].reshape(nm, nm, order="F")
        assert np.abs(d - d.conj().T).max() < 1e-14
        w = np.linalg.eigvalsh(d)
        assert w.min() > -1e-12 * w.max()
    # q = 0: three zero modes (uniform translations)
    w0 = np.linalg.eigvalsh(Dq[0].reshape(nm, nm, order="F"))
    assert np.all(np.abs(w0[:3]) < 1e-12 * w0.max()) and w0[3] > 1e-6


def test_generator_is_deterministic_and_columnwise():
    m1 = M.make_model(nw=2, natoms=1, box_hd=3, box_g=3, seed=5)
    m2 = M.make_model(nw=2, natoms=1, box_hd=3, box_g=3, seed=5)
    assert np.array_equal(m1.ham_r, m2.ham_r) and np.array_equal(m1.epmat_ph, m2.epmat_ph)
    col = M.eph_column(5, 2, 3, m1.irvec_el, m1.irvec_ep[4], 4, 0.8, np.complex128)
    assert np.array_equal(col, m1.epmat_ph[4])
    m3 = M.make_model(nw=2, natoms=1, box_hd=3, box_g=3, seed=6)
    assert not np.array_equal(m1.ham_r, m3.ham_r)


def test_eph_decay_envelope():
    m = M.make_model(nw=2, natoms=1, box_hd=3, box_g=5, seed=7)
    g = np.abs(m.epmat_ph.reshape(len(m.irvec_ep), len(m.irvec_el), -1)).mean(axis=2)
    ne = np.linalg.norm(m.irvec_el, axis=1)
    npn = np.linalg.norm(m.irvec_ep, axis=1)
    env = np.exp(-(ne[None, :] + npn[:, None]) / 0.8)
    ratio = g / env
    assert 0.5 < ratio.mean() < 1.5 and ratio.std() / ratio.mean() < 0.5


def test_planted_degeneracy_fraction_and_pairs():
    """cfg3's planting: exact two-fold pairs on model.planted_degenerate (19.8% of a 48^3 grid: the planes
    4 k2 in Z or 6 k3 in Z), split pairs elsewhere; H(R) stays Hermitian, and the time-reversal variant
    keeps E(-k) = E(k)."""
    for tr in (False, True):
        m, g = M.make_config_model("cfg3", with_g=False, time_reversal=tr)
        for i in range(len(m.irvec_h)):
            j = int(np.flatnonzero((m.irvec_h == -m.irvec_h[i]).all(1))[0])
            a = m.ham_r[i].reshape(m.nw, m.nw, order="F")
            b = m.ham_r[j].reshape(m.nw, m.nw, order="F")
            assert np.allclose(b, a.conj().T, atol=1e-14)
        xk = M.kpoints_grid((1, 48, 48))
        E, _ = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, xk))
        mask = M.planted_degenerate(xk)
        assert abs(mask.mean() - (1 - (44 / 48) * (42 / 48))) < 1e-12   # 19.8%
        split = np.abs(E[:, 1::2] - E[:, 0::2]).max(1)
        assert split[mask].max() < 1e-11
        assert split[~mask].min() > 1e-3
        if tr:
            Em, _ = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, -xk))
            assert np.abs(Em - E).max() < 1e-11
    assert abs(M.planted_degenerate(M.kpoints_grid((48, 48, 48))).mean() - 0.1979) < 1e-3
