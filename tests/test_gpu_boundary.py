"""Backend methods with the reference's own argument lists, through the C-ABI:

* eph_apply_rotations! (src/wannier_to_bloch_batched.jl:442-455; ext/ElectronPhononCUDAExt.jl:177-201) on a
  strided g view, with and without the fused g2, FP64 and FP32;
* eph_apply_rotations_rqkq! (src/wannier_to_bloch_batched.jl:360-377; ext/ElectronPhononCUDAExt.jl:241-258);
* bte_window_accumulate! (src/boltzmann/boltzmann_calculator.jl:336-372; ext/ElectronPhononCUDAExt.jl:323-372),
  all six occupation methods, both smearings, an out-of-window k band, k+q states outside the window and
  modes below the cutoff, against the oracle's golden-pinned increments summed in plain Python;
  plus its dropped-entry report for a tile that does not hold the rows.

The rotations are checked against numpy einsum in FP64 (an independent contraction): relative error
<= 1e-12 (FP64) and <= 2e-6 (FP32, max-norm)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("with_g2", [True, False])
def test_eph_apply_rotations(backend, fp64, with_g2):
    from ephip import core
    dev = backend.torch_device
    rng = np.random.default_rng(41)
    nw, nbk, nbkq, nm, nq = 8, 3, 5, 6, 17
    cdt, rdt = (torch.complex128, torch.float64) if fp64 else (torch.complex64, torch.float32)
    # g as a strided view: iw is not the unit-stride axis, and n is a slice
    big = _c(rng, (nq, nm, nw, nbk + 2))
    g_t = torch.from_numpy(big).to(dev, cdt).permute(0, 1, 3, 2)[:, :, 1:nbk + 1, :]
    assert g_t.stride()[3] != 1 and not g_t.is_contiguous()
    ukq = _c(rng, (nq, nbkq, nw))
    uph = _c(rng, (nq, nm, nm))
    wq = rng.uniform(0.01, 0.05, (nq, nm))
    ep = torch.zeros((nq, nm, nbk, nbkq), dtype=cdt, device=dev)
    g2 = torch.zeros((nq, nm, nbk, nbkq), dtype=rdt, device=dev) if with_g2 else None
    core.eph_apply_rotations(ep, g_t, torch.from_numpy(ukq).to(dev, cdt), torch.from_numpy(uph).to(dev, cdt),
                             g2_out=g2, wq=torch.from_numpy(wq).to(dev, rdt) if with_g2 else None)
    torch.cuda.synchronize()
    g_np = np.transpose(big, (0, 1, 3, 2))[:, :, 1:nbk + 1, :]   # [q, nu, n, iw]
    ref = np.einsum("qmi,qvni,qwv->qwnm", ukq.conj(), g_np, uph)  # [q, nu', n, m]
    got = ep.cpu().numpy()
    tol = 1e-12 if fp64 else 2e-6
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    if with_g2:
        g2_ref = np.abs(ref) ** 2 / (2 * wq[:, :, None, None])
        assert np.abs(g2.cpu().numpy() - g2_ref).max() <= 2 * tol * g2_ref.max()


@pytest.mark.parametrize("fp64", [True, False])
def test_eph_apply_rotations_rqkq(backend, fp64):
    from ephip import core
    dev = backend.torch_device
    rng = np.random.default_rng(42)
    nw, nbk, nbkq, nm, nk = 7, 4, 2, 5, 23
    cdt = torch.complex128 if fp64 else torch.complex64
    g = _c(rng, (nk, nm, nw, nw))      # [k, nu, jw, iw]
    uk = _c(rng, (nk, nbk, nw))
    ukq = _c(rng, (nk, nbkq, nw))
    ep = torch.zeros((nk, nm, nbk, nbkq), dtype=cdt, device=dev)
    core.eph_apply_rotations_rqkq(ep, torch.from_numpy(g).to(dev, cdt), torch.from_numpy(uk).to(dev, cdt),
                                  torch.from_numpy(ukq).to(dev, cdt))
    torch.cuda.synchronize()
    ref = np.einsum("kmi,kvji,knj->kvnm", ukq.conj(), g, uk)
    tol = 1e-12 if fp64 else 2e-6
    assert np.abs(ep.cpu().numpy() - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("method", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("smearing", [1, 2])
def test_bte_window_accumulate(backend, method, smearing):
    from ephip import core
    dev = backend.torch_device
    rng = np.random.default_rng(100 + 10 * method + smearing)
    nbkq, nbk, nm, nq, nT = 3, 2, 6, 5, 2
    ni, nf, nk, ld_f, i0 = 9, 30, 8, 4, 3
    imap_i = np.array([5, 0], dtype=np.int64)            # the second k band is outside the window
    imap_f = np.zeros((nk, ld_f), dtype=np.int64)          # (nk, ld) = column-major (ld, nk)
    # distinct final states for every (k+q, band), as in the reference (each (i, f) is written once)
    imap_f[:, :nbkq] = (rng.permutation(nf)[:nk * nbkq] + 1).reshape(nk, nbkq)
    imap_f[2, 1] = 0                                       # a k+q band outside the window
    ikqs = np.array([3, 1, 6, 8, 2], dtype=np.int64)      # 1-based, distinct
    e_i = rng.uniform(-0.1, 0.1, ni)
    e_f = rng.uniform(-0.1, 0.1, nf)
    wf = rng.uniform(0.5, 1.5, nf) / nq
    wq = rng.uniform(0.005, 0.04, (nq, nm))
    wq[:, 0] = 1e-7                                        # acoustic modes below the cutoff
    g2 = rng.uniform(0.0, 1e-3, (nq, nm, nbk, nbkq))
    mus, Ts, etas = np.array([0.0, 0.01]), np.array([0.02585, 0.05]), np.array([0.02, 0.03])
    w_cut = 1e-5
    S_out = torch.zeros((nT, ni), dtype=torch.float64, device=dev)
    S_in = torch.full((nT, nf, 4), -1.0, dtype=torch.float64, device=dev)  # tile rows i0+1 .. i0+4
    t = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    core.bte_window_accumulate(S_out, S_in, t(g2), t(wq), t(imap_i, torch.int64), t(imap_f, torch.int64),
                               t(ikqs, torch.int64), t(e_i), t(e_f), t(wf), t(mus), t(Ts), t(etas), method, w_cut, i0,
                               smearing=smearing)
    torch.cuda.synchronize()
    so_ref = np.zeros((nT, ni))
    si_ref = np.full((nT, nf, 4), -1.0)
    for j in range(nq):
        for n in range(nbk):
            i = imap_i[n]
            if i <= 0:
                continue
            for m in range(nbkq):
                f = imap_f[ikqs[j] - 1, m]
                if f <= 0:
                    continue
                for it in range(nT):
                    so = si = 0.0
                    for nu in range(nm):
                        if wq[j, nu] < w_cut:
                            continue
                        a, b = O.bte_increments(method, e_i[i - 1], e_f[f - 1], wq[j, nu], g2[j, nu, n, m],
                                                wf[f - 1], mus[it], Ts[it], smearing, etas[it])
                        so += a
                        si += b
                    so_ref[it, i - 1] += so
                    si_ref[it, f - 1, i - i0 - 1] = si
    got_o, got_i = S_out.cpu().numpy(), S_in.cpu().numpy()
    assert np.abs(got_o - so_ref).max() <= 1e-12 * max(np.abs(so_ref).max(), 1e-300)
    assert np.abs(got_i - si_ref).max() <= 1e-12 * max(np.abs(si_ref).max(), 1e-300)
    assert np.abs(so_ref).max() > 0 and (si_ref > 0).any()


def test_bte_window_accumulate_reports_rows_outside_the_tile(backend):
    from ephip import _lib as L
    from ephip import core
    dev = backend.torch_device
    z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
    i64 = lambda a: torch.tensor(a, dtype=torch.int64, device=dev)
    S_out, S_in = z(1, 6), z(1, 4, 2)                        # tile of 2 rows from i0 + 1 = 1
    with pytest.raises(L.EphipError, match="fell outside"):
        core.bte_window_accumulate(S_out, S_in, z(1, 2, 1, 1) + 1e-4, z(1, 2) + 0.02, i64([5]),
                                   i64([[1], [2]]), i64([2]), z(6), z(4), z(4) + 1, z(1), z(1) + 0.025,
                                   z(1) + 0.02, 5, 1e-5, 0)


@pytest.mark.parametrize("variant", [0, 1], ids=["kq-valu", "kq-mfma"])
def test_qbatched_payload_from_outer_k(backend, variant):
    """The outer-k loop's batch payload laid out as the reference's EPDataQBatched (QBatchedAdapter over
    ephip_units_to_qbatched, the complex elements from the staged kR->kq kernel): for sampled k, every
    listed k+q: eps against the FP64 oracle evaluated with the device's own eigenvectors (so the gauge
    matches; FP32 bound 1e-4 of the largest element), g2s = |eps|^2 / (2 omega), zeros outside the k and
    k+q windows, wqs = omega(q), ikqs = the 1-based positions of the kept k+q states."""
    from ephip import _lib as L
    from ephip import eph, tune
    from ephip import model as M
    m = M.make_model(nw=16, natoms=4, box_hd=3, box_g=3, seed=4242, g_dtype=np.complex128)
    kgrid, qgrid, w = (6, 6, 6), (3, 3, 3), 0.6
    st = eph.setup_outer_q(m, kgrid, qgrid, (-w, w), backend)
    old = tune.get_variant(tune.KIND_GRIDOPT_KQ)
    tune.set_variant(tune.KIND_GRIDOPT_KQ, variant)
    try:
        loop = eph.OuterKLoop(st, nk_batch=40, precision=L.PREC_FP32, with_ep=True)
        assert loop.gridopt_q is not None and loop.ep is not None
        got = []

        class Rec(eph.AbstractCalculator):
            def run(self, p):
                got.append((p.ik, p.ibandk_offset, p.ikqs.cpu().numpy(), p.eps.cpu().numpy(), p.g2s.cpu().numpy(),
                            p.wqs.cpu().numpy()))

        adapter = eph.QBatchedAdapter(st, [Rec()], nq_batch=7)   # tiles of 7 k+q: a k spans several payloads
        loop.run_kbatch(0, 40, [adapter])
        torch.cuda.synchronize()
    finally:
        tune.set_variant(tune.KIND_GRIDOPT_KQ, old)
    ks = st.ksurv.cpu().numpy()
    U = st.el.U.cpu().numpy().astype(np.complex128)            # (nk, nw, nw) column-major blocks
    bmin, bmax = st.el.filt.band_min.cpu().numpy(), st.el.filt.band_max.cpu().numpy()
    xk, xq = M.kpoints_grid(kgrid), M.kpoints_grid(qgrid)
    om = st.ph.omega.cpu().numpy()
    uph = st.ph.u.cpu().numpy().astype(np.complex128)
    nw, nm, nbk = m.nw, m.nmodes, st.nbw
    assert sorted({g[0] for g in got}) == list(range(1, 41))
    cache, checked = {}, 0
    for ik, off, ikqs, eps, g2s, wqs in got:
        if ik % 6 != 1:
            continue
        k = int(ks[ik - 1])
        assert off == min(max(int(bmin[k]) - 1, 0), nw - nbk)
        for j, pkq in enumerate(ikqs):
            kq = int(ks[pkq - 1])
            iq = int(O.hash_xk(xk[kq] - xk[k], qgrid)[0])
            assert np.allclose(xk[k] + xq[iq] - xk[kq], np.round(xk[k] + xq[iq] - xk[kq]))
            assert np.array_equal(wqs[j], om[iq])
            if iq not in cache:
                cache[iq] = O.eph_RR_to_Rq(nw, nm, m.epmat_ph, m.irvec_el, m.irvec_ep, xq[iq], uph[iq].ravel())
            uk = U[k].reshape(-1).reshape(nw, nw, order="F")[:, off:off + nbk]   # column-major block
            ukq = U[kq].reshape(-1).reshape(nw, nw, order="F")
            ref = O.eph_Rq_to_kq(nw, nm, cache[iq], m.irvec_el, xk[k], uk, ukq)   # (nm, nbk, nw) = [nu, n, m]
            mask = np.zeros((nbk, nw), bool)
            kin = np.arange(off, off + nbk)
            mask[(kin >= bmin[k] - 1) & (kin < bmax[k])] = True
            mask[:, :bmin[kq] - 1] = False
            mask[:, bmax[kq]:] = False
            e = eps[j]                                         # (nm, nbk, nw)
            assert np.abs(e[:, ~mask]).max(initial=0.0) == 0.0 and np.abs(g2s[j][:, ~mask]).max(initial=0.0) == 0.0
            assert np.abs(e[:, mask] - ref[:, mask]).max() <= 1e-4 * np.abs(ref[:, mask]).max()
            g2r = np.abs(e) ** 2 / (2 * om[iq][:, None, None])
            assert np.abs(g2s[j] - g2r).max() <= 1e-5 * g2r.max()
            checked += 1
    assert checked >= 20, checked
