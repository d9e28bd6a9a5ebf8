"""End-to-end parity of the device outer-q e-ph loop against the FP64 oracle on a full small model
(the plain method of the reference's CPU path, cfg1 shape: nw 8, 2 atoms, 5^3 boxes, 8^3 grids).

Gauge-invariant comparison: g2 summed over degenerate subspaces (epstate_g2_degenerate_average!,
src/EPState.jl:205-238) with the 0.01 meV cutoff, modes below the acoustic cutoff excluded (the
reference's consumers skip them, src/common/constants.jl:5). Tolerance: norm-wise relative error
of |g|^2 <= 1e-4 (the e-ph bound) for FP32 and split-BF16."""
import numpy as np
import pytest

import oracle as O
from ephip import model as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cfg1():
    m, g = M.make_config_model("cfg1")
    return m, g


@pytest.fixture(params=[(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 0, 0), (0, 0, 1, 0), (1, 0, 1, 0), (0, 0, 0, 1),
                        (1, 0, 0, 1), (0, 0, 1, 1)],
                ids=["kq-valu", "kq-mfma", "kq-mfma-xcd", "kq-valu-general", "kq-mfma-general", "kq-valu-pre",
                     "kq-mfma-pre", "kq-valu-general-pre"])
def kq_variant(request):
    """The staged kR->kq kernel form (tuner kind 6): 0 = stage 3 and the k+q-window rotation on the vector
    ALUs, 1 = on FP32 MFMAs; its workgroup order (kind 7): 0 = plain, 1 = XCD-grouped; its stage-2/3 sums
    (kind 8): 0 = folded on symmetric Rp boxes, 1 = always the general form; and where stage 1 runs: 0 = in
    each workgroup, 1 = by the pre-pass (cube boxes; others keep the in-kernel stage 1)."""
    from ephip import eph, tune
    kinds = (tune.KIND_GRIDOPT_KQ, tune.KIND_KQ_ORDER, tune.KIND_KQ_FOLD)
    old = [tune.get_variant(k) for k in kinds]
    old_pre = eph._GridOptKQ.prepass_default
    for k, v in zip(kinds, request.param):
        tune.set_variant(k, v)
    eph._GridOptKQ.prepass_default = bool(request.param[3])
    yield request.param[0]
    for k, v in zip(kinds, old):
        tune.set_variant(k, v)
    eph._GridOptKQ.prepass_default = old_pre


def _oracle_states(m, g):
    xk = M.kpoints_grid(g["kgrid"])
    E, U = O.heev_batched(O.fourier(m.ham_r, m.irvec_h, xk))
    keep, bmin, bmax, _ = O.window_filter(E, -g["window"], g["window"])
    xq = M.kpoints_grid(g["qgrid"])
    Dq = O.fourier(m.dyn_r, m.irvec_d, xq)
    ph = [O.ph_eigen(Dq[i], m.mass) for i in range(len(xq))]
    return xk, E, U, keep, bmin, bmax, xq, ph


def _outer_q_parity(backend, m, g, prec, iqs):
    """Outer-q tiles of the q in iqs against the oracle's RR->Rq->kq: k lists exact, |g|^2 within 1e-4
    (degenerate-averaged, max-norm relative)."""
    from ephip import eph
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    xk, E, U, keep, bmin, bmax, xq, ph = _oracle_states(m, g)
    assert np.array_equal(st.el.filt.keep.cpu().numpy(), keep)
    assert np.array_equal(st.ksurv.cpu().numpy(), np.flatnonzero(keep))
    loop = eph.OuterQLoop(st, nk_batch=100, nq_tile=7, precision=prec)
    rec = eph.G2Recorder()
    for iq in iqs:
        loop.run_qtile(iq, 1, [rec])
    nw, nm, nbw = m.nw, m.nmodes, st.nbw
    cutoff = 1e-5  # 0.01 meV in eV
    errs = []
    for iq in iqs:
        om, uph = ph[iq]
        eRpq = O.eph_RR_to_Rq(nw, nm, m.epmat_ph, m.irvec_el, m.irvec_ep, xq[iq], uph)
        blocks = [b for b in rec.blocks if b[0] == iq]
        ik = np.concatenate([b[1] for b in blocks])
        ikq = np.concatenate([b[2] for b in blocks])
        g2 = np.concatenate([b[3] for b in blocks])  # (nl, nm, nbw, nbw) = [nu, n, m]
        assert np.array_equal(ik, np.flatnonzero(keep))
        assert np.array_equal(ikq, O.hash_xk(xk[ik] + xq[iq], g["kgrid"]))
        ep_ref = O.outer_q_block(nw, nm, eRpq, m.irvec_el, xk, U, bmin, bmax, ik, ikq, nbw)
        ok = np.abs(om) >= 1e-5
        for t in range(len(ik)):
            a, b = ik[t], ikq[t]
            if not keep[b]:
                assert not g2[t].any()
                continue
            nbk, nbkq = bmax[a] - bmin[a] + 1, bmax[b] - bmin[b] + 1
            ek, ekq = E[a, bmin[a] - 1:bmax[a]], E[b, bmin[b] - 1:bmax[b]]
            r = O.g2(ep_ref[t][:, :nbk, :nbkq], om)
            gg = g2[t][:, :nbk, :nbkq].astype(np.float64)
            r = O.g2_degenerate_average(r, ekq, ek, cutoff)[ok]
            gg = O.g2_degenerate_average(gg, ekq, ek, cutoff)[ok]
            assert not g2[t][:, nbk:, :].any() and not g2[t][:, :, nbkq:].any()
            errs.append((np.abs(gg - r).max(), np.abs(r).max()))
    e = np.array(errs)
    rel = e[:, 0].max() / e[:, 1].max()
    assert rel < 1e-4, rel


@pytest.mark.parametrize("prec_name", ["fp32", "bf16x3"])
def test_outer_q_loop_g2_parity(backend, cfg1, prec_name):
    from ephip import _lib as L
    m, g = cfg1
    prec = L.PREC_FP32 if prec_name == "fp32" else L.PREC_BF16X3
    _outer_q_parity(backend, m, g, prec, [0, 1, 9, 77, 300, 511])


@pytest.mark.parametrize("nw", [12, 20])
def test_outer_q_loop_other_band_counts(backend, nw):
    """The plain outer-q loop at Wannier band counts other than 8 and 16 (FP32), against the oracle."""
    from ephip import _lib as L
    m = M.make_model(nw=nw, natoms=2, box_hd=5, box_g=3, seed=3200 + nw, g_dtype=np.complex128)
    g = {"kgrid": (6, 6, 6), "qgrid": (3, 3, 3), "window": 0.5}
    _outer_q_parity(backend, m, g, L.PREC_FP32, [0, 1, 13, 26])


def _outer_k_vs_oracle(m, g, st, rec, qstep=23):
    """Compare recorded outer-k payloads with the oracle's RR->Rq->kq on the same pairs: kept-q lists
    exact, |g|^2 within 1e-4 norm-wise after degenerate averaging. Returns the relative error."""
    xk, E, U, keep, bmin, bmax, xq, ph = _oracle_states(m, g)
    nw, nm = m.nw, m.nmodes
    per_kq = {}
    for b in rec.batches:
        for u in range(len(b["unit_state"])):
            k, n = int(b["unit_state"][u]), int(b["unit_band"][u])
            expect = [c for c in range(len(xq)) if keep[O.hash_xk(xk[k] + xq[c], g["kgrid"])[0]]]
            assert list(b["cols"][u]) == expect
            assert np.array_equal(b["ikq"][u], O.hash_xk(xk[k] + xq[b["cols"][u]], g["kgrid"]))
            for c, q in enumerate(b["cols"][u]):
                per_kq.setdefault((k, int(q)), {})[n] = b["g2"][u][c].reshape(st.nbw, nm)  # [m, nu]
    qsel = set(range(0, len(xq), qstep)) | {0, 1}
    cutoff, errs = 1e-5, []
    eRpq_cache = {}
    for (k, q), bands in per_kq.items():
        if q not in qsel:
            continue
        om, uph = ph[q]
        if q not in eRpq_cache:
            eRpq_cache[q] = O.eph_RR_to_Rq(nw, nm, m.epmat_ph, m.irvec_el, m.irvec_ep, xq[q], uph)
        kq = O.hash_xk(xk[k] + xq[q], g["kgrid"])[0]
        Uk = U[k].reshape(nw, nw, order="F")[:, bmin[k] - 1:bmax[k]]
        Ukq = U[kq].reshape(nw, nw, order="F")[:, bmin[kq] - 1:bmax[kq]]
        ep = O.eph_Rq_to_kq(nw, nm, eRpq_cache[q], m.irvec_el, xk[k], Uk, Ukq)  # [nu, n, m]
        nbkq = Ukq.shape[1]
        assert sorted(bands) == list(range(bmin[k] - 1, bmax[k]))
        got = np.stack([bands[n][:nbkq, :].T for n in range(bmin[k] - 1, bmax[k])], axis=1)  # [nu, n, m]
        assert not np.any([bands[n][nbkq:, :] for n in bands])
        ref = O.g2(ep, om)
        ok = np.abs(om) >= 1e-5
        ek, ekq = E[k, bmin[k] - 1:bmax[k]], E[kq, bmin[kq] - 1:bmax[kq]]
        r = O.g2_degenerate_average(ref, ekq, ek, cutoff)[ok]
        gg = O.g2_degenerate_average(got.astype(np.float64), ekq, ek, cutoff)[ok]
        errs.append((np.abs(gg - r).max(), np.abs(r).max()))
    assert len(errs) > 20
    e = np.array(errs)
    return e[:, 0].max() / e[:, 1].max()


@pytest.mark.parametrize("prec_name", ["fp32", "bf16x3"])
def test_outer_k_loop_g2_parity(backend, cfg1, prec_name):
    """Outer-k loop with the k-side window projection (RR->kR->kq) against the oracle's RR->Rq->kq
    on the same pairs: the kept-q lists are exact (q kept iff k+q has an in-window band), |g|^2 within
    1e-4 norm-wise after degenerate averaging."""
    from ephip import _lib as L
    from ephip import eph
    m, g = cfg1
    prec = L.PREC_FP32 if prec_name == "fp32" else L.PREC_BF16X3
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    loop = eph.OuterKLoop(st, nk_batch=37, precision=prec)
    rec = eph.UnitRecorder()
    loop.run_kbatch(0, 37, [rec])
    loop.run_kbatch(200, 37, [rec])
    rel = _outer_k_vs_oracle(m, g, st, rec)
    assert rel < 1e-4, rel


@pytest.fixture(scope="module")
def model16():
    """nw 16, 4 atoms (nm 12): the e-ph unit shape of the larger workloads, on small boxes and grids."""
    m = M.make_model(nw=16, natoms=4, box_hd=5, box_g=3, seed=2024, g_dtype=np.complex128)
    return m, {"kgrid": (8, 8, 8), "qgrid": (4, 4, 4), "window": 0.5}


@pytest.mark.parametrize("variant,nq,window", [
    (0, 4, 0.5), (1, 4, 0.5), (2, 4, 0.5), (3, 4, 0.5), (4, 4, 0.5), (5, 4, 0.5), (6, 4, 0.5), (7, 4, 0.5),
    (8, 4, 0.5), (9, 4, 0.5), (10, 4, 0.5), (11, 4, 0.5), (12, 4, 0.5), (0, 8, 0.5), (5, 8, 0.5), (6, 8, 0.5),
    (8, 8, 0.5), (10, 8, 0.5), (12, 8, 0.5),
    (0, 4, 0.8), (5, 4, 0.8), (10, 4, 0.8), (10, 8, 0.8), (12, 4, 0.8),
    (0, 4, 1.5), (4, 4, 1.5), (5, 4, 1.5), (10, 4, 1.5), (12, 4, 1.5)])
def test_outer_k_fused_rotation(backend, model16, variant, nq, window):
    """Split-BF16 outer-k with the kR->kq Fourier fused to the rotations (nw 16, nm 12), every kernel
    variant: same kept-q lists as the unfused kernels, g2 within 1e-5 of them, and within 1e-4 of the
    oracle. The 8^3 q-grid gives units several column tiles of both widths (64 and 128); windows of
    0.5, 0.8 and 1.5 eV give nbw = 2, 3 and 4 (the kernels' window-width instantiations)."""
    from ephip import _lib as L
    from ephip import eph, tune
    tune.set_variant(tune.KIND_FUSED_ROT, variant)
    m, g = model16
    g = dict(g, qgrid=(nq, nq, nq), window=window)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    assert st.nbw == {0.5: 2, 0.8: 3, 1.5: 4}[window]
    fused = eph.OuterKLoop(st, nk_batch=40, precision=L.PREC_BF16X3, gridopt=False)
    plain = eph.OuterKLoop(st, nk_batch=40, precision=L.PREC_BF16X3, fuse_rotation=False, gridopt=False)
    assert fused.fused and not plain.fused and fused.Gq is None
    rf, rp = eph.UnitRecorder(), eph.UnitRecorder()
    nks = st.ksurv.shape[0]
    for i0 in (0, max(0, nks - 40)):
        fused.run_kbatch(i0, 40, [rf])
        plain.run_kbatch(i0, 40, [rp])
    num = den = 0.0
    for bf, bp in zip(rf.batches, rp.batches):
        assert np.array_equal(bf["unit_state"], bp["unit_state"])
        for u in range(len(bf["cols"])):
            assert np.array_equal(bf["cols"][u], bp["cols"][u])
            assert np.array_equal(bf["ikq"][u], bp["ikq"][u])
            num = max(num, float(np.abs(bf["g2"][u] - bp["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bp["g2"][u]).max(initial=0.0)))
    assert den > 0 and num / den < 1e-5, num / den
    if variant >= 5:
        # eight-wave variants: a grid of one column tile (workgroups loop over the unit's tiles) gives the
        # same g2 bit for bit as the grid sized by the kept-q bound
        assert fused.max_kept >= max(len(c) for b in rf.batches for c in b["cols"])
        narrow = eph.OuterKLoop(st, nk_batch=40, precision=L.PREC_BF16X3, gridopt=False)
        narrow.max_kept = 1
        rn = eph.UnitRecorder()
        for i0 in (0, max(0, nks - 40)):
            narrow.run_kbatch(i0, 40, [rn])
        for bf, bn in zip(rf.batches, rn.batches):
            for u in range(len(bf["cols"])):
                assert np.array_equal(bf["g2"][u], bn["g2"][u])
    rel = _outer_k_vs_oracle(m, g, st, rf, qstep=5 if nq == 4 else 37)
    tune.set_variant(tune.KIND_FUSED_ROT, 0)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("nq,variant", [(4, 0), (8, 0), (4, 1), (8, 1)])
def test_outer_k_fused_projection(backend, model16, nq, variant):
    """Split-BF16 outer-k with RR->kR fused to the k-side window projection (nw 16): the unit operators
    A and the g2 payloads agree with the separate Fourier + projection kernels to 1e-5, and g2 with the
    oracle to 1e-4. Batches of 40 and 150 k give one and two column tiles of 128 k, with a partial tile."""
    from ephip import _lib as L
    from ephip import eph, tune
    m, g = model16
    g = dict(g, qgrid=(nq, nq, nq))
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    tune.set_variant(tune.KIND_FUSED_PROJ, variant)
    for nb in (40, 150):
        fused = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_BF16X3, gridopt=False)
        plain = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_BF16X3, fuse_projection=False, gridopt=False)
        assert fused.fused_proj and not plain.fused_proj and fused.G is None
        rf, rp = eph.UnitRecorder(), eph.UnitRecorder()
        for i0 in (0, max(0, nks - nb)):
            fused.run_kbatch(i0, nb, [rf])
            plain.run_kbatch(i0, nb, [rp])
            nu = int(fused.k_unit_start[min(i0 + nb, nks)] - fused.k_unit_start[i0])
            a, b = fused.A[:nu].cpu().numpy(), plain.A[:nu].cpu().numpy()
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
        for bf, bp in zip(rf.batches, rp.batches):
            assert np.array_equal(bf["unit_state"], bp["unit_state"])
            for u in range(len(bf["cols"])):
                assert np.array_equal(bf["cols"][u], bp["cols"][u])
                d = np.abs(bf["g2"][u] - bp["g2"][u]).max(initial=0.0)
                assert d <= 1e-5 * max(np.abs(bp["g2"][u]).max(initial=0.0), 1e-30)
    tune.set_variant(tune.KIND_FUSED_PROJ, 0)
    rel = _outer_k_vs_oracle(m, g, st, rf, qstep=5 if nq == 4 else 37)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("which,nb,mfma", [("model16", 40, True), ("model16", 150, True), ("model16", 40, False),
                                          ("model16", 150, False), ("cfg1", 64, True), ("cfg1", 7, True)])
def test_outer_k_gridopt(backend, model16, cfg1, which, nb, mfma):
    """RR->kR as the staged (GridOpt) transform with the projection fused (FP32; for nw 16 stages 2/3 on
    the matrix cores or on the vector ALUs): the unit operators A agree with the direct FP32 Fourier +
    projection kernels to 1e-5 and g2 with the oracle to 1e-4. Batches of 7 to 150 k cut lines and k1
    planes; the batches are run out of order and one twice, so cached plane sums are reused and evicted."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16 if which == "model16" else cfg1
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt_mfma=mfma)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt is not None and direct.gridopt is None
    assert staged.gridopt.mfma == (mfma and m.nw == 16)
    rs = eph.UnitRecorder()
    starts = [0, max(0, nks - nb), nb % nks, 0, min(2 * nb, max(0, nks - 1))]
    for i0 in starts:
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb)
        nu = int(staged.k_unit_start[min(i0 + nb, nks)] - staged.k_unit_start[i0])
        a, b = staged.A[:nu].cpu().numpy(), direct.A[:nu].cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), (i0, np.linalg.norm(a - b) / np.linalg.norm(b))
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=5)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("nq,window,nb", [(4, 0.5, 40), (8, 0.5, 150), (4, 0.8, 40), (8, 1.5, 64), (8, 0.15, 40)])
def test_outer_k_gridopt_kq(backend, model16, nq, window, nb, kq_variant):
    """kR->kq as the staged (GridOpt) transform over the uniform q grid with the k+q/phonon rotations and
    g2 fused (FP32): kept-q lists and g2 agree with the direct FP32 Fourier + rotation kernels (g2 to
    1e-5) and with the oracle (1e-4), for nbw = 1 (one column through the two-column kernel), 2, 3 and 4
    and q grids of 4^3 and 8^3."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, qgrid=(nq, nq, nq), window=window)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt_q is not None and staged.gridopt is not None and direct.gridopt_q is None
    assert st.nbw == {0.15: 1, 0.5: 2, 0.8: 3, 1.5: 4}[window]
    rs, rd = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb, [rd])
    num = den = 0.0
    for bs, bd in zip(rs.batches, rd.batches):
        assert np.array_equal(bs["unit_state"], bd["unit_state"])
        for u in range(len(bs["cols"])):
            assert np.array_equal(bs["cols"][u], bd["cols"][u])
            num = max(num, float(np.abs(bs["g2"][u] - bd["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bd["g2"][u]).max(initial=0.0)))
    assert den > 0 and num / den < 1e-5, num / den
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=5 if nq == 4 else 37)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("box_g", [5])
def test_outer_k_gridopt_box_faces(backend, box_g, kq_variant):
    """The staged kernels on 5^3 e-ph R boxes (other instantiations of the stage-2/3 and kR->kq kernels
    than the 3^3 and 7^3 boxes elsewhere): unit operators and g2 agree with the direct FP32 kernels to
    1e-5 and g2 with the oracle to 1e-4."""
    from ephip import _lib as L
    from ephip import eph
    m = M.make_model(nw=16, natoms=4, box_hd=5, box_g=box_g, seed=2025 + box_g, g_dtype=np.complex128)
    g = {"kgrid": (6, 6, 6), "qgrid": (3, 3, 3), "window": 0.5}
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    nb = 50
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt is not None and staged.gridopt.mfma and staged.gridopt_q is not None
    assert staged.gridopt.np == max(box_g, 3)
    rs, rd = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb, [rd])
        nu = int(staged.k_unit_start[min(i0 + nb, nks)] - staged.k_unit_start[i0])
        a, b = staged.A[:nu].cpu().numpy(), direct.A[:nu].cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    num = den = 0.0
    for bs, bd in zip(rs.batches, rd.batches):
        for u in range(len(bs["cols"])):
            assert np.array_equal(bs["cols"][u], bd["cols"][u])
            num = max(num, float(np.abs(bs["g2"][u] - bd["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bd["g2"][u]).max(initial=0.0)))
    assert den > 0 and num / den < 1e-5, num / den
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=7)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("natoms,window", [(3, 0.5), (3, 0.8), (5, 0.5), (5, 0.8)])
def test_outer_k_gridopt_mode_counts(backend, natoms, window, kq_variant):
    """The staged kernels at mode counts other than 12 (nm 9 and 15: the kR->kq kernel's run-time mode
    loop, the MFMA projection over nm blocks of 16 rows; nm 15 fills 240 of 256 threads), window widths
    2 (column pairs) and 3 (odd path): unit operators and g2 agree with the direct FP32 kernels to 1e-5
    and g2 with the oracle to 1e-4."""
    from ephip import _lib as L
    from ephip import eph
    m = M.make_model(nw=16, natoms=natoms, box_hd=5, box_g=3, seed=3031, g_dtype=np.complex128)
    assert m.nmodes == 3 * natoms
    g = {"kgrid": (6, 6, 6), "qgrid": (4, 4, 4), "window": window}
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    assert st.nbw == {0.5: 2, 0.8: 3}[window]
    nks = st.ksurv.shape[0]
    nb = 60
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt is not None and staged.gridopt.mfma and staged.gridopt_q is not None
    rs, rd = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb, [rd])
        nu = int(staged.k_unit_start[min(i0 + nb, nks)] - staged.k_unit_start[i0])
        a, b = staged.A[:nu].cpu().numpy(), direct.A[:nu].cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    num = den = 0.0
    for bs, bd in zip(rs.batches, rd.batches):
        assert np.array_equal(bs["unit_state"], bd["unit_state"])
        for u in range(len(bs["cols"])):
            assert np.array_equal(bs["cols"][u], bd["cols"][u])
            num = max(num, float(np.abs(bs["g2"][u] - bd["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bd["g2"][u]).max(initial=0.0)))
    assert den > 0 and num / den < 1e-5, num / den
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=7)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("window", [0.5, 0.8])
def test_outer_k_gridopt_kq_long_lines(backend, model16, window, kq_variant):
    """Staged kR->kq on a q grid whose lines hold more kept q than one chunk of the kernel (32): k and q
    grids (2, 4, 40), so the kernel splits lines; g2 agrees with the direct FP32 kernels and the oracle."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, kgrid=(2, 4, 40), qgrid=(2, 4, 40), window=window)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    nb = 40
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt_q is not None and staged.gridopt is not None
    rs, rd = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb, [rd])
    longest = 0
    num = den = 0.0
    for bs, bd in zip(rs.batches, rd.batches):
        for u in range(len(bs["cols"])):
            cols = np.asarray(bs["cols"][u])
            assert np.array_equal(cols, bd["cols"][u])
            longest = max(longest, int(np.bincount(cols // 40).max(initial=0)))
            num = max(num, float(np.abs(bs["g2"][u] - bd["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bd["g2"][u]).max(initial=0.0)))
    assert longest > 32, longest  # some (unit, q1, q2) line is split across chunks
    assert den > 0 and num / den < 1e-5, num / den
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=37)
    assert rel < 1e-4, rel


def test_outer_k_wide_window_takes_direct_kernels(backend, model16):
    """A window wider than the staged kR->kq kernel covers (more than 4 bands at k+q: nbw 5 here) runs
    the outer-k loop on the direct FP32 kernels; g2 against the oracle (1e-4)."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, qgrid=(4, 4, 4), window=3.0)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    assert st.nbw > 4, st.nbw
    loop = eph.OuterKLoop(st, nk_batch=64, precision=L.PREC_FP32)
    assert loop.gridopt_q is None
    rec = eph.UnitRecorder()
    nks = st.ksurv.shape[0]
    for i0 in (0, max(0, nks - 64)):
        loop.run_kbatch(i0, 64, [rec])
    rel = _outer_k_vs_oracle(m, g, st, rec, qstep=5)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("nw", [12, 20, 32])
def test_outer_k_other_band_counts(backend, nw):
    """Wannier band counts other than 8 and 16 (which the staged transforms cover) run the outer-k loop on
    the direct FP32 kernels; kept lists exact and g2 against the oracle (1e-4)."""
    from ephip import _lib as L
    from ephip import eph
    m = M.make_model(nw=nw, natoms=2, box_hd=5, box_g=3, seed=3100 + nw, g_dtype=np.complex128)
    g = {"kgrid": (6, 6, 6), "qgrid": (3, 3, 3), "window": 0.5}
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    loop = eph.OuterKLoop(st, nk_batch=48, precision=L.PREC_FP32)
    assert loop.gridopt is None and loop.gridopt_q is None
    rec = eph.UnitRecorder()
    nks = st.ksurv.shape[0]
    for i0 in (0, max(0, nks - 48)):
        loop.run_kbatch(i0, 48, [rec])
    rel = _outer_k_vs_oracle(m, g, st, rec, qstep=4)
    assert rel < 1e-4, rel


def test_outer_k_many_modes_take_direct_kr_kq(backend):
    """More phonon modes than the staged kR->kq kernel covers (6 atoms, 18 modes > 16): the RR->kR side
    stays staged, kR->kq runs on the direct FP32 kernels; g2 against the oracle (1e-4)."""
    from ephip import _lib as L
    from ephip import eph
    m = M.make_model(nw=16, natoms=6, box_hd=5, box_g=3, seed=3033, g_dtype=np.complex128)
    g = {"kgrid": (6, 6, 6), "qgrid": (3, 3, 3), "window": 0.5}
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    assert m.nmodes == 18
    loop = eph.OuterKLoop(st, nk_batch=48, precision=L.PREC_FP32)
    assert loop.gridopt_q is None
    rec = eph.UnitRecorder()
    nks = st.ksurv.shape[0]
    for i0 in (0, max(0, nks - 48)):
        loop.run_kbatch(i0, 48, [rec])
    rel = _outer_k_vs_oracle(m, g, st, rec, qstep=4)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("window", [0.5, 0.8])
def test_outer_k_gridopt_kq_segmented_slab(backend, model16, window):
    """Staged kR->kq on a q grid whose (unit, q1) slabs hold more than the kernel's 1024-entry LDS tables
    (n2 n3 = 48 x 48 = 2304, about 1500 and 2000 of them kept at these windows): the slab is taken in
    segments of whole q lines; kept lists exact, g2 against the direct FP32 kernels (1e-5) and the oracle
    (1e-4)."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, kgrid=(2, 48, 48), qgrid=(2, 48, 48), window=window)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    nb = 24
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt_q is not None and staged.gridopt is not None
    rs, rd = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb, [rd])
    most = 0
    num = den = 0.0
    for bs, bd in zip(rs.batches, rd.batches):
        for u in range(len(bs["cols"])):
            cols = np.asarray(bs["cols"][u])
            assert np.array_equal(cols, bd["cols"][u])
            most = max(most, int(np.bincount(cols // (48 * 48)).max(initial=0)))
            num = max(num, float(np.abs(bs["g2"][u] - bd["g2"][u]).max(initial=0.0)))
            den = max(den, float(np.abs(bd["g2"][u]).max(initial=0.0)))
    assert most > 1024, most  # some (unit, q1) slab spans more than one segment
    assert den > 0 and num / den < 1e-5, num / den
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=331)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("mfma", [True, False])
def test_outer_k_gridopt_long_k_lines(backend, model16, mfma):
    """Staged RR->kR on a k grid whose (k1, k2) lines hold more surviving k than the projection kernels take
    per line (64; k grid (2, 2, 100), the shape of k3 grids above 64 such as 96^3): the lines are cut into
    pieces of at most 64 k; unit operators agree with the direct FP32 kernels to 1e-5 and g2 with the
    oracle to 1e-4."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, kgrid=(2, 2, 100), qgrid=(2, 2, 4), window=0.5)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    ks = st.ksurv.cpu().numpy()
    per_line = np.bincount(ks // 100)
    assert per_line.max() > 64, per_line            # some (k1, k2) line outgrows one piece
    nks = len(ks)
    nb = 150
    staged = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt_mfma=mfma)
    direct = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32, gridopt=False)
    assert staged.gridopt is not None and staged.gridopt.max_line_k <= 64
    rs = eph.UnitRecorder()
    for i0 in (0, max(0, nks - nb)):
        staged.run_kbatch(i0, nb, [rs])
        direct.run_kbatch(i0, nb)
        nu = int(staged.k_unit_start[min(i0 + nb, nks)] - staged.k_unit_start[i0])
        a, b = staged.A[:nu].cpu().numpy(), direct.A[:nu].cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), (i0, np.linalg.norm(a - b) / np.linalg.norm(b))
    rel = _outer_k_vs_oracle(m, g, st, rs, qstep=3)
    assert rel < 1e-4, rel


@pytest.mark.parametrize("gridopt", [True, False])
def test_outer_k_pipelined(backend, model16, gridopt):
    """Batches run on two streams (RR->kR of the next batch beside kR->kq of the current one, double-
    buffered unit operators) give the same payloads, bit for bit, as the batches run one by one."""
    from ephip import _lib as L
    from ephip import eph
    m, g = model16
    g = dict(g, qgrid=(8, 8, 8))
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)
    nks = st.ksurv.shape[0]
    nb = 37
    starts = list(range(0, nks, nb)) + [0, nb]  # every batch, then two again (buffers reused)
    prec = L.PREC_FP32 if gridopt else L.PREC_BF16X3
    seq = eph.OuterKLoop(st, nk_batch=nb, precision=prec, gridopt=gridopt)
    pip = eph.OuterKLoop(st, nk_batch=nb, precision=prec, gridopt=gridopt)
    rs, rp = eph.UnitRecorder(), eph.UnitRecorder()
    for i0 in starts:
        seq.run_kbatch(i0, nb, [rs])
    pip.run_pipelined(starts, nb, [rp])
    assert len(rs.batches) == len(rp.batches) == len(starts)
    for bs, bp in zip(rs.batches, rp.batches):
        assert np.array_equal(bs["unit_state"], bp["unit_state"])
        for u in range(len(bs["cols"])):
            assert np.array_equal(bs["cols"][u], bp["cols"][u])
            assert np.array_equal(bs["g2"][u], bp["g2"][u])
