"""GPU parity of each libephip kernel against the FP64 oracle (called through the C-ABI).

Error metric for floating point: norm-wise max|got - ref| / max|ref| (the reference's own test
style, test/test_gpu.jl:508,516). Tolerances are written in each test."""
import numpy as np
import pytest

import oracle as O
from ephip import model as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


# ------------------------------------------------------------------------------ Fourier
@pytest.mark.parametrize("ndata,box,nk", [(1, 3, 1), (77, 5, 130), (256, 9, 257), (1000, 7, 64)])
def test_fourier_parity(backend, ndata, box, nk):
    from ephip import core
    from ephip._lib import PREC_BF16X3, PREC_FP32
    rng = np.random.default_rng(ndata + nk)
    irvec = M.rbox(box)
    nr = len(irvec)
    decay = np.exp(-np.linalg.norm(irvec, axis=1) / 1.5)[:, None]
    op = (rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))) * decay
    xk = rng.random((nk, 3)) * 2 - 0.5
    ref = O.fourier(op, irvec, xk)
    xk_d = _dev(xk)
    obj64 = core.WannierObject(irvec, _dev(op))
    got64 = core.fourier_batched(obj64, xk_d).cpu().numpy()
    assert _rel(got64, ref) < 1e-13, "fp64"
    obj32 = core.WannierObject(irvec, _dev(op.astype(np.complex64)))
    got32 = core.fourier_batched(obj32, xk_d, precision=PREC_FP32).cpu().numpy()
    assert _rel(got32, ref) < 2e-6, "fp32"
    gotb = core.fourier_batched(obj32, xk_d, precision=PREC_BF16X3).cpu().numpy()
    assert _rel(gotb, ref) < 2e-5, "split-bf16"


@pytest.mark.parametrize("shift", [(1, -2, 0), (-3, 0, 4)])
def test_fourier_offset_rvectors(backend, shift):
    """R ranges that are not centred on 0 (per-axis phase tables start at the smallest R_a), with
    k-points far outside the first zone."""
    from ephip import core
    from ephip._lib import PREC_BF16X3, PREC_FP32
    rng = np.random.default_rng(sum(shift) + 50)
    irvec = M.rbox(5) + np.array(shift, dtype=np.int32)
    nr, ndata, nk = len(irvec), 90, 150
    op = rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))
    xk = rng.random((nk, 3)) * 40 - 20
    ref = O.fourier(op, irvec, xk)
    obj32 = core.WannierObject(irvec, _dev(op.astype(np.complex64)))
    assert _rel(core.fourier_batched(obj32, _dev(xk), precision=PREC_FP32).cpu().numpy(), ref) < 2e-6
    assert _rel(core.fourier_batched(obj32, _dev(xk), precision=PREC_BF16X3).cpu().numpy(), ref) < 2e-5


def test_fourier_leading_dimensions(backend):
    from ephip import core
    rng = np.random.default_rng(3)
    irvec = M.rbox(3)
    op = rng.standard_normal((27, 40)) + 1j * rng.standard_normal((27, 40))
    xk = rng.random((70, 3))
    ref = O.fourier(op, irvec, xk, ndata=33)
    obj = core.WannierObject(irvec, _dev(op.astype(np.complex64)), ndata=33)
    out = torch.full((70, 50), 7.0 + 7j, dtype=torch.complex64, device="cuda")
    core.fourier_batched(obj, _dev(xk), out=out)
    o = out.cpu().numpy()
    assert _rel(o[:, :33], ref) < 2e-6
    assert np.all(o[:, 33:] == 7.0 + 7j)  # padding beyond ndata untouched


def test_fourier_phase_parity(backend):
    from ephip import core
    rng = np.random.default_rng(4)
    irvec = M.rbox(5)
    xk = rng.random((33, 3))
    got = core.fourier_phase(_dev(irvec), _dev(xk)).cpu().numpy()
    assert np.abs(got - O.fourier_phase(irvec, xk)).max() < 1e-14


# ------------------------------------------------------------------------------ eigensolver
def _planted(rng, n, nb, degenerate_frac=0.3, split=0.0):
    hs = []
    for b in range(nb):
        q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        w = np.sort(rng.uniform(-5, 5, n))
        if n >= 3 and b < degenerate_frac * nb:
            w[1] = w[2] = w[1] if split == 0 else w[1]
            w[2] = w[1] + split
            w = np.sort(w)
        hs.append((q * w) @ q.conj().T)
    return np.stack(hs)


@pytest.mark.parametrize("n", [1, 2, 3, 6, 8, 12, 16, 17, 24, 32])
def test_heev_c32_parity(backend, n):
    from ephip import core
    rng = np.random.default_rng(100 + n)
    nb = 203
    a = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    h = a + np.conj(np.transpose(a, (0, 2, 1)))
    hc = np.ascontiguousarray(np.transpose(h, (0, 2, 1)))  # column-major blocks
    E, U = core.eigen_batched(_dev(hc.astype(np.complex64)))
    E, U = E.cpu().numpy(), U.cpu().numpy()
    Ev = core.eigvals_batched(_dev(hc)).cpu().numpy()  # c64 input, values only
    for b in range(nb):
        wl = np.linalg.eigvalsh(h[b])
        scale = np.abs(wl).max()
        assert np.abs(E[b] - wl).max() < 3e-6 * scale
        assert np.abs(Ev[b] - wl).max() < 3e-6 * scale
        V = U[b].T  # column-major block -> matrix
        assert np.abs(h[b] @ V - V * E[b]).max() < 5e-6 * scale
        assert np.abs(V.conj().T @ V - np.eye(n)).max() < 3e-7 * n


@pytest.mark.parametrize("n,split", [(8, 0.0), (16, 0.0), (16, 3e-6), (12, 5e-5), (32, 0.0), (32, 3e-6)])
def test_heev_refine_fp64(backend, n, split):
    """FP32 solve + FP64 refinement (Loewdin orthonormalisation, projection, parallel Jacobi) reaches
    FP64-grade eigenvalues, also on exactly and nearly degenerate pairs; vectors are orthonormal to
    FP32 rounding."""
    from ephip import core
    rng = np.random.default_rng(7 + n)
    h = _planted(rng, n, 150, degenerate_frac=0.4, split=split)
    hc = np.ascontiguousarray(np.transpose(h, (0, 2, 1)))
    E, U = core.eigen_batched_refined(_dev(hc))
    E, U = E.cpu().numpy(), U.cpu().numpy()
    for b in range(len(h)):
        wl = np.linalg.eigvalsh(h[b])
        assert np.abs(E[b] - wl).max() < 1e-13 * np.abs(wl).max()
        V = U[b].T.astype(np.complex128)
        assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-6
        assert np.abs(h[b] @ V - V * E[b]).max() < 1e-6 * np.abs(wl).max()


@pytest.mark.parametrize("n", [8, 16, 31, 32])
def test_degenerate_gauge_fix_parity(backend, n):
    """ephip_degenerate_gauge_fix after the refined eigensolve against the oracle's solve_eigen_el! gauge
    fixing (src/diagonalize.jl:47-85) on matrices with planted 2- and 3-fold groups, pairs split by 2e-6
    (inside the 1e-5 cutoff), a chain cut by the walk and non-degenerate spectra: every column equals the
    oracle's up to a phase (|overlap| >= 1 - 1e-6: FP32 vectors); matrices without a close pair are left
    bit-for-bit unchanged."""
    from ephip import core
    rng = np.random.default_rng(100 + n)
    cut = 1e-5
    spectra = []
    for b in range(96):
        e = np.sort(rng.uniform(-2.0, 2.0, n))
        kind = b % 4
        if kind == 1:
            i = int(rng.integers(0, n - 1))
            e[i + 1] = e[i]
        elif kind == 2 and n >= 6:
            i = int(rng.integers(0, n - 5))
            e[i + 1] = e[i + 2] = e[i]
            e[i + 4] = e[i + 3] + 2e-6
        elif kind == 3 and n >= 4:
            i = int(rng.integers(0, n - 3))
            e[i + 1], e[i + 2] = e[i] + 0.6e-5, e[i] + 1.2e-5
        e = np.sort(e)
        if kind == 0:   # keep non-degenerate spectra well separated from the cutoff
            e += np.arange(n) * 1e-3
        spectra.append(e)
    hs = []
    for e in spectra:
        q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        h = q @ np.diag(e) @ q.conj().T
        hs.append(0.5 * (h + h.conj().T))
    hcol = np.stack([h.ravel(order="F") for h in hs])
    E_ref, U_ref = O.heev_batched(hcol, degen_cutoff=cut)
    E, U = core.eigen_batched_refined(_dev(hcol.reshape(-1, n, n)))
    U_raw = U.clone()
    core.degenerate_gauge_fix(E, U, cut)
    E, U, U_raw = E.cpu().numpy(), U.cpu().numpy(), U_raw.cpu().numpy()
    assert np.abs(E - E_ref).max() < 1e-12
    worst, fixed = 0.0, 0
    for b in range(len(hs)):
        close = bool((np.diff(E[b]) < cut).any())
        if not close:
            assert np.array_equal(U[b], U_raw[b])
            continue
        fixed += 1
        assert np.bincount(O.degenerate_groups(E_ref[b], cut)).max() <= 3
        V = U[b].reshape(-1).reshape(n, n, order="F").astype(np.complex128)
        Vr = U_ref[b].reshape(n, n, order="F")
        ov = np.abs(np.sum(V.conj() * Vr, axis=0))
        worst = max(worst, float(np.abs(ov - 1).max()))
    print(f"gauge fix n={n}: {fixed} of {len(hs)} matrices fixed, worst 1 - |overlap| {worst:.2e}")
    assert fixed >= len(hs) // 2 and worst <= 1e-6


# ------------------------------------------------------------------------------ states + filter
def test_electron_states_and_filter_bit_exact(backend):
    from ephip import core, eph
    m, g = M.make_config_model("cfg1", with_g=False)
    xk = M.kpoints_grid(g["kgrid"])
    Hk = O.fourier(m.ham_r, m.irvec_h, xk)
    E_ref, _ = O.heev_batched(Hk)
    w = g["window"]
    keep_r, bmin_r, bmax_r, nb_r = O.window_filter(E_ref, -w, w)
    ham = core.WannierObject(m.irvec_h, m.ham_r).to_device(backend)
    st = eph.compute_electron_states(ham, _dev(xk), (-w, w), chunk=200)
    E = st.E.cpu().numpy()
    assert np.abs(E - E_ref).max() < 1e-6  # eV (the eigenvalue bound); FP64 refinement gives far less
    assert np.abs(E - E_ref).max() < 1e-13 * np.abs(E_ref).max()
    f = st.filt
    assert np.array_equal(f.keep.cpu().numpy(), keep_r)
    assert np.array_equal(f.band_min.cpu().numpy(), bmin_r)
    assert np.array_equal(f.band_max.cpu().numpy(), bmax_r)
    assert np.array_equal(f.nbelow.cpu().numpy(), nb_r)
    # the window kernel itself on identical input is bit-exact
    keep2 = core.window_filter(_dev(E_ref), -w, w)
    assert np.array_equal(keep2.band_min.cpu().numpy(), bmin_r)


def test_window_filter_edges(backend):
    from ephip import core
    E = np.array([[-1.0, -0.5, -0.5, 0.0, 0.5, 0.5, 1.0, 2.0],
                  [-3, -2, -1, -0.6, 0.6, 1, 2, 3],
                  [0.5, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1]], dtype=np.float64)
    f = core.window_filter(_dev(E), -0.5, 0.5)
    ref = O.window_filter(E, -0.5, 0.5)
    for got, r in zip((f.keep, f.band_min, f.band_max, f.nbelow), ref):
        assert np.array_equal(got.cpu().numpy(), r)


def test_kq_index_bit_exact(backend):
    from ephip import core
    for ng in [(4, 4, 4), (3, 4, 5), (48, 48, 48)]:
        xk = M.kpoints_grid(ng)
        rng = np.random.default_rng(sum(ng))
        sel = rng.choice(len(xk), size=min(len(xk), 5000), replace=False)
        for dq in [(0.0, 0.0, 0.0), (1 / ng[0], 2 / ng[1], -3 / ng[2]), (0.5, -0.25, 0.75)]:
            got = core.kq_index(_dev(xk[sel]), dq, ng).cpu().numpy()
            ref = O.hash_xk(xk[sel] + np.array(dq), ng)
            assert np.array_equal(got, ref)


def test_phonon_states_parity(backend):
    from ephip import core, eph
    m, g = M.make_config_model("cfg1", with_g=False)
    xq = M.kpoints_grid(g["qgrid"])
    Dq = O.fourier(m.dyn_r, m.irvec_d, xq)
    om_ref = np.array([O.ph_eigen(Dq[i], m.mass)[0] for i in range(len(xq))])
    dyn = core.WannierObject(m.irvec_d, m.dyn_r).to_device(backend)
    ph = eph.compute_phonon_states(dyn, m.mass, _dev(xq))
    om = ph.omega.cpu().numpy()
    sel = np.abs(om_ref) > 1e-5  # acoustic modes at Gamma are ~0 and sign-ambiguous
    assert np.abs(om - om_ref)[sel].max() < 1e-9
    # eigenvectors: D(q) u = w^2 M u  with u rows scaled by 1/sqrt(m)
    U = ph.u.cpu().numpy()
    for i in range(0, len(xq), 37):
        d = Dq[i].reshape(m.nmodes, m.nmodes, order="F")
        u = U[i].T.astype(np.complex128)
        lhs = d @ u
        rhs = u * (np.sign(om_ref[i]) * om_ref[i] ** 2)
        assert np.abs(lhs - rhs).max() < 1e-6 * np.abs(d).max()


def test_rotate_modes_parity(backend):
    from ephip import _lib as L
    rng = np.random.default_rng(9)
    nw, nm, nr_el, nq = 3, 6, 27, 5
    nb = nw * nw
    g = rng.standard_normal((nq, nb * nm * nr_el)) + 1j * rng.standard_normal((nq, nb * nm * nr_el))
    u = rng.standard_normal((nq, nm, nm)) + 1j * rng.standard_normal((nq, nm, nm))  # column-major blocks
    gd = _dev(g.astype(np.complex64))
    ud = _dev(u.astype(np.complex64))
    out = torch.empty_like(gd)
    L.call("ephip_eph_rotate_modes", nb, nm, nr_el, nq, L.ptr(gd), gd.stride(0), L.ptr(ud), L.ptr(out),
           out.stride(0), L.stream_handle())
    got = out.cpu().numpy().reshape(nq, nr_el, nm, nb)
    gg = g.reshape(nq, nr_el, nm, nb)
    ref = np.einsum("qrnb,qmn->qrmb", gg, u)  # u[q] stored transposed: u[q][nu', nu] = uph[nu, nu']
    assert _rel(got, ref) < 1e-6


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_fourier_variants_parity(backend, variant):
    """Every tuner variant of the FP32/FP64 kernels gives the same contraction (odd sizes, both tile
    widths)."""
    from ephip import core, tune
    rng = np.random.default_rng(50 + variant)
    irvec = M.rbox(7)
    nr = len(irvec)
    old = [tune.get_variant(k) for k in range(3)]
    try:
        for k in range(3):
            tune.set_variant(k, variant)
        for ndata, nk in [(333, 200), (70, 33)]:
            op = rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))
            xk = rng.random((nk, 3))
            ref = O.fourier(op, irvec, xk)
            got64 = core.fourier_batched(core.WannierObject(irvec, _dev(op)), _dev(xk)).cpu().numpy()
            assert _rel(got64, ref) < 1e-13
            got32 = core.fourier_batched(core.WannierObject(irvec, _dev(op.astype(np.complex64))), _dev(xk))
            assert _rel(got32.cpu().numpy(), ref) < 2e-6
    finally:
        for k in range(3):
            tune.set_variant(k, old[k])


@pytest.mark.parametrize("prec_name", ["fp32", "bf16x3"])
def test_fourier_batched_units(backend, prec_name):
    """Batched entry: per-unit operands, gathered column lists and per-unit counts; columns beyond a
    unit's count are left untouched."""
    from ephip import _lib as L
    prec = L.PREC_FP32 if prec_name == "fp32" else L.PREC_BF16X3
    rng = np.random.default_rng(60)
    irvec = M.rbox(5)
    nr, nd, nu, npts, cs = len(irvec), 48, 5, 300, 260
    op = rng.standard_normal((nu, nr, nd)) + 1j * rng.standard_normal((nu, nr, nd))
    xk = rng.random((npts, 3))
    counts = np.array([260, 0, 1, 129, 77], np.int32)
    cols = np.zeros((nu, cs), np.int32)
    for u in range(nu):
        cols[u, :counts[u]] = np.sort(rng.choice(npts, counts[u], replace=False))
    out = torch.full((nu, cs, nd), 3.0 + 0j, dtype=torch.complex64, device="cuda")
    opd, xkd, irv = _dev(op.astype(np.complex64)), _dev(xk), _dev(irvec)
    cd, cnt = _dev(cols), _dev(counts)
    L.call("ephip_fourier_batched_c32", nd, nr, L.ptr(opd), nd, nr * nd, L.ptr(irv), L.ptr(xkd), nu, cs,
           L.ptr(cnt), L.ptr(cd), cs, L.ptr(out), nd, cs * nd, prec, L.stream_handle())
    o = out.cpu().numpy()
    tol = 2e-6 if prec_name == "fp32" else 2e-5
    for u in range(nu):
        c = counts[u]
        if c:
            ref = O.fourier(op[u], irvec, xk[cols[u, :c]])
            assert _rel(o[u, :c], ref) < tol
        assert np.all(o[u, c:] == 3.0)


@pytest.mark.parametrize("ndata,box,nk,shift", [(1, 3, 1, (0, 0, 0)), (77, 5, 130, (0, 0, 0)),
                                                (192, 7, 257, (1, -2, 0)), (1000, 7, 64, (0, 0, 0)),
                                                (261, 5, 129, (-3, 0, 4))])
def test_fourier_packed_parity(backend, ndata, box, nk, shift):
    """Pre-packed split-BF16 contraction (both row-tile widths, ragged rows/columns, R count not a
    multiple of 16, R ranges off centre) against the FP64 oracle."""
    from ephip import core
    rng = np.random.default_rng(ndata + nk + box)
    irvec = M.rbox(box) + np.array(shift, dtype=np.int32)
    nr = len(irvec)
    decay = np.exp(-np.linalg.norm(irvec, axis=1) / 1.5)[:, None]
    op = (rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))) * decay
    xk = rng.random((nk, 3)) * 4 - 2
    ref = O.fourier(op, irvec, xk)
    apk = core.ft_pack_operand(_dev(op.astype(np.complex64)), ndata, nr, ndata)
    bpk = core.ft_pack_phases(_dev(irvec), _dev(xk))
    out = torch.full((nk, ndata + 3), 5.0 + 0j, dtype=torch.complex64, device="cuda")
    core.fourier_packed(apk, ndata, nr, bpk, nk, nk, out, ndata + 3)
    o = out.cpu().numpy()
    assert _rel(o[:, :ndata], ref) < 2e-5
    assert np.all(o[:, ndata:] == 5.0)


def test_fourier_packed_units(backend):
    """Packed batched entry: per-unit operands, columns gathered from a point table, per-unit counts;
    columns past a unit's count untouched."""
    from ephip import core
    rng = np.random.default_rng(61)
    irvec = M.rbox(5)
    nr, nd, nu, npts, cs = len(irvec), 192, 5, 300, 260
    op = rng.standard_normal((nu, nr, nd)) + 1j * rng.standard_normal((nu, nr, nd))
    xk = rng.random((npts, 3))
    counts = np.array([260, 0, 1, 129, 77], np.int32)
    cols = np.zeros((nu, cs), np.int32)
    for u in range(nu):
        cols[u, :counts[u]] = np.sort(rng.choice(npts, counts[u], replace=False))
    out = torch.full((nu, cs, nd), 3.0 + 0j, dtype=torch.complex64, device="cuda")
    apk = core.ft_pack_operand(_dev(op.astype(np.complex64)), nd, nr, nd, nunits=nu, op_stride=nr * nd)
    bpk = core.ft_pack_phases(_dev(irvec), _dev(xk))
    core.fourier_packed(apk, nd, nr, bpk, npts, cs, out, nd, nunits=nu, ncol=_dev(counts), col_index=_dev(cols),
                        col_stride=cs, out_stride=cs * nd)
    o = out.cpu().numpy()
    for u in range(nu):
        c = counts[u]
        if c:
            assert _rel(o[u, :c], O.fourier(op[u], irvec, xk[cols[u, :c]])) < 2e-5
        assert np.all(o[u, c:] == 3.0)


def test_fourier_packed_argument_checks(backend):
    from ephip import _lib as L
    x = torch.zeros(16, device="cuda")
    with pytest.raises(L.EphipError):   # more columns than table points without a column list
        L.call("ephip_fourier_packed_bf16x3", 4, 3, L.ptr(x), 1, L.ptr(x), 2, 3, None, None, 0, L.ptr(x), 4, 0,
               L.stream_handle())
    with pytest.raises(L.EphipError):   # ldout < ndata
        L.call("ephip_fourier_packed_bf16x3", 4, 3, L.ptr(x), 1, L.ptr(x), 8, 3, None, None, 0, L.ptr(x), 3, 0,
               L.stream_handle())


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7])
def test_fourier_packed_variants(backend, variant):
    """Every tile shape of the packed split-BF16 kernel (kind 3 of the tuner) on an operand whose row
    count suits all three (384 = 3 x 128 = 2 x 192)."""
    from ephip import core, tune
    old = tune.get_variant(tune.KIND_PACKED_BF16)
    rng = np.random.default_rng(70 + variant)
    irvec = M.rbox(7)
    nr, ndata, nk = len(irvec), 384 * 9, 200   # 18 row tiles of 192: two XCD groups of 8 and a tail
    op = rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))
    xk = rng.random((nk, 3))
    try:
        tune.set_variant(tune.KIND_PACKED_BF16, variant)
        apk = core.ft_pack_operand(_dev(op.astype(np.complex64)), ndata, nr, ndata)
        bpk = core.ft_pack_phases(_dev(irvec), _dev(xk))
        out = torch.empty((nk, ndata), dtype=torch.complex64, device="cuda")
        core.fourier_packed(apk, ndata, nr, bpk, nk, nk, out, ndata)
        assert _rel(out.cpu().numpy(), O.fourier(op, irvec, xk)) < 2e-5
    finally:
        tune.set_variant(tune.KIND_PACKED_BF16, old)


@pytest.mark.parametrize("nr", [10, 16, 17, 33])
def test_fourier_packed_eight_wave_stages(backend, nr):
    """The eight-wave 192 x 128 tile (kind 3, variant 7) with one, one full, two and three R stages
    (its operand and phase tiles are double-buffered one stage ahead), with a partial last column tile."""
    from ephip import core, tune
    old = tune.get_variant(tune.KIND_PACKED_BF16)
    rng = np.random.default_rng(90 + nr)
    irvec = M.rbox(5)[:nr]
    ndata, nk = 192 * 3, 131
    op = rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))
    xk = rng.random((nk, 3))
    try:
        tune.set_variant(tune.KIND_PACKED_BF16, 7)
        apk = core.ft_pack_operand(_dev(op.astype(np.complex64)), ndata, nr, ndata)
        bpk = core.ft_pack_phases(_dev(irvec), _dev(xk))
        out = torch.full((nk, ndata), 7.0, dtype=torch.complex64, device="cuda")
        core.fourier_packed(apk, ndata, nr, bpk, nk, nk, out, ndata)
        assert _rel(out.cpu().numpy(), O.fourier(op, irvec, xk)) < 2e-5
    finally:
        tune.set_variant(tune.KIND_PACKED_BF16, old)


def test_pack_operand_swapped(backend):
    """The j-fastest packing is ft_pack_operand of the row-permuted operand, bit for bit."""
    from ephip import _lib as L
    from ephip import core
    rng = np.random.default_rng(95)
    nb, rest, nr = 4, 6, 21
    ndata = nb * nb * rest
    op = (rng.standard_normal((nr, ndata)) + 1j * rng.standard_normal((nr, ndata))).astype(np.complex64)
    d = np.arange(ndata)
    j, i, r = d % nb, (d // nb) % nb, d // (nb * nb)
    perm = op[:, i + nb * (j + nb * r)]
    a = core.ft_pack_operand_swapped(_dev(op), ndata, nr, nb, ndata)
    b = core.ft_pack_operand(_dev(np.ascontiguousarray(perm)), ndata, nr, ndata)
    assert torch.equal(a.view(torch.int16).reshape(-1), b.view(torch.int16).reshape(-1))
    with pytest.raises(L.EphipError):   # ndata not a multiple of nb^2
        core.ft_pack_operand_swapped(_dev(op), ndata - 1, nr, nb, ndata)


@pytest.mark.parametrize("nr", [10, 16, 17, 33])
def test_fourier_packed_few_stages(backend, nr):
    """Operands with one, one full, and two or three R stages (the loads run up to two stages ahead)."""
    from ephip import core
    rng = np.random.default_rng(80 + nr)
    irvec = M.rbox(5)[:nr]
    nd, nu, npts, cs = 192, 2, 90, 70
    op = rng.standard_normal((nu, nr, nd)) + 1j * rng.standard_normal((nu, nr, nd))
    xk = rng.random((npts, 3))
    counts = np.array([70, 33], np.int32)
    cols = np.stack([np.sort(rng.choice(npts, cs, replace=False)) for _ in range(nu)]).astype(np.int32)
    out = torch.zeros((nu, cs, nd), dtype=torch.complex64, device="cuda")
    apk = core.ft_pack_operand(_dev(op.astype(np.complex64)), nd, nr, nd, nunits=nu, op_stride=nr * nd)
    bpk = core.ft_pack_phases(_dev(irvec), _dev(xk))
    core.fourier_packed(apk, nd, nr, bpk, npts, cs, out, nd, nunits=nu, ncol=_dev(counts), col_index=_dev(cols),
                        col_stride=cs, out_stride=cs * nd)
    o = out.cpu().numpy()
    for u in range(nu):
        c = counts[u]
        assert _rel(o[u, :c], O.fourier(op[u], irvec, xk[cols[u, :c]])) < 2e-5


@pytest.mark.parametrize("nb", [1, 16])
def test_gridopt_stage1_planes(backend, nb):
    """Stage 1 of the staged RR->kR for several k1 planes in one pass equals one pass per plane bit for
    bit, and the FP64 sum over r1 of the box to 1e-5 (absent box points contribute zero, faces padded
    with zeros, rows j-fastest for nb > 1); more than 8 planes is refused."""
    import ctypes as C
    from ephip import _lib as L
    rng = np.random.default_rng(41)
    n1, n2, n3, np_, r1min, nd = 5, 3, 4, 4, -2, 16 * 16 * 3
    nr, ldop = 80, 16 * 16 * 3 + 7
    rmap = rng.permutation(nr)[:n1 * n2 * n3].astype(np.int32)
    assert rmap.size == n1 * n2 * n3
    rmap[rng.random(rmap.size) < 0.2] = -1
    op = (rng.standard_normal((nr, ldop)) + 1j * rng.standard_normal((nr, ldop))).astype(np.complex64)
    x1 = [0.0, 0.125, 0.7]
    dmap, dop = _dev(rmap), _dev(op)
    outs = [torch.zeros(np_ * np_ * nd, dtype=torch.complex64, device="cuda") for _ in x1]
    singles = [torch.zeros_like(o) for o in outs]
    xa = (C.c_double * len(x1))(*x1)
    pa = (C.c_void_p * len(x1))(*[L.ptr(o) for o in outs])
    L.call("ephip_gridopt_stage1_planes", nd, n1, n2, n3, np_, r1min, L.ptr(dmap), L.ptr(dop), ldop, len(x1),
           C.addressof(xa), nb, C.addressof(pa), L.stream_handle())
    for x, o in zip(x1, singles):
        L.call("ephip_gridopt_stage1", nd, n1, n2, n3, np_, r1min, L.ptr(dmap), L.ptr(dop), ldop, x, nb, L.ptr(o),
               L.stream_handle())
    torch.cuda.synchronize()
    d = np.arange(nd)
    dd = d if nb <= 1 else (d // nb) % nb + nb * (d % nb + nb * (d // (nb * nb)))
    opd = op.astype(np.complex128)
    for p, x in enumerate(x1):
        assert torch.equal(outs[p].view(torch.float32), singles[p].view(torch.float32))
        ref = np.zeros((np_ * np_, nd), np.complex128)
        for c in range(n3):
            for b in range(n2):
                for a in range(n1):
                    ir = rmap[a + n1 * (b + n2 * c)]
                    if ir >= 0:
                        ref[b + np_ * c, dd] += np.exp(2j * np.pi * x * (r1min + a)) * opd[ir, :nd]
        got = outs[p].cpu().numpy().reshape(np_ * np_, nd)
        assert _rel(got, ref) < 1e-5
    pa9 = (C.c_void_p * 9)(*([L.ptr(outs[0])] * 9))
    xa9 = (C.c_double * 9)(*([0.0] * 9))
    with pytest.raises(L.EphipError):
        L.call("ephip_gridopt_stage1_planes", nd, n1, n2, n3, np_, r1min, L.ptr(dmap), L.ptr(dop), ldop, 9,
               C.addressof(xa9), nb, C.addressof(pa9), L.stream_handle())


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batched_gemm(backend, dtype):
    """batched_gemm! (alpha 1, beta 0) for every op pair of 'N', 'T', 'C' on ragged shapes against
    numpy's matrix product; inner size 0 gives zeros, empty batches are a no-op."""
    from ephip import core
    dev = backend.torch_device
    tdt = getattr(torch, dtype)
    rng = np.random.default_rng(11)
    m, n, k, nb = 5, 7, 3, 4
    tol = 1e-12 if dtype == "complex128" else 2e-6

    def op(x, t):  # x: (rows, cols) matrix
        return x if t == "N" else (x.T if t == "T" else x.conj().T)

    for ta in "NTC":
        for tb in "NTC":
            ra, ca = (m, k) if ta == "N" else (k, m)
            rb, cb = (k, n) if tb == "N" else (n, k)
            a = rng.standard_normal((nb, ra, ca)) + 1j * rng.standard_normal((nb, ra, ca))
            b = rng.standard_normal((nb, rb, cb)) + 1j * rng.standard_normal((nb, rb, cb))
            # column-major (rows, cols, batch) arrays as (batch, cols, rows) tensors
            A = torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).to(dev, tdt)
            B = torch.from_numpy(np.ascontiguousarray(b.transpose(0, 2, 1))).to(dev, tdt)
            C = torch.full((nb, n, m), 7 + 7j, dtype=tdt, device=dev)
            core.batched_gemm(ta, tb, A, B, C)
            got = C.cpu().numpy().transpose(0, 2, 1)
            ref = np.stack([op(a[i], ta) @ op(b[i], tb) for i in range(nb)])
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (ta, tb)
    A = torch.zeros((2, 0, m), dtype=tdt, device=dev)
    B = torch.zeros((2, n, 0), dtype=tdt, device=dev)
    C = torch.full((2, n, m), 1.0, dtype=tdt, device=dev)
    core.batched_gemm("N", "N", A, B, C)
    assert not C.cpu().numpy().any()
    core.batched_gemm("N", "N", torch.zeros((0, 3, m), dtype=tdt, device=dev), torch.zeros((0, n, 3), dtype=tdt,
                      device=dev), torch.zeros((0, n, m), dtype=tdt, device=dev))
    with pytest.raises(Exception):
        core.batched_gemm("X", "N", A, B, C)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_eph_window_scatter(backend, dtype):
    """eph_window_scatter! on random injective state maps with out-of-window zeros, a per-batch buffer
    (i0 > 0) and k+q indices repeated across q: every in-window entry lands at the mode-fastest slot
    (nu, i - i0 - 1, f - 1) of an (nm, ni_stride, nf) array (np.ravel_multi_index), untouched slots keep
    their sentinel, and no two entries share a slot."""
    from ephip import core
    dev = backend.torch_device
    tdt = getattr(torch, dtype)
    rng = np.random.default_rng(12)
    nbkq, nbk, nm, nq, nk, ld_f = 3, 4, 5, 6, 9, 4
    ni_stride, i0, nf = 6, 10, 40
    imap_i = np.array([11, 0, 14, 16], dtype=np.int64)               # global outer states, 0 = outside
    imap_f = np.zeros((nk, ld_f), dtype=np.int64)                     # column-major (ld_f, nk): [ikq, m]
    fvals = rng.permutation(np.arange(1, nf + 1))[: nk * nbkq].reshape(nk, nbkq)
    fvals[rng.random((nk, nbkq)) < 0.3] = 0
    imap_f[:, :nbkq] = fvals
    ikqs = rng.permutation(np.arange(1, nk + 1))[:nq].astype(np.int64)  # distinct k+q per q
    g2 = rng.random((nq, nm, nbk, nbkq))                                 # column-major (nbkq, nbk, nm, nq)
    w = rng.random((nq, nm))
    nout = nm * ni_stride * nf
    out = torch.full((nout,), -1.0, dtype=tdt, device=dev)
    wout = torch.full((nout,), -1.0, dtype=tdt, device=dev)
    t = lambda x, d=tdt: torch.from_numpy(np.ascontiguousarray(x)).to(dev, d)  # noqa: E731
    core.eph_window_scatter(out, wout, t(g2), t(imap_i, torch.int64), t(imap_f, torch.int64),
                            t(ikqs, torch.int64), t(w), nbkq, nbk, nm, nq, ni_stride, i0)
    exp = np.full(nout, -1.0)
    expw = np.full(nout, -1.0)
    slots = []
    for iq in range(nq):
        for nu in range(nm):
            for n in range(nbk):
                for mm in range(nbkq):
                    i, f = imap_i[n], imap_f[ikqs[iq] - 1, mm]
                    if i > 0 and f > 0:
                        s = np.ravel_multi_index((f - 1, i - i0 - 1, nu), (nf, ni_stride, nm))
                        slots.append(s)
                        exp[s], expw[s] = g2[iq, nu, n, mm], w[iq, nu]
    assert len(slots) > 20 and len(set(slots)) == len(slots)
    tol = 0.0 if dtype == "float64" else 1e-7
    assert np.abs(out.cpu().numpy() - exp).max() <= tol
    assert np.abs(wout.cpu().numpy() - expw).max() <= tol
