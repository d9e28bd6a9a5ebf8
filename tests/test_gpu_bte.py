"""Device Boltzmann transport against the FP64 oracle (oracle/bte_ref.py).

* bte_scattering_increments on the device against the reference's golden vectors (rtol 1e-12);
* on the cfg1 shape with a time-reversal-symmetric model (nw 8, 2 atoms, 8^3 grids, window +-1 eV,
  Gaussian smearing 100 meV, occupation convention 5, two temperatures): in-window state set
  (exact), band velocities, S_out, S_in, and the SERTA and iterative-BTE conductivities, all within
  1e-5 relative (Frobenius norm) of the oracle; mobility = sigma / n inherits the bound;
* the one-call driver (run_transport) from the model and grids alone, mobilities against the oracle;
* the compressed scattering-in product against the dense matrix, and independence of the k batching;
* the same quantities on the default nw-16 e-ph path (matrix-core RR->kR projection, staged kR->kq)
  for a 4-atom time-reversal-symmetric model on 6^3 grids, window widths 2 and 3.
"""
import json
import os

import numpy as np
import pytest

import bte_ref as B
import oracle as O
from ephip import model as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))

T_LIST = np.array([0.025852, 0.05])     # 300 K and 580 K in eV
MU_LIST = np.array([0.0, 0.02])
ETA = 0.1
METHOD = 5


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_increments_golden(backend):
    from ephip import _lib as L
    with open(os.path.join(HERE, "golden", "bte_increments.json")) as f:
        gold = json.load(f)
    i = gold["input"]
    row = [i["ek"], i["ekq"], i["wq"], i["g2"], i["wtq"], i["mu"], i["T"], i["eta"]]
    u = gold["underflow_case"]
    urow = [u["ek"], u["ekq"], u["wq"], u["g2"], u["wtq"], u["mu"], u["T"], u["eta"]]
    dev = backend.torch_device
    for meth, (so, si) in gold["expected"].items():
        x = torch.tensor([row, row], dtype=torch.float64, device=dev)
        out = torch.empty((2, 2), dtype=torch.float64, device=dev)
        L.call("ephip_bte_increments", 2, int(meth), 1, L.ptr(x), L.ptr(out), L.stream_handle())
        o = out.cpu().numpy()
        assert o[0, 0] == pytest.approx(so, rel=gold["rtol"]) and o[0, 1] == pytest.approx(si, rel=gold["rtol"])
        assert np.array_equal(o[0], o[1])
    x = torch.tensor([urow], dtype=torch.float64, device=dev)
    out = torch.empty((1, 2), dtype=torch.float64, device=dev)
    L.call("ephip_bte_increments", 1, u["method"], 1, L.ptr(x), L.ptr(out), L.stream_handle())
    assert tuple(out.cpu().numpy()[0]) == tuple(u["expected"])
    # Lorentzian delta against the oracle on a spread of inputs, every method
    rng = np.random.default_rng(4)
    rows = np.column_stack([rng.uniform(-0.2, 0.2, 64), rng.uniform(-0.2, 0.2, 64), rng.uniform(0.01, 0.1, 64),
                            rng.random(64), np.full(64, 0.01), np.zeros(64), np.full(64, 0.03),
                            np.full(64, 0.02)])
    x = torch.from_numpy(rows).to(dev)
    out = torch.empty((64, 2), dtype=torch.float64, device=dev)
    for meth in range(1, 7):
        L.call("ephip_bte_increments", 64, meth, 2, L.ptr(x), L.ptr(out), L.stream_handle())
        ref = np.array([O.bte_increments(meth, *r[:7], 2, r[7]) for r in rows])
        assert np.allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=0)
    with pytest.raises(Exception):
        L.call("ephip_bte_increments", 1, 7, 1, L.ptr(x), L.ptr(out), L.stream_handle())


@pytest.fixture(scope="module")
def transport(backend):
    from ephip import _lib as L
    from ephip import bte, core, eph
    m, g = M.make_config_model("cfg1", time_reversal=True)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend)

    def run(nk_batch):
        loop = eph.OuterKLoop(st, nk_batch=nk_batch, precision=L.PREC_FP32)
        params = bte.ElectronTransportParams(Tlist=T_LIST, nlist=[0.1, 0.1], volume=1.0, mulist=MU_LIST)
        calc = bte.BoltzmannCalculator(params, smearing=("Gaussian", ETA), occupation_method=METHOD)
        return bte.run_boltzmann(loop, calc)

    calc = run(64)
    vel = core.WannierObject(m.irvec_h, bte.velocity_operator(m.irvec_h, m.ham_r).op_r).to_device(
        backend, dtype=torch.complex128)
    bte.compute_velocities(vel, st.el, calc.bs)
    res = bte.solve_electron_bte(calc, max_iter=600, rtol=1e-10)
    # oracle
    s = B.states(m, g["kgrid"], (-w, w))
    v = B.velocities(m, s)
    So, Si = B.scattering(m, s, g["qgrid"], g["kgrid"], T_LIST, MU_LIST, np.full(2, ETA), method=METHOD)
    ss, sb, its = B.solve_bte(So, Si, s["es"], v, s["weight"], T_LIST, MU_LIST, max_iter=600, rtol=1e-10)
    return dict(m=m, g=g, st=st, calc=calc, res=res, s=s, v=v, So=So, Si=Si, ss=ss, sb=sb, its=its, run=run)


def test_states_and_velocities(transport):
    t = transport
    bs, s = t["calc"].bs, t["s"]
    assert np.array_equal(bs.ks.cpu().numpy(), s["ks"])
    assert np.array_equal(bs.ibands.cpu().numpy(), s["ib"])
    assert np.abs(bs.es.cpu().numpy() - s["es"]).max() <= 1e-6
    v = bs.vs.cpu().numpy()
    err = np.abs(v - t["v"]).max() / np.abs(t["v"]).max()
    print(f"velocity max rel err {err:.2e}")
    assert err <= 1e-6


def test_scattering_rates(transport):
    t = transport
    So_d = t["calc"].S_out.cpu().numpy()
    for it in range(len(T_LIST)):
        e_out = _rel(So_d[it], t["So"][it])
        e_in = _rel(t["calc"].scattering_in_dense(it), t["Si"][it])
        print(f"T={T_LIST[it]}: S_out rel {e_out:.2e}  S_in rel {e_in:.2e}")
        assert e_out <= 1e-5 and e_in <= 1e-5


def test_conductivity_and_mobility(transport):
    from ephip import bte
    t = transport
    res = t["res"]
    for it in range(len(T_LIST)):
        es, eb = _rel(res.sigma_serta[it], t["ss"][it]), _rel(res.sigma[it], t["sb"][it])
        print(f"T={T_LIST[it]}: sigma SERTA rel {es:.2e}, BTE rel {eb:.2e}, iterations {res.niter[it]} "
              f"(oracle {t['its'][it]})")
        assert res.converged[it]
        assert es <= 1e-5 and eb <= 1e-5
    p = t["calc"].params
    mob = bte.transport_mobility(res.sigma, p)
    ref = bte.transport_mobility(t["sb"], p)
    assert _rel(mob, ref) <= 1e-5


def test_run_transport_driver(transport, backend):
    """The one-call driver (bte.run_transport: states and window filter, phonons, the outer-k e-ph loop
    into a BoltzmannCalculator, velocities, SERTA and the iterative BTE) from the model and grids alone
    gives the oracle's SERTA and BTE mobilities within 1e-5 relative (cfg1 shape)."""
    from ephip import _lib as L
    from ephip import bte
    t = transport
    m, g = t["m"], t["g"]
    w = g["window"]
    params = bte.ElectronTransportParams(Tlist=T_LIST, nlist=[0.1, 0.1], volume=1.0, mulist=MU_LIST)
    run = bte.run_transport(m, g["kgrid"], g["qgrid"], (-w, w), params, backend, precision=L.PREC_FP32,
                            smearing=("Gaussian", ETA), occupation_method=METHOD, nk_batch=64, max_iter=600,
                            rtol=1e-10)
    assert set(run.seconds) == {"setup", "eph_and_scattering", "velocities", "solve"}
    assert all(run.result.converged)
    ref_serta = bte.transport_mobility(t["ss"], params)
    ref_bte = bte.transport_mobility(t["sb"], params)
    e_s, e_b = _rel(run.mobility_serta, ref_serta), _rel(run.mobility, ref_bte)
    print(f"run_transport mobility rel err: SERTA {e_s:.2e}, BTE {e_b:.2e}")
    assert e_s <= 1e-5 and e_b <= 1e-5


def test_spmv_matches_dense(transport):
    t = transport
    calc = t["calc"]
    n = calc.bs.n
    x = torch.randn((n, 3), dtype=torch.float64, device=calc.S_out.device)
    y = torch.zeros_like(x)
    for it in range(len(T_LIST)):
        calc.spmv(it, x, y)
        ref = calc.scattering_in_dense(it) @ x.cpu().numpy()
        assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_batching_independent(transport):
    t = transport
    other = t["run"](200)
    a, b = other.S_out.cpu().numpy(), t["calc"].S_out.cpu().numpy()
    assert _rel(a, b) <= 1e-12
    assert np.abs(other.scattering_in_dense(0) - t["calc"].scattering_in_dense(0)).max() <= \
        1e-12 * np.abs(t["Si"][0]).max()


def test_mu_from_density(transport, backend):
    """Chemical potential from the carrier count on the device state energies agrees with the
    oracle's bisection on the oracle energies."""
    from ephip import bte
    t = transport
    bs = t["calc"].bs
    s = t["s"]
    base = s["nbelow"].sum() / len(s["xk"])
    p = bte.ElectronTransportParams(Tlist=[0.025852], nlist=[2 * (0.5 + base)], volume=1.0)   # 0.5 in the window
    bte.bte_compute_mu(p, bs)
    ref = B.find_mu(0.5, 0.025852, s["es"], s["weight"])
    assert abs(p.mulist[0] - ref) <= 1e-9


@pytest.mark.parametrize("window", [0.3, 0.5])
def test_staged_nw16_transport(backend, window):
    """Transport on the default nw-16 e-ph path (RR->kR projection on the matrix cores, staged kR->kq
    with the rotations fused; window widths 2 and 3) against the oracle: S_out, S_in, the SERTA and BTE
    conductivities and the mobility within 1e-5 relative (nw 16, 4 atoms, 6^3 grids)."""
    from ephip import _lib as L
    from ephip import bte, core, eph
    m = M.make_model(nw=16, natoms=4, box_hd=5, box_g=3, seed=2024, g_dtype=np.complex128, time_reversal=True)
    kgrid = qgrid = (6, 6, 6)
    st = eph.setup_outer_q(m, kgrid, qgrid, (-window, window), backend)
    assert st.nbw == {0.3: 2, 0.5: 3}[window]
    loop = eph.OuterKLoop(st, nk_batch=64, precision=L.PREC_FP32)
    assert loop.gridopt is not None and loop.gridopt.mfma and loop.gridopt_q is not None
    params = bte.ElectronTransportParams(Tlist=T_LIST, nlist=[0.1, 0.1], volume=1.0, mulist=MU_LIST)
    calc = bte.BoltzmannCalculator(params, smearing=("Gaussian", ETA), occupation_method=METHOD)
    calc = bte.run_boltzmann(loop, calc)
    vel = core.WannierObject(m.irvec_h, bte.velocity_operator(m.irvec_h, m.ham_r).op_r).to_device(
        backend, dtype=torch.complex128)
    bte.compute_velocities(vel, st.el, calc.bs)
    res = bte.solve_electron_bte(calc, max_iter=600, rtol=1e-10)
    s = B.states(m, kgrid, (-window, window))
    v = B.velocities(m, s)
    So, Si = B.scattering(m, s, qgrid, kgrid, T_LIST, MU_LIST, np.full(2, ETA), method=METHOD)
    ss, sb, its = B.solve_bte(So, Si, s["es"], v, s["weight"], T_LIST, MU_LIST, max_iter=600, rtol=1e-10)
    assert np.array_equal(calc.bs.ks.cpu().numpy(), s["ks"])
    So_d = calc.S_out.cpu().numpy()
    for it in range(len(T_LIST)):
        e_out, e_in = _rel(So_d[it], So[it]), _rel(calc.scattering_in_dense(it), Si[it])
        es, eb = _rel(res.sigma_serta[it], ss[it]), _rel(res.sigma[it], sb[it])
        print(f"T={T_LIST[it]}: S_out {e_out:.2e} S_in {e_in:.2e} SERTA {es:.2e} BTE {eb:.2e}")
        assert res.converged[it]
        assert e_out <= 1e-5 and e_in <= 1e-5 and es <= 1e-5 and eb <= 1e-5
    assert _rel(bte.transport_mobility(res.sigma, params), bte.transport_mobility(sb, params)) <= 1e-5


def test_velocity_matrices_berry_connection(backend):
    """Full velocity matrices (U^H dH U, and with the Berry-connection term i (e_m - e_n) rbar) against
    mathematics rather than a restated formula: with Wannier centres tau (position operator diag(tau)
    at R = 0), the Berry-connection velocity must equal U'^H dH'(k) U' in the tau gauge, where
    H'(k)_ij = sum_R e^{2 pi i k.(R + tau_j - tau_i)} H_ij(R) and U' = diag(e^{-2 pi i k.tau}) U, built
    here in FP64 numpy (nw 8 and 16, 40 random k, 1e-6 relative). Also: the diagonal of the plain
    rotation equals ephip_band_velocity, both outputs are Hermitian and tr(rbar) = sum tau."""
    from ephip import _lib as L
    from ephip import bte, core, eph
    dev = backend.torch_device
    rng = np.random.default_rng(7)
    for nw in (8, 16):
        m = M.make_model(nw=nw, natoms=2, box_hd=5, box_g=3, seed=1001 + nw, with_g=False)
        irvec, ham_r = m.irvec_h, m.ham_r
        nr = irvec.shape[0]
        tau = rng.uniform(-0.5, 0.5, (nw, 3))
        pos = np.zeros((nr, 3 * nw * nw), dtype=np.complex128)
        r0 = int(np.flatnonzero((irvec == 0).all(1))[0])
        for a in range(3):
            pos[r0, a * nw * nw + np.arange(nw) * (nw + 1)] = tau[:, a]
        xk = rng.random((40, 3))
        xk_dev = torch.from_numpy(xk).to(dev)
        ham = core.WannierObject(irvec, ham_r).to_device(backend, dtype=torch.complex128)
        el = eph.compute_electron_states(ham, xk_dev)
        vel = core.WannierObject(irvec, bte.velocity_operator(irvec, ham_r).op_r).to_device(backend)
        pos_dev = core.WannierObject(irvec, pos).to_device(backend)
        kidx = torch.arange(40, device=dev)
        rbar = bte.el_velocity_direct_batched(pos_dev, el, kidx)
        v0 = bte.el_velocity_direct_batched(vel, el, kidx)
        vb = bte.el_velocity_direct_batched(vel, el, kidx, rbar=rbar)
        rbar_h, v0_h, vb_h = rbar.cpu().numpy(), v0.cpu().numpy(), vb.cpu().numpy()
        # diagonal of the plain rotation = the band-diagonal velocity kernel
        dH = core.fourier_batched(vel, xk_dev)
        vd = torch.empty((40 * nw, 3), dtype=torch.float64, device=dev)
        col = torch.arange(40, device=dev).repeat_interleave(nw)
        band = torch.arange(nw, device=dev, dtype=torch.int32).repeat(40)
        L.call("ephip_band_velocity", nw, 40 * nw, L.ptr(col), L.ptr(band), L.ptr(col), L.ptr(dH), dH.stride(0),
               L.ptr(el.U), L.ptr(vd), L.stream_handle())
        diag = np.stack([np.diagonal(v0_h[:, a], axis1=1, axis2=2).real for a in range(3)], axis=-1)  # (k, n, a)
        assert np.abs(diag.reshape(-1, 3) - vd.cpu().numpy()).max() <= 1e-12 * np.abs(diag).max()
        for x in (rbar_h, v0_h, vb_h):
            assert np.abs(x - np.conj(np.swapaxes(x, -1, -2))).max() <= 1e-12 * np.abs(x).max()
        assert np.abs(np.trace(rbar_h, axis1=-2, axis2=-1) - tau.sum(0)).max() <= 1e-6
        # tau-gauge derivative, FP64, from H(R) directly
        U = el.U.cpu().numpy().astype(np.complex128).transpose(0, 2, 1)  # [k, i, n]
        Hr = ham_r.reshape(nr, nw, nw).transpose(0, 2, 1)                 # [R, i, j]
        shift = tau[None, :, :] - tau[:, None, :]                          # [i, j, a] = tau_j - tau_i
        num = den = 0.0
        for k in range(40):
            arg = irvec[:, None, None, :] + shift[None]                      # [R, i, j, a]
            ph = np.exp(2j * np.pi * (arg @ xk[k]))                          # [R, i, j]
            D = np.exp(2j * np.pi * (tau @ xk[k]))
            Ut = np.conj(D)[:, None] * U[k]
            for a in range(3):
                dHt = np.sum(1j * arg[..., a] * ph * Hr, axis=0)
                ref = np.conj(Ut).T @ dHt @ Ut
                num = max(num, float(np.abs(vb_h[k, a] - ref).max()))
                den = max(den, float(np.abs(ref).max()))
        print(f"nw {nw}: Berry-connection velocity vs tau-gauge derivative, max rel {num / den:.2e}")
        assert num / den <= 1e-6, num / den


@pytest.mark.parametrize("mode", ["block", "full", "auto-small-memory"])
def test_qbatched_boltzmann_tiled_output(transport, mode):
    """The reference's device BoltzmannCalculator shape: EPDataQBatched payloads (eph.QBatchedAdapter) into
    ephip_bte_window_accumulate with S_in held in a TiledDeviceOutput (block mode: one outer-k tile per
    batch streamed to the host; full mode: the whole matrix on the device; auto with a stub free-bytes
    value below the full size: block). S_out and the dense S_in agree with the compressed-row calculator
    (pinned against the oracle above) to 1e-12 and with the oracle to 1e-5."""
    from ephip import _lib as L
    from ephip import bte, eph
    t = transport
    st = t["st"]
    loop = eph.OuterKLoop(st, nk_batch=64, precision=L.PREC_FP32)
    params = bte.ElectronTransportParams(Tlist=T_LIST, nlist=[0.1, 0.1], volume=1.0, mulist=MU_LIST)
    kw = {"block": dict(force_block=True), "full": dict(force_block=False),
          "auto-small-memory": dict(free_bytes=lambda: 1000)}[mode]
    qc = bte.BoltzmannQBatchedCalculator(params, smearing=("Gaussian", ETA), occupation_method=METHOD, **kw)
    qc.setup(loop)
    adapter = eph.QBatchedAdapter(st, [qc], nq_batch=100, n_batch_max=64)
    loop.run(calculators=[adapter])
    qc.postprocess()
    assert qc.tiled.ni_cap == 0  # freed
    calc = t["calc"]
    so_q, so_c = qc.S_out.cpu().numpy(), calc.S_out.cpu().numpy()
    assert np.abs(so_q - so_c).max() <= 1e-12 * np.abs(so_c).max()
    for iT in range(len(T_LIST)):
        a, b = qc.scattering_in_dense(iT), calc.scattering_in_dense(iT)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
        assert _rel(a, t["Si"][iT]) < 1e-5
    assert _rel(so_q, t["So"]) < 1e-5
