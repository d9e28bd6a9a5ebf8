"""Parity on the full workload shapes, against the FP64 oracle:

* cfg2 (nw 16, nm 12, H/D 9^3, g 7^3 x 7^3, k 48^3, q 24^3, +-0.3 eV) and cfg4 (cfg2's sizes with on-site
  energies in [-3, 3] eV, k 64^3, q 32^3): the window filter of the WHOLE fine grid (110 592 / 262 144 k)
  bit for bit (keep, band_min, band_max, hence K_surv) against the oracle's FP64 eigenvalues of every k,
  and the band energies within 1e-6 eV;
* gauge-invariant |g|^2 of sampled (k, q) pairs drawn from a production-size outer-k batch (1152 k,
  the staged FP32 kernels the benchmark times; 384 k for the split-BF16 MFMA kernels), averaged over
  degenerate bands at k and k+q: norm-wise relative error <= 1e-4 (the metric's bound).

The measured values are printed (pytest -s).
"""
import numpy as np
import pytest

import eph_checks as C
from ephip import model as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", params=["cfg2", "cfg4"])
def shape(request, backend):
    from ephip import core, eph
    name = request.param
    m, g = M.make_config_model(name, with_g=False)
    m.epmat_ph = M.build_epmat(m)
    ep_dev = core.WannierObject(m.irvec_ep, torch.from_numpy(m.epmat_ph).to(backend.torch_device), check=False)
    w = g["window"]
    st = eph.setup_outer_q(m, g["kgrid"], g["qgrid"], (-w, w), backend, epmat_dev=ep_dev)
    yield name, m, g, st
    del st, ep_dev
    torch.cuda.empty_cache()


def test_full_grid_filter_bit_exact(shape):
    name, m, g, st = shape
    w = g["window"]
    xk = M.kpoints_grid(g["kgrid"])
    E_o = C.full_grid_states(m, xk)
    f = st.el.filt
    r = C.filter_compare(E_o, f.keep.cpu().numpy(), f.band_min.cpu().numpy(), f.band_max.cpu().numpy(),
                         st.el.E.cpu().numpy(), -w, w)
    print(f"{name} full-grid filter: {r}")
    assert r["nk"] == int(np.prod(g["kgrid"]))
    assert r["keep_equal"] and r["band_min_equal"] and r["band_max_equal"]
    assert r["K_surv_device"] == r["K_surv_oracle"] == int(st.ksurv.shape[0])
    assert r["max_abs_dE_eV"] <= 1e-6


@pytest.mark.parametrize("prec_name", ["fp32", "bf16x3"])
def test_production_batch_g2(shape, prec_name):
    from ephip import _lib as L
    from ephip import eph
    name, m, g, st = shape
    if prec_name == "fp32":   # the default staged kernels at the benchmark's batch
        loop = eph.OuterKLoop(st, nk_batch=1152, precision=L.PREC_FP32)
        assert loop.gridopt is not None and loop.gridopt_q is not None
    else:                     # the split-BF16 MFMA kernels (packed RR->kR + projection, fused kR->kq + rotations)
        loop = eph.OuterKLoop(st, nk_batch=384, precision=L.PREC_BF16X3, gridopt=False)
        assert loop.fused and loop.fused_proj
    nks = int(st.ksurv.shape[0])
    i0 = (nks // 2 // loop.nk_batch) * loop.nk_batch   # a batch from the middle of the surviving list
    payload = loop.run_kbatch(i0, loop.nk_batch)
    torch.cuda.synchronize()
    nq = int(np.prod(g["qgrid"]))
    qsel = [int(q) for q in np.random.default_rng(11).choice(nq, 2, replace=False)]
    got = C.sample_units_g2(payload, qsel, 48, torch, seed=12)
    bmin = st.el.filt.band_min.cpu().numpy()
    bmax = st.el.filt.band_max.cpu().numpy()
    r = C.g2_parity(m, m.epmat_ph, g["kgrid"], g["qgrid"], got, bmin, bmax, st.nbw)
    print(f"{name} |g|^2 {prec_name}, batch of {loop.nk_batch} k at {i0}, q {qsel}: {r}")
    del loop, payload
    torch.cuda.empty_cache()
    assert r["pairs"] >= 20
    assert r["rel_err"] <= 1e-4
