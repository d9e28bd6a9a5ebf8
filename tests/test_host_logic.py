"""Host-side logic that needs no GPU: batch planning (the reference's stub-backend tests,
test/test_gpu.jl:994-1062) and byte accounting."""
import pytest

from ephip import core
from ephip._lib import EphipError


def test_plan_batch_unbounded_is_cap():
    assert core.plan_batch(None, 1600, 1600, 42, what="cpu") == 42


def test_plan_batch_memory_bound():
    per_point, committed, free = 1600, 1600, 1_000_000
    expect = min(1000, max(1, ((free - committed) // 10 * 7) // per_point))
    assert core.plan_batch(free, per_point, committed, 1000) == expect
    assert expect < 1000
    per_point, committed, cap = 1000, 10 ** 6, 500
    assert core.plan_batch(committed + 20 * per_point, per_point, committed, cap) == 14


def test_plan_batch_fails_early():
    with pytest.raises(EphipError, match="exceeds free device memory"):
        core.plan_batch(1000, 1600, 16000, 100, what="stub-oom")


def test_outer_k_bytes_hand_formula():
    nw, nm, nre, nrp, nk, nq, nbw = 16, 12, 343, 343, 110592, 13824, 2
    per_k, committed = core.outer_k_staging_bytes(nw, nm, nre, nrp, nk, nq, nbw)
    g = 8 * nw * nw * nm * nrp                     # RR->kR output column, 8.4 MB
    unit = 8 * nw * nm * nrp + 8 * nq + 8 * nw * nm * nq + 4 * nbw * nm * nq + 4
    assert per_k == g + nbw * unit
    assert committed == 2 * 8 * nw * nw * nm * nre * nrp + (8 * nw + 8 * nw * nw + 49) * nk + (
        8 * nm + 8 * nm * nm + 24) * nq
    assert 40e6 < per_k < 60e6 and 5.5e9 < committed < 6.5e9


@pytest.mark.parametrize("kgrid,qgrid", [((8, 8, 8), (4, 4, 4)), ((6, 4, 9), (3, 2, 3)), ((4, 4, 4), (4, 4, 4)),
                                         ((6, 6, 6), (2, 3, 1))])
def test_kept_q_bound_brute_force(kgrid, qgrid):
    """The coset bound equals the largest kept-q count found by adding every q to every surviving k
    on the grids (unshifted, third index fastest)."""
    import numpy as np

    from ephip import eph
    rng = np.random.default_rng(sum(kgrid) + sum(qgrid))
    nk = int(np.prod(kgrid))
    surv = np.sort(rng.choice(nk, size=max(1, nk // 3), replace=False))
    keep = np.zeros(nk, bool)
    keep[surv] = True
    n = np.array(kgrid)
    r = n // np.array(qgrid)
    qs = np.array([(a, b, c) for a in range(qgrid[0]) for b in range(qgrid[1]) for c in range(qgrid[2])])
    best = 0
    for k in surv:
        i = np.array([k // (n[1] * n[2]), (k // n[2]) % n[1], k % n[2]])
        j = (i + qs * r) % n
        h = (j[:, 0] * n[1] + j[:, 1]) * n[2] + j[:, 2]
        best = max(best, int(keep[h].sum()))
    assert eph.kept_q_bound(surv, kgrid, qgrid) == best


def test_tiled_device_output_residency_and_tiles():
    """TiledDeviceOutput (src/calculator/tiled_device_output.jl:32-151) on the host, with a stub free-bytes
    value as the reference tests its residency logic (test/test_gpu.jl:994-1062): block mode when
    headroom x full bytes exceed the free bytes (and full mode otherwise), the tile capacity equal to the
    largest k batch's state count (brute force), every batch's offset / length, the zeroing of the
    active tile region, the download, and force_block overriding the decision."""
    import warnings

    import numpy as np
    import torch
    from ephip import tiled
    assert tiled.residency_use_block(1000, 900, 1.2) and not tiled.residency_use_block(1100, 900, 1.2)
    rng = np.random.default_rng(7)
    nk_grid = 60
    nb = rng.integers(0, 4, nk_grid)                       # states per grid k (0 = no window band)
    off = np.concatenate([[0], np.cumsum(nb)])
    ks = np.flatnonzero(nb > 0)
    n, nT, bmax = int(off[-1]), 2, 7
    dims = (n, n, nT)
    full = 8 * n * n * nT
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        t = tiled.TiledDeviceOutput(dims, 0, (off, ks), dtype=torch.float64, device="cpu",
                                    free_bytes=lambda: full, headroom=1.2)
        t.tile_begin(0, bmax - 1, bmax)
        assert t.is_block() and any("block residency" in str(x.message) for x in w)
    cap = max(int(off[ks[min(a + bmax, len(ks)) - 1] + 1] - off[ks[a]]) for a in range(0, len(ks), bmax))
    assert t.ni_cap == cap and t.tile_stride() == cap
    assert tuple(t.device_array(0).shape) == (nT, n, cap)
    for a in range(0, len(ks), bmax):
        b = min(a + bmax, len(ks)) - 1
        t.device_array(0).fill_(7.0)
        t.tile_begin(a, b, bmax)
        i0, ni = int(off[ks[a]]), int(off[ks[b] + 1] - off[ks[a]])
        assert (t.tile_offset(), t.tile_length()) == (i0, ni)
        d = t.device_array(0)
        assert float(d[:, :, :ni].abs().max()) == 0.0 and (ni == cap or float(d[:, :, ni:].min()) == 7.0)
        d[:, :, :ni] = float(a + 1)
        t.tile_download()
        assert float(t.host_array(0)[:, :, :ni].min()) == float(a + 1)
    t.tile_free()
    assert not t.is_allocated()
    t2 = tiled.TiledDeviceOutput(dims, 0, (off, ks), device="cpu", free_bytes=lambda: 2 * full)
    t2.tile_begin(0, bmax - 1, bmax)
    assert not t2.is_block() and t2.tile_offset() == 0 and tuple(t2.device_array(0).shape) == (nT, n, n)
    assert float(t2.device_array(0).abs().max()) == 0.0
    t3 = tiled.TiledDeviceOutput(dims, 0, (off, ks), device="cpu", force_block=True, free_bytes=lambda: 10 * full)
    t3.tile_begin(0, bmax - 1, bmax)
    assert t3.is_block()


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU only: --gpus N > 1 exits with status 2 and says so, before any device work."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
