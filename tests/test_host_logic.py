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
