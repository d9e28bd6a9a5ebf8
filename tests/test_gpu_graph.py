"""The outer-k batch loop captured as one HIP graph (OuterKLoop.capture) against the same batches run
eagerly: several replays, with the program's own eager work (other batches, buffer clears) and
synchronisations between them, each checked against the eager result."""
import numpy as np
import pytest

from ephip import model as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


class _G2Copy:
    """Device-only consumer: copies each batch's g2 and kept-q counts into its own preallocated buffers
    (the b-th call of the capture writes buffer b)."""

    def __init__(self, shapes):
        self.g2 = [torch.empty(s_g2, dtype=torch.float32, device="cuda") for s_g2, _ in shapes]
        self.count = [torch.empty(s_c, dtype=torch.int32, device="cuda") for _, s_c in shapes]
        self.b = 0

    def run(self, p):
        self.g2[self.b].copy_(p.g2)
        self.count[self.b].copy_(p.count)
        self.b += 1

    def clear(self):
        for a in self.g2:
            a.fill_(-1.0)
        for a in self.count:
            a.fill_(-1)


def _valid(g2, count):
    """g2 entries of the kept columns only (columns past a unit's count are not written)."""
    g2, count = g2.cpu().numpy(), count.cpu().numpy()
    return [g2[u, :count[u]] for u in range(len(count))], count


@pytest.mark.parametrize("kq_form", [0, 1], ids=["kq-valu", "kq-mfma"])
def test_outer_k_graph_replays_match_eager(backend, kq_form):
    """nw 16, 4 atoms on a (24, 4, 4) k grid: three groups of eight k1 planes, so the batches below make the
    stage-1 plane cache evict within the captured sequence (groups 0, 1, 2, 0). Every replay reproduces the
    eager g2 and kept-q counts of all four batches (bound 1e-6 relative; the same kernels on the same
    inputs), with eager batches of other k run between the replays."""
    from ephip import _lib as L
    from ephip import eph, tune
    m = M.make_model(nw=16, natoms=4, box_hd=5, box_g=3, seed=2024, g_dtype=np.complex128)
    kgrid, qgrid, w = (24, 4, 4), (4, 4, 4), 0.5
    st = eph.setup_outer_q(m, kgrid, qgrid, (-w, w), backend)
    old = tune.get_variant(tune.KIND_GRIDOPT_KQ)
    tune.set_variant(tune.KIND_GRIDOPT_KQ, kq_form)
    try:
        nb = 16
        loop = eph.OuterKLoop(st, nk_batch=nb, precision=L.PREC_FP32)
        assert loop.gridopt is not None and loop.gridopt_q is not None
        go = loop.gridopt
        ks = st.ksurv.cpu().numpy()
        plane_of = ks // (kgrid[1] * kgrid[2])
        group = np.array([go.plane_pos[int(p)] // go.PLANES for p in plane_of])
        assert group.max() >= 2, "the k grid must span three stage-1 plane groups"
        first = [int(np.flatnonzero(group == g)[0]) for g in (0, 1, 2)]
        starts = [first[0], first[1], first[2], first[0] + nb]
        print(f"graph test: {len(ks)} surviving k, batches at {starts} (plane groups "
              f"{[int(group[i]) for i in starts]}), kq form {kq_form}", flush=True)
        # eager reference
        loop.gridopt.invalidate()
        ref = []
        for i0 in starts:
            p = loop.run_kbatch(i0, nb)
            ref.append((p.g2.clone(), p.count.clone()))
        torch.cuda.synchronize()
        ref = [_valid(g2, c) for g2, c in ref]
        # capture
        sink = _G2Copy([(tuple(p_g2.shape), tuple(p_c.shape)) for p_g2, p_c in
                        ((loop.g2[:len(r[1])], loop.count[:len(r[1])]) for r in ref)])
        graph = loop.capture(starts, nb, [sink])
        assert sink.b == len(starts)
        for rep in range(3):
            sink.clear()
            loop.run_kbatch((first[2] + 2 * nb * rep) % max(1, len(ks) - nb), nb)   # eager work between replays
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            worst = 0.0
            for b in range(len(starts)):
                got, cnt = _valid(sink.g2[b], sink.count[b])
                assert np.array_equal(cnt, ref[b][1]), (rep, b)
                num = max(float(np.abs(x - y).max(initial=0.0)) for x, y in zip(got, ref[b][0]))
                den = max(float(np.abs(y).max(initial=0.0)) for y in ref[b][0])
                worst = max(worst, num / den)
            print(f"replay {rep}: max relative g2 difference {worst:.2e}", flush=True)
            assert worst <= 1e-6
        # eager batches after the capture recompute their planes (the cache was invalidated)
        p = loop.run_kbatch(starts[1], nb)
        got, cnt = _valid(p.g2, p.count)
        assert np.array_equal(cnt, ref[1][1])
        assert max(float(np.abs(x - y).max(initial=0.0)) for x, y in zip(got, ref[1][0])) <= \
            1e-6 * max(float(np.abs(y).max(initial=0.0)) for y in ref[1][0])
    finally:
        tune.set_variant(tune.KIND_GRIDOPT_KQ, old)
