"""The allocation-free eigenvalue entry (ephip_eigvals_batched_noalloc) against the entry with a caller
workspace (ephip_eigvals_batched), both called through the C-ABI on the same FP64 Hermitian stacks: two
calls in a row, the second with fewer matrices (the scratch array keeps the first call's contents), and a
third at the largest batch the scratch holds. The two entries run the same kernels on the same inputs, so
the eigenvalues must agree bit for bit; they are also checked against numpy's LAPACK eigvalsh (1e-12 eV
relative to the spectrum's scale)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _hermitian(nb, n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    return (a + np.conj(np.transpose(a, (0, 2, 1)))) / 2


def test_eigvals_noalloc_matches_workspace_entry(backend):
    from ephip import _lib as L
    from ephip import core
    dev = backend.torch_device
    n = 16
    cap = (16 << 20) // (8 * n * n + 4 * n)
    for nb, seed in ((3000, 1), (517, 2), (cap, 3)):
        print(f"noalloc: n={n} nb={nb} seed={seed}", flush=True)
        h = _hermitian(nb, n, seed)
        # column-major blocks: the (nk, n, n) tensor holds H[k].T, whose eigenvalues are those of H[k]
        H = torch.from_numpy(h).to(dev).contiguous()
        e_ref = core.eigvals_batched_refined(H).cpu().numpy()
        E = torch.full((nb, n), float("nan"), dtype=torch.float64, device=dev)
        info = torch.full((nb,), -7, dtype=torch.int32, device=dev)
        core.eigvals_batched_noalloc(H, E, info)
        torch.cuda.synchronize()
        e = E.cpu().numpy()
        assert int(info.abs().max()) == 0
        assert np.array_equal(e, e_ref), float(np.abs(e - e_ref).max())
        lap = np.linalg.eigvalsh(np.transpose(h, (0, 2, 1)))
        assert np.abs(e - lap).max() <= 1e-12 * np.abs(lap).max()
    with pytest.raises(L.EphipError, match="scratch"):
        core.eigvals_batched_noalloc(torch.zeros((cap + 1, n, n), dtype=torch.complex128, device=dev),
                                     torch.zeros((cap + 1, n), dtype=torch.float64, device=dev))
