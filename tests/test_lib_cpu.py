"""C-ABI checks that need no GPU: the library loads, exports exactly what include/ephip.h declares,
and argument errors are reported synchronously with a message (no kernel is launched)."""
import ctypes as C
import os
import re

import pytest

from ephip import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ephip.h")).read()
    return set(re.findall(r"^(?:const char \*|int )(ephip_\w+)\(", src, flags=re.M))


def test_header_and_binding_agree():
    assert _declared() == set(L.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_empty_error():
    lib = L.load()
    assert lib.ephip_version() >= 100


@pytest.mark.parametrize("n", [0, 33, 64])
def test_heev_rejects_bad_n(n):
    with pytest.raises(L.EphipError, match="outside 1..32"):
        L.call("ephip_heev_c32", n, 4, 1, 0, 1, 1, None, None)


def test_fourier_rejects_bad_leading_dimension():
    with pytest.raises(L.EphipError, match="leading dimension"):
        L.call("ephip_fourier_c32", 100, 7, 1, 50, 1, 3, 1, 1, 100, L.PREC_FP32, None)


def test_fourier_rejects_unknown_precision():
    with pytest.raises(L.EphipError, match="precision"):
        L.call("ephip_fourier_c32", 10, 7, 1, 10, 1, 3, 1, 1, 10, 9, None)


def test_zero_sized_calls_are_noops():
    # nothing to do -> OK without touching the device
    L.call("ephip_fourier_c32", 0, 7, None, 0, None, 3, None, None, 0, L.PREC_FP32, None)
    L.call("ephip_heev_c32", 8, 0, None, 0, None, None, None, None)
    L.call("ephip_window_filter", 8, 0, None, 0.0, 1.0, None, None, None, None, None)


def test_rotation_rejects_oversized_window_block():
    with pytest.raises(L.EphipError, match="nbw"):
        L.call("ephip_eph_rotate_rqkq_g2", 16, 12, 4, 1, 16 * 16 * 12, 1, 1, 1, 1, 1, 1, 17, 1, None, None)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_LIB", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libephip.so"))
    with pytest.raises(L.EphipError, match="no CPU fallback"):
        L.load()


def test_generic_backend_methods_validate_arguments():
    # batched_gemm!: op characters and leading dimensions; empty work is a no-op
    with pytest.raises(L.EphipError, match="'N', 'T' or 'C'"):
        L.call("ephip_batched_gemm", 0, ord("X"), ord("N"), 2, 2, 2, 1, 1, 2, 4, 1, 2, 4, 1, 2, 4, None)
    with pytest.raises(L.EphipError, match="leading dimension"):
        L.call("ephip_batched_gemm", 1, ord("T"), ord("N"), 4, 2, 3, 1, 1, 2, 12, 1, 3, 6, 1, 4, 8, None)
    L.call("ephip_batched_gemm", 1, ord("N"), ord("C"), 4, 2, 3, 0, None, 4, 12, None, 2, 6, None, 4, 8, None)
    # eph_window_scatter!: imap_f's leading dimension must hold the k+q bands
    with pytest.raises(L.EphipError, match="ld_f"):
        L.call("ephip_eph_window_scatter", 1, 1, 1, 10, 1, 1, 1, 2, 5, 1, 1, 3, 2, 2, 1, 4, 0, None, None)
    L.call("ephip_eph_window_scatter", 0, None, None, 0, None, None, None, 3, 5, None, None, 3, 2, 2, 0, 4, 0, None,
           None)
    # velocity matrices: nw <= 32, energies and rbar together
    with pytest.raises(L.EphipError, match="nw=33"):
        L.call("ephip_el_velocity_direct_batched", 33, 1, 1, 3 * 33 * 33, 1, None, None, 0, 1, 3 * 33 * 33, None)
    with pytest.raises(L.EphipError, match="come together"):
        L.call("ephip_el_velocity_direct_batched", 8, 1, 1, 192, 1, 1, None, 192, 1, 192, None)
    L.call("ephip_el_velocity_direct_batched", 8, 0, None, 192, None, None, None, 0, None, 192, None)


def test_velocity_wrappers_reject_wrong_dtypes_before_launch():
    """The velocity kernels read the operator as FP64 complex and U as complex64; a complex64 operator
    (what to_device gives without a dtype) would be read past its end, so the wrappers refuse it first."""
    torch = pytest.importorskip("torch")
    import numpy as np
    from ephip import bte, core, eph
    nw, nk = 4, 3
    irvec = np.zeros((1, 3), np.int32)
    el = eph.ElectronStates(torch.zeros((nk, 3), dtype=torch.float64), torch.zeros((nk, nw), dtype=torch.float64),
                            torch.zeros((nk, nw, nw), dtype=torch.complex64), None, None)
    op32 = core.WannierObject(irvec, torch.zeros((1, 3 * nw * nw), dtype=torch.complex64))
    with pytest.raises(L.EphipError, match="complex128"):
        bte.el_velocity_direct_batched(op32, el, [0, 1])
    bs = bte.BandStates.__new__(bte.BandStates)
    with pytest.raises(L.EphipError, match="complex128"):
        bte.compute_velocities(op32, el, bs)
    op64 = core.WannierObject(irvec, torch.zeros((1, 3 * nw * nw), dtype=torch.complex128))
    el_u64 = eph.ElectronStates(el.xk, el.E, el.U.to(torch.complex128), None, None)
    with pytest.raises(L.EphipError, match="complex64"):
        bte.el_velocity_direct_batched(op64, el_u64, [0])


def test_binding_arity_matches_header():
    """Every ctypes signature has exactly as many arguments as the C declaration (an extra argument would
    be converted by ctypes' defaults, truncating a 64-bit stream handle to an int)."""
    src = open(os.path.join(ROOT, "include", "ephip.h")).read()
    decls = re.findall(r"^int (ephip_\w+)\(([^;]*)\);", src, flags=re.M)
    assert decls
    for name, params in decls:
        params = " ".join(params.split())
        n = 0 if params in ("", "void") else params.count(",") + 1
        if name in L.SIGNATURES:
            assert len(L.SIGNATURES[name]) == n, (name, len(L.SIGNATURES[name]), n)


def test_eigen_entries_validate_arguments():
    with pytest.raises(L.EphipError, match="n=33 outside 1..32"):
        L.call("ephip_eigen_batched", 33, 4, 1, 1, 1, 1, None, None)
    with pytest.raises(L.EphipError, match="null device pointer"):
        L.call("ephip_eigvals_batched", 16, 4, 1, 1, None, None, None)
    L.call("ephip_eigvals_batched", 16, 0, None, None, None, None, None)
    L.call("ephip_eigen_batched", 8, 0, None, None, None, None, None, None)


def test_degenerate_gauge_fix_validates_arguments():
    with pytest.raises(L.EphipError, match="n=33 outside 1..32"):
        L.call("ephip_degenerate_gauge_fix", 33, 4, 1, 1, 1e-5, None)
    with pytest.raises(L.EphipError, match="negative cutoff"):
        L.call("ephip_degenerate_gauge_fix", 8, 4, 1, 1, -1.0, None)
    with pytest.raises(L.EphipError, match="null device pointer"):
        L.call("ephip_degenerate_gauge_fix", 8, 4, None, None, 1e-5, None)
    L.call("ephip_degenerate_gauge_fix", 8, 0, None, None, 1e-5, None)


def test_eigvals_noalloc_validates_before_any_device_call():
    """The allocation-free entry rejects sizes beyond its scratch array (16 MiB per GPU) and bad arguments
    before touching the device; nb = 0 is a no-op."""
    with pytest.raises(L.EphipError, match="scratch"):
        L.call("ephip_eigvals_batched_noalloc", 16, 7944, 1, 1, None, None)
    with pytest.raises(L.EphipError, match="n=33 outside 1..32"):
        L.call("ephip_eigvals_batched_noalloc", 33, 4, 1, 1, None, None)
    with pytest.raises(L.EphipError, match="null device pointer"):
        L.call("ephip_eigvals_batched_noalloc", 16, 7943, None, None, None, None)
    L.call("ephip_eigvals_batched_noalloc", 16, 0, None, None, None, None)


def test_staged_entries_validate_arguments():
    """The staged (GridOpt) entries reject shapes their kernels do not cover before any device call, and
    treat empty work as a no-op."""
    q = (C.c_int32 * 3)(24, 24, 24)
    q_long = (C.c_int32 * 3)(24, 80, 24)      # n2 > 64
    q_slab = (C.c_int32 * 3)(24, 40, 40)      # n2 n3 > 1024: taken in segments, accepted
    qp, qlp, qsp = (C.cast(a, C.c_void_p).value for a in (q, q_long, q_slab))

    def kq(nw=16, nm=12, nbw=2, nunits=1, n1=7, np_=7, qgrid=qp):
        L.call("ephip_gridopt_kq_rotate", nw, nm, nbw, nunits, n1, np_, -3, -3, -3, None, qgrid, None,
               16 * 12 * 343, 16 * 12, None, None, None, 576, None, None, None, None, None, None, None, None,
               None, None)

    for kw, msg in [(dict(nw=8), "nw must be 16"), (dict(nm=17), "nw must be 16"), (dict(nbw=5), "outside 1..4"),
                    (dict(np_=9), "Rp box"), (dict(n1=33), "Rp box"), (dict(qgrid=qlp), "q grid"),
                    (dict(qgrid=qsp), "null device pointer"), (dict(qgrid=None), "q grid"),
                    (dict(), "null device pointer")]:
        with pytest.raises(L.EphipError, match=msg):
            kq(**kw)
    kq(nunits=0)

    def proj(nw=16, nbw=2, np_=7, max_line_k=48, k0=0, k1=8):
        L.call("ephip_gridopt_kR_project_mfma", nw, 12, 343, np_, -3, None, k0, k1, 0, 4, 0, None, max_line_k, 0,
               None, None, None, None, nbw, None, 16 * 12 * 343, None)

    for kw, msg in [(dict(nw=8), "nw must be 16"), (dict(nbw=5), "bad sizes"), (dict(np_=2), "padded box face"),
                    (dict(max_line_k=65), "k on a line"), (dict(), "null device pointer")]:
        with pytest.raises(L.EphipError, match=msg):
            proj(**kw)
    proj(k1=0)
