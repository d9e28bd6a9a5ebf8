"""Device-side mirror of ElectronPhonon.jl's batched Wannier->Bloch API, over libephip.

Names follow the reference (``src/common/gpu_utils.jl``, ``src/wannier/batched_fourier_core.jl``,
``src/wannier_to_bloch_batched.jl``, ``src/filter.jl``). Device arrays are torch CUDA tensors; every
computation runs in libephip's HIP kernels (torch only allocates memory and orders the stream).

Layout convention: a column-major array of reference shape ``(a, b, c)`` is stored as a C-ordered
tensor of shape ``(c, b, a)`` -- the same memory. So ``op_r (ndata x nr)`` is a ``(nr, ndata)``
tensor, a stack of Hamiltonians ``(nw, nw, nk)`` is ``(nk, nw, nw)`` holding the transposes, etc.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib as L

# Degeneracy threshold of the reference, 0.01 meV (src/common/constants.jl:6), in eV.
ELECTRON_DEGEN_CUTOFF_EV = 0.01e-3
# Phonon modes below this are skipped by consumers (src/common/constants.jl:5), in eV.
OMEGA_ACOUSTIC_EV = 0.01e-3


def _torch():
    import torch
    return torch


# ------------------------------------------------------------------------------------------------
# Backend primitives (src/common/gpu_utils.jl:16-78)
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass(frozen=True)
class GPUBackend:
    device: int = 0

    @property
    def torch_device(self):
        return _torch().device("cuda", self.device)


def gpu_backend(device: int = 0) -> GPUBackend:
    """gpu_backend() (src/common/gpu_utils.jl:29; ext/ElectronPhononCUDAExt.jl:34)."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise L.EphipError("no GPU visible: the HIP path has no CPU fallback")
    L.load()
    return GPUBackend(device)


def free_bytes(backend: GPUBackend) -> int:
    """free_bytes (src/common/gpu_utils.jl:51)."""
    import ctypes as C
    f, t = C.c_int64(), C.c_int64()
    L.call("ephip_free_bytes", C.addressof(f), C.addressof(t))
    return f.value


def synchronize(backend: GPUBackend | None = None) -> None:
    """synchronize (src/common/gpu_utils.jl:60): drain the current stream, surface async faults."""
    L.call("ephip_synchronize", L.stream_handle())


def plan_batch(free: int | None, per_point: int, committed: int, cap: int, headroom=(7, 10), what: str = "") -> int:
    """plan_batch (src/calculator/eph_device_staging.jl:29-47): batch width that fits
    ``(free - committed) // den * num`` bytes at ``per_point`` bytes each, clamped to [1, cap]; ``free``
    None means unbounded (batch = cap). Fails early if the whole-run commitments exceed free memory."""
    num, den = headroom
    if free is not None and committed > free:
        raise L.EphipError(f"batched {what}: committed device memory ({committed / 1e9:.2f} GB, whole-run "
                           f"stacks) exceeds free device memory ({free / 1e9:.2f} GB). Reduce the batch cap or "
                           "the grid size.")
    nb_mem = int(cap) if free is None else max(1, ((free - committed) // den * num) // per_point)
    return min(int(cap), nb_mem)


def outer_k_staging_bytes(nw: int, nm: int, nr_el: int, nr_ep: int, nk: int, nq: int, nbw: int) -> tuple[int, int]:
    """Device bytes of the outer-k loop: (per k of a batch, whole-run committed). Per k: the RR->kR
    Fourier output (8 nw^2 nm nr_ep) and, for each of up to nbw window bands, the projected operator
    (8 nw nm nr_ep), kept-q lists (8 nq), the kR->kq Fourier output (8 nw nm nq) and the g2 block
    (4 nbw nm nq). Committed: both e-ph tensor layouts, fine-grid states and phonons."""
    per_k = 8 * nw * nw * nm * nr_ep + nbw * (8 * nw * nm * nr_ep + 8 * nq + 8 * nw * nm * nq + 4 * nbw * nm * nq + 4)
    committed = (2 * 8 * nw * nw * nm * nr_el * nr_ep + (8 * nw + 8 * nw * nw + 49) * nk
                 + (8 * nm + 8 * nm * nm + 24) * nq)
    return per_k, committed


def alloc(backend: GPUBackend, dtype, *dims_colmajor):
    """alloc(backend, T, dims...) (src/common/gpu_utils.jl:41-42) for column-major dims."""
    torch = _torch()
    return torch.empty(tuple(reversed(dims_colmajor)), dtype=dtype, device=backend.torch_device)


def to_device(backend: GPUBackend, x):
    """to_device (src/common/gpu_utils.jl:39; ext/ElectronPhononCUDAExt.jl:51-54)."""
    torch = _torch()
    if isinstance(x, WannierObject):
        return x.to_device(backend)
    if isinstance(x, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(x)).to(backend.torch_device)
    return x.to(backend.torch_device)


def batched_gemm(transA: str, transB: str, A, B, C):
    """batched_gemm! (src/common/gpu_utils.jl:60-78): C[:, :, b] = op(A[:, :, b]) op(B[:, :, b]) with
    alpha 1, beta 0 and op 'N' (the matrix), 'T' (transpose) or 'C' (adjoint), for column-major
    (rows, cols, batch) arrays held as contiguous tensors of shape (batch, cols, rows) (as ``alloc``
    makes them); complex64 or complex128."""
    torch = _torch()
    if transA not in ("N", "T", "C") or transB not in ("N", "T", "C"):
        raise L.EphipError(f"batched_gemm: trans must be 'N', 'T' or 'C' (got {transA!r}, {transB!r})")
    for X in (A, B, C):
        if X.dim() != 3 or not X.is_contiguous():
            raise L.EphipError("batched_gemm: operands must be contiguous 3-d tensors (batch, cols, rows)")
    if not (A.dtype == B.dtype == C.dtype) or A.dtype not in (torch.complex64, torch.complex128):
        raise L.EphipError(f"batched_gemm: complex64 or complex128 operands of one type (got {A.dtype}, "
                           f"{B.dtype}, {C.dtype})")
    nb = C.shape[0]
    ra, ca, rb, cb = A.shape[2], A.shape[1], B.shape[2], B.shape[1]
    m, k = (ra, ca) if transA == "N" else (ca, ra)
    k2, n = (rb, cb) if transB == "N" else (cb, rb)
    if A.shape[0] != nb or B.shape[0] != nb or k != k2 or C.shape[2] != m or C.shape[1] != n:
        raise L.EphipError(f"batched_gemm: shapes do not match (op(A) {m}x{k}, op(B) {k2}x{n}, "
                           f"C {C.shape[2]}x{C.shape[1]}, batches {A.shape[0]}, {B.shape[0]}, {nb})")
    L.call("ephip_batched_gemm", int(A.dtype == torch.complex128), ord(transA), ord(transB), m, n, k, nb,
           L.ptr(A), max(ra, 1), ra * ca, L.ptr(B), max(rb, 1), rb * cb, L.ptr(C), max(m, 1), m * n,
           L.stream_handle())
    return C


def eph_window_scatter(g2_out, w_out, g2vals, imap_i_col, imap_f, ikqs, w, nbandkq: int, nbandk: int, nm: int,
                       nq_batch: int, ni_stride: int, i0: int, check: bool = True):
    """eph_window_scatter! (src/calculator/calculator_utils.jl:6-52): for every (m, n, nu, iq) entry of
    g2vals (nbandkq, nbandk, nm, nq_batch) column-major, i = imap_i_col[n] and f = imap_f[m, ikqs[iq]]
    (1-based states, 0 = outside the window); when both are positive, the mode-fastest slot
    nu + nm (i - i0 - 1) + nm ni_stride (f - 1) of the flat g2_out / w_out takes the value and
    w[nu, iq]. imap_f is a column-major (ld, nk) int64 matrix held as a contiguous (nk, ld) tensor;
    ikqs are 1-based k indices. float32 or float64 values, all of one type. Entries whose k+q index or
    output slot falls outside the buffers are not written; with ``check`` (default) their count is read
    back after the launch (a stream synchronisation) and a non-zero count raises EphipError."""
    torch = _torch()
    vals = (g2_out, w_out, g2vals, w)
    if not all(x.dtype == g2vals.dtype and x.is_contiguous() for x in vals) or \
            g2vals.dtype not in (torch.float32, torch.float64):
        raise L.EphipError("eph_window_scatter: contiguous float32 or float64 values of one type")
    maps = (imap_i_col, imap_f, ikqs)
    if not all(x.dtype == torch.int64 and x.is_contiguous() for x in maps) or imap_f.dim() != 2:
        raise L.EphipError("eph_window_scatter: contiguous int64 maps, imap_f 2-d (nk, ld)")
    nk_f, ld_f = imap_f.shape
    if (g2vals.numel() < nbandkq * nbandk * nm * nq_batch or w.numel() < nm * nq_batch or
            imap_i_col.numel() < nbandk or ikqs.numel() < nq_batch or ld_f < nbandkq):
        raise L.EphipError("eph_window_scatter: an input is smaller than its declared size")
    dropped = torch.zeros(1, dtype=torch.int32, device=g2vals.device) if check else None
    L.call("ephip_eph_window_scatter", int(g2vals.dtype == torch.float64), L.ptr(g2_out), L.ptr(w_out),
           min(g2_out.numel(), w_out.numel()), L.ptr(g2vals), L.ptr(imap_i_col), L.ptr(imap_f), ld_f, nk_f,
           L.ptr(ikqs), L.ptr(w), nbandkq, nbandk, nm, nq_batch, ni_stride, i0, L.ptr(dropped), L.stream_handle())
    if check:
        n = int(dropped.item())
        if n:
            raise L.EphipError(f"eph_window_scatter: {n} entries fell outside the output (k+q index or slot out of "
                               f"range: check i0={i0}, ni_stride={ni_stride} and ikqs)")


def _cplx_pair(*xs):
    torch = _torch()
    dt = xs[0].dtype
    if dt not in (torch.complex64, torch.complex128) or any(x.dtype != dt for x in xs):
        raise L.EphipError("complex64 or complex128 operands, all of one type")
    return int(dt == torch.complex128), (torch.float64 if dt == torch.complex128 else torch.float32)


def eph_apply_rotations(ep, g, ukqs, u_phs, g2_out=None, wq=None):
    """eph_apply_rotations! (src/wannier_to_bloch_batched.jl:442-455; CUDA method ext/ElectronPhononCUDAExt.jl:
    177-201) through ephip_eph_apply_rotations. Arrays in the reference's column-major index order held as
    torch tensors with the axes reversed: ep (nq, nm, nbk, nbkq) contiguous; g (nq, nm, nbk, nw) any strided
    view (its strides are passed); ukqs (nq, nbkq, nw), u_phs (nq, nm, nm) contiguous; optional g2_out
    (nq, nm, nbk, nbkq) and wq (nq, nm) real of the matching precision. Launch only."""
    fp64, rdt = _cplx_pair(ep, g, ukqs, u_phs)
    nq, nm, nbk, nbkq = ep.shape
    nw = g.shape[3]
    if tuple(g.shape) != (nq, nm, nbk, nw) or tuple(ukqs.shape) != (nq, nbkq, nw) or \
            tuple(u_phs.shape) != (nq, nm, nm):
        raise L.EphipError("eph_apply_rotations: shapes do not match ep (nq, nm, nbk, nbkq)")
    if not (ep.is_contiguous() and ukqs.is_contiguous() and u_phs.is_contiguous()):
        raise L.EphipError("eph_apply_rotations: ep, ukqs and u_phs must be dense")
    if g2_out is not None:
        if wq is None or g2_out.dtype != rdt or wq.dtype != rdt or tuple(g2_out.shape) != tuple(ep.shape) or \
                tuple(wq.shape) != (nq, nm) or not (g2_out.is_contiguous() and wq.is_contiguous()):
            raise L.EphipError("eph_apply_rotations: g2_out (as ep) and wq (nq, nm) dense, real of ep's precision")
    st = g.stride()
    L.call("ephip_eph_apply_rotations", fp64, nw, nbkq, nbk, nm, nq, L.ptr(ep), L.ptr(g), st[3], st[2], st[1], st[0],
           L.ptr(ukqs), L.ptr(u_phs), L.ptr(g2_out) if g2_out is not None else None,
           L.ptr(wq) if g2_out is not None else None, L.stream_handle())
    return ep


def eph_apply_rotations_rqkq(ep, g, uks, ukqs):
    """eph_apply_rotations_rqkq! (src/wannier_to_bloch_batched.jl:360-377; CUDA method ext/ElectronPhononCUDAExt.jl:
    241-258) through ephip_eph_apply_rotations_rqkq: ep (nk, nm, nbk, nbkq), g (nk, nm, nw, nw) (the reference's
    (nw, nw, nm, nk), iw fastest), uks (nk, nbk, nw), ukqs (nk, nbkq, nw), all contiguous. Launch only."""
    fp64, _ = _cplx_pair(ep, g, uks, ukqs)
    nk, nm, nbk, nbkq = ep.shape
    nw = uks.shape[2]
    if tuple(g.shape) != (nk, nm, nw, nw) or tuple(uks.shape) != (nk, nbk, nw) or tuple(ukqs.shape) != (nk, nbkq, nw):
        raise L.EphipError("eph_apply_rotations_rqkq: shapes do not match ep (nk, nm, nbk, nbkq)")
    if not all(x.is_contiguous() for x in (ep, g, uks, ukqs)):
        raise L.EphipError("eph_apply_rotations_rqkq: dense operands")
    L.call("ephip_eph_apply_rotations_rqkq", fp64, nw, nbkq, nbk, nm, nk, L.ptr(ep), L.ptr(g), L.ptr(uks),
           L.ptr(ukqs), L.stream_handle())
    return ep


def bte_window_accumulate(S_out, S_in, g2vals, wqmat, imap_i_at_k, imap_f, ikqs, e_i, e_f, wf, mus, Ts, etas,
                          method: int, w_cut: float, i0: int, smearing: int = 1, check: bool = True):
    """bte_window_accumulate! (src/boltzmann/boltzmann_calculator.jl:336-372; CUDA method
    ext/ElectronPhononCUDAExt.jl:323-372) through ephip_bte_window_accumulate. Column-major arrays held with
    the axes reversed: S_out (nT, ni) float64, S_in (nT, nf, ni_tile) float64 (the tile of outer rows
    i0 + 1 .. i0 + ni_tile), g2vals (nq, nm, nbk, nbkq) and wqmat (nq, nm) float32 or float64, imap_i_at_k
    (nbk) int64, imap_f (nk, ld) int64, ikqs (nq) int64 (1-based), e_i, e_f, wf float64 per state; mus, Ts,
    etas (nT) float64; smearing 1 = Gaussian, 2 = Lorentzian. With ``check`` the count of entries that fell
    outside the outputs is read back (a synchronisation) and a non-zero count raises EphipError."""
    torch = _torch()
    if g2vals.dtype not in (torch.float32, torch.float64) or wqmat.dtype != g2vals.dtype:
        raise L.EphipError("bte_window_accumulate: g2vals and wqmat float32 or float64 of one type")
    f64 = (S_out, S_in, e_i, e_f, wf, mus, Ts, etas)
    if not all(x.dtype == torch.float64 and x.is_contiguous() for x in f64):
        raise L.EphipError("bte_window_accumulate: contiguous float64 S_out, S_in, energies, weights, mus, Ts, etas")
    if not all(x.dtype == torch.int64 and x.is_contiguous() for x in (imap_i_at_k, imap_f, ikqs)) or imap_f.dim() != 2:
        raise L.EphipError("bte_window_accumulate: contiguous int64 maps, imap_f 2-d (nk, ld)")
    nq, nm, nbk, nbkq = g2vals.shape
    nT = mus.numel()
    nk_f, ld_f = imap_f.shape
    if S_out.dim() != 2 or S_in.dim() != 3 or S_out.shape[0] != nT or S_in.shape[0] != nT or \
            tuple(wqmat.shape) != (nq, nm) or ikqs.numel() != nq or imap_i_at_k.numel() != nbk or ld_f < nbkq:
        raise L.EphipError("bte_window_accumulate: shapes do not match g2vals (nq, nm, nbk, nbkq)")
    dropped = torch.zeros(1, dtype=torch.int32, device=g2vals.device) if check else None
    L.call("ephip_bte_window_accumulate", int(g2vals.dtype == torch.float64), L.ptr(S_out), S_out.shape[1],
           L.ptr(S_in), S_in.shape[2], S_in.shape[1], L.ptr(g2vals.contiguous()), L.ptr(wqmat.contiguous()),
           L.ptr(imap_i_at_k), L.ptr(imap_f), ld_f, nk_f, L.ptr(ikqs), L.ptr(e_i), L.ptr(e_f), L.ptr(wf), nT,
           L.ptr(mus), L.ptr(Ts), L.ptr(etas), int(method), int(smearing), float(w_cut), nbkq, nbk, nm, nq, int(i0),
           L.ptr(dropped), L.stream_handle())
    if check:
        n = int(dropped.item())
        if n:
            raise L.EphipError(f"bte_window_accumulate: {n} entries fell outside S_in or the k+q range (i0={i0})")


# ------------------------------------------------------------------------------------------------
# Wannier objects and the phase-sum Fourier transform
# ------------------------------------------------------------------------------------------------
class WannierObject:
    """WannierObject (src/wannier/WannierObject.jl:22-38): ``op_r`` (ndata x nr) with sorted irvec.

    ``op_r`` is a ``(nr, ld)`` array (numpy on the host, torch on the device); only the first
    ``ndata`` rows of each column are transformed. ``irvec`` is a host ``(nr, 3)`` int array."""

    def __init__(self, irvec, op_r, ndata: int | None = None, irvec_next=None, check: bool = True):
        irvec = np.ascontiguousarray(irvec, dtype=np.int32)
        if check:
            if op_r.shape[0] != irvec.shape[0]:
                raise ValueError(f"op_r has {op_r.shape[0]} columns but irvec has {irvec.shape[0]} vectors")
            key = irvec[:, ::-1]
            if any(tuple(key[i]) > tuple(key[i + 1]) for i in range(len(key) - 1)):
                raise ValueError("irvec must be sorted by reversed R (check_wannierobject)")
        self.irvec = irvec
        self.op_r = op_r
        self.nr = irvec.shape[0]
        self.ndata = op_r.shape[1] if ndata is None else int(ndata)
        self.irvec_next = irvec_next
        self._irvec_dev = None

    @property
    def on_device(self) -> bool:
        return not isinstance(self.op_r, np.ndarray)

    def irvec_dev(self):
        if self._irvec_dev is None:
            torch = _torch()
            self._irvec_dev = torch.from_numpy(self.irvec).to(self.op_r.device)
        return self._irvec_dev

    def to_device(self, backend: GPUBackend, dtype=None) -> "WannierObject":
        torch = _torch()
        op = self.op_r
        if isinstance(op, np.ndarray):
            op = torch.from_numpy(np.ascontiguousarray(op))
        if dtype is not None:
            op = op.to(dtype)
        op = op.to(backend.torch_device)
        return WannierObject(self.irvec, op, self.ndata, self.irvec_next, check=False)


def fourier_phase(irvec_dev, xk_dev):
    """fourier_phase! (src/wannier/batched_fourier_core.jl:24-30): (nk, nr) complex128 tensor
    ``[j, r] = exp(2 pi i R_r . x_j)``."""
    torch = _torch()
    nr, nk = irvec_dev.shape[0], xk_dev.shape[0]
    out = torch.empty((nk, nr), dtype=torch.complex128, device=xk_dev.device)
    L.call("ephip_fourier_phase_c64", nr, L.ptr(irvec_dev), nk, L.ptr(xk_dev), L.ptr(out), L.stream_handle())
    return out


def fourier_batched(parent: WannierObject, xk_dev, out=None, precision: int = L.PREC_FP32):
    """fourier_batched! (src/wannier/batched_fourier_core.jl:117-139): ``(nk, ndata)`` tensor with
    ``out[j] = op_r[1:ndata, :] * phase[:, j]``. complex64 op -> FP32/split-BF16 MFMA,
    complex128 op -> FP64 MFMA."""
    torch = _torch()
    op = parent.op_r
    nk = xk_dev.shape[0]
    if xk_dev.dtype != torch.float64 or not xk_dev.is_contiguous():
        raise L.EphipError("k-points must be a contiguous float64 (nk, 3) device tensor")
    if out is None:
        out = torch.empty((nk, parent.ndata), dtype=op.dtype, device=op.device)
    ldout = out.stride(0) if nk > 0 else parent.ndata
    if out.dim() != 2 or out.stride(1) != 1 or out.shape[0] < nk or out.shape[1] < parent.ndata:
        raise L.EphipError("fourier_batched: out must be (nk, >=ndata) with unit inner stride")
    if op.dtype == torch.complex64:
        L.call("ephip_fourier_c32", parent.ndata, parent.nr, L.ptr(op), op.stride(0), L.ptr(parent.irvec_dev()),
               nk, L.ptr(xk_dev), L.ptr(out), ldout, precision, L.stream_handle())
    elif op.dtype == torch.complex128:
        L.call("ephip_fourier_c64", parent.ndata, parent.nr, L.ptr(op), op.stride(0), L.ptr(parent.irvec_dev()),
               nk, L.ptr(xk_dev), L.ptr(out), ldout, L.stream_handle())
    else:
        raise L.EphipError(f"unsupported op_r dtype {op.dtype}")
    return out


def ft_stages(nr: int) -> int:
    """LDS/MFMA stages of 16 R vectors in the packed split-BF16 layout."""
    return (int(nr) + 15) // 16


def ft_pack_operand(op, ndata: int, nr: int, ldop: int, nunits: int = 1, op_stride: int = 0, out=None):
    """Split complex64 operand(s) into the packed bf16 layout of ephip_fourier_packed_bf16x3:
    (nunits, nrs, ndata, 64) bfloat16 (see include/ephip.h)."""
    torch = _torch()
    shape = (nunits, ft_stages(nr), ndata, 64)
    if out is None:
        out = torch.empty(shape, dtype=torch.bfloat16, device=op.device)
    elif out.dtype != torch.bfloat16 or out.numel() < math.prod(shape) or not out.is_contiguous():
        raise L.EphipError("ft_pack_operand: out must be a contiguous bfloat16 buffer of (nunits, nrs, ndata, 64)")
    L.call("ephip_ft_pack_operand", ndata, nr, L.ptr(op), ldop, nunits, op_stride, L.ptr(out), L.stream_handle())
    return out


def ft_pack_operand_swapped(op, ndata: int, nr: int, nb: int, ldop: int, out=None):
    """As ft_pack_operand for one operand, with the two fastest row indices (each of size nb) swapped:
    packed row j + nb (i + nb rest) holds operand row i + nb (j + nb rest)."""
    torch = _torch()
    shape = (ft_stages(nr), ndata, 64)
    if out is None:
        out = torch.empty(shape, dtype=torch.bfloat16, device=op.device)
    elif out.dtype != torch.bfloat16 or out.numel() < math.prod(shape) or not out.is_contiguous():
        raise L.EphipError("ft_pack_operand_swapped: out must be a contiguous bfloat16 buffer of (nrs, ndata, 64)")
    L.call("ephip_ft_pack_operand_swapped", ndata, nr, nb, L.ptr(op), ldop, L.ptr(out), L.stream_handle())
    return out


def ft_pack_phases(irvec_dev, xk_dev, out=None):
    """Packed split-BF16 phase table of the points xk_dev: (nrs, npts, 64) bfloat16."""
    torch = _torch()
    nr, npts = irvec_dev.shape[0], xk_dev.shape[0]
    shape = (ft_stages(nr), npts, 64)
    if out is None:
        out = torch.empty(shape, dtype=torch.bfloat16, device=xk_dev.device)
    elif out.dtype != torch.bfloat16 or out.numel() < math.prod(shape) or not out.is_contiguous():
        raise L.EphipError("ft_pack_phases: out must be a contiguous bfloat16 buffer of (nrs, npts, 64)")
    if xk_dev.dtype != torch.float64 or not xk_dev.is_contiguous():
        raise L.EphipError("k-points must be a contiguous float64 (nk, 3) device tensor")
    L.call("ephip_ft_pack_phases", nr, L.ptr(irvec_dev), npts, L.ptr(xk_dev), L.ptr(out), L.stream_handle())
    return out


def fourier_packed(apk, ndata: int, nr: int, bpk, npts: int, nk_max: int, out, ldout: int, nunits: int = 1,
                   ncol=None, col_index=None, col_stride: int = 0, out_stride: int = 0):
    """Split-BF16 Fourier contraction of packed operands against a packed phase table."""
    L.call("ephip_fourier_packed_bf16x3", ndata, nr, L.ptr(apk), nunits, L.ptr(bpk), npts, nk_max, L.ptr(ncol),
           L.ptr(col_index), col_stride, L.ptr(out), ldout, out_stride, L.stream_handle())
    return out


class BatchedWannierInterpolator:
    """BatchedWannierInterpolator (src/wannier/batched_interpolator.jl:39-87): chunks the batched
    Fourier transform at ``batch_size`` points per call."""

    def __init__(self, parent: WannierObject, batch_size: int = 32768, precision: int = L.PREC_FP32):
        if not parent.on_device:
            raise L.EphipError("BatchedWannierInterpolator needs a device WannierObject (use to_device)")
        self.parent = parent
        self.batch_size = int(batch_size)
        self.precision = precision


def get_fourier_batched(out, itp: BatchedWannierInterpolator, xk_dev):
    """get_fourier_batched! (src/wannier/batched_interpolator.jl:200-214)."""
    nk = xk_dev.shape[0]
    for s in range(0, nk, itp.batch_size):
        e = min(nk, s + itp.batch_size)
        fourier_batched(itp.parent, xk_dev[s:e], out=out[s:e], precision=itp.precision)
    return out


# ------------------------------------------------------------------------------------------------
# Batched Hermitian eigensolves (src/wannier_to_bloch_batched.jl:30-65)
# ------------------------------------------------------------------------------------------------
def _as_stack(Hk):
    if Hk.dim() == 2:
        n = int(round(Hk.shape[1] ** 0.5))
        Hk = Hk.view(Hk.shape[0], n, n)
    return Hk


def eigvals_batched(Hk):
    """eigvals_batched: ``Hk`` (nk, n, n) (column-major blocks) -> E (nk, n) float32 ascending."""
    return _heev(Hk, vectors=False)[0]


def eigen_batched(Hk):
    """eigen_batched: -> (E (nk, n) float32, U (nk, n, n) complex64 column-major blocks)."""
    return _heev(Hk, vectors=True)


def _heev(Hk, vectors: bool):
    torch = _torch()
    Hk = _as_stack(Hk).contiguous()
    nb, n = Hk.shape[0], Hk.shape[1]
    E = torch.empty((nb, n), dtype=torch.float32, device=Hk.device)
    U = torch.empty((nb, n, n), dtype=torch.complex64, device=Hk.device) if vectors else None
    info = torch.zeros(nb, dtype=torch.int32, device=Hk.device)
    is64 = Hk.dtype == torch.complex128
    if not is64 and Hk.dtype != torch.complex64:
        raise L.EphipError(f"eigen: unsupported dtype {Hk.dtype}")
    L.call("ephip_heev_c32", n, nb, L.ptr(Hk), int(is64), L.ptr(U), L.ptr(E), L.ptr(info), L.stream_handle())
    if nb and int(info.max()) != 0:
        raise L.EphipError(f"eigen: {int((info != 0).sum())} matrices did not converge")
    return E, U


def eigen_batched_refined(H64, tol: float = 0.0):
    """FP32 eigensolve of the FP64 stack ``H64`` followed by the FP64 Rayleigh-Ritz / Jacobi
    refinement (``tol`` = 0: off-diagonal norm below 1e-15 of the norm). Returns
    (E64 (nk, n) float64, U (nk, n, n) complex64)."""
    torch = _torch()
    H64 = _as_stack(H64).contiguous()
    if H64.dtype != torch.complex128:
        raise L.EphipError("eigen_batched_refined needs complex128 matrices")
    nb, n = H64.shape[0], H64.shape[1]
    dev = H64.device
    E64 = torch.empty((nb, n), dtype=torch.float64, device=dev)
    U = torch.empty((nb, n, n), dtype=torch.complex64, device=dev)
    if tol == 0.0:   # the reference-shaped entry (ephip_eigen_batched)
        work = torch.empty((nb, n), dtype=torch.float32, device=dev)
        info = torch.zeros(nb, dtype=torch.int32, device=dev)
        L.call("ephip_eigen_batched", n, nb, L.ptr(H64), L.ptr(E64), L.ptr(U), L.ptr(work), L.ptr(info),
               L.stream_handle())
    else:
        E32, U = _heev(H64, vectors=True)
        info = None
        L.call("ephip_heev_refine", n, nb, L.ptr(H64), L.ptr(U), L.ptr(E32), L.ptr(E64), float(tol),
               L.stream_handle())
    if info is not None and nb and int(info.max()) != 0:
        raise L.EphipError(f"eigen: {int((info != 0).sum())} matrices did not converge")
    return E64, U


def degenerate_gauge_fix(E64, U, cutoff: float = ELECTRON_DEGEN_CUTOFF_EV):
    """In place: the degenerate-band gauge fixing of solve_eigen_el! (src/diagonalize.jl:47-85) on the
    eigenpairs E64 (nk, n) float64 ascending, U (nk, n, n) complex64 column-major blocks, through
    ephip_degenerate_gauge_fix. Returns U."""
    torch = _torch()
    if E64.dtype != torch.float64 or U.dtype != torch.complex64:
        raise L.EphipError("degenerate_gauge_fix needs float64 E and complex64 U")
    nb, n = E64.shape
    if tuple(U.shape) != (nb, n, n) or not (E64.is_contiguous() and U.is_contiguous()):
        raise L.EphipError(f"degenerate_gauge_fix: shapes {tuple(E64.shape)}, {tuple(U.shape)} (contiguous) expected")
    L.call("ephip_degenerate_gauge_fix", n, nb, L.ptr(E64), L.ptr(U), float(cutoff), L.stream_handle())
    return U


def eigvals_batched_refined(H64):
    """eigvals_batched (src/wannier_to_bloch_batched.jl:30-41) at FP64 accuracy through the C-ABI entry
    ephip_eigvals_batched: E64 (nk, n) float64 ascending of the complex128 stack ``H64``; the FP32
    eigenvectors the refinement needs live in a workspace freed on return."""
    torch = _torch()
    H64 = _as_stack(H64).contiguous()
    if H64.dtype != torch.complex128:
        raise L.EphipError("eigvals_batched_refined needs complex128 matrices")
    nb, n = H64.shape[0], H64.shape[1]
    dev = H64.device
    E64 = torch.empty((nb, n), dtype=torch.float64, device=dev)
    work = torch.empty(max(1, nb * (8 * n * n + 4 * n)), dtype=torch.uint8, device=dev)
    info = torch.zeros(nb, dtype=torch.int32, device=dev)
    L.call("ephip_eigvals_batched", n, nb, L.ptr(H64), L.ptr(E64), L.ptr(work), L.ptr(info), L.stream_handle())
    if nb and int(info.max()) != 0:
        raise L.EphipError(f"eigvals: {int((info != 0).sum())} matrices did not converge")
    return E64


def eigvals_batched_noalloc(H64, E64, info=None):
    """ephip_eigvals_batched_noalloc: the eigenvalues of ``eigvals_batched_refined`` written into the
    caller's float64 (nk, n) tensor ``E64`` (and int32 ``info`` per matrix, optional) with no device
    allocation: the workspace is the library's 16 MiB per-GPU scratch array, so nk <= 16 MiB / (8 n^2 + 4 n).
    Launch-only (no synchronisation, no convergence check)."""
    torch = _torch()
    if H64.dtype != torch.complex128 or E64.dtype != torch.float64:
        raise L.EphipError("eigvals_batched_noalloc needs complex128 matrices and a float64 output")
    if not (H64.is_contiguous() and E64.is_contiguous()):
        raise L.EphipError("eigvals_batched_noalloc needs contiguous tensors")
    nb, n = H64.shape[0], H64.shape[1]
    if tuple(E64.shape) != (nb, n) or (info is not None and (info.dtype != torch.int32 or info.numel() != nb)):
        raise L.EphipError("eigvals_batched_noalloc: E64 must be (nk, n), info int32 (nk)")
    L.call("ephip_eigvals_batched_noalloc", n, nb, L.ptr(H64), L.ptr(E64), L.ptr(info) if info is not None else None,
           L.stream_handle())


# ------------------------------------------------------------------------------------------------
# Window filter (src/filter.jl:19, 40-100, 195-225)
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class FilterResult:
    keep: object        # (nk,) uint8
    band_min: object    # (nk,) int64, 1-based, 0 if dropped
    band_max: object
    nbelow: object      # (nk,) int64 = searchsortedfirst - 1

    def nelec_below_window(self, weights) -> float:
        return float((self.nbelow.double() * weights).sum())


def window_filter(E64, wmin: float, wmax: float) -> FilterResult:
    torch = _torch()
    E64 = E64.contiguous()
    nk, nw = E64.shape
    dev = E64.device
    keep = torch.empty(nk, dtype=torch.uint8, device=dev)
    bmin, bmax, nb = (torch.empty(nk, dtype=torch.int64, device=dev) for _ in range(3))
    L.call("ephip_window_filter", nw, nk, L.ptr(E64), float(wmin), float(wmax), L.ptr(keep), L.ptr(bmin),
           L.ptr(bmax), L.ptr(nb), L.stream_handle())
    return FilterResult(keep, bmin, bmax, nb)


def kq_index(xk_dev, dx, ngrid, shift=(0.0, 0.0, 0.0), table=None):
    """Grid index of x_k + dx (``_hash_xk`` / ``xk_to_ik``, src/common/kpoints.jl:614-656)."""
    import ctypes as C
    torch = _torch()
    nk = xk_dev.shape[0]
    out = torch.empty(nk, dtype=torch.int64, device=xk_dev.device)
    d = (C.c_double * 3)(*map(float, dx))
    g = (C.c_int32 * 3)(*map(int, ngrid))
    s = (C.c_double * 3)(*map(float, shift))
    L.call("ephip_kq_index", nk, L.ptr(xk_dev), C.addressof(d), C.addressof(g), C.addressof(s), L.ptr(table),
           L.ptr(out), L.stream_handle())
    return out
