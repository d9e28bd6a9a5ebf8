"""Electron/phonon states and the outer-q e-ph loop on the device.

Mirrors the reference's batched outer-q driver ``run_eph_over_q_and_k`` with ``EPDataKBatched``
payloads (src/calculator/run_eph_over_q_and_k.jl:28-54, :428-613) and the device states branches
(src/compute_states.jl:135-215, :285-311). The k+q eigenpairs are looked up in the fine-grid states
(the q-grid divides the k-grid, so every k+q is a fine-grid point) instead of being re-solved per
(q, k-batch); this gives the same eigen-decomposition as the reference's per-batch solve (its
gauge is arbitrary too), and makes k and k+q states of one grid point identical.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib as L
from . import core
from .core import WannierObject


def _torch():
    import torch
    return torch


@dataclasses.dataclass
class ElectronStates:
    """Full-band states of a k list: E (nk, nw) float64 ascending (FP64-refined), U (nk, nw, nw)
    complex64 column-major eigenvector blocks, and the window filter."""
    xk: object
    E: object
    U: object
    filt: core.FilterResult | None
    window: tuple[float, float] | None

    @property
    def nw(self) -> int:
        return self.E.shape[1]


def compute_electron_states(ham_dev: WannierObject, xk_dev, window=None, chunk: int = 32768,
                            degen_cutoff: float | None = core.ELECTRON_DEGEN_CUTOFF_EV) -> ElectronStates:
    """H(R) -> H(k) in FP64 on the matrix cores, FP32 eigensolve, FP64 refinement, window filter.
    Device counterpart of compute_electron_states + filter_electron_states
    (src/compute_states.jl:135-178, src/filter.jl:46-72). degen_cutoff: the eigenvectors of bands closer
    than this get the per-k path's degenerate gauge fixing (get_el_eigen! -> solve_eigen_el!,
    src/diagonalize.jl:47-85), so per-band quantities at degenerate k (|g|^2, band velocities) are those of
    the reference's CPU path and do not depend on the solver; None keeps the batched path's raw gauge
    (eigen_batched applies no fixing)."""
    torch = _torch()
    nk = xk_dev.shape[0]
    nw = int(round(math.sqrt(ham_dev.ndata)))
    if nw * nw != ham_dev.ndata:
        raise L.EphipError(f"ndata={ham_dev.ndata} is not a perfect square")
    dev = xk_dev.device
    E = torch.empty((nk, nw), dtype=torch.float64, device=dev)
    U = torch.empty((nk, nw, nw), dtype=torch.complex64, device=dev)
    H64 = torch.empty((min(chunk, max(nk, 1)), nw * nw), dtype=torch.complex128, device=dev)
    for s in range(0, nk, chunk):
        e = min(nk, s + chunk)
        h = H64[: e - s]
        core.fourier_batched(ham_dev, xk_dev[s:e], out=h)
        Ec, Uc = core.eigen_batched_refined(h)
        if degen_cutoff is not None:
            core.degenerate_gauge_fix(Ec, Uc, degen_cutoff)
        E[s:e] = Ec
        U[s:e] = Uc
    filt = core.window_filter(E, window[0], window[1]) if window is not None else None
    return ElectronStates(xk_dev, E, U, filt, window)


def compute_electron_states_pruned(ham_dev: WannierObject, xk_dev, window, surrogate, chunk: int = 32768):
    """compute_electron_states with the band-energy surrogate as a pre-filter: BF16 inference marks the
    k-points whose predicted bands come within the surrogate's margin of the window; only those get the
    exact FP64 Fourier + eigensolve + window filter. Returns (states, candidates): states are full-grid
    arrays in which pruned points have keep = 0, band ranges 0, E = NaN, U = 0 and nbelow = -1 (not
    computed). With zero misses the keep / band-range arrays equal the exact filter's."""
    torch = _torch()
    nk = xk_dev.shape[0]
    nw = int(round(math.sqrt(ham_dev.ndata)))
    dev = xk_dev.device
    keep_s, _ = surrogate.infer_prune(xk_dev, window, surrogate.margin)
    cand = torch.nonzero(keep_s, as_tuple=False).view(-1)
    E = torch.full((nk, nw), float("nan"), dtype=torch.float64, device=dev)
    U = torch.zeros((nk, nw, nw), dtype=torch.complex64, device=dev)
    keep = torch.zeros(nk, dtype=torch.uint8, device=dev)
    bmin = torch.zeros(nk, dtype=torch.int64, device=dev)
    bmax = torch.zeros(nk, dtype=torch.int64, device=dev)
    nbelow = torch.full((nk,), -1, dtype=torch.int64, device=dev)
    if cand.numel():
        sub = compute_electron_states(ham_dev, xk_dev.index_select(0, cand).contiguous(), window, chunk)
        E.index_copy_(0, cand, sub.E)
        U.index_copy_(0, cand, sub.U)
        keep.index_copy_(0, cand, sub.filt.keep)
        bmin.index_copy_(0, cand, sub.filt.band_min)
        bmax.index_copy_(0, cand, sub.filt.band_max)
        nbelow.index_copy_(0, cand, sub.filt.nbelow)
    return ElectronStates(xk_dev, E, U, core.FilterResult(keep, bmin, bmax, nbelow), window), cand


@dataclasses.dataclass
class PhononStates:
    xq: object
    omega: object   # (nq, nm) float64, sign(w2) sqrt(|w2|)
    u: object       # (nq, nm, nm) complex64 column-major, rows divided by sqrt(mass)


def compute_phonon_states(dyn_dev: WannierObject, mass, xq_dev) -> PhononStates:
    """Device phonon states (src/compute_states.jl:285-311): D(q) in FP64, mass scaling, eigensolve
    with FP64 refinement, omega = sign(w2) sqrt(|w2|), u /= sqrt(m)."""
    torch = _torch()
    nq = xq_dev.shape[0]
    nm = int(round(math.sqrt(dyn_dev.ndata)))
    dev = xq_dev.device
    mass_d = torch.as_tensor(np.asarray(mass, dtype=np.float64), device=dev)
    D = core.fourier_batched(dyn_dev, xq_dev)
    L.call("ephip_ph_mass_scale", nm, nq, L.ptr(D), L.ptr(mass_d), L.stream_handle())
    w2, U = core.eigen_batched_refined(D)
    omega = torch.empty_like(w2)
    L.call("ephip_ph_finalize", nm, nq, L.ptr(w2), L.ptr(omega), L.ptr(U), L.ptr(mass_d), L.stream_handle())
    return PhononStates(xq_dev, omega, U)


def get_eph_RR_to_Rq_batched(epmat_dev: WannierObject, nw: int, nm: int, xq_dev, uph, precision=L.PREC_FP32,
                             work=None, out=None):
    """get_eph_RR_to_Rq! (src/wannier_to_bloch.jl:277-300) for a batch of q: Fourier over Rp of the
    "ph"-layout tensor (rows (iw, jw, nu, Re)), then the phonon-mode rotation by u_ph(q).
    Returns (nq, nw*nw*nm*nr_el) complex64: per q, the op_r of the electron-Wannier /
    phonon-Bloch object, column Re of nw*nw*nm rows."""
    torch = _torch()
    nq = xq_dev.shape[0]
    nd = epmat_dev.ndata
    nr_el = nd // (nw * nw * nm)
    if work is None:
        work = torch.empty((nq, nd), dtype=torch.complex64, device=xq_dev.device)
    if out is None:
        out = torch.empty((nq, nd), dtype=torch.complex64, device=xq_dev.device)
    core.fourier_batched(epmat_dev, xq_dev, out=work[:nq], precision=precision)
    uph = uph.contiguous()
    L.call("ephip_eph_rotate_modes", nw * nw, nm, nr_el, nq, L.ptr(work), work.stride(0), L.ptr(uph),
           L.ptr(out), out.stride(0), L.stream_handle())
    return out[:nq]


def eph_rotate_rqkq_g2(g, ik, ikq, states: ElectronStates, omega_q, nbw: int, g2=None, ep=None):
    """eph_apply_rotations_rqkq! + g2 (src/wannier_to_bloch_batched.jl:360-377, :451-453) on the
    in-window bands; returns g2 (ncol, nm, nbw, nbw) float32 (column-major (nbw, nbw, nm) blocks)."""
    torch = _torch()
    ncol = ik.shape[0]
    nw = states.nw
    nm = g.shape[1] // (nw * nw)
    if g2 is None:
        g2 = torch.empty((ncol, nm, nbw, nbw), dtype=torch.float32, device=g.device)
    L.call("ephip_eph_rotate_rqkq_g2", nw, nm, ncol, L.ptr(g), g.stride(0), L.ptr(ik), L.ptr(ikq),
           L.ptr(states.U), L.ptr(states.filt.band_min), L.ptr(states.filt.band_max), L.ptr(omega_q), nbw,
           L.ptr(g2), L.ptr(ep), L.stream_handle())
    return g2


# ------------------------------------------------------------------------------------------------
# Calculators (payload side of src/calculator/AbstractCalculator.jl:137-223)
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class EPDataKWindowed:
    """Device payload of one (q, k-batch): the windowed counterpart of EPDataKBatched
    (src/EPData.jl:105-114). g2[t] is the (nbw, nbw, nm) window block of pair (ik[t], iq);
    slot (ms, ns) is band (bmin[ikq[t]] - 1 + ms, bmin[ik[t]] - 1 + ns)."""
    iq: int
    ik: object
    ikq: object
    g2: object
    ep: object | None
    nbw: int


class AbstractCalculator:
    def setup(self, ctx):
        pass

    def begin_q(self, iq: int):
        pass

    def run(self, payload: EPDataKWindowed):
        raise NotImplementedError

    def end_q(self, iq: int):
        pass

    def postprocess(self):
        pass


class G2Recorder(AbstractCalculator):
    """Records every payload to the host (test helper; the reference's _RecordCalcOuterQ pattern,
    test/test_gpu.jl:600-654)."""

    def __init__(self, keep_ep: bool = False):
        self.blocks = []
        self.keep_ep = keep_ep

    def run(self, p: EPDataKWindowed):
        self.blocks.append((p.iq, p.ik.cpu().numpy(), p.ikq.cpu().numpy(), p.g2.cpu().numpy(),
                            None if p.ep is None else p.ep.cpu().numpy()))


class G2Sum(AbstractCalculator):
    """Gauge-invariant per-q scalar sum_k sum |g|^2 / (2w) on the device (modes above the acoustic
    cutoff only) -- the _RecordCalcOuterQ reduction of test/test_gpu.jl:600-654 without weights."""

    def __init__(self, nq: int, device):
        torch = _torch()
        self.acc = torch.zeros(nq, dtype=torch.float64, device=device)
        self.mode_mask = None

    def run(self, p: EPDataKWindowed):
        g2 = p.g2
        if self.mode_mask is not None:
            g2 = g2 * self.mode_mask[p.iq].view(1, -1, 1, 1)
        self.acc[p.iq] += g2.sum(dtype=_torch().float64)


# ------------------------------------------------------------------------------------------------
# Outer-q loop
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class OuterQSetup:
    el: ElectronStates
    ph: PhononStates
    ksurv: object          # (nks,) int64 fine-grid indices of k with an in-window band
    nbw: int               # max in-window band count
    epmat: WannierObject   # device, "ph" layout, complex64
    irvec_el_dev: object
    nw: int
    nm: int
    kgrid: tuple
    qgrid: tuple


def setup_outer_q(model, kgrid, qgrid, window, backend: core.GPUBackend, epmat_dev=None,
                  el_chunk: int = 32768) -> OuterQSetup:
    """Setup of run_eph_over_q_and_k (src/calculator/run_eph_over_q_and_k.jl:144-277): fine-grid
    electron states + window filter, phonon states on the q-grid, e-ph tensor on the device."""
    torch = _torch()
    from .model import kpoints_grid
    dev = backend.torch_device
    nw, nm = model.nw, model.nmodes
    ham = WannierObject(model.irvec_h, model.ham_r).to_device(backend, dtype=torch.complex128)
    dyn = WannierObject(model.irvec_d, model.dyn_r).to_device(backend, dtype=torch.complex128)
    xk = torch.from_numpy(kpoints_grid(kgrid)).to(dev)
    xq = torch.from_numpy(kpoints_grid(qgrid)).to(dev)
    el = compute_electron_states(ham, xk, window, chunk=el_chunk)
    ph = compute_phonon_states(dyn, model.mass, xq)
    ksurv = torch.nonzero(el.filt.keep, as_tuple=False).view(-1)
    width = (el.filt.band_max - el.filt.band_min + 1) * (el.filt.band_min > 0)
    nbw = max(1, int(width.max()))
    if epmat_dev is None:
        epmat_dev = WannierObject(model.irvec_ep, model.epmat_ph, check=False).to_device(backend,
                                                                                          dtype=torch.complex64)
    irvec_el_dev = torch.from_numpy(np.ascontiguousarray(model.irvec_el, dtype=np.int32)).to(dev)
    return OuterQSetup(el, ph, ksurv, nbw, epmat_dev, irvec_el_dev, nw, nm, tuple(kgrid), tuple(qgrid))


class OuterQLoop:
    """The per-q body of _loop_eph_over_q_and_k_batched (src/calculator/run_eph_over_q_and_k.jl:
    526-610): RR->Rq for a tile of q, then for each q and k-batch: Fourier over Re at the surviving
    k, k+q lookup, windowed rotations and g2. Buffers are allocated once."""

    def __init__(self, st: OuterQSetup, nk_batch: int = 8192, nq_tile: int = 16,
                 precision: int = L.PREC_FP32, keep_ep: bool = False):
        torch = _torch()
        self.st = st
        self.nk_batch = int(nk_batch)
        self.nq_tile = int(nq_tile)
        self.precision = precision
        self.keep_ep = keep_ep
        dev = st.el.xk.device
        nd_q = st.nw * st.nw * st.nm
        self.nr_el = st.epmat.ndata // nd_q
        self.gq_work = torch.empty((self.nq_tile, st.epmat.ndata), dtype=torch.complex64, device=dev)
        self.eRpq = torch.empty((self.nq_tile, st.epmat.ndata), dtype=torch.complex64, device=dev)
        self.gk = torch.empty((self.nk_batch, nd_q), dtype=torch.complex64, device=dev)
        self.g2 = torch.empty((self.nk_batch, st.nm, st.nbw, st.nbw), dtype=torch.float32, device=dev)
        self.ep = (torch.empty((self.nk_batch, st.nm, st.nbw, st.nbw), dtype=torch.complex64, device=dev)
                   if keep_ep else None)
        self.xk_surv = st.el.xk.index_select(0, st.ksurv).contiguous()
        self.xq_host = st.ph.xq.cpu().numpy()
        # optional list collecting (start event, end event, columns) around each Rq->kq Fourier launch;
        # the events are recorded on the current stream, which is the stream the kernels run on
        self.ft_timer = None

    def run_qtile(self, iq0: int, nq: int, calculators=()):
        st = self.st
        xq = st.ph.xq[iq0:iq0 + nq]
        uph = st.ph.u[iq0:iq0 + nq]
        get_eph_RR_to_Rq_batched(st.epmat, st.nw, st.nm, xq, uph, self.precision, work=self.gq_work,
                                 out=self.eRpq)
        nd_q = st.nw * st.nw * st.nm
        nks = st.ksurv.shape[0]
        for t in range(nq):
            iq = iq0 + t
            obj = _RqObject(self.eRpq[t].view(self.nr_el, nd_q), st.irvec_el_dev, nd_q)
            for c in calculators:
                c.begin_q(iq)
            for s in range(0, nks, self.nk_batch):
                e = min(nks, s + self.nk_batch)
                ik = st.ksurv[s:e]
                ikq = core.kq_index(self.xk_surv[s:e], self.xq_host[iq], st.kgrid)
                if self.ft_timer is not None:
                    torch = _torch()
                    ev0 = torch.cuda.Event(enable_timing=True)
                    ev1 = torch.cuda.Event(enable_timing=True)
                    ev0.record()
                core.fourier_batched(obj, self.xk_surv[s:e], out=self.gk[: e - s], precision=self.precision)
                if self.ft_timer is not None:
                    ev1.record()
                    self.ft_timer.append((ev0, ev1, e - s))
                g2 = eph_rotate_rqkq_g2(self.gk[: e - s], ik, ikq, st.el, st.ph.omega[iq], st.nbw,
                                        g2=self.g2[: e - s], ep=None if self.ep is None else self.ep[: e - s])
                payload = EPDataKWindowed(iq, ik, ikq, g2, None if self.ep is None else self.ep[: e - s], st.nbw)
                for c in calculators:
                    c.run(payload)
            for c in calculators:
                c.end_q(iq)

    def run(self, iq_start: int = 0, iq_stop: int | None = None, calculators=()):
        nq = self.st.ph.xq.shape[0]
        iq_stop = nq if iq_stop is None else min(nq, iq_stop)
        for iq0 in range(iq_start, iq_stop, self.nq_tile):
            self.run_qtile(iq0, min(self.nq_tile, iq_stop - iq0), calculators)


class _RqObject(WannierObject):
    """Device electron-Wannier / phonon-Bloch object at one q (op_r (nw*nw*nm) x nr_el); irvec lives
    on the device already."""

    def __init__(self, op_r, irvec_dev, ndata):
        self.op_r = op_r
        self.nr = op_r.shape[0]
        self.ndata = ndata
        self.irvec_next = None
        self._irvec_dev = irvec_dev
        self.irvec = None


# ------------------------------------------------------------------------------------------------
# Outer-k loop with the k-side window projection
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class EPDataUnits:
    """Device payload of one k batch of the outer-k loop: the windowed counterpart of
    EPDataQBatched (src/EPData.jl:66-73) for a whole k batch. Unit u is the pair (k state
    ``unit_state[u]``, in-window band ``unit_band[u]`` (0-based)); its kept q are
    ``col_index[u, :count[u]]`` (q-grid indices) with k+q states ``ikq[u, :count[u]]``, and
    ``g2[u, c]`` holds (nbw, nm) values [m, nu] for k+q window band bmin(k+q) - 1 + m. q not listed
    for a unit have no in-window band at k+q: their g2 is exactly zero."""
    k_range: tuple
    unit_state: object
    unit_band: object
    count: object
    col_index: object
    ikq: object
    g2: object
    nbw: int
    ep: object = None   # complex64 (units, cs, nbw * nm): the rotated elements behind g2, when requested


@dataclasses.dataclass
class EPDataQBatched:
    """The reference's batched outer-k payload EPDataQBatched (src/EPData.jl:66-73): one outer k with a
    batch of k+q. Column-major arrays of the reference held as torch tensors with the axes reversed:
    eps (nq, nm, nbandk, nw) complex64 = eps[m, n, nu, j] over the full k+q band range m and the k-side
    window (physical band ibandk_offset + n, 1-based n), zero outside both energy windows; g2s the same
    shape, float32 (|ep|^2 / (2 omega)); wqs (nq, nm) float64 = omega of each column's q; ik the 1-based
    position of k in the filtered state list; ikqs (nq) int64, the 1-based positions of the k+q states
    (k+q with no window band are not listed: their elements are exactly zero)."""
    eps: object
    g2s: object
    wqs: object
    ik: int
    ikqs: object
    ibandk_offset: int


class QBatchedAdapter(AbstractCalculator):
    """Runs calculators written against EPDataQBatched on the outer-k loop's batch payloads: for each k of
    a batch (its units are consecutive) and tiles of at most nq_batch kept k+q, ephip_units_to_qbatched lays
    the k's g2 (and eps, when the loop was built with with_ep=True) out as the reference's payload
    (src/calculator/run_eph_over_k_and_kq.jl:843-876: full band at k+q, the nbandk_max-wide k-side window
    at offset clamp(first band - 1, 0, nw - nbandk_max)), then calls calc.run(payload) for each calculator."""

    def __init__(self, st, calculators, nq_batch: int = 4096, n_batch_max: int | None = None):
        torch = _torch()
        self.st = st
        self.calculators = list(calculators)
        self.nq_batch = int(nq_batch)
        self.n_batch_max = n_batch_max   # the loop's k batch (sizes calculators' outer-k tiles)
        dev = st.el.xk.device
        self.nbandk = int(st.nbw)
        nk_f = st.el.xk.shape[0]
        pos = np.zeros(nk_f, dtype=np.int64)
        pos[st.ksurv.cpu().numpy()] = np.arange(1, st.ksurv.shape[0] + 1)
        self.pos_host = pos
        self.state_pos = torch.from_numpy(pos).to(dev)
        self.bmin_host = st.el.filt.band_min.cpu().numpy()
        nw, nm, nb = st.nw, st.nm, self.nq_batch
        self.eps = torch.empty((nb, nm, self.nbandk, nw), dtype=torch.complex64, device=dev)
        self.g2s = torch.empty((nb, nm, self.nbandk, nw), dtype=torch.float32, device=dev)
        self.wqs = torch.empty((nb, nm), dtype=torch.float64, device=dev)
        self.ikqs = torch.empty(nb, dtype=torch.int64, device=dev)

    def run(self, p: EPDataUnits):
        st = self.st
        nw, nm = st.nw, st.nm
        ustate = p.unit_state.cpu().numpy()
        count = p.count.cpu().numpy()
        cs = p.g2.shape[1]
        pos = self.pos_host
        stream = L.stream_handle()
        nbm = self.n_batch_max or (p.k_range[1] - p.k_range[0])
        for calc in self.calculators:   # calculator_begin!(calc, OuterIterationBatch(), ctx)
            if hasattr(calc, "begin_batch"):
                calc.begin_batch(p.k_range, nbm)
        u = 0
        while u < len(ustate):
            k = int(ustate[u])
            ub = u
            while ub + 1 < len(ustate) and ustate[ub + 1] == k:
                ub += 1
            nunits = ub - u + 1
            band0 = int(self.bmin_host[k]) - 1
            off = min(max(band0, 0), nw - self.nbandk)
            c = int(count[u])
            for j0 in range(0, c, self.nq_batch):
                nq = min(self.nq_batch, c - j0)
                L.call("ephip_units_to_qbatched", nw, nm, p.nbw, self.nbandk, cs, L.ptr(p.col_index[u]),
                       L.ptr(p.ikq[u]), j0, nq, nunits, band0, off,
                       L.ptr(p.ep[u]) if p.ep is not None else None, L.ptr(p.g2[u]), L.ptr(st.el.filt.band_min),
                       L.ptr(st.el.filt.band_max), L.ptr(st.ph.omega), L.ptr(self.state_pos),
                       L.ptr(self.eps) if p.ep is not None else None, L.ptr(self.g2s), L.ptr(self.wqs),
                       L.ptr(self.ikqs), stream)
                payload = EPDataQBatched(self.eps[:nq] if p.ep is not None else None, self.g2s[:nq], self.wqs[:nq],
                                         int(pos[k]), self.ikqs[:nq], off)
                for calc in self.calculators:
                    calc.run(payload)
            u = ub + 1
        for calc in self.calculators:   # calculator_end!(calc, OuterIterationBatch(), ctx)
            if hasattr(calc, "end_batch"):
                calc.end_batch()


@dataclasses.dataclass
class BatchGraph:
    """A captured sequence of outer-k batches (OuterKLoop.capture): ``replay()`` launches the recorded
    graph on the current stream. ``payloads`` are the batches' payload views into the loop's buffers as
    the capture built them (after a replay, the last batch's is current)."""
    graph: object
    starts: list
    n: int
    payloads: list

    def replay(self):
        self.graph.replay()


class OuterKLoop:
    """Outer-k e-ph loop on the device (the reference's _loop_eph_over_k_and_kq_batched,
    src/calculator/run_eph_over_k_and_kq.jl:584-895, with the k-side window projection of :657-659,
    and the q-grid as the inner list). Per k batch:
      1. RR->kR: Fourier over Re of the "el"-layout tensor at the batch's k (matrix cores), then the
         projection onto each k's in-window bands -> one operator g(k)[i, n, nu, Rp] per unit (k, n);
      2. per unit, the q whose k+q has an in-window band (k+q looked up on the fine grid);
      3. kR->kq: one batched Fourier over Rp for all units at their kept q (matrix cores);
      4. k+q-window and phonon rotations, g2 = |ep|^2 / (2 omega).
    Pairs (k, q) with no in-window band at k+q are exactly zero and are not computed."""

    def __init__(self, st: OuterQSetup, nk_batch: int | None = 128, precision: int = L.PREC_FP32,
                 nk_batch_cap: int = 512, fuse_rotation: bool = True, fuse_projection: bool = True,
                 gridopt: bool = True, gridopt_mfma: bool = True, with_ep: bool = False):
        torch = _torch()
        self.st = st
        if nk_batch is None:  # memory-adaptive width (plan_batch), the e-ph tensor already resident
            per_k, _ = core.outer_k_staging_bytes(st.nw, st.nm, st.irvec_el_dev.shape[0], st.epmat.nr,
                                                  st.el.xk.shape[0], st.ph.xq.shape[0], st.nbw)
            nre_b = 8 * st.epmat.op_r.numel()  # the "el" copy is allocated below
            nk_batch = core.plan_batch(core.free_bytes(None), per_k, nre_b, nk_batch_cap, what="outer-k")
        self.nk_batch = int(nk_batch)
        self.precision = precision
        dev = st.el.xk.device
        nw, nm = st.nw, st.nm
        self.nre = st.irvec_el_dev.shape[0]
        self.nrp = st.epmat.nr
        self.nd_el = nw * nw * nm * self.nrp
        # "el" layout copy of the e-ph tensor: rows (i, j, nu, Rp), one column per Re
        self.epmat_el = torch.empty((self.nre, self.nd_el), dtype=torch.complex64, device=dev)
        L.call("ephip_eph_ph_to_el", nw * nw * nm, self.nre, self.nrp, L.ptr(st.epmat.op_r), L.ptr(self.epmat_el),
               L.stream_handle())
        self.el_obj = _RqObject(self.epmat_el, st.irvec_el_dev, self.nd_el)
        self.irvec_ep_dev = st.epmat.irvec_dev()
        # RR->kR as the staged (GridOpt) transform with the projection fused, in FP32 for either precision:
        # k1 planes and (k1, k2) lines of the grid share their partial sums
        ks_host = st.ksurv.cpu().numpy()
        self.gridopt = _GridOptKR.build(st, ks_host, self.nd_el, gridopt_mfma) if (gridopt and nw in (8, 16)) \
            else None
        # kR->kq as the staged transform over the uniform q grid with the rotations and g2 fused (FP32)
        self.gridopt_q = _GridOptKQ.build(st) if (gridopt and nw == 16 and nm <= 16 and st.nbw <= 4) else None
        # split-BF16: the contractions that are not staged run on pre-packed operands. The "el" tensor is
        # split once (same bytes as complex64) and its complex64 copy released; the q-grid phase table is
        # built once; k phases and unit operands are packed per batch.
        self.packed = precision == L.PREC_BF16X3 and (self.gridopt is None or self.gridopt_q is None)
        # split-BF16 with nw = 16, nm = 12, nbw <= 4: kR->kq and the rotations run as one kernel that
        # writes g2 only (no g(k, q) staging buffer)
        self.fused = (bool(fuse_rotation) and self.packed and nw == 16 and nm == 12 and st.nbw <= 4
                      and self.gridopt_q is None)
        # split-BF16 with nw = 16: RR->kR and the k-side window projection run as one kernel on the "el"
        # operand packed with j fastest; g(k, Rp) is never stored
        self.fused_proj = (bool(fuse_projection) and self.packed and nw == 16 and self.nd_el % 192 == 0
                           and self.gridopt is None)
        # arithmetic each contraction runs in: the staged kernels are FP32 whatever ``precision`` asks for
        self.precision_requested = precision
        self.stage_precision = {
            "RR->kR": L.PREC_FP32 if self.gridopt is not None else precision,
            "kR->kq": L.PREC_FP32 if self.gridopt_q is not None else precision,
        }
        self.precision_overridden = any(p != precision for p in self.stage_precision.values())
        import ctypes
        self.qgrid_c = (ctypes.c_int32 * 3)(*map(int, st.qgrid))  # for the on-the-fly q phases
        if self.gridopt is not None:
            self.el_obj = None   # the complex64 "el" tensor stays: stage 1 streams it
        elif self.fused_proj:
            self.apk_el = core.ft_pack_operand_swapped(self.epmat_el, self.nd_el, self.nre, nw, self.nd_el)
        elif self.packed:
            self.apk_el = core.ft_pack_operand(self.epmat_el, self.nd_el, self.nre, self.nd_el)
        if self.packed:
            if self.gridopt is None:
                self.epmat_el = None
                self.el_obj = None
                self.bpk_k = torch.empty(core.ft_stages(self.nre) * self.nk_batch * 64, dtype=torch.bfloat16,
                                         device=dev)
            if self.gridopt_q is None:
                self.bpk_q = core.ft_pack_phases(self.irvec_ep_dev, st.ph.xq)
        # units of every surviving k, in k order
        bmin = st.el.filt.band_min.cpu().numpy()
        bmax = st.el.filt.band_max.cpu().numpy()
        ks = st.ksurv.cpu().numpy()
        nb = (bmax[ks] - bmin[ks] + 1).astype(np.int64)
        self.k_unit_start = np.concatenate([[0], np.cumsum(nb)])
        self.k_unit_start_dev = torch.from_numpy(self.k_unit_start.astype(np.int32)).to(dev)
        self.max_units_per_k = int(nb.max()) if len(nb) else 1
        self.unit_kpos = torch.from_numpy(np.repeat(np.arange(len(ks)), nb).astype(np.int32)).to(dev)
        self.unit_state = torch.from_numpy(np.repeat(ks, nb).astype(np.int64)).to(dev)
        band = np.concatenate([np.arange(bmin[k] - 1, bmax[k]) for k in ks]) if len(ks) else np.zeros(0)
        self.unit_band = torch.from_numpy(band.astype(np.int32)).to(dev)
        self.xk_surv = st.el.xk.index_select(0, st.ksurv).contiguous()
        self.nq = st.ph.xq.shape[0]
        self.max_kept = kept_q_bound(ks, st.kgrid, st.qgrid)
        self.cs = self.nq
        # units of the largest run of nk_batch consecutive surviving k: a batch may start at any position
        w = min(self.nk_batch, len(ks))
        max_units = int((self.k_unit_start[w:] - self.k_unit_start[:len(ks) + 1 - w]).max()) if len(ks) else 1
        self.max_units = max_units
        nd2 = nw * nm
        self.G = None if (self.fused_proj or self.gridopt is not None) else torch.empty(
            (self.nk_batch, self.nd_el), dtype=torch.complex64, device=dev)
        self.A = torch.empty((max_units, nd2 * self.nrp), dtype=torch.complex64, device=dev)
        self.col_index = torch.empty((max_units, self.cs), dtype=torch.int32, device=dev)
        self.ikq = torch.empty((max_units, self.cs), dtype=torch.int32, device=dev)
        self.count = torch.empty(max_units, dtype=torch.int32, device=dev)
        self.Gq = None if (self.fused or self.gridopt_q is not None) else torch.empty(
            (max_units, self.cs, nd2), dtype=torch.complex64, device=dev)
        if self.gridopt_q is not None:
            self.line_off = torch.empty((max_units, self.gridopt_q.nlines + 1), dtype=torch.int32, device=dev)
        self.g2 = torch.empty((max_units, self.cs, st.nbw * nm), dtype=torch.float32, device=dev)
        # the complex elements behind g2 (the EPDataQBatched eps), from the staged or the unfused rotation kernels
        if with_ep and self.fused:
            raise ValueError("with_ep: the fused split-BF16 kR->kq kernel writes g2 only; use gridopt=True or "
                             "fuse_rotation=False")
        self.ep = torch.empty((max_units, self.cs, st.nbw * nm), dtype=torch.complex64, device=dev) if with_ep else None
        if self.packed and self.gridopt_q is None:
            self.apk_units = torch.empty(max_units * core.ft_stages(self.nrp) * nd2 * 64, dtype=torch.bfloat16,
                                         device=dev)
        self.uph = st.ph.u.contiguous()
        self.omega = st.ph.omega.contiguous()
        self.ft_timer = None   # optional list of (event, event, label, flops) around the two Fourier launches

    def _timed(self, label, flops, fn):
        if self.ft_timer is None:
            return fn()
        torch = _torch()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        r = fn()
        b.record()
        self.ft_timer.append((a, b, label, flops))
        return r

    def _batch_units(self, i0: int, n: int):
        """(n, u0, u1) of the batch of surviving k i0 .. i0 + n - 1 (n clipped to the list); the batch
        buffers hold nk_batch k and max_units units, so a longer batch is refused."""
        if n > self.nk_batch or i0 < 0:
            raise ValueError(f"k batch of {n} at {i0}: the loop's buffers hold batches of at most {self.nk_batch}")
        n = min(n, self.st.ksurv.shape[0] - i0)
        if n <= 0:
            return 0, 0, 0
        u0, u1 = int(self.k_unit_start[i0]), int(self.k_unit_start[i0 + n])
        assert u1 - u0 <= self.max_units
        return n, u0, u1

    def run_kbatch(self, i0: int, n: int, calculators=()):
        n, u0, u1 = self._batch_units(i0, n)
        if n <= 0:
            return None
        stream = L.stream_handle()
        self._stage_rrkr(i0, n, u0, u1, self.A, stream)
        return self._stage_krkq(i0, n, u0, u1, self.A, calculators, stream)

    def capture(self, starts, n: int, calculators=(), warmup: bool = True) -> "BatchGraph":
        """Capture the device work of the k batches starting at the surviving-k positions ``starts`` (n k
        each), with the consumers' device work, as one HIP graph on the current device (the loop of
        src/calculator/run_eph_over_k_and_kq.jl:584-895 over those batches, launch for launch). Replaying
        it redoes exactly those launches into the loop's buffers; nothing is read back between batches
        (the staged path sizes every launch on the host from tables fixed at setup), so a capture sees the
        whole batch sequence. The stage-1 plane cache is invalidated before and after the capture, so the
        graph computes every k1 plane it reads and eager batches after it do not trust planes the graph
        only recorded. Consumers must do device work only (no host copies, allocations or
        synchronisation). ``warmup``: run the first batch eagerly first (first launches of each kernel and
        the lazily allocated tables happen outside the capture)."""
        torch = _torch()
        starts = [int(i0) for i0 in starts]
        if warmup and starts:
            self.run_kbatch(starts[0], n)
        if self.gridopt is not None:
            self.gridopt.invalidate()
        timer, self.ft_timer = self.ft_timer, None
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        try:
            with torch.cuda.graph(graph):
                payloads = [self.run_kbatch(i0, n, calculators) for i0 in starts]
        finally:
            self.ft_timer = timer
            if self.gridopt is not None:
                self.gridopt.invalidate()
        return BatchGraph(graph, starts, n, payloads)

    def run_pipelined(self, starts, n: int, calculators=()):
        """Run the k batches starting at the surviving-k positions ``starts`` (n k each) on two streams of
        default priority: RR->kR of batch b+1 overlaps kR->kq (and the consumers) of batch b, with two
        unit-operator buffers and events between the streams. Consumers run on the kR->kq stream in batch
        order; the caller's stream waits for both at the end."""
        torch = _torch()
        if getattr(self, "A2", None) is None:
            self.A2 = torch.empty_like(self.A)
        cur = torch.cuda.current_stream()
        s_rr, s_kq = torch.cuda.Stream(), torch.cuda.Stream()
        s_rr.wait_stream(cur)
        s_kq.wait_stream(cur)
        bufs, free = (self.A, self.A2), [None, None]
        out = []
        for b, i0 in enumerate(starts):
            nb, u0, u1 = self._batch_units(i0, n)
            if nb <= 0:
                continue
            A = bufs[b % 2]
            with torch.cuda.stream(s_rr):
                if free[b % 2] is not None:
                    s_rr.wait_event(free[b % 2])
                self._stage_rrkr(i0, nb, u0, u1, A, L.stream_handle())
                ready = torch.cuda.Event()
                ready.record(s_rr)
            with torch.cuda.stream(s_kq):
                s_kq.wait_event(ready)
                out.append(self._stage_krkq(i0, nb, u0, u1, A, calculators, L.stream_handle()))
                free[b % 2] = torch.cuda.Event()
                free[b % 2].record(s_kq)
        cur.wait_stream(s_rr)
        cur.wait_stream(s_kq)
        return out

    def _stage_rrkr(self, i0, n, u0, u1, A, stream):
        """RR->kR over Re for the batch and the k-side projection: writes the unit operators into A."""
        st = self.st
        nw, nm = st.nw, st.nm
        nu = u1 - u0
        # 1. RR->kR Fourier over Re for the batch
        xk_b = self.xk_surv[i0:i0 + n]
        if self.gridopt is not None:  # staged RR->kR with the projection fused: writes the unit operators A
            self._timed("RR->kR", 8.0 * self.nd_el * self.nre * n,
                        lambda: self.gridopt.run(self, i0, n, u0, u1, A, stream))
        elif self.fused_proj:  # RR->kR with the projection in its epilogue: writes the unit operators A
            bpk = core.ft_pack_phases(st.irvec_el_dev, xk_b, out=self.bpk_k)
            self._timed("RR->kR", 8.0 * self.nd_el * self.nre * n, lambda: L.call(
                "ephip_eph_RR_to_kR_project_packed_bf16x3", nw, nm, self.nrp, self.nre, L.ptr(self.apk_el), L.ptr(bpk),
                n, n, L.ptr(self.k_unit_start_dev[i0:]), L.ptr(self.unit_state[u0:u1]), L.ptr(self.unit_band[u0:u1]),
                L.ptr(st.el.U), self.max_units_per_k, L.ptr(A), A.stride(0), stream))
        elif self.packed:
            bpk = core.ft_pack_phases(st.irvec_el_dev, xk_b, out=self.bpk_k)
            self._timed("RR->kR", 8.0 * self.nd_el * self.nre * n, lambda: core.fourier_packed(
                self.apk_el, self.nd_el, self.nre, bpk, n, n, self.G, self.G.stride(0)))
        else:
            self._timed("RR->kR", 8.0 * self.nd_el * self.nre * n,
                        lambda: core.fourier_batched(self.el_obj, xk_b, out=self.G[:n], precision=self.precision))
        if not self.fused_proj and self.gridopt is None:
            # projection; the unit's column in G is its k position minus i0 (base pointer shifted by -i0 columns)
            g_base = L.ptr(self.G) - i0 * self.G.stride(0) * 8
            L.call("ephip_eph_project_kR", nw, nm, self.nrp, nu, g_base, self.G.stride(0),
                   L.ptr(self.unit_kpos[u0:u1]), L.ptr(self.unit_state[u0:u1]), L.ptr(self.unit_band[u0:u1]),
                   L.ptr(st.el.U), L.ptr(A), stream)

    def _stage_krkq(self, i0, n, u0, u1, A, calculators, stream):
        """Kept q per unit, kR->kq over Rp with the rotations and g2 from the unit operators in A, and the
        consumers of the batch payload."""
        import ctypes as C
        st = self.st
        nw, nm = st.nw, st.nm
        nu = u1 - u0
        # 2. kept q per unit
        g = (C.c_int32 * 3)(*map(int, st.kgrid))
        L.call("ephip_kq_compact", nu, L.ptr(self.unit_state[u0:u1]), L.ptr(st.el.xk), self.nq, L.ptr(st.ph.xq),
               C.addressof(g), L.ptr(st.el.filt.band_min), self.cs, L.ptr(self.col_index), L.ptr(self.ikq),
               L.ptr(self.count), stream)
        # 3. kR->kq batched Fourier over Rp at the kept q
        nd2 = nw * nm
        if self.gridopt_q is not None:
            self._timed("kR->kq", None, lambda: self.gridopt_q.run(self, nu, A, stream))
        elif self.fused:
            core.ft_pack_operand(A, nd2, self.nrp, nd2, nu, nd2 * self.nrp, out=self.apk_units)
            self._timed("kR->kq", None, lambda: L.call(
                "ephip_eph_kq_rotate_packed_bf16x3", nw, nm, st.nbw, self.nrp, L.ptr(self.apk_units), nu,
                L.ptr(self.bpk_q), self.nq, self.cs, self.max_kept, L.ptr(self.count), L.ptr(self.col_index), L.ptr(self.ikq),
                L.ptr(st.el.U), L.ptr(st.el.filt.band_min), L.ptr(st.el.filt.band_max), L.ptr(self.uph),
                L.ptr(self.omega), L.ptr(self.g2), L.ptr(self.irvec_ep_dev), C.addressof(self.qgrid_c), stream))
        elif self.packed:
            def krkq():
                core.ft_pack_operand(A, nd2, self.nrp, nd2, nu, nd2 * self.nrp, out=self.apk_units)
                core.fourier_packed(self.apk_units, nd2, self.nrp, self.bpk_q, self.nq, self.cs, self.Gq, nd2,
                                    nunits=nu, ncol=self.count, col_index=self.col_index, col_stride=self.cs,
                                    out_stride=self.cs * nd2)
            self._timed("kR->kq", None, krkq)
        else:
            self._timed("kR->kq", None, lambda: L.call(
                "ephip_fourier_batched_c32", nd2, self.nrp, L.ptr(A), nd2, nd2 * self.nrp,
                L.ptr(self.irvec_ep_dev), L.ptr(st.ph.xq), nu, self.nq, L.ptr(self.count), L.ptr(self.col_index),
                self.cs, L.ptr(self.Gq), nd2, self.cs * nd2, self.precision, stream))
        # 4. rotations + g2 (already done by the fused kernels)
        if not self.fused and self.gridopt_q is None:
            L.call("ephip_eph_rotate_kq_g2", nw, nm, st.nbw, nu, self.cs, L.ptr(self.count),
                   L.ptr(self.col_index), L.ptr(self.ikq), L.ptr(self.Gq), nd2, self.cs * nd2, L.ptr(st.el.U),
                   L.ptr(st.el.filt.band_min), L.ptr(st.el.filt.band_max), L.ptr(self.uph), L.ptr(self.omega),
                   L.ptr(self.g2), L.ptr(self.ep) if self.ep is not None else None, stream)
        payload = EPDataUnits((i0, i0 + n), self.unit_state[u0:u1], self.unit_band[u0:u1], self.count[:nu],
                              self.col_index[:nu], self.ikq[:nu], self.g2[:nu], st.nbw,
                              self.ep[:nu] if self.ep is not None else None)
        for c in calculators:
            c.run(payload)
        return payload

    def kernel_flops(self, i0: int, n: int, kept: int, kept_slots: int | None = None) -> tuple[float, float]:
        """Algorithmic flops (8 per complex multiply-add) of the batch's RR->kR and kR->kq launches for the
        kernels in use; kept = the batch's kept (unit, q) pairs, kept_slots = their (unit, q, k+q window
        band) slots (default: nbw per pair). Direct Fourier kernels: the full phase sums. Staged (GridOpt)
        kernels: their own stages, the per-plane stage 1 of RR->kR counted once per surviving k of its
        plane, the stage 2 of kR->kq for every q line of every unit. Rotations: per k+q window band."""
        st = self.st
        nw, nm = st.nw, st.nm
        n = min(n, st.ksurv.shape[0] - i0)
        nu = int(self.k_unit_start[i0 + n] - self.k_unit_start[i0])
        nd2 = nw * nm
        slots = st.nbw * kept if kept_slots is None else kept_slots
        rot = 8.0 * (nw * nm + nm * nm) * slots
        if self.gridopt is not None:
            go = self.gridopt
            planes = go.line_plane[go.line_of_k[i0:i0 + n]]
            per_plane = np.bincount(go.line_plane[go.line_of_k], minlength=int(planes.max()) + 1)
            # stage 2 per line: n2 n3 complex multiply-adds, or in the folded form h pairs (one multiply-add's
            # worth each) per c over the padded np x np face, the fold's adds counted once per workgroup
            s2 = go.np * (go.np // 2) if go.folded() else go.n2 * go.n3
            f_rr = 8.0 * self.nd_el * (self.nre * float(np.sum(1.0 / per_plane[planes]))
                                       + s2 * (int(go.line_of_k[i0 + n - 1]) - int(go.line_of_k[i0]) + 1)
                                       + go.n3 * n + nu)
        else:
            f_rr = 8.0 * self.nd_el * self.nre * n
        if self.gridopt_q is not None:
            gq = self.gridopt_q
            n2n3 = int((gq.rmap.view(gq.np, gq.n1, 8) >= 0).any(1).sum().item())  # (c, b) columns present
            from . import tune as T
            mfma_kq = T.get_variant(T.KIND_GRIDOPT_KQ) == 1
            # stage 1: nrp complex multiply-adds per (unit, q1) and row in the workgroup, or by the pre-pass
            # np^2 sums of np terms (h pairs when r1 spans -h .. h) per (unit, q1) and row
            s1 = self.nrp
            if gq.prepass and gq.can_prepass():
                s1 = gq.np * gq.np * ((gq.np // 2) if gq.prepass_folded() else gq.np)
            if gq.folded():
                # stages 2 (and 3 in the VALU form) as pair sums over opposite offsets: h = np // 2 pairs of
                # 4 FMAs (one complex multiply-add's worth) per sum, plus the folds (4 adds per pair) of the
                # stage-1 sums once per workgroup: over b, and in the VALU form also over c (ahead of stage 2)
                h = gq.np // 2
                fold = 0.5 * gq.np * h * gq.m[0] * nu
                if mfma_kq:
                    f_kq = (8.0 * nd2 * (s1 * gq.m[0] * nu + fold + gq.np * h * gq.nlines * nu + gq.np * slots)
                            + 8.0 * nm * nm * slots)
                else:
                    f_kq = 8.0 * nd2 * (s1 * gq.m[0] * nu + 2 * fold + gq.np * h * gq.nlines * nu + h * kept) + rot
            elif mfma_kq:
                # stage 3 and the k+q-window rotation as one contraction over (i, c) per window band slot
                f_kq = (8.0 * nd2 * (s1 * gq.m[0] * nu + n2n3 * gq.nlines * nu + gq.np * slots)
                        + 8.0 * nm * nm * slots)
            else:
                f_kq = 8.0 * nd2 * (s1 * gq.m[0] * nu + n2n3 * gq.nlines * nu + gq.np * kept) + rot
        else:
            f_kq = 8.0 * nd2 * self.nrp * kept + (rot if self.fused else 0.0)
        return f_rr, f_kq

    def tune(self, reps: int = 3) -> dict:
        """Pick the Fourier kernel variant (FP32: LDS staging; split-BF16: packed tile shape) with the
        shortest time for one full k batch."""
        from . import tune as T
        n = min(self.nk_batch, self.st.ksurv.shape[0])
        if self.gridopt is not None and self.gridopt_q is not None:
            # both contractions staged: the staged kR->kq kernel's form (vector ALUs or MFMAs), then where its
            # stage 1 runs (in each workgroup or by the pre-pass)
            res = {"gridopt_kq": T.tune_fourier(T.KIND_GRIDOPT_KQ, lambda: self.run_kbatch(0, n), reps=reps,
                                                variants=T.GRIDOPT_KQ_VARIANTS)}
            gq = self.gridopt_q
            if gq.can_prepass():
                import time
                torch = _torch()
                times = {}
                for v in T.KQ_STAGE1_VARIANTS:
                    gq.prepass = bool(v)
                    self.run_kbatch(0, n)
                    torch.cuda.synchronize()
                    ts = []
                    for _ in range(reps):
                        t0 = time.perf_counter()
                        self.run_kbatch(0, n)
                        torch.cuda.synchronize()
                        ts.append(1e3 * (time.perf_counter() - t0))
                    times[v] = float(np.median(ts))
                best = min(times, key=times.get)
                gq.prepass = bool(best)
                if not gq.prepass:
                    gq.b1 = None  # the trial's workspace is not kept
                res["kq_stage1"] = {**times, "best": best}
            return res
        if self.packed:
            # the fused RR->kR + projection kernel has its own variants; tile shapes apply to the unfused one
            if self.gridopt is not None:
                res = {}
            elif self.fused_proj and len(T.PROJ_TUNE) == 1:
                T.set_variant(T.KIND_FUSED_PROJ, T.PROJ_TUNE[0])
                res = {}
            elif self.fused_proj:
                res = {"proj": T.tune_fourier(T.KIND_FUSED_PROJ, lambda: self.run_kbatch(0, n), reps=reps,
                                              variants=T.PROJ_TUNE)}
            else:
                res = {"packed": T.tune_fourier(T.KIND_PACKED_BF16, lambda: self.run_kbatch(0, n), reps=reps,
                                                variants=T.PACKED_VARIANTS)}
            if self.fused and self.gridopt_q is None:
                res["fused"] = T.tune_fourier(T.KIND_FUSED_ROT, lambda: self.run_kbatch(0, n), reps=reps,
                                              variants=T.FUSED_TUNE)
            return res
        return T.tune_fourier(T.KIND_FP32_WIDE, lambda: self.run_kbatch(0, n), reps=reps)

    def run(self, k_start: int = 0, k_stop: int | None = None, calculators=()):
        nks = self.st.ksurv.shape[0]
        k_stop = nks if k_stop is None else min(nks, k_stop)
        for i0 in range(k_start, k_stop, self.nk_batch):
            self.run_kbatch(i0, min(self.nk_batch, k_stop - i0), calculators)


class _GridOptKR:
    """Staged RR->kR over Re (ephip_gridopt_stage1_planes + ephip_gridopt_kR_project[_mfma]), the device
    GridOpt of src/wannier/GridOpt.jl: Re embedded in its bounding box, the surviving k (grid order,
    third index fastest) grouped into k1 planes and (k1, k2) lines. Stage 1 runs for a group of PLANES
    consecutive planes at once (one pass over the operand for all of them); the two most recent groups
    are kept, so a pass over the k batches in order computes each plane once."""

    N1_MAX, N23_MAX, KLINE_MAX, NBW_MAX = 32, 8, 64, 4
    PLANES = 8

    @classmethod
    def build(cls, st, ks, nd, mfma=True):
        r = np.asarray(st.irvec_el_dev.cpu().numpy(), dtype=np.int64)
        lo, hi = r.min(0), r.max(0)
        n = hi - lo + 1
        kg = np.asarray(st.kgrid, dtype=np.int64)
        if (len(ks) == 0 or n[0] > cls.N1_MAX or n[1] > cls.N23_MAX or n[2] > cls.N23_MAX
                or st.nbw > cls.NBW_MAX):
            return None
        return cls(st, ks, nd, r, lo, n, kg, mfma)

    def __init__(self, st, ks, nd, r, lo, n, kg, mfma):
        torch = _torch()
        dev = st.el.xk.device
        self.nd = nd
        # stages 2/3 on the matrix cores for nw = 16 (op1 rows in natural order); otherwise the VALU form,
        # whose lane reduction wants op1 rows j-fastest
        self.mfma = bool(mfma) and st.nw == 16
        self._nb = 1 if self.mfma else st.nw
        self.n1, self.n2, self.n3 = (int(x) for x in n)
        self.np = max(self.n2, self.n3, 3)
        self.rmin = [int(x) for x in lo]
        rmap = -np.ones((self.n3, self.n2, self.n1), np.int32)
        rmap[r[:, 2] - lo[2], r[:, 1] - lo[1], r[:, 0] - lo[0]] = np.arange(len(r), dtype=np.int32)
        self.rmap = torch.from_numpy(rmap.ravel()).to(dev)
        i1, i2 = ks // (kg[1] * kg[2]), (ks // kg[2]) % kg[1]
        change = np.flatnonzero((np.diff(i1) != 0) | (np.diff(i2) != 0)) + 1
        starts = np.concatenate([[0], change, [len(ks)]]).astype(np.int64)
        # a (k1, k2) line with more than KLINE_MAX surviving k (k3 grids above 64) is cut into pieces of at
        # most KLINE_MAX k; each piece forms its own stage-2 sums (the same k2 phases)
        if np.diff(starts).max() > self.KLINE_MAX:
            pieces = [np.arange(a, b, self.KLINE_MAX) for a, b in zip(starts[:-1], starts[1:])]
            starts = np.concatenate(pieces + [[len(ks)]]).astype(np.int64)
        self.line_start_host = starts
        self.line_start = torch.from_numpy(starts.astype(np.int32)).to(dev)
        self.line_plane = i1[starts[:-1]]
        self.line_of_k = np.repeat(np.arange(len(starts) - 1), np.diff(starts))
        self.max_line_k = int(np.diff(starts).max())
        nb_k = (st.el.filt.band_max.cpu().numpy()[ks] - st.el.filt.band_min.cpu().numpy()[ks] + 1).astype(np.int64)
        cum = np.concatenate([[0], np.cumsum(nb_k)])
        self.max_line_units = int((cum[starts[1:]] - cum[starts[:-1]]).max())  # units of the fullest line
        x1 = st.el.xk.index_select(0, st.ksurv).cpu().numpy()[:, 0]
        # planes with surviving k in order, their k1 coordinates, and the stage-1 group of each
        first_line = np.flatnonzero(np.concatenate([[True], np.diff(self.line_plane) != 0]))
        self.planes = [int(p) for p in self.line_plane[first_line]]
        self.plane_x1 = {p: float(x1[starts[l]]) for p, l in zip(self.planes, first_line)}
        self.plane_pos = {p: i for i, p in enumerate(self.planes)}
        self.cache = []  # [group, {plane: buffer}, buffers], most recent last, at most two

    def folded(self) -> bool:
        """The MFMA projection kernel's folded stage-2 sums apply: the Re box spans -h .. h along axis 2
        (odd np) and tuner kind 8 is at its default."""
        from . import tune as T
        return (self.mfma and self.np % 2 == 1 and self.np <= 7 and self.rmin[1] == -(self.np // 2)
                and T.get_variant(T.KIND_KQ_FOLD) == 0)

    def invalidate(self):
        """Forget which k1 planes the cached stage-1 buffers hold (the buffers are kept for reuse): the next
        batch recomputes every plane it needs."""
        for entry in self.cache:
            entry[0], entry[1] = None, {}

    def _plane_buffer(self, plane, epmat_el, stream):
        import ctypes as C
        torch = _torch()
        for entry in self.cache:
            if plane in entry[1]:
                return entry[1][plane]
        g = self.plane_pos[plane] // self.PLANES
        members = self.planes[g * self.PLANES:(g + 1) * self.PLANES]
        if len(self.cache) < 2:
            bufs = [torch.empty((self.np * self.np, self.nd), dtype=torch.complex64, device=epmat_el.device)
                    for _ in range(self.PLANES)]
        else:  # reuse the older group's buffers (stream order: its projections are already queued)
            bufs = self.cache.pop(0)[2]
        self.cache.append([g, dict(zip(members, bufs)), bufs])
        n = len(members)
        x1 = (C.c_double * n)(*[self.plane_x1[p] for p in members])
        out = (C.c_void_p * n)(*[L.ptr(b) for b in bufs[:n]])
        L.call("ephip_gridopt_stage1_planes", self.nd, self.n1, self.n2, self.n3, self.np, self.rmin[0],
               L.ptr(self.rmap), L.ptr(epmat_el), epmat_el.stride(0), n, C.addressof(x1), self._nb,
               C.addressof(out), stream)
        return self.cache[-1][1][plane]

    def _tables(self, loop, dev):
        """Per-batch tables of the MFMA form (U column of each unit, stage-2/3 phases), allocated once."""
        torch = _torch()
        if getattr(self, "ucol", None) is None:
            self.ucol = torch.empty((loop.max_units, loop.st.nw), dtype=torch.complex64, device=dev)
            self.ph3 = torch.empty((loop.nk_batch, self.np), dtype=torch.complex64, device=dev)
            self.ph2 = torch.empty((loop.nk_batch, self.np), dtype=torch.complex64, device=dev)  # lines <= k
        return self.ucol, self.ph2, self.ph3

    def run(self, loop, i0, n, u0, u1, A, stream):
        st = loop.st
        l_first, l_last = int(self.line_of_k[i0]), int(self.line_of_k[i0 + n - 1]) + 1
        if self.mfma:
            ucol, ph2, ph3 = self._tables(loop, A.device)
            L.call("ephip_gridopt_kR_tables", st.nw, self.np, self.rmin[1], self.rmin[2], L.ptr(loop.xk_surv), i0,
                   i0 + n, l_first, l_last, L.ptr(self.line_start), u1 - u0, L.ptr(loop.unit_state[u0:u1]),
                   L.ptr(loop.unit_band[u0:u1]), L.ptr(st.el.U), L.ptr(ucol), L.ptr(ph2), L.ptr(ph3), stream)
        l = l_first
        while l < l_last:
            plane = self.line_plane[l]
            l_end = l
            while l_end < l_last and self.line_plane[l_end] == plane:
                l_end += 1
            buf = self._plane_buffer(int(plane), loop.epmat_el, stream)
            if self.mfma:
                L.call("ephip_gridopt_kR_project_mfma", st.nw, st.nm, loop.nrp, self.np, self.rmin[1], L.ptr(buf),
                       i0, i0 + n, l,
                       l_end, l_first, L.ptr(self.line_start), self.max_line_k, self.max_line_units,
                       L.ptr(loop.k_unit_start_dev),
                       L.ptr(ucol), L.ptr(ph2), L.ptr(ph3), loop.max_units_per_k, L.ptr(A), A.stride(0), stream)
            else:
                L.call("ephip_gridopt_kR_project", st.nw, st.nm, loop.nrp, self.np, self.rmin[1], self.rmin[2],
                       L.ptr(buf), L.ptr(loop.xk_surv), i0, i0 + n, l, l_end, L.ptr(self.line_start),
                       self.max_line_k, L.ptr(loop.k_unit_start_dev), L.ptr(loop.unit_state[u0:u1]),
                       L.ptr(loop.unit_band[u0:u1]), L.ptr(st.el.U), loop.max_units_per_k, L.ptr(A), A.stride(0),
                       stream)
            l = l_end


class _GridOptKQ:
    """Staged kR->kq over Rp for the uniform q grid (ephip_kq_line_offsets + ephip_gridopt_kq_rotate):
    the unit operators' Rp set embedded in its bounding box, the q of each unit's kept list grouped
    into (q1, q2) lines; writes g2 in the layout of the fused Fourier/rotation kernel."""

    prepass_default = False  # where stage 1 runs in new instances (in each workgroup, or by the pre-pass)

    @classmethod
    def build(cls, st):
        r = np.asarray(st.epmat.irvec_dev().cpu().numpy(), dtype=np.int64)
        lo, hi = r.min(0), r.max(0)
        n = hi - lo + 1
        m = np.asarray(st.qgrid, dtype=np.int64)
        if n[0] > 32 or n[1] > 8 or n[2] > 8 or m[1] > 64 or m[2] > 64 or m[0] > 65535:
            return None
        return cls(st, r, lo, n, m)

    def __init__(self, st, r, lo, n, m):
        torch = _torch()
        dev = st.el.xk.device
        self.n1 = int(n[0])
        self.np = max(int(n[1]), int(n[2]), 3)
        self.pmin = [int(x) for x in lo]
        rmap = -np.ones((self.np, self.n1, 8), np.int32)  # index b + 8 (a + n1 c)
        rmap[r[:, 2] - lo[2], r[:, 0] - lo[0], r[:, 1] - lo[1]] = np.arange(len(r), dtype=np.int32)
        self.rmap = torch.from_numpy(rmap.ravel()).to(dev)
        import ctypes
        # dense box (n1 = np, every Rp present, index affine in (a, b, c)): the kernel skips the index map
        self.dense = None
        box = rmap[:, :, :self.np]  # [c, a, b]
        if self.n1 == self.np and (box >= 0).all() and self.np > 1:
            s1, s2, s3 = int(box[0, 1, 0] - box[0, 0, 0]), int(box[0, 0, 1] - box[0, 0, 0]), int(box[1, 0, 0] - box[0, 0, 0])
            c, a, b = np.meshgrid(*(np.arange(self.np),) * 3, indexing="ij")
            if box[0, 0, 0] == 0 and min(s1, s2, s3) >= 0 and np.array_equal(box, s1 * a + s2 * b + s3 * c):
                self.dense = (ctypes.c_int32 * 3)(s1, s2, s3)
        self.m = [int(x) for x in m]
        self.nlines = self.m[0] * self.m[1]
        self.prepass = _GridOptKQ.prepass_default  # stage 1 by the pre-pass (OuterKLoop.tune picks per workload)
        import ctypes
        self.qgrid_c = (ctypes.c_int32 * 3)(*self.m)

    def folded(self) -> bool:
        """The kernel's folded stage-2/3 sums apply: the Rp box spans -h .. h along axes 2 and 3 (odd np)
        and tuner kind 8 is at its default."""
        from . import tune as T
        h = self.np // 2
        return (self.np % 2 == 1 and self.pmin[1] == -h and self.pmin[2] == -h
                and T.get_variant(T.KIND_KQ_FOLD) == 0)

    def can_prepass(self) -> bool:
        """The stage-1 pre-pass applies: a cube box with odd np <= 7 (n1 = np) and m1 <= 64."""
        return self.n1 == self.np and self.np % 2 == 1 and self.np <= 7 and self.m[0] <= 64

    def prepass_folded(self) -> bool:
        return self.pmin[0] == -(self.np // 2)

    def run(self, loop, nu, A, stream):
        import ctypes as C
        torch = _torch()
        st = loop.st
        b1 = None
        if self.prepass and self.can_prepass():  # stage 1 of every (unit, q1) by the pre-pass (workspace b1)
            need = loop.max_units * self.m[0] * self.np * self.np * st.nw * st.nm
            if getattr(self, "b1", None) is None or self.b1.numel() < need:
                self.b1 = torch.empty(need, dtype=torch.complex64, device=A.device)
            b1 = self.b1
        L.call("ephip_kq_line_offsets", nu, L.ptr(loop.count), L.ptr(loop.col_index), loop.cs, self.m[2],
               self.nlines, L.ptr(loop.line_off), stream)
        L.call("ephip_gridopt_kq_rotate", st.nw, st.nm, st.nbw, nu, self.n1, self.np, self.pmin[0], self.pmin[1],
               self.pmin[2], L.ptr(self.rmap), C.addressof(self.qgrid_c), L.ptr(A), A.stride(0),
               st.nw * st.nm, L.ptr(loop.line_off), L.ptr(loop.col_index), L.ptr(loop.ikq), loop.cs,
               L.ptr(st.el.U), L.ptr(st.el.filt.band_min), L.ptr(st.el.filt.band_max), L.ptr(loop.uph),
               L.ptr(loop.omega), C.addressof(self.dense) if self.dense is not None else None, L.ptr(loop.g2),
               L.ptr(loop.ep) if loop.ep is not None else None, L.ptr(b1) if b1 is not None else None, stream)


def kept_q_bound(ksurv, kgrid, qgrid) -> int:
    """Largest number of kept q of any surviving k: k + q runs over the coset of k in the k grid modulo
    the q grid (the q grid divides the k grid, both unshifted, third index fastest), so k keeps exactly
    the surviving k of its coset. Sizes the fused kR->kq grid (a hint: its workgroups loop over tiles)."""
    n = np.asarray(kgrid, dtype=np.int64)
    m = np.asarray(qgrid, dtype=np.int64)
    ks = np.asarray(ksurv, dtype=np.int64)
    if len(ks) == 0:
        return 1
    if np.any(n % m):
        return int(np.prod(m))
    r = n // m
    i1, rest = ks // (n[1] * n[2]), ks % (n[1] * n[2])
    i2, i3 = rest // n[2], rest % n[2]
    coset = ((i1 % r[0]) * r[1] + i2 % r[1]) * r[2] + i3 % r[2]
    return int(np.bincount(coset, minlength=int(np.prod(r))).max())


class UnitRecorder(AbstractCalculator):
    """Copies every outer-k payload to the host (test helper)."""

    def __init__(self):
        self.batches = []

    def run(self, p: EPDataUnits):
        cnt = p.count.cpu().numpy()
        ci = p.col_index.cpu().numpy()
        kq = p.ikq.cpu().numpy()
        g2 = p.g2.cpu().numpy()
        self.batches.append(dict(unit_state=p.unit_state.cpu().numpy(), unit_band=p.unit_band.cpu().numpy(),
                                 cols=[ci[u, :cnt[u]] for u in range(len(cnt))],
                                 ikq=[kq[u, :cnt[u]] for u in range(len(cnt))],
                                 g2=[g2[u, :cnt[u]] for u in range(len(cnt))]))
