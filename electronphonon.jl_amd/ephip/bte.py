"""Electron Boltzmann transport on the device, fed by the outer-k e-ph loop.

Mirrors the reference's transport pipeline on a uniform grid without symmetry:

* ``band_states``: the in-window (k, band) states in k-major order (electron_states_to_BandStates,
  src/boltzmann/BTStates.jl:52-93), with per-state energies, bands, grid k and the k weight 1/nk;
* ``compute_velocities``: band-diagonal velocities Re(u^H dH/dk u) with dH/dk = sum_R i R_cart H(R)
  e^{2 pi i k.R} (wannier_object_multiply_R, src/wannier/WannierObject.jl:78-87, and
  get_el_velocity_diag_berry_connection!, src/wannier_to_bloch.jl:70-81), FP64 Fourier on the matrix
  cores and one kernel for the diagonal products;
* ``BoltzmannCalculator``: consumes every ``EPDataUnits`` payload and accumulates the scattering-out
  rates S_out (inverse SERTA lifetimes) and the scattering-in matrix S_in with the six occupation
  conventions of bte_scattering_increments (src/boltzmann/boltzmann_calculator.jl:51-99, :336-393);
  S_in is kept on the device in compressed rows (one row per kept (k state, q), k+q window slots as
  columns) instead of the reference's dense host matrix;
* ``bte_compute_mu``: chemical potential of a metal by bisection (bte_compute_μ!,
  src/boltzmann/electron_serta.jl:19-52, find_chemical_potential, src/common/fermi_energy.jl:62-69);
* ``solve_electron_bte``: SERTA occupations and conductivity and the fixed-point iteration of the
  linearised BTE (src/boltzmann/electron_lbte.jl:43-50, :116-190) with the scattering-in product as a
  device sparse product.

Units are the model's: energies and temperatures in eV (k_B T), lengths in units of the cubic cell
edge, so sigma is in e^2/hbar per cell edge with e = hbar = 1. The phonon-momentum weight is 1/nq,
the reference's per-final-state weight on a uniform k+q grid equal to the k grid.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib as L
from . import core
from .core import WannierObject

SMEARING_KIND = {"Gaussian": 1, "Lorentzian": 2}


def _torch():
    import torch
    return torch


@dataclasses.dataclass
class ElectronTransportParams:
    """Subset of ElectronTransportParams / ElectronOccupationParams (src/transport_electron.jl:28-40):
    temperatures, carrier numbers per cell relative to nband_valence filled bands (spin included),
    cell volume, spin degeneracy and chemical potentials (NaN until computed)."""
    Tlist: np.ndarray
    nlist: np.ndarray
    volume: float
    spin_degeneracy: int = 2
    nband_valence: int = 0
    mulist: np.ndarray | None = None

    def __post_init__(self):
        self.Tlist = np.atleast_1d(np.asarray(self.Tlist, dtype=np.float64))
        self.nlist = np.atleast_1d(np.asarray(self.nlist, dtype=np.float64))
        if self.mulist is None:
            self.mulist = np.full(len(self.Tlist), np.nan)
        self.mulist = np.atleast_1d(np.asarray(self.mulist, dtype=np.float64))
        if not (len(self.Tlist) == len(self.nlist) == len(self.mulist)):
            raise ValueError("Tlist, nlist and mulist must have one entry per temperature")
        if np.any(self.Tlist <= 0):
            raise ValueError("temperatures must be positive")


# ------------------------------------------------------------------------------------------------
# States and velocities
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class BandStates:
    """In-window electron states of a grid, k-major then band (device tensors).
    ``state_offset[k]`` is the first state of grid point k (nk + 1 entries); state s is band
    ``ibands[s]`` (0-based) of grid point ``ks[s]`` with energy ``es[s]``; every state has weight
    1/nk. ``nstates_base`` counts the electrons per cell (spin excluded) of the bands below the
    window."""
    n: int
    nk: int
    es: object
    ks: object
    ibands: object
    state_offset: object
    weight: float
    nstates_base: float
    vs: object | None = None


def band_states(el) -> BandStates:
    """States of an ``ElectronStates`` with a window filter (electron_states_to_BandStates)."""
    torch = _torch()
    f = el.filt
    if f is None:
        raise ValueError("band_states needs electron states with a window filter")
    nk = el.E.shape[0]
    nb = ((f.band_max - f.band_min + 1) * (f.band_min > 0)).to(torch.int64)
    off = torch.zeros(nk + 1, dtype=torch.int64, device=nb.device)
    off[1:] = torch.cumsum(nb, 0)
    n = int(off[-1])
    ks = torch.repeat_interleave(torch.arange(nk, device=nb.device), nb)
    ib = (torch.arange(n, device=nb.device) - off[ks] + f.band_min[ks] - 1).to(torch.int32)
    es = el.E[ks, ib.to(torch.int64)].contiguous()
    base = float(f.nbelow.clamp(min=0).sum()) / nk
    return BandStates(n, nk, es, ks, ib, off, 1.0 / nk, base)


def velocity_operator(irvec: np.ndarray, ham_r: np.ndarray, lattice=None) -> WannierObject:
    """Host operator i R_cart H(R) with the direction the slowest data index: ndata = 3 nw^2,
    entry [a * nw^2 + i + nw j] of column R (wannier_object_multiply_R)."""
    lat = np.eye(3) if lattice is None else np.asarray(lattice, dtype=np.float64)
    rc = irvec.astype(np.float64) @ lat.T          # R_cart[a] = sum_j lattice[a, j] R_j
    op = np.concatenate([1j * rc[:, a:a + 1] * ham_r for a in range(3)], axis=1)
    return WannierObject(irvec, np.ascontiguousarray(op, dtype=np.complex128))


def _check_velocity_dtypes(op_dev: WannierObject, el):
    """The velocity kernels read the Fourier-transformed operator as FP64 complex (double2) and the
    eigenvectors as complex64 (float2): a complex64 operator would be read past its end. Checked before
    any launch."""
    torch = _torch()
    if op_dev.op_r.dtype != torch.complex128:
        raise L.EphipError(f"the velocity/position operator must be complex128 on the device (got {op_dev.op_r.dtype}); "
                           "use to_device(backend, dtype=torch.complex128)")
    if el.U.dtype != torch.complex64:
        raise L.EphipError(f"the electron eigenvectors must be complex64 (got {el.U.dtype})")


def compute_velocities(vel_dev: WannierObject, el, bs: BandStates, chunk: int = 16384):
    """Band-diagonal velocities (bs.n, 3) float64 of every in-window state; fills bs.vs."""
    torch = _torch()
    nw = el.nw
    if vel_dev.ndata != 3 * nw * nw:
        raise L.EphipError(f"velocity operator has ndata={vel_dev.ndata}, expected 3*nw^2={3 * nw * nw}")
    _check_velocity_dtypes(vel_dev, el)
    dev = el.E.device
    v = torch.zeros((bs.n, 3), dtype=torch.float64, device=dev)
    if bs.n == 0:
        bs.vs = v
        return v
    nb = bs.state_offset[1:] - bs.state_offset[:-1]
    klist = torch.nonzero(nb, as_tuple=False).view(-1)
    off_h = bs.state_offset.cpu().numpy()
    kl_h = klist.cpu().numpy()
    for c0 in range(0, len(kl_h), chunk):
        c1 = min(len(kl_h), c0 + chunk)
        kc = klist[c0:c1]
        dH = core.fourier_batched(vel_dev, el.xk.index_select(0, kc).contiguous())
        s0, s1 = int(off_h[kl_h[c0]]), int(off_h[kl_h[c1 - 1] + 1])
        # column of each state: its k position within the chunk
        col = torch.repeat_interleave(torch.arange(c1 - c0, device=dev), nb[kc])
        L.call("ephip_band_velocity", nw, s1 - s0, L.ptr(col), L.ptr(bs.ibands[s0:s1]), L.ptr(bs.ks[s0:s1]),
               L.ptr(dH), dH.stride(0), L.ptr(el.U), L.ptr(v[s0:s1]), L.stream_handle())
    bs.vs = v
    return v


def el_velocity_direct_batched(op_dev: WannierObject, el, kidx, rbar=None):
    """U_k^H M_a(k) U_k for the k of ``kidx`` (indices into ``el``): complex128 (nk, 3, nw, nw) indexed
    [k, a, m, n], with M the FP64 Fourier of a 3 nw^2-data operator (``velocity_operator`` for the
    velocity, the position operator for rbar). With ``rbar`` (this function's output for the position
    operator, same k), the Berry-connection term i (e_m - e_n) rbar[m, n] is added
    (get_el_velocity_direct_batched, src/wannier_to_bloch_batched.jl:121-148; the Berry-connection
    update of _compute_electron_states_device!, src/compute_states.jl:149-168)."""
    torch = _torch()
    nw = el.nw
    if op_dev.ndata != 3 * nw * nw:
        raise L.EphipError(f"operator has ndata={op_dev.ndata}, expected 3*nw^2={3 * nw * nw}")
    _check_velocity_dtypes(op_dev, el)
    dev = el.E.device
    kidx = torch.as_tensor(kidx, dtype=torch.int64, device=dev).reshape(-1)
    nk = kidx.shape[0]
    out = torch.empty((nk, 3 * nw * nw), dtype=torch.complex128, device=dev)
    if nk == 0:
        return out.view(0, 3, nw, nw)
    M = core.fourier_batched(op_dev, el.xk.index_select(0, kidx).contiguous())
    U = el.U.index_select(0, kidx).contiguous()
    E = r = None
    if rbar is not None:
        if tuple(rbar.shape) != (nk, 3, nw, nw) or rbar.dtype != torch.complex128:
            raise L.EphipError(f"rbar must be complex128 ({nk}, 3, {nw}, {nw}), got {tuple(rbar.shape)} {rbar.dtype}")
        E = el.E.index_select(0, kidx).contiguous()
        r = rbar.transpose(-1, -2).contiguous().view(nk, 3 * nw * nw)  # back to column-major blocks
    L.call("ephip_el_velocity_direct_batched", nw, nk, L.ptr(M), M.stride(0), L.ptr(U), L.ptr(E), L.ptr(r),
           3 * nw * nw, L.ptr(out), out.stride(0), L.stream_handle())
    return out.view(nk, 3, nw, nw).transpose(-1, -2)


# ------------------------------------------------------------------------------------------------
# Chemical potential (host; a handful of bisection steps over the state energies)
# ------------------------------------------------------------------------------------------------
def _occ(x):
    return 0.5 * (1.0 - np.tanh(0.5 * x))   # 1 / (exp(x) + 1) without overflow


def compute_ncarrier(mu: float, T: float, es: np.ndarray, w: float) -> float:
    """Electrons per cell (spin excluded) in the states es of weight w (compute_ncarrier,
    src/common/fermi_energy.jl:36-38)."""
    return float(w * _occ((es - mu) / T).sum())


def find_chemical_potential(ncarrier: float, T: float, es: np.ndarray, w: float, tol: float = 1e-13) -> float:
    """Bisection for compute_ncarrier(mu) = ncarrier (find_chemical_potential)."""
    if not (0 < ncarrier < w * len(es)):
        raise ValueError(f"target carrier count {ncarrier} outside (0, {w * len(es)}) for the window states")
    lo, hi = float(es.min()) - 50 * T, float(es.max()) + 50 * T
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if compute_ncarrier(mid, T, es, w) > ncarrier:
            hi = mid
        else:
            lo = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)


def bte_compute_mu(params: ElectronTransportParams, bs: BandStates) -> np.ndarray:
    """Metal branch of bte_compute_μ!: target = n / spin + nband_valence - nstates_base."""
    es = bs.es.cpu().numpy()
    for i, T in enumerate(params.Tlist):
        target = params.nlist[i] / params.spin_degeneracy + params.nband_valence - bs.nstates_base
        params.mulist[i] = find_chemical_potential(target, T, es, bs.weight)
    return params.mulist


# ------------------------------------------------------------------------------------------------
# Scattering accumulation
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class SinBlock:
    """Compressed scattering-in rows: unit u (initial state unit_index[u]) owns rows
    rowptr[u]..rowptr[u+1]-1; row r holds nbw x nT values for the final states f0 .. f0 + nb - 1 (the
    k+q window slots) with row_code[r] = f0 | (nb << 27)."""
    unit_index: object
    rowptr: object
    row_code: object
    S_in: object
    nrows: int


class BoltzmannCalculator:
    """Accumulates S_out (n x nT) and compressed S_in on the device from outer-k payloads
    (BoltzmannCalculator, src/boltzmann/boltzmann_calculator.jl:51-99). ``smearing`` is
    ("Gaussian" | "Lorentzian", eta) for every temperature or a list of etas per temperature
    with one kind."""

    def __init__(self, params: ElectronTransportParams, smearing=("Gaussian", 0.1), occupation_method: int = 5,
                 omega_cutoff: float = core.OMEGA_ACOUSTIC_EV):
        if not 1 <= int(occupation_method) <= 6:
            raise ValueError(f"occupation_method must be an integer in 1:6, got {occupation_method}")
        kind, eta = smearing
        if kind not in SMEARING_KIND:
            raise ValueError(f"smearing kind must be one of {sorted(SMEARING_KIND)}, not {kind!r}")
        etas = np.broadcast_to(np.asarray(eta, dtype=np.float64), params.Tlist.shape).copy()
        if np.any(etas <= 0):
            raise ValueError("smearing widths must be positive")
        self.params = params
        self.kind = SMEARING_KIND[kind]
        self.etas = etas
        self.method = int(occupation_method)
        self.omega_cutoff = float(omega_cutoff)
        self.blocks: list[SinBlock] = []
        self.done = False
        self.bs = None

    def setup(self, loop):
        """Device buffers for an OuterKLoop (setup_calculator!): states, mu, S_out zeros."""
        torch = _torch()
        if self.done:
            raise ValueError("this BoltzmannCalculator has already been run; reconstruct it")
        st = loop.st
        self.loop = loop
        self.bs = band_states(st.el)
        if np.any(np.isnan(self.params.mulist)):
            bte_compute_mu(self.params, self.bs)
        dev = st.el.E.device
        nT = len(self.params.Tlist)
        self.nT = nT
        self.S_out = torch.zeros((nT, self.bs.n), dtype=torch.float64, device=dev)
        self.mus = torch.as_tensor(self.params.mulist, device=dev)
        self.Ts = torch.as_tensor(self.params.Tlist, device=dev)
        self.etas_d = torch.as_tensor(self.etas, device=dev)
        self.wtq = 1.0 / st.ph.xq.shape[0]
        return self

    def run(self, p):
        torch = _torch()
        st = self.loop.st
        bs = self.bs
        nu = p.count.shape[0]
        if nu == 0:
            return
        f = st.el.filt
        usi = (bs.state_offset[p.unit_state] + p.unit_band.to(torch.int64) - (f.band_min[p.unit_state] - 1))
        rowptr = torch.zeros(nu + 1, dtype=torch.int64, device=usi.device)
        rowptr[1:] = torch.cumsum(p.count, 0)
        nrows = int(rowptr[-1])
        S_in = torch.empty((max(nrows, 1), p.nbw, self.nT), dtype=torch.float64, device=usi.device)
        row_code = torch.empty(max(nrows, 1), dtype=torch.int32, device=usi.device)
        cs = p.col_index.stride(0)
        L.call("ephip_bte_accumulate_units", nu, cs, p.nbw, st.nm, L.ptr(p.count), L.ptr(p.col_index),
               L.ptr(p.ikq), L.ptr(p.g2), L.ptr(usi), L.ptr(bs.state_offset), L.ptr(f.band_min),
               L.ptr(f.band_max), L.ptr(bs.es), L.ptr(self.loop.omega), self.wtq, self.nT, L.ptr(self.mus),
               L.ptr(self.Ts), L.ptr(self.etas_d), self.method, self.kind, self.omega_cutoff, bs.n,
               L.ptr(rowptr), L.ptr(self.S_out), L.ptr(S_in), L.ptr(row_code), L.stream_handle())
        self.blocks.append(SinBlock(usi, rowptr, row_code, S_in[:nrows], nrows))

    def postprocess(self):
        """Merge the per-batch compressed rows into one block (one product launch over every state)."""
        torch = _torch()
        self.done = True
        if len(self.blocks) > 1:
            off, parts = 0, []
            for b in self.blocks:
                parts.append(b.rowptr[:-1] + off)
                off += b.nrows
            rowptr = torch.cat(parts + [torch.tensor([off], dtype=torch.int64, device=self.S_out.device)])
            merged = SinBlock(torch.cat([b.unit_index for b in self.blocks]), rowptr,
                              torch.cat([b.row_code[:b.nrows] for b in self.blocks]),
                              torch.cat([b.S_in for b in self.blocks]), off)
            self.blocks = [merged]
            torch.cuda.empty_cache()

    def begin_q(self, iq):
        pass

    def end_q(self, iq):
        pass

    # ---------------------------------------------------------------------------------------------
    def scattering_in_dense(self, iT: int = 0) -> np.ndarray:
        """Host dense S_in (n x n) for one temperature (tests and small grids)."""
        n = self.bs.n
        out = np.zeros((n, n))
        for b in self.blocks:
            ui = b.unit_index.cpu().numpy()
            rp = b.rowptr.cpu().numpy()
            code = b.row_code.cpu().numpy().view(np.uint32)
            rf = (code & 0x7FFFFFF).astype(np.int64)
            rn = (code >> 27).astype(np.int64)
            S = b.S_in.cpu().numpy()
            for u in range(len(ui)):
                for r in range(rp[u], rp[u + 1]):
                    out[ui[u], rf[r]:rf[r] + rn[r]] += S[r, :rn[r], iT]
        return out

    def spmv(self, iT: int, x, y):
        """y = S_in(iT) x for (n, 3) float64 device vectors over the in-window states."""
        for b in self.blocks:
            nu = b.unit_index.shape[0]
            L.call("ephip_bte_spmv", nu, b.S_in.shape[1], self.nT, iT, L.ptr(b.rowptr), L.ptr(b.row_code),
                   L.ptr(b.unit_index), L.ptr(b.S_in), L.ptr(x), L.ptr(y), L.stream_handle())
        return y


class BoltzmannQBatchedCalculator:
    """The reference's device BoltzmannCalculator on EPDataQBatched payloads
    (src/boltzmann/boltzmann_calculator.jl:213-386 with the CUDA method of bte_window_accumulate!,
    ext/ElectronPhononCUDAExt.jl:323-372): S_out (nT x n) accumulates on the device; S_in is a
    TiledDeviceOutput (n_i, n_f, nT) tiled over the outer states: begin_batch zeroes the batch's tile
    (calculator_begin!), every k payload accumulates into it through ephip_bte_window_accumulate with
    tile-local rows, end_batch streams the tile into the host dense S_in (calculator_end!) in block
    mode. The residency is decided on the first batch from the free device memory (force_block /
    free_bytes override it, as TiledDeviceOutput's); in full mode the whole S_in stays on the device and
    is downloaded once by postprocess. Feed it through eph.QBatchedAdapter. Host dense S_in: n^2 nT
    float64 (small grids, as in the reference)."""

    def __init__(self, params: ElectronTransportParams, smearing=("Gaussian", 0.1), occupation_method: int = 5,
                 omega_cutoff: float = core.OMEGA_ACOUSTIC_EV, force_block=None, free_bytes=None,
                 headroom: float = 1.2):
        kind, eta = smearing
        if kind not in SMEARING_KIND or not 1 <= int(occupation_method) <= 6:
            raise ValueError("smearing kind must be Gaussian or Lorentzian and occupation_method in 1:6")
        self.params = params
        self.kind = SMEARING_KIND[kind]
        self.etas = np.broadcast_to(np.asarray(eta, dtype=np.float64), params.Tlist.shape).copy()
        self.method = int(occupation_method)
        self.omega_cutoff = float(omega_cutoff)
        self.force_block, self.free_bytes, self.headroom = force_block, free_bytes, headroom

    def setup(self, loop):
        from .tiled import TiledDeviceOutput
        torch = _torch()
        st = loop.st
        self.loop = loop
        self.bs = bs = band_states(st.el)
        if np.any(np.isnan(self.params.mulist)):
            bte_compute_mu(self.params, bs)
        dev = st.el.E.device
        self.nT = nT = len(self.params.Tlist)
        n = bs.n
        self.S_out = torch.zeros((nT, n), dtype=torch.float64, device=dev)
        self.mus = torch.as_tensor(self.params.mulist, device=dev)
        self.Ts = torch.as_tensor(self.params.Tlist, device=dev)
        self.etas_d = torch.as_tensor(self.etas, device=dev)
        nq = st.ph.xq.shape[0]
        self.wf = torch.full((n,), 1.0 / nq, dtype=torch.float64, device=dev)   # per-final-state weight
        # state map of the filtered k list: imap[p, band] = 1-based state of band (0-based) at surviving k p
        ks = st.ksurv.cpu().numpy()
        off = bs.state_offset.cpu().numpy()
        bmin = st.el.filt.band_min.cpu().numpy()
        bmax = st.el.filt.band_max.cpu().numpy()
        imap = np.zeros((len(ks), st.nw), dtype=np.int64)   # (nks, nw) = column-major (nw, nks)
        for p_, k in enumerate(ks):
            nb = int(bmax[k] - bmin[k] + 1)
            imap[p_, bmin[k] - 1:bmax[k]] = off[k] + 1 + np.arange(nb)
        self.imap = torch.from_numpy(imap).to(dev)
        self.tiled = TiledDeviceOutput((n, n, nT), 0, (off, ks), dtype=torch.float64, device=dev,
                                       force_block=self.force_block, headroom=self.headroom,
                                       free_bytes=self.free_bytes)
        self.S_in = np.zeros((nT, n, n))                    # host dense, [iT, f, i] = reference Si[iT][i, f]
        return self

    def begin_batch(self, k_range, n_batch_max):
        self.tiled.tile_begin(int(k_range[0]), int(k_range[1]) - 1, int(n_batch_max))

    def run(self, p):
        from . import core as C_
        t = self.tiled
        nbk = p.g2s.shape[2]
        imap_i = self.imap[p.ik - 1, p.ibandk_offset:p.ibandk_offset + nbk].contiguous()
        # FP64 increments from the payload's g2 and its FP64 frequencies (the entry takes both in one type)
        g2 = p.g2s.to(self.S_out.dtype)
        C_.bte_window_accumulate(self.S_out, t.device_array(0), g2, p.wqs.contiguous(), imap_i, self.imap, p.ikqs,
                                 self.bs.es, self.bs.es, self.wf, self.mus, self.Ts, self.etas_d, self.method,
                                 self.omega_cutoff, t.tile_offset(), smearing=self.kind, check=False)

    def end_batch(self):
        t = self.tiled
        if t.is_block() and t.tile_length() > 0:
            t.tile_download()
            i0, ni = t.tile_offset(), t.tile_length()
            self.S_in[:, :, i0:i0 + ni] = t.host_array(0)[:, :, :ni].numpy()

    def postprocess(self):
        t = self.tiled
        if t.is_allocated() and not t.is_block():
            self.S_in[:] = t.device_array(0).cpu().numpy()
        t.tile_free()

    def scattering_in_dense(self, iT: int = 0) -> np.ndarray:
        """Host dense S_in (n x n) [i, f] for one temperature."""
        return self.S_in[iT].T.copy()


def run_boltzmann(loop, calc: BoltzmannCalculator, k_start: int = 0, k_stop: int | None = None):
    """setup -> every outer-k batch -> postprocess (the run_eph_over_k_and_kq calculator lifecycle)."""
    calc.setup(loop)
    loop.run(k_start, k_stop, calculators=[calc])
    calc.postprocess()
    return calc


# ------------------------------------------------------------------------------------------------
# Conductivity and the iterative solve
# ------------------------------------------------------------------------------------------------
def _dfermi(e, T):
    """-df/de of the Fermi-Dirac occupation (device tensors)."""
    torch = _torch()
    return 1.0 / (2.0 + torch.exp(e / T) + torch.exp(-e / T)) / T


def occupation_to_conductivity(df, bs: BandStates, params: ElectronTransportParams):
    """sigma[a, b] = spin / V * sum_i w df[i, a] v[i, b] (src/boltzmann/electron_lbte.jl:43-50)."""
    return (df.T @ bs.vs) * (bs.weight * params.spin_degeneracy / params.volume)


@dataclasses.dataclass
class BTEResult:
    sigma_serta: np.ndarray   # (nT, 3, 3)
    sigma: np.ndarray         # (nT, 3, 3) iterative BTE
    niter: list
    converged: list
    sigma_iter: list          # per T, list of (3, 3) after each iteration (index 0 = SERTA)


def solve_electron_bte(calc: BoltzmannCalculator, max_iter: int = 100, rtol: float = 1e-10,
                       mixing: float = 1.0) -> BTEResult:
    """SERTA and iterative BTE conductivities (solve_electron_bte without symmetry):
    dF_serta = -df/de / S_out * v, then dF <- mixing ((S_in dF) / S_out + dF_serta) + (1 - mixing) dF
    until the conductivity changes by less than rtol (relative Frobenius norm)."""
    torch = _torch()
    bs, params = calc.bs, calc.params
    if bs.vs is None:
        raise ValueError("velocities are not set: call compute_velocities first")
    nT = len(params.Tlist)
    out_s = np.zeros((nT, 3, 3))
    out = np.zeros((nT, 3, 3))
    niters, conv, hist = [], [], []
    for iT in range(nT):
        mu, T = float(params.mulist[iT]), float(params.Tlist[iT])
        inv_tau = calc.S_out[iT]
        rhs = (_dfermi(bs.es - mu, T) / inv_tau).unsqueeze(1) * bs.vs
        s_serta = occupation_to_conductivity(rhs, bs, params)
        df = rhs.clone()
        tmp = torch.empty_like(df)
        obs = s_serta
        h = [s_serta.cpu().numpy()]
        it, ok = 0, False
        for it in range(1, max_iter + 1):
            calc.spmv(iT, df, tmp)
            new = tmp / inv_tau.unsqueeze(1) + rhs
            df = new * mixing + df * (1.0 - mixing)
            obs_old, obs = obs, occupation_to_conductivity(df, bs, params)
            h.append(obs.cpu().numpy())
            if float(torch.linalg.norm(obs - obs_old) / torch.linalg.norm(obs)) < rtol:
                ok = True
                break
        out_s[iT] = s_serta.cpu().numpy()
        out[iT] = obs.cpu().numpy()
        niters.append(it)
        conv.append(ok)
        hist.append(h)
    return BTEResult(out_s, out, niters, conv, hist)


def transport_mobility(sigma: np.ndarray, params: ElectronTransportParams) -> np.ndarray:
    """Mobility sigma / |carrier density| per temperature (transport_print_mobility, in model units:
    carrier density n / V per cell volume)."""
    dens = np.abs(params.nlist) / params.volume
    return sigma / dens[:, None, None]


@dataclasses.dataclass
class TransportRun:
    result: BTEResult
    mobility_serta: np.ndarray
    mobility: np.ndarray
    calc: BoltzmannCalculator
    setup: object
    seconds: dict


def run_transport(model, kgrid, qgrid, window, params: ElectronTransportParams, backend,
                  precision: int = L.PREC_FP32, smearing=("Gaussian", 0.01), occupation_method: int = 5,
                  nk_batch: int | None = 128, max_iter: int = 100, rtol: float = 1e-10, epmat_dev=None,
                  setup=None) -> TransportRun:
    """Transport on one GPU, uniform grids, no symmetry: fine-grid states and window filter, phonons,
    the outer-k e-ph loop feeding a BoltzmannCalculator, band velocities, then SERTA and the
    iterative BTE (the flow of run_eph_over_k_and_kq with a BoltzmannCalculator followed by
    solve_electron_bte, src/calculator/run_eph_over_k_and_kq.jl:584-895,
    src/boltzmann/electron_lbte.jl:153-190). ``setup`` reuses an OuterQSetup of the same grids."""
    import time
    from . import eph
    torch = _torch()
    t = {}
    t0 = time.perf_counter()
    st = setup if setup is not None else eph.setup_outer_q(model, kgrid, qgrid, window, backend,
                                                           epmat_dev=epmat_dev)
    core.synchronize(backend)
    t["setup"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    loop = eph.OuterKLoop(st, nk_batch=nk_batch, precision=precision)
    calc = BoltzmannCalculator(params, smearing=smearing, occupation_method=occupation_method)
    run_boltzmann(loop, calc)
    core.synchronize(backend)
    t["eph_and_scattering"] = time.perf_counter() - t0
    del loop
    t0 = time.perf_counter()
    vel = velocity_operator(model.irvec_h, model.ham_r).to_device(backend, dtype=torch.complex128)
    compute_velocities(vel, st.el, calc.bs)
    core.synchronize(backend)
    t["velocities"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = solve_electron_bte(calc, max_iter=max_iter, rtol=rtol)
    t["solve"] = time.perf_counter() - t0
    return TransportRun(res, transport_mobility(res.sigma_serta, params), transport_mobility(res.sigma, params),
                        calc, st, t)
