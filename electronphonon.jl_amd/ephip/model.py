"""Seeded synthetic tight-binding Wannier models (host side, numpy).

A model is the set of real-space operators the Wannier->Bloch path consumes, stored in the
layouts of the reference's ``WannierObject`` (``src/wannier/WannierObject.jl:22-38``):

* ``op_r`` is ``(ndata, nr)`` column-major (numpy: shape ``(nr, ndata)`` C-order, i.e. each
  R-vector owns one contiguous column of ``ndata`` values);
* ``irvec`` is ``(nr, 3)`` int, sorted by reversed R (R1 fastest, R3 slowest), which is the order
  ``check_wannierobject`` enforces (``src/wannier/WannierObject.jl:14``);
* the e-ph tensor uses the "ph" outer-momentum layout ``(iw, jw, nu, Re) x Rp``
  (``src/model_new.jl:551-565``): ``ndata = nw*nw*nmodes*nr_el`` rows, one column per Rp.

Energies are in eV, masses are 1 and the cell is cubic with a = 1, so every k-point is in
crystal coordinates. The model definition follows the workload description: H(R) blocks are
complex Gaussian matrices scaled by exp(-|R|/1.0) with H(-R) = H(R)^dagger and a uniform on-site
diagonal; D(R) is positive semi-definite and obeys the acoustic sum rule; g(Re, Rp) is complex
Gaussian scaled by exp(-(|Re|+|Rp|)/0.8).

Random numbers come from numpy's PCG64 seeded through ``SeedSequence`` with a per-column key, so
any column of g can be regenerated on its own and the tensor can be built in chunks.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Scale of the dynamical matrix (eV^2 at unit mass). Chosen so phonon energies sit in the
# 10-100 meV range, where an electron window of a few tenths of an eV and room-temperature
# occupations behave like a real metal.
PH_SCALE_EV2 = 2.5e-4


def rbox(n: int) -> np.ndarray:
    """All integer R in the cube [-(n//2), n//2]^3, sorted by reversed R. Shape (n**3, 3)."""
    if n % 2 != 1:
        raise ValueError("R-box edge must be odd")
    h = n // 2
    r = np.arange(-h, h + 1)
    r3, r2, r1 = np.meshgrid(r, r, r, indexing="ij")
    return np.stack([r1.ravel(), r2.ravel(), r3.ravel()], axis=1).astype(np.int32)


def _rng(seed: int, *key: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([int(seed), *map(int, key)])))


def _cgauss(g: np.random.Generator, shape, dtype=np.float64) -> np.ndarray:
    re = g.standard_normal(shape, dtype=dtype)
    im = g.standard_normal(shape, dtype=dtype)
    z = np.empty(shape, dtype=np.complex128 if dtype == np.float64 else np.complex64)
    z.real = re
    z.imag = im
    z *= np.asarray(1.0 / math.sqrt(2.0), dtype=dtype)
    return z


def hermitian_operator(seed: int, n: int, irvec: np.ndarray, decay: float,
                       onsite: tuple[float, float] | None, key: int) -> np.ndarray:
    """Complex (nr, n*n) array: column ir holds the n x n block of R = irvec[ir], flattened
    column-major, with block(-R) = block(R)^dagger (so the Fourier sum is Hermitian)."""
    nr = irvec.shape[0]
    out = np.zeros((nr, n * n), dtype=np.complex128)
    g = _rng(seed, key)
    for i in range(nr):
        j = nr - 1 - i  # -R sits at the mirrored position of a symmetric box
        assert np.all(irvec[j] == -irvec[i])
        if i > j:
            continue
        a = _cgauss(g, (n, n)) * math.exp(-float(np.linalg.norm(irvec[i])) / decay)
        if i == j:
            a = 0.5 * (a + a.conj().T)
            if onsite is not None:
                np.fill_diagonal(a, g.uniform(onsite[0], onsite[1], size=n))
        out[i] = a.ravel(order="F")
        out[j] = a.conj().T.ravel(order="F")
    return out


def dynamical_matrix(seed: int, natoms: int, irvec_d: np.ndarray, decay: float = 1.0,
                     scale: float = PH_SCALE_EV2) -> np.ndarray:
    """Real (nr, nm*nm) D(R) on the box ``irvec_d`` with D(q) = B(q)^T B(q) positive
    semi-definite for every q and the acoustic sum rule sum_R sum_j D(R)[i a, j b] = 0.

    B(R) lives on the half-width box; D(R) = scale * sum_R1 B(R1)^T B(R1 + R)."""
    nm = 3 * natoms
    nd = round(irvec_d.shape[0] ** (1 / 3))
    nb = nd // 2 + 1  # B box edge: autocorrelation doubles the half-width
    if nb % 2 == 0:
        nb += 1
    irvec_b = rbox(nb)
    g = _rng(seed, 2)
    b = np.stack([g.standard_normal((nm, nm)) * math.exp(-float(np.linalg.norm(r)) / decay)
                  for r in irvec_b])
    # Acoustic sum rule on B: a uniform translation t (same on every atom) is annihilated,
    # i.e. sum_R sum_j B(R)[:, 3j+beta] = 0 for each direction beta. Absorb the total into
    # atom 0 of the on-site block.
    i0 = irvec_b.shape[0] // 2
    for beta in range(3):
        tot = b[:, :, beta::3].sum(axis=(0, 2))
        b[i0][:, beta] -= tot
    pos = {tuple(r): i for i, r in enumerate(irvec_b)}
    out = np.zeros((irvec_d.shape[0], nm * nm))
    for i, r in enumerate(irvec_d):
        acc = np.zeros((nm, nm))
        for i1, r1 in enumerate(irvec_b):
            j1 = pos.get((r1[0] + r[0], r1[1] + r[1], r1[2] + r[2]))
            if j1 is not None:
                acc += b[i1].T @ b[j1]
        out[i] = (scale * acc).ravel(order="F")
    return out


def eph_column(seed: int, nw: int, nm: int, irvec_el: np.ndarray, rp: np.ndarray, ip: int,
               decay: float, dtype) -> np.ndarray:
    """One column (fixed Rp) of the "ph" layout e-ph tensor: rows (iw, jw, nu, Re), iw fastest."""
    nr_el = irvec_el.shape[0]
    real = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    g = _rng(seed, 3, ip)
    col = _cgauss(g, (nr_el, nm, nw * nw), dtype=real)
    w = np.exp(-(np.linalg.norm(irvec_el, axis=1) + float(np.linalg.norm(rp))) / decay).astype(real)
    col *= w[:, None, None]
    return col.reshape(-1)


@dataclasses.dataclass
class Model:
    """Real-space operators of one synthetic model (see the module docstring for layouts)."""
    nw: int
    natoms: int
    irvec_h: np.ndarray          # (nr_h, 3) int32
    ham_r: np.ndarray            # (nr_h, nw*nw) complex128
    irvec_d: np.ndarray          # (nr_d, 3)
    dyn_r: np.ndarray            # (nr_d, nm*nm) complex128
    mass: np.ndarray             # (nm,) float64, one entry per mode
    irvec_el: np.ndarray         # (nr_el, 3) electron R of the e-ph tensor
    irvec_ep: np.ndarray         # (nr_ep, 3) phonon R of the e-ph tensor
    epmat_ph: np.ndarray | None  # (nr_ep, nw*nw*nm*nr_el) "ph" layout, complex64/complex128
    seed: int
    g_decay: float

    @property
    def nmodes(self) -> int:
        return 3 * self.natoms

    def epmat_col(self, ip: int, dtype=None) -> np.ndarray:
        if self.epmat_ph is not None:
            return self.epmat_ph[ip]
        return eph_column(self.seed, self.nw, self.nmodes, self.irvec_el, self.irvec_ep[ip], ip,
                          self.g_decay, dtype or np.complex64)


# The planted splitting f(k) = sin(2 pi 2 k2) sin(2 pi 3 k3) vanishes where 4 k2 or 6 k3 is an integer
# (k2 in {0, 1/4, 1/2, 3/4}, k3 in {0, 1/6, ..., 5/6}): 1 - (44/48)(42/48) = 19.8% of a 48^3 grid.
PLANT_SPLIT_R = (2, 3)  # the splitting's harmonics along k2 and k3


def planted_degenerate(xk: np.ndarray) -> np.ndarray:
    """Mask of the k (crystal coordinates, (nk, 3)) where the planted pairs stay exactly degenerate."""
    a, b = PLANT_SPLIT_R
    return (np.abs(np.sin(2 * np.pi * a * xk[:, 1])) < 1e-9) | (np.abs(np.sin(2 * np.pi * b * xk[:, 2])) < 1e-9)


def plant_pair_degeneracy(ham: np.ndarray, irvec: np.ndarray, nw: int, seed: int, split: float = 0.5) -> np.ndarray:
    """Two-fold degeneracies on a chosen fraction of the k-points. With J = i sigma_y (x) 1 on the two
    halves of the orbitals (real, J^2 = -1), H(R) -> (H(R) + J H(R)^T J^-1) / 2 makes the antiunitary
    J K commute with every H(k), so every band is exactly two-fold degenerate at every k; Hermiticity,
    and H(-R) = H(R)* when H(R) is real, are kept. A splitting f(k) S is then added, S = (S0 - J S0 J^-1)/2
    for a random real symmetric S0 (so J S^T J^-1 = -S lifts the pairs) and f(k) = sin(2 pi 2 k2)
    sin(2 pi 3 k3), a real Fourier series on R = (0, +-2, +-3) with coefficients -s2 s3 / 4: the pairs
    stay exactly degenerate where f = 0 (``planted_degenerate``) and split elsewhere, with near-
    degenerate clusters next to those planes. Needs |R2| >= 2, |R3| >= 3 in the box."""
    if nw % 2:
        raise ValueError("planted degeneracies need an even number of Wannier functions")
    n = nw // 2
    J = np.zeros((nw, nw))
    J[:n, n:] = np.eye(n)
    J[n:, :n] = -np.eye(n)          # i sigma_y (x) 1 = [[0, 1], [-1, 0]] (x) 1
    Jinv = -J
    index = {tuple(r): i for i, r in enumerate(irvec.tolist())}
    out = np.empty_like(ham)
    for ir in range(irvec.shape[0]):
        h = ham[ir].reshape(nw, nw, order="F")
        out[ir] = (0.5 * (h + J @ h.T @ Jinv)).ravel(order="F")
    rng = np.random.default_rng(seed + 7919)
    s0 = rng.standard_normal((nw, nw))
    s0 = 0.5 * (s0 + s0.T)
    S = 0.5 * (s0 - J @ s0 @ Jinv)
    S *= split / np.abs(np.linalg.eigvalsh(S)).max()
    a, b = PLANT_SPLIT_R
    for s2 in (1, -1):
        for s3 in (1, -1):
            ir = index.get((0, a * s2, b * s3))
            if ir is None:
                raise ValueError("the R box does not hold the planted splitting's harmonics")
            out[ir] += (-0.25 * s2 * s3 * S).astype(out.dtype).ravel(order="F")
    return out


def hermitian_eph(ep: np.ndarray, irvec: np.ndarray, nw: int, nm: int) -> np.ndarray:
    """Symmetrise a "ph"-layout e-ph tensor (Re and Rp on the same box ``irvec``) so that the vertex is
    the matrix element of a Hermitian perturbation: g_ij,nu(Re, Rp) = conj(g_ji,nu(-Re, Rp - Re)),
    which makes g_mn,nu(k, q) = conj(g_nm,nu(k + q, -q)) in the Wannier gauge. Pairs whose partner
    Rp - Re falls outside the box are set to zero. Returns a new array of ep's dtype."""
    import concurrent.futures as cf
    nr = irvec.shape[0]
    index = {tuple(r): i for i, r in enumerate(irvec.tolist())}
    neg = np.array([index[tuple(-x for x in r)] for r in irvec.tolist()])
    # column layout (numpy C order of the Julia column (i, j, nu, Re)): (Re, nu, j, i)
    g = ep.reshape(nr, nr, nm, nw, nw)
    out = np.zeros_like(g)

    def column(ip):  # writes out[ip] only (numpy releases the GIL on these block copies)
        part = [index.get(tuple(irvec[ip] - irvec[ie])) for ie in range(nr)]
        ok = np.array([p is not None for p in part])
        ies = np.flatnonzero(ok)
        ips = np.array([part[ie] for ie in ies], dtype=np.int64)
        partner = g[ips, neg[ies]]                      # (n, nm, j', i') at (-Re, Rp - Re)
        out[ip, ies] = 0.5 * (g[ip, ies] + np.conj(np.swapaxes(partner, -1, -2)))

    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        list(ex.map(column, range(nr)))
    return out.reshape(ep.shape)


def make_model(nw: int, natoms: int, box_hd: int, box_g: int, seed: int,
               onsite=(-2.0, 2.0), g_dtype=np.complex128, with_g: bool = True,
               h_decay: float = 1.0, g_decay: float = 0.8, hermitian_g: bool = False,
               time_reversal: bool = False, planted_degeneracy: bool = False) -> Model:
    """Build a model. ``box_hd`` is the R-box edge of H and D, ``box_g`` that of both g boxes.
    ``hermitian_g`` symmetrises the e-ph tensor (``hermitian_eph``), which balances scattering in and
    out in the linearised BTE. ``time_reversal`` keeps the real parts of H(R) and of the symmetrised
    g (so E(k) = E(-k), v(-k) = -v(k)); the BTE right-hand side is then odd in k and the fixed-point
    iteration converges. ``planted_degeneracy`` makes every band exactly two-fold degenerate on the
    k-points of ``planted_degenerate`` (19.8% of a 48^3 grid; ``plant_pair_degeneracy``)."""
    hermitian_g = hermitian_g or time_reversal
    irvec_h = rbox(box_hd)
    ham = hermitian_operator(seed, nw, irvec_h, h_decay, onsite, key=1)
    if time_reversal:
        ham = ham.real.astype(ham.dtype)
    if planted_degeneracy:
        ham = plant_pair_degeneracy(ham, irvec_h, nw, seed)
    dyn = dynamical_matrix(seed, natoms, irvec_h).astype(np.complex128)
    nm = 3 * natoms
    irvec_g = rbox(box_g)
    ep = None
    if with_g:
        ep = np.empty((irvec_g.shape[0], nw * nw * nm * irvec_g.shape[0]), dtype=g_dtype)
        for ip in range(irvec_g.shape[0]):
            ep[ip] = eph_column(seed, nw, nm, irvec_g, irvec_g[ip], ip, g_decay, g_dtype)
        if hermitian_g:
            ep = hermitian_eph(ep, irvec_g, nw, nm)
        if time_reversal:
            ep = ep.real.astype(ep.dtype)
    return Model(nw=nw, natoms=natoms, irvec_h=irvec_h, ham_r=ham, irvec_d=irvec_h.copy(),
                 dyn_r=dyn, mass=np.ones(nm), irvec_el=irvec_g, irvec_ep=irvec_g.copy(),
                 epmat_ph=ep, seed=seed, g_decay=g_decay)


# Workloads. ``kgrid``/``qgrid`` are the fine electron grid and the phonon grid; the window is
# (E_F - w, E_F + w) in eV with E_F = 0.
CONFIGS = {
    "cfg1": dict(nw=8, natoms=2, box_hd=5, box_g=5, seed=1001, onsite=(-2.0, 2.0),
                 g_dtype=np.complex128, kgrid=(8, 8, 8), qgrid=(8, 8, 8), window=1.0),
    "cfg2": dict(nw=16, natoms=4, box_hd=9, box_g=7, seed=1002, onsite=(-2.0, 2.0),
                 g_dtype=np.complex64, kgrid=(48, 48, 48), qgrid=(24, 24, 24), window=0.3),
    # cfg2's model with exact two-fold degeneracies planted on 19.8% of the k-points (precision form)
    "cfg3": dict(nw=16, natoms=4, box_hd=9, box_g=7, seed=1002, onsite=(-2.0, 2.0), planted_degeneracy=True,
                 g_dtype=np.complex64, kgrid=(48, 48, 48), qgrid=(24, 24, 24), window=0.3),
    "cfg4": dict(nw=16, natoms=4, box_hd=9, box_g=7, seed=1004, onsite=(-3.0, 3.0),
                 g_dtype=np.complex64, kgrid=(64, 64, 64), qgrid=(32, 32, 32), window=0.3),
    # cfg4's model on a 96^3 fine grid (learned-surrogate form: the band-energy network prunes the grid)
    "cfg5": dict(nw=16, natoms=4, box_hd=9, box_g=7, seed=1004, onsite=(-3.0, 3.0),
                 g_dtype=np.complex64, kgrid=(96, 96, 96), qgrid=(32, 32, 32), window=0.3),
}


def build_epmat(m: Model, workers: int = 16, dtype=np.complex64) -> np.ndarray:
    """The "ph"-layout e-ph tensor of ``m`` (its seed and boxes), generated column by column in parallel
    threads (numpy releases the GIL); the same values as ``make_model(with_g=True)`` without symmetrisation."""
    import concurrent.futures as cf
    nd = m.nw * m.nw * m.nmodes * len(m.irvec_el)
    ep = np.empty((len(m.irvec_ep), nd), dtype=dtype)

    def col(ip):
        ep[ip] = eph_column(m.seed, m.nw, m.nmodes, m.irvec_el, m.irvec_ep[ip], ip, m.g_decay, dtype)

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(col, range(len(m.irvec_ep))))
    return ep


def make_config_model(name: str, with_g: bool = True, **overrides) -> tuple[Model, dict]:
    c = dict(CONFIGS[name])
    c.update(overrides)
    grids = {k: c.pop(k) for k in ("kgrid", "qgrid", "window")}
    return make_model(with_g=with_g, **c), grids


def kpoints_grid(ngrid, shift=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Regular grid in the reference's order, third index fastest
    (``kpoints_grid_range``, ``src/common/kpoints.jl:135-153``). Shape (n, 3) float64."""
    n1, n2, n3 = ngrid
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    x = np.stack([i.ravel() / n1, j.ravel() / n2, k.ravel() / n3], axis=1)
    return x + np.asarray(shift, dtype=np.float64)
