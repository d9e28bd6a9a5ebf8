"""ctypes wrapper over ``libephor.so``, the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this module.
The functions mirror the reference's per-k CPU routines; each docstring names the C routine, and
the C file cites the reference source it restates.

Layout conversion: the C side is column-major. A numpy array of shape ``(ncol, nrow)`` in C order
is the same memory as an ``(nrow, ncol)`` column-major matrix, so every wrapper takes and returns
arrays in that "transposed" numpy shape (e.g. op_r is ``(nr, ndata)``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64, _i32, _d = C.c_int64, C.c_int, C.c_double


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libephor.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libephor.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "ephor_fourier_phase": [_i32, _ip, _i32, _dp, _zp],
            "ephor_fourier": [_i64, _i32, _zp, _i64, _ip, _i32, _dp, _zp, _i64],
            "ephor_heev": [_i32, _zp, _i32, _dp, C.c_void_p, _i32],
            "ephor_heev_batched": [_i32, _i64, _zp, _dp, C.c_void_p],
            "ephor_degenerate_gauge_fix": [_i32, _i64, _dp, _zp, _d],
            "ephor_degenerate_groups": [_i32, _dp, _d, C.c_void_p],
            "ephor_ph_eigen": [_i32, _zp, _dp, _dp, _zp],
            "ephor_window_filter": [_i32, _i64, _dp, _d, _d, _bp, _lp, _lp, _lp],
            "ephor_kpoints_grid": [_i32, _i32, _i32, _dp, _dp],
            "ephor_hash_xk": [_i64, _dp, _ip, _dp, _lp],
            "ephor_eph_RR_to_Rq": [_i32, _i32, _i32, _i32, _zp, _ip, _dp, _zp, _zp],
            "ephor_eph_Rq_to_kq": [_i32, _i32, _i32, _zp, _ip, _dp, _zp, _i32, _zp, _i32, _zp],
            "ephor_eph_RR_to_kR": [_i32, _i32, _i32, _i32, _zp, _ip, _dp, _zp, _i32, _zp],
            "ephor_eph_kR_to_kq": [_i32, _i32, _i32, _i32, _zp, _ip, _dp, _zp, _zp, _i32, _zp],
            "ephor_g2": [_i64, _i32, _zp, _dp, _dp],
            "ephor_g2_degenerate_average": [_i32, _i32, _i32, _dp, _dp, _dp, _d],
            "ephor_outer_q_block": [_i32, _i32, _i32, _zp, _ip, _dp, _zp, _lp, _lp, _i64, _lp, _lp,
                                    _i32, _zp],
            "ephor_bte_accumulate_q": [_i32, _i32, _i64, _lp, _lp, _dp, _lp, _lp, _lp, _dp, _dp, _d, _i32, _dp,
                                       _dp, _dp, _i32, _i32, _d, _i64, _dp, _dp],
            "ephor_bte_increments": [_i32, _d, _d, _d, _d, _d, _d, _d, _i32, _d, _dp],
            "ephor_num_threads": [],
            "ephor_gridopt_new": [_i32, _ip, _i64, _i32],
            "ephor_gridopt_free": [C.c_void_p],
            "ephor_gridopt_reset": [C.c_void_p],
            "ephor_gridopt_get": [C.c_void_p, _zp, _i64, _dp, _zp],
            "ephor_gridopt_nr23": [C.c_void_p],
            "ephor_gridopt_nr3": [C.c_void_p],
            "ephor_outer_q_gridopt": [_i32, _i32, _i32, _i32, _zp, _ip, _ip, _dp, _zp, _dp, _i32, _lp, _dp, _ip,
                                      _zp, _lp, _lp, _i64, _lp, _i32, _dp, _i64, _lp, C.c_void_p],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int if name in ("ephor_heev", "ephor_heev_batched", "ephor_degenerate_gauge_fix",
                                            "ephor_degenerate_groups", "ephor_ph_eigen",
                                            "ephor_num_threads", "ephor_gridopt_nr23", "ephor_gridopt_nr3") else None
        L.ephor_gridopt_new.restype = C.c_void_p
        L.ephor_outer_q_gridopt.restype = C.c_int64
        _LIB = L
    return _LIB


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def fourier_phase(irvec, xk) -> np.ndarray:
    """(nk, nr) complex: [j, r] = exp(2 pi i R_r . x_j)."""
    irvec, xk = _c(irvec, np.int32), _c(xk, np.float64)
    out = np.empty((xk.shape[0], irvec.shape[0]), np.complex128)
    lib().ephor_fourier_phase(irvec.shape[0], irvec, xk.shape[0], xk, out)
    return out


def fourier(op_r, irvec, xk, ndata=None) -> np.ndarray:
    """op_r (nr, ld) -> (nk, ndata): out[j] = sum_r op_r[r, :ndata] exp(2 pi i R_r.x_j)."""
    op_r = _c(op_r, np.complex128)
    irvec, xk = _c(irvec, np.int32), _c(np.atleast_2d(xk), np.float64)
    ld = op_r.shape[1]
    nd = ld if ndata is None else ndata
    out = np.empty((xk.shape[0], nd), np.complex128)
    lib().ephor_fourier(nd, irvec.shape[0], op_r, ld, irvec, xk.shape[0], xk, out, nd)
    return out


def heev(h, vectors=True):
    """Hermitian eigen of (n, n) numpy matrix h (upper triangle read). Returns (w, V) with V
    a normal (row, col) numpy matrix whose columns are eigenvectors."""
    n = h.shape[0]
    a = _c(np.asarray(h).T, np.complex128)  # column-major copy
    w = np.empty(n)
    v = np.empty((n, n), np.complex128)
    st = lib().ephor_heev(n, a, n, w, v.ctypes.data if vectors else None, n)
    if st != 0:
        raise RuntimeError("ephor_heev did not converge")
    return (w, v.T.copy()) if vectors else w


def heev_batched(hs, vectors=True, degen_cutoff=None):
    """hs: (nb, n*n) column-major blocks (Fourier output layout). Returns E (nb, n) and
    U (nb, n*n) column-major eigenvector blocks. degen_cutoff: apply the per-k path's degenerate-band
    gauge fixing (degenerate_gauge_fix) to the eigenvectors."""
    hs = np.array(hs, dtype=np.complex128, order="C", copy=True)
    nb = hs.shape[0]
    n = int(round(np.sqrt(hs.shape[1])))
    w = np.empty((nb, n))
    v = np.empty((nb, n * n), np.complex128) if vectors else None
    st = lib().ephor_heev_batched(n, nb, hs, w, v.ctypes.data if vectors else None)
    if st != 0:
        raise RuntimeError("ephor_heev_batched did not converge")
    if vectors and degen_cutoff is not None:
        degenerate_gauge_fix(w, v, degen_cutoff)
    return (w, v) if vectors else w


def degenerate_groups(w, cutoff):
    """First band of each band's degenerate group, for ascending eigenvalues w (n,), grouped as
    solve_eigen_el! groups them (src/diagonalize.jl:58-84)."""
    w = _c(w, np.float64)
    g = np.empty(len(w), np.int32)
    lib().ephor_degenerate_groups(len(w), w, float(cutoff), g.ctypes.data)
    return g


def degenerate_gauge_fix(E, U, cutoff):
    """In place on E (nb, n) ascending and U (nb, n*n) complex128 column-major blocks (C-contiguous): the
    degenerate-band gauge fixing of solve_eigen_el! (src/diagonalize.jl:47-85). Returns U."""
    E = _c(E, np.float64)
    if U.dtype != np.complex128 or not U.flags.c_contiguous:
        raise ValueError("degenerate_gauge_fix needs a C-contiguous complex128 U")
    nb, n = E.shape
    if lib().ephor_degenerate_gauge_fix(n, nb, E, U, float(cutoff)) != 0:
        raise RuntimeError("ephor_degenerate_gauge_fix: a projected eigensolve did not converge")
    return U


def ph_eigen(dq_colmajor, mass):
    """One q: dq is (nm*nm,) column-major D(q). Returns omega (nm,), U (nm*nm,) column-major."""
    nm = len(mass)
    d = np.array(dq_colmajor, dtype=np.complex128, copy=True)
    om = np.empty(nm)
    u = np.empty(nm * nm, np.complex128)
    if lib().ephor_ph_eigen(nm, d, _c(mass, np.float64), om, u) != 0:
        raise RuntimeError("ephor_ph_eigen did not converge")
    return om, u


def window_filter(E, wmin, wmax):
    """E (nk, nw) ascending. Returns keep (nk,) uint8, bmin, bmax (1-based, 0 if dropped), nbelow."""
    E = _c(E, np.float64)
    nk, nw = E.shape
    keep = np.empty(nk, np.uint8)
    bmin, bmax, nb = (np.empty(nk, np.int64) for _ in range(3))
    lib().ephor_window_filter(nw, nk, E, wmin, wmax, keep, bmin, bmax, nb)
    return keep, bmin, bmax, nb


def kpoints_grid(ngrid, shift=(0.0, 0.0, 0.0)):
    n = int(np.prod(ngrid))
    x = np.empty((n, 3))
    lib().ephor_kpoints_grid(ngrid[0], ngrid[1], ngrid[2], _c(shift, np.float64), x)
    return x


def hash_xk(xk, ngrid, shift=(0.0, 0.0, 0.0)):
    xk = _c(np.atleast_2d(xk), np.float64)
    h = np.empty(xk.shape[0], np.int64)
    lib().ephor_hash_xk(xk.shape[0], xk, _c(ngrid, np.int32), _c(shift, np.float64), h)
    return h


def eph_RR_to_Rq(nw, nm, epmat_ph, irvec_el, irvec_ep, xq, uph_colmajor):
    """epmat_ph (nr_ep, nw*nw*nm*nr_el) -> (nr_el, nw*nw*nm) (the eRpq op_r)."""
    nr_el, nr_ep = len(irvec_el), len(irvec_ep)
    out = np.empty((nr_el, nw * nw * nm), np.complex128)
    lib().ephor_eph_RR_to_Rq(nw, nm, nr_el, nr_ep, _c(epmat_ph, np.complex128), _c(irvec_ep, np.int32),
                             _c(xq, np.float64), _c(uph_colmajor, np.complex128), out)
    return out


def eph_Rq_to_kq(nw, nm, eRpq, irvec_el, xk, uk, ukq):
    """uk (nw, nbk), ukq (nw, nbkq) normal numpy matrices. Returns ep (nm, nbk, nbkq) C-order,
    i.e. column-major (nbkq, nbk, nm)."""
    nbk, nbkq = uk.shape[1], ukq.shape[1]
    ep = np.empty((nm, nbk, nbkq), np.complex128)
    lib().ephor_eph_Rq_to_kq(nw, nm, len(irvec_el), _c(eRpq, np.complex128), _c(irvec_el, np.int32),
                             _c(xk, np.float64), _c(np.asarray(uk).T, np.complex128), nbk,
                             _c(np.asarray(ukq).T, np.complex128), nbkq, ep)
    return ep


def eph_RR_to_kR(nw, nm, epmat_el, irvec_el, irvec_ep, xk, uk):
    """epmat_el (nr_el, nw*nw*nm*nr_ep) "el" layout -> (nr_ep, nw*nbk*nm)."""
    nbk = uk.shape[1]
    out = np.empty((len(irvec_ep), nw * nbk * nm), np.complex128)
    lib().ephor_eph_RR_to_kR(nw, nm, len(irvec_el), len(irvec_ep), _c(epmat_el, np.complex128),
                             _c(irvec_el, np.int32), _c(xk, np.float64),
                             _c(np.asarray(uk).T, np.complex128), nbk, out)
    return out


def eph_kR_to_kq(nw, nbk, nm, ekpR, irvec_ep, xq, uph_colmajor, ukq):
    nbkq = ukq.shape[1]
    ep = np.empty((nm, nbk, nbkq), np.complex128)
    lib().ephor_eph_kR_to_kq(nw, nbk, nm, len(irvec_ep), _c(ekpR, np.complex128), _c(irvec_ep, np.int32),
                             _c(xq, np.float64), _c(uph_colmajor, np.complex128),
                             _c(np.asarray(ukq).T, np.complex128), nbkq, ep)
    return ep


def g2(ep, omega):
    """ep (nm, nbk, nbkq) -> g2 same shape, |ep|^2/(2 omega_nu)."""
    ep = _c(ep, np.complex128)
    nm = ep.shape[0]
    out = np.empty(ep.shape)
    lib().ephor_g2(ep.size // nm, nm, ep, _c(omega, np.float64), out)
    return out


def g2_degenerate_average(g2v, ekq, ek, cutoff):
    g = np.array(g2v, dtype=np.float64, order="C", copy=True)
    nm, nbk, nbkq = g.shape
    lib().ephor_g2_degenerate_average(nbkq, nbk, nm, g, _c(ekq, np.float64), _c(ek, np.float64), cutoff)
    return g


def outer_q_block(nw, nm, eRpq, irvec_el, xk_grid, U, bmin, bmax, ik_list, ikq_list, nbw):
    """U (nk_grid, nw*nw) column-major blocks. Returns ep (nlist, nm, nbw, nbw) C-order."""
    nl = len(ik_list)
    out = np.empty((nl, nm, nbw, nbw), np.complex128)
    lib().ephor_outer_q_block(nw, nm, len(irvec_el), _c(eRpq, np.complex128), _c(irvec_el, np.int32),
                              _c(xk_grid, np.float64), _c(U, np.complex128), _c(bmin, np.int64),
                              _c(bmax, np.int64), nl, _c(ik_list, np.int64), _c(ikq_list, np.int64),
                              nbw, out)
    return out


def bte_increments(method, ek, ekq, wq, g2v, wtq, mu, T, smearing, eta):
    out = np.empty(2)
    lib().ephor_bte_increments(method, ek, ekq, wq, g2v, wtq, mu, T, smearing, eta, out)
    return out[0], out[1]


def bte_accumulate_q(nm, nbw, ik, ikq, g2v, bmin, bmax, off, e_state, omega_q, wtq, mus, Ts, etas, method,
                     smearing, w_cut, S_out, S_in):
    """Adds one q's pairs into S_out (nT, ns) and writes S_in (nT, ns, ns) in place (C-contiguous float64)."""
    assert S_out.flags.c_contiguous and S_in.flags.c_contiguous and S_out.dtype == S_in.dtype == np.float64
    ns = S_out.shape[1]
    lib().ephor_bte_accumulate_q(nm, nbw, len(ik), _c(ik, np.int64), _c(ikq, np.int64), _c(g2v, np.float64),
                                 _c(bmin, np.int64), _c(bmax, np.int64), _c(off, np.int64), _c(e_state, np.float64),
                                 _c(omega_q, np.float64), wtq, len(mus), _c(mus, np.float64), _c(Ts, np.float64),
                                 _c(etas, np.float64), method, smearing, w_cut, ns, S_out, S_in)


def num_threads() -> int:
    return lib().ephor_num_threads()


class GridOpt:
    """The GridOpt staged transform of one operator (ephor_gridopt_*): per-k queries reuse the k1-plane
    and (k1, k2)-line partial sums while k1 (k2) stays the same."""

    def __init__(self, op_r, irvec, threads=False):
        self.op = _c(op_r, np.complex128)  # (nr, ld)
        self.irvec = _c(irvec, np.int32)
        self.nd = self.op.shape[1]
        self.h = lib().ephor_gridopt_new(self.irvec.shape[0], self.irvec, self.nd, int(threads))

    @property
    def groups(self):
        return lib().ephor_gridopt_nr23(self.h), lib().ephor_gridopt_nr3(self.h)

    def get(self, xk) -> np.ndarray:
        out = np.empty(self.nd, np.complex128)
        lib().ephor_gridopt_get(self.h, self.op, self.nd, _c(xk, np.float64), out)
        return out

    def reset(self):
        lib().ephor_gridopt_reset(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ephor_gridopt_free(self.h)
            self.h = None


def outer_q_gridopt(nw, nm, epmat_ph, irvec_el, irvec_ep, xq, uph_colmajor, omega, iq_list, xk_grid, kgrid, U,
                    bmin, bmax, ksurv, nbw, nout=0):
    """The reference's CPU outer-q loop in GridOpt mode over the q of iq_list (see ephor_outer_q_gridopt).
    uph_colmajor (nq, nm*nm) column-major blocks, omega (nq, nm). Returns (pairs, g2sum per q, ik of the
    recorded pairs of iq_list[0], their ep (nout, nm, nbw, nbw) C-order)."""
    iq_list = _c(iq_list, np.int64)
    g2sum = np.empty(len(iq_list))
    ik_out = np.empty(max(nout, 1), np.int64)
    ep_out = np.empty((max(nout, 1), nm, nbw, nbw), np.complex128)
    n = lib().ephor_outer_q_gridopt(nw, nm, len(irvec_el), len(irvec_ep), _c(epmat_ph, np.complex128),
                                    _c(irvec_el, np.int32), _c(irvec_ep, np.int32), _c(xq, np.float64),
                                    _c(uph_colmajor, np.complex128), _c(omega, np.float64), len(iq_list), iq_list,
                                    _c(xk_grid, np.float64), _c(kgrid, np.int32), _c(U, np.complex128),
                                    _c(bmin, np.int64), _c(bmax, np.int64), len(ksurv), _c(ksurv, np.int64), nbw,
                                    g2sum, nout, ik_out, ep_out.ctypes.data if nout > 0 else None)
    return int(n), g2sum, ik_out[:nout], ep_out[:nout]
