"""Band-energy surrogate: a small MLP predicting the nw band energies of k, used to skip the exact
H(k) Fourier + eigensolve at fine-grid k-points whose bands all lie far outside the energy window.

Network: periodic Fourier features of k (cos/sin(2 pi m.k) for the nf2 shortest integer vectors m, one
per +-m pair) -> Linear(2 nf2, hidden) -> SiLU -> Linear(hidden, hidden) -> SiLU -> Linear(hidden, nout).
nout is nw rounded up to a multiple of 16; the extra outputs are trained to the top band.

Training runs on the GPU through libephip's kernels (features, tiled FP32 GEMMs, SiLU, MSE gradient,
split-K weight gradients, Adam) on exact eigenvalues the engine computes itself. Inference runs in BF16
on the matrix cores, fused with the pruning test. The safety margin is twice the maximum error of the
BF16 inference on the training points. The exact window filter (src/filter.jl:40-100) decides the
final selection on the surviving points; the surrogate never adds or removes an exact survivor by
itself -- it only chooses which points get an exact solve.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib as L


def _torch():
    import torch
    return torch


def fourier_vectors(nf2: int) -> np.ndarray:
    """The nf2 shortest nonzero integer vectors, one per +-m pair, by |m|^2 then lexicographic order."""
    cand = []
    r = 4
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            for c in range(-r, r + 1):
                m = (a, b, c)
                if m != (0, 0, 0) and tuple(-x for x in m) not in cand:
                    cand.append(m)
    cand.sort(key=lambda m: (m[0] ** 2 + m[1] ** 2 + m[2] ** 2, m))
    return np.array(cand[:nf2], dtype=np.int32)


@dataclasses.dataclass
class TrainReport:
    steps: int
    loss: list
    max_err_f32: float
    max_err_bf16: float
    margin: float


class BandSurrogate:
    def __init__(self, nw: int, device, nf2: int = 32, hidden: int = 128, seed: int = 0):
        torch = _torch()
        self.nw, self.nf2, self.hidden = nw, nf2, hidden
        self.nf = 2 * nf2
        self.nout = ((nw + 15) // 16) * 16
        self.dev = device
        self.mvec = torch.from_numpy(fourier_vectors(nf2)).to(device)
        H, F, O = hidden, self.nf, self.nout
        self.shapes = {"W1": (H, F), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (O, H), "b3": (O,)}
        self.offs = {}
        off = 0
        for k, sh in self.shapes.items():
            self.offs[k] = off
            off += int(np.prod(sh))
        self.nparam = off
        rng = np.random.default_rng(seed)
        flat = np.zeros(off, np.float32)
        for k, sh in self.shapes.items():
            if k.startswith("W"):
                w = rng.uniform(-1, 1, size=sh) * np.sqrt(3.0 / sh[1])
                flat[self.offs[k]:self.offs[k] + w.size] = w.ravel()
        self.p = torch.from_numpy(flat).to(device)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step_count = 0
        self.bf16 = None
        self.margin = None

    def view(self, buf, name):
        o = self.offs[name]
        return buf[o:o + int(np.prod(self.shapes[name]))]

    # ---------------------------------------------------------------------------------- FP32 path
    def _alloc_acts(self, n):
        torch = _torch()
        H, F, O = self.hidden, self.nf, self.nout
        mk = lambda c: torch.empty((n, c), dtype=torch.float32, device=self.dev)
        return dict(f=mk(F), z1=mk(H), a1=mk(H), z2=mk(H), a2=mk(H), y=mk(O))

    def _gemm_nt(self, X, W, bias, out):
        """out (n x o) = X (n x i) W^T (W: o x i) + bias."""
        n, i = X.shape
        o = W.numel() // i
        L.call("ephip_gemm_f32", 0, 1, n, o, i, L.ptr(X), i, L.ptr(W), i, L.ptr(out), o, L.ptr(bias), 0.0,
               L.stream_handle())

    def forward(self, xk, acts=None):
        n = xk.shape[0]
        a = acts if acts is not None else self._alloc_acts(n)
        st = L.stream_handle()
        L.call("ephip_mlp_features", n, L.ptr(xk), self.nf2, L.ptr(self.mvec), L.ptr(a["f"]), st)
        self._gemm_nt(a["f"], self.view(self.p, "W1"), self.view(self.p, "b1"), a["z1"])
        L.call("ephip_silu_fwd", a["z1"].numel(), L.ptr(a["z1"]), L.ptr(a["a1"]), st)
        self._gemm_nt(a["a1"], self.view(self.p, "W2"), self.view(self.p, "b2"), a["z2"])
        L.call("ephip_silu_fwd", a["z2"].numel(), L.ptr(a["z2"]), L.ptr(a["a2"]), st)
        self._gemm_nt(a["a2"], self.view(self.p, "W3"), self.view(self.p, "b3"), a["y"])
        return a

    def _wgrad(self, gz, x, out_name, part):
        """grad W (o x i) = gz^T (o x n) x (n x i), split over n, reduced by column sums."""
        n, o = gz.shape
        i = x.shape[1]
        nsplit = max(1, min(64, n // 256))
        st = L.stream_handle()
        L.call("ephip_gemm_f32_splitk", 1, 0, o, i, n, L.ptr(gz), o, L.ptr(x), i, L.ptr(part), i, nsplit, st)
        L.call("ephip_colsum", nsplit, o * i, L.ptr(part), o * i, L.ptr(self.view(self.g, out_name)), st)

    def _bgrad(self, gz, out_name, part, ones):
        """grad b (o) = gz^T (o x n) 1 (n), split over n like the weight gradients."""
        n, o = gz.shape
        nsplit = max(1, min(64, n // 256))
        st = L.stream_handle()
        L.call("ephip_gemm_f32_splitk", 1, 0, o, 1, n, L.ptr(gz), o, L.ptr(ones), 1, L.ptr(part), 1, nsplit, st)
        L.call("ephip_colsum", nsplit, o, L.ptr(part), o, L.ptr(self.view(self.g, out_name)), st)

    def loss_and_grad(self, xk, target, acts=None, grads=None):
        """Mean squared error over points and outputs; fills self.g. Returns (loss tensor (device), acts)."""
        torch = _torch()
        n = xk.shape[0]
        a = self.forward(xk, acts)
        st = L.stream_handle()
        O, H = self.nout, self.hidden
        gw = grads if grads is not None else dict(gy=torch.empty((n, O), dtype=torch.float32, device=self.dev),
                                                  ga=torch.empty((n, H), dtype=torch.float32, device=self.dev),
                                                  gz=torch.empty((n, H), dtype=torch.float32, device=self.dev),
                                                  part=torch.empty(64 * H * max(H, self.nf), dtype=torch.float32,
                                                                   device=self.dev),
                                                  partial=torch.empty((n * O + 255) // 256, dtype=torch.float64,
                                                                      device=self.dev),
                                                  ones=torch.ones(n, dtype=torch.float32, device=self.dev))
        L.call("ephip_mse_grad", n * O, L.ptr(a["y"]), L.ptr(target), 1.0 / (n * O), L.ptr(gw["gy"]),
               L.ptr(gw["partial"]), st)
        # layer 3
        self._wgrad(gw["gy"], a["a2"], "W3", gw["part"])
        self._bgrad(gw["gy"], "b3", gw["part"], gw["ones"])
        L.call("ephip_gemm_f32", 0, 0, n, H, O, L.ptr(gw["gy"]), O, L.ptr(self.view(self.p, "W3")), H,
               L.ptr(gw["ga"]), H, None, 0.0, st)
        L.call("ephip_silu_bwd", n * H, L.ptr(a["z2"]), L.ptr(gw["ga"]), L.ptr(gw["gz"]), st)
        # layer 2
        self._wgrad(gw["gz"], a["a1"], "W2", gw["part"])
        self._bgrad(gw["gz"], "b2", gw["part"], gw["ones"])
        L.call("ephip_gemm_f32", 0, 0, n, H, H, L.ptr(gw["gz"]), H, L.ptr(self.view(self.p, "W2")), H,
               L.ptr(gw["ga"]), H, None, 0.0, st)
        L.call("ephip_silu_bwd", n * H, L.ptr(a["z1"]), L.ptr(gw["ga"]), L.ptr(gw["gz"]), st)
        # layer 1
        self._wgrad(gw["gz"], a["f"], "W1", gw["part"])
        self._bgrad(gw["gz"], "b1", gw["part"], gw["ones"])
        return gw["partial"], a, gw

    def targets(self, E):
        """FP32 targets (n, nout) from FP64 band energies (n, nw); padded outputs repeat the top band."""
        torch = _torch()
        E = E.to(torch.float32)
        if self.nout > self.nw:
            E = torch.cat([E, E[:, -1:].expand(-1, self.nout - self.nw)], dim=1)
        return E.contiguous()

    def train(self, xk, E, steps: int = 3000, lr: float = 3e-3, lr_final: float = 1e-4, log_every: int = 250):
        """Full-batch Adam on (xk (n,3) float64, E (n, nw) float64) device tensors."""
        torch = _torch()
        tgt = self.targets(E)
        if self.step_count == 0:  # start the output layer at the mean band energies
            self.view(self.p, "b3").copy_(tgt.mean(dim=0))
        acts = self._alloc_acts(xk.shape[0])
        grads = None
        losses = []
        for s in range(steps):
            partial, acts, grads = self.loss_and_grad(xk, tgt, acts, grads)
            self.step_count += 1
            lr_s = lr * (lr_final / lr) ** (s / max(1, steps - 1))
            L.call("ephip_adam_step", self.nparam, L.ptr(self.p), L.ptr(self.g), L.ptr(self.m), L.ptr(self.v),
                   lr_s, 0.9, 0.999, 1e-8, self.step_count, L.stream_handle())
            if log_every and (s % log_every == 0 or s == steps - 1):
                losses.append((s, float(partial.sum()) / tgt.numel()))
        self.bf16 = None
        return losses

    # ---------------------------------------------------------------------------------- BF16 path
    def _to_bf16(self):
        torch = _torch()
        if self.bf16 is None:
            self.bf16 = {}
            for k in ("W1", "W2", "W3"):
                src = self.view(self.p, k)
                dst = torch.empty(src.numel(), dtype=torch.bfloat16, device=self.dev)
                L.call("ephip_f32_to_bf16", src.numel(), L.ptr(src), L.ptr(dst), L.stream_handle())
                self.bf16[k] = dst
        return self.bf16

    def infer_prune(self, xk, window, margin: float, want_y: bool = False):
        """BF16 inference + pruning on (n,3) float64 device k-points. Returns (keep uint8 (n,), y or None)."""
        torch = _torch()
        w = self._to_bf16()
        n = xk.shape[0]
        keep = torch.empty(n, dtype=torch.uint8, device=self.dev)
        y = torch.empty((n, self.nout), dtype=torch.float32, device=self.dev) if want_y else None
        L.call("ephip_mlp_bf16_prune", n, L.ptr(xk), self.nf2, L.ptr(self.mvec), self.hidden, self.nout,
               L.ptr(w["W1"]), L.ptr(self.view(self.p, "b1")), L.ptr(w["W2"]), L.ptr(self.view(self.p, "b2")),
               L.ptr(w["W3"]), L.ptr(self.view(self.p, "b3")), float(window[0]), float(window[1]), float(margin),
               L.ptr(y), L.ptr(keep), L.stream_handle())
        return keep, y

    @staticmethod
    def window_distance(E, window):
        """Distance (n,) of the band nearest to the energy window [wmin, wmax] (0 inside it)."""
        lo, hi = float(window[0]), float(window[1])
        return ((lo - E).clamp(min=0.0) + (E - hi).clamp(min=0.0)).min(dim=1).values

    def calibrate(self, xk, E, factor: float = 2.0, window=None) -> tuple[float, float]:
        """Max |error| of the FP32 and BF16 predictions over (xk, E); margin = factor x the BF16 error.
        With ``window``, the error that sets the margin is the one the prune decision uses: of the distance
        of the nearest predicted band to the window (a point is pruned when that distance exceeds the
        margin, so a point with an exact band in the window is kept whenever the BF16 error of the
        distance is below the margin); bands far from the window no longer widen it."""
        a = self.forward(xk)
        e32 = float((a["y"][:, :self.nw].double() - E).abs().max())
        _, y = self.infer_prune(xk, (0.0, 0.0), 0.0, want_y=True)
        yb = y[:, :self.nw].double()
        e16 = float((yb - E).abs().max())
        if window is None:
            self.margin = factor * e16
        else:
            ed = float((self.window_distance(yb, window) - self.window_distance(E, window)).abs().max())
            self.margin = factor * ed
            self.margin_error = ed
        return e32, e16
