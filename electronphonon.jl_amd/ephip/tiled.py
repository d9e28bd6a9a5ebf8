"""Outer-k tiled device output (TiledDeviceOutput, src/calculator/tiled_device_output.jl:32-151).

A device-resident calculator output whose axis ``i_axis`` runs over outer (k-side) states, held either
whole on the device ("full": one download at the end) or as one outer-k tile at a time ("block": the
tile of the current k batch is zeroed when the batch begins and streamed to pinned host memory when it
ends), decided once, on the first batch, from the free device memory: block when
``headroom * full bytes > free bytes`` (residency_use_block). The scatter stays with the calculator,
which writes global state i at tiled index ``i - tile_offset()`` with the buffer's extent
``tile_stride()`` along that axis.

Arrays are torch tensors in the reference's column-major index order with the axes reversed, so
``dims`` is given in the reference's order and the tiled axis is counted from there (0-based here).
"""
from __future__ import annotations

import warnings

import numpy as np


def _torch():
    import torch
    return torch


def residency_use_block(free_bytes: int, full_bytes: int, headroom: float = 1.2) -> bool:
    """True when full_bytes of device-resident output would not fit free_bytes with the headroom."""
    return headroom * full_bytes > free_bytes


def state_range_for_k_range(state_offset: np.ndarray, ksurv: np.ndarray, k_first: int, k_last: int) -> tuple:
    """Half-open range of outer states of the surviving-k positions k_first .. k_last (inclusive): the
    states of a k batch are contiguous (k-major ordering), ind_range_for_k_range of
    src/band_states.jl:398-428."""
    return int(state_offset[ksurv[k_first]]), int(state_offset[ksurv[k_last] + 1])


class TiledDeviceOutput:
    """narr arrays of shape dims (the reference's order) whose axis i_axis is tiled over outer states.
    states: (state_offset (nk_grid + 1), ksurv) host arrays giving each surviving k's states.
    free_bytes: callable returning the free device bytes (ephip_free_bytes by default) -- the decision
    input, replaceable for tests. force_block: True / False to fix the mode, None to decide."""

    def __init__(self, dims, i_axis: int, states, dtype=None, device=None, narr: int = 1,
                 force_block: bool | None = None, headroom: float = 1.2, free_bytes=None, pin_host: bool = True):
        torch = _torch()
        self.dims = [int(d) for d in dims]
        if not 0 <= i_axis < len(self.dims):
            raise ValueError(f"i_axis {i_axis} out of range for dims {self.dims}")
        self.i_axis = int(i_axis)
        self.n_full = self.dims[i_axis]
        self.narr = int(narr)
        self.state_offset = np.asarray(states[0], dtype=np.int64)
        self.ksurv = np.asarray(states[1], dtype=np.int64)
        self.dtype = dtype if dtype is not None else torch.float64
        self.device = device
        self.force_block = force_block
        self.headroom = float(headroom)
        self.free_bytes = free_bytes
        self.pin_host = pin_host
        self.dev, self.host = [], []
        self.decided = False
        self.block = False
        self.ni_cap = 0
        self.tile_i0 = 0
        self.tile_ni = 0

    # ------------------------------------------------------------------ layout helpers
    def _torch_axis(self) -> int:
        return len(self.dims) - 1 - self.i_axis   # reversed axes

    def _shape(self, ni: int):
        d = list(self.dims)
        d[self.i_axis] = ni
        return tuple(reversed(d))

    def is_allocated(self) -> bool:
        return bool(self.dev)

    def is_block(self) -> bool:
        return self.block

    def device_array(self, k: int = 0):
        return self.dev[k]

    def host_array(self, k: int = 0):
        return self.host[k]

    def tile_offset(self) -> int:
        return self.tile_i0

    def tile_length(self) -> int:
        return self.tile_ni

    def tile_stride(self) -> int:
        return self.dev[0].shape[self._torch_axis()]

    def full_bytes(self) -> int:
        torch = _torch()
        return self.narr * torch.empty((), dtype=self.dtype).element_size() * int(np.prod(self.dims))

    # ------------------------------------------------------------------ batches
    def tile_begin(self, k_first: int, k_last: int, n_batch_max: int):
        """Begin the k batch of surviving-k positions k_first .. k_last (inclusive): on the first batch,
        decide the residency and allocate; in block mode record the batch's state range and zero the
        tile's active region."""
        torch = _torch()
        if not self.decided:
            if self.force_block is not None:
                self.block = bool(self.force_block)
            else:
                free = self.free_bytes() if self.free_bytes is not None else _device_free_bytes()
                self.block = residency_use_block(free, self.full_bytes(), self.headroom)
                if self.block:
                    warnings.warn(f"full device-resident output ({self.narr} x "
                                  f"{self.full_bytes() / self.narr / 1e9:.2f} GB) does not fit free device memory; "
                                  "falling back to per-tile block residency (per-batch device-to-host streaming)")
            self.decided = True
            if self.block:
                nk = len(self.ksurv)
                self.ni_cap = max(b - a for a, b in (state_range_for_k_range(
                    self.state_offset, self.ksurv, ks, min(ks + n_batch_max, nk) - 1) for ks in range(0, nk, n_batch_max)))
                shape = self._shape(self.ni_cap)
                self.dev = [torch.empty(shape, dtype=self.dtype, device=self.device) for _ in range(self.narr)]
                pin = self.pin_host and torch.cuda.is_available() and str(self.device).startswith("cuda")
                self.host = [torch.empty(shape, dtype=self.dtype, pin_memory=pin) for _ in range(self.narr)]
            else:
                shape = self._shape(self.n_full)
                self.dev = [torch.zeros(shape, dtype=self.dtype, device=self.device) for _ in range(self.narr)]
        if self.block:
            a, b = state_range_for_k_range(self.state_offset, self.ksurv, k_first, k_last)
            if b - a > self.ni_cap:
                raise ValueError(f"k batch {k_first}..{k_last} holds {b - a} states, above the tile's {self.ni_cap}")
            self.tile_i0, self.tile_ni = a, b - a
            ax = self._torch_axis()
            for A in self.dev:
                A.narrow(ax, 0, self.tile_ni).zero_()
        return self

    def tile_download(self):
        """Block mode: one contiguous device-to-host copy per array into the host mirrors (synchronous)."""
        for d, h in zip(self.dev, self.host):
            h.copy_(d)
        return self

    def tile_free(self):
        self.dev, self.host = [], []
        self.decided = self.block = False
        self.ni_cap = self.tile_i0 = self.tile_ni = 0
        return self


def _device_free_bytes() -> int:
    from . import core
    return core.free_bytes(None)
