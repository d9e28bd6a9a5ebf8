"""ctypes binding of libephip.so (the C-ABI declared in include/ephip.h).

The library is loaded from the package directory (``electronphonon.jl_amd/libephip.so``). There is
no fallback: if it is missing, every device entry point raises ``EphipError``. Device buffers are
torch CUDA tensors; their raw pointers and the current stream handle are passed through.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libephip.so")

EPHIP_OK, EPHIP_ERR_ARG, EPHIP_ERR_HIP, EPHIP_ERR_NOCONV = 0, 1, 2, 3
PREC_FP32, PREC_BF16X3 = 0, 1

_vp, _i32, _i64, _d = C.c_void_p, C.c_int32, C.c_int64, C.c_double

# name -> argtypes (restype is always int status, except ephip_last_error / ephip_version)
SIGNATURES = {
    "ephip_free_bytes": [_vp, _vp],
    "ephip_synchronize": [_vp],
    "ephip_fourier_phase_c64": [_i32, _vp, _i32, _vp, _vp, _vp],
    "ephip_fourier_c32": [_i64, _i32, _vp, _i64, _vp, _i32, _vp, _vp, _i64, _i32, _vp],
    "ephip_fourier_c64": [_i64, _i32, _vp, _i64, _vp, _i32, _vp, _vp, _i64, _vp],
    "ephip_heev_c32": [_i32, _i64, _vp, _i32, _vp, _vp, _vp, _vp],
    "ephip_heev_refine": [_i32, _i64, _vp, _vp, _vp, _vp, _d, _vp],
    "ephip_eigen_batched": [_i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_degenerate_gauge_fix": [_i32, _i64, _vp, _vp, _d, _vp],
    "ephip_eigvals_batched": [_i32, _i64, _vp, _vp, _vp, _vp, _vp],
    "ephip_eigvals_batched_noalloc": [_i32, _i64, _vp, _vp, _vp, _vp],
    "ephip_window_filter": [_i32, _i64, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp],
    "ephip_ph_mass_scale": [_i32, _i64, _vp, _vp, _vp],
    "ephip_ph_finalize": [_i32, _i64, _vp, _vp, _vp, _vp, _vp],
    "ephip_kq_index": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_eph_rotate_modes": [_i64, _i32, _i32, _i32, _vp, _i64, _vp, _vp, _i64, _vp],
    "ephip_eph_rotate_rqkq_g2": [_i32, _i32, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp,
                                 _vp, _vp],
    "ephip_fourier_batched_c32": [_i64, _i32, _vp, _i64, _i64, _vp, _vp, _i32, _i32, _vp, _vp, _i64, _vp,
                                  _i64, _i64, _i32, _vp],
    "ephip_eph_ph_to_el": [_i64, _i32, _i32, _vp, _vp, _vp],
    "ephip_eph_project_kR": [_i32, _i32, _i32, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_kq_compact": [_i32, _vp, _vp, _i32, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "ephip_mlp_features": [_i64, _vp, _i32, _vp, _vp, _vp],
    "ephip_gemm_f32": [_i32, _i32, _i32, _i32, _i32, _vp, _i64, _vp, _i64, _vp, _i64, _vp, C.c_float, _vp],
    "ephip_gemm_f32_splitk": [_i32, _i32, _i32, _i32, _i32, _vp, _i64, _vp, _i64, _vp, _i64, _i32, _vp],
    "ephip_silu_fwd": [_i64, _vp, _vp, _vp],
    "ephip_silu_bwd": [_i64, _vp, _vp, _vp, _vp],
    "ephip_mse_grad": [_i64, _vp, _vp, C.c_float, _vp, _vp, _vp],
    "ephip_colsum": [_i32, _i32, _vp, _i64, _vp, _vp],
    "ephip_adam_step": [_i64, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float, C.c_float, _i32, _vp],
    "ephip_f32_to_bf16": [_i64, _vp, _vp, _vp],
    "ephip_bte_increments": [_i64, _i32, _i32, _vp, _vp, _vp],
    "ephip_bte_accumulate_units": [_i32, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d,
                                   _i32, _vp, _vp, _vp, _i32, _i32, _d, _i64, _vp, _vp, _vp, _vp, _vp],
    "ephip_band_velocity": [_i32, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "ephip_el_velocity_direct_batched": [_i32, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _i64, _vp],
    "ephip_batched_gemm": [_i32, _i32, _i32, _i32, _i32, _i32, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                           _i64, _i64, _vp],
    "ephip_eph_apply_rotations": [_i32, _i32, _i32, _i32, _i32, _i64, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp,
                                  _vp, _vp, _vp],
    "ephip_eph_apply_rotations_rqkq": [_i32, _i32, _i32, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp],
    "ephip_bte_window_accumulate": [_i32, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp,
                                    _vp, _i32, _vp, _vp, _vp, _i32, _i32, _d, _i32, _i32, _i32, _i64, _i64, _vp,
                                    _vp],
    "ephip_units_to_qbatched": [_i32, _i32, _i32, _i32, _i64, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_eph_window_scatter": [_i32, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _i32, _i32, _i32,
                                 _i64, _i64, _i64, _vp, _vp],
    "ephip_bte_spmv": [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_mlp_bf16_prune": [_i64, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_float, C.c_float,
                             C.c_float, _vp, _vp, _vp],
    "ephip_ft_pack_operand": [_i64, _i32, _vp, _i64, _i32, _i64, _vp, _vp],
    "ephip_ft_pack_operand_swapped": [_i64, _i32, _i32, _vp, _i64, _vp, _vp],
    "ephip_gridopt_stage1": [_i64, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, C.c_double, _i32, _vp, _vp],
    "ephip_gridopt_stage1_planes": [_i64, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _vp, _i32, _vp, _vp],
    "ephip_gridopt_kR_project": [_i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _i64, _i32, _i32, _vp,
                                 _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp],
    "ephip_gridopt_kR_tables": [_i32, _i32, _i32, _i32, _vp, _i64, _i64, _i32, _i32, _vp, _i32, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp],
    "ephip_gridopt_kR_project_mfma": [_i32, _i32, _i32, _i32, _i32, _vp, _i64, _i64, _i32, _i32, _i32, _vp, _i32,
                                      _i32,
                                      _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp],
    "ephip_ft_pack_phases": [_i32, _vp, _i64, _vp, _vp, _vp],
    "ephip_kq_line_offsets": [_i32, _vp, _vp, _i64, _i32, _i32, _vp, _vp],
    "ephip_gridopt_kq_rotate": [_i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i64, _i64,
                                _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_eph_RR_to_kR_project_packed_bf16x3": [_i32, _i32, _i32, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp,
                                                 _i32, _vp, _i64, _vp],
    "ephip_fourier_packed_bf16x3": [_i64, _i32, _vp, _i32, _vp, _i64, _i32, _vp, _vp, _i64, _vp, _i64, _i64, _vp],
    "ephip_eph_kq_rotate_packed_bf16x3": [_i32, _i32, _i32, _i32, _vp, _i32, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp,
                                          _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ephip_ft_set_variant": [_i32, _i32],
    "ephip_ft_get_variant": [_i32],
    "ephip_eph_rotate_kq_g2": [_i32, _i32, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp,
                               _vp, _vp, _vp, _vp, _vp],
}


class EphipError(RuntimeError):
    pass


_LIB = None


def load():
    """Load libephip.so (once). Raises EphipError if the library has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise EphipError(f"libephip.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
                         "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    lib.ephip_last_error.restype = C.c_char_p
    lib.ephip_last_error.argtypes = []
    lib.ephip_version.restype = C.c_int
    lib.ephip_version.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    return ["ephip_last_error", "ephip_version", *SIGNATURES]


def call(name: str, *args) -> None:
    lib = load()
    if name in SIGNATURES and len(args) != len(SIGNATURES[name]):
        raise EphipError(f"{name}: {len(args)} arguments given, the C-ABI takes {len(SIGNATURES[name])}")
    st = getattr(lib, name)(*args)
    if st != EPHIP_OK:
        msg = lib.ephip_last_error().decode(errors="replace")
        raise EphipError(f"{name}: {msg} (status {st})")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise EphipError("expected a device (CUDA/HIP) tensor")
    return t.data_ptr()


def stream_handle() -> int | None:
    import torch
    return torch.cuda.current_stream().cuda_stream or None
