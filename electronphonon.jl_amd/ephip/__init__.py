"""ephip: MI355X-native Wannier->Bloch engine for ElectronPhonon.jl's e-ph hot path.

Host mirror of the reference's batched device API over the libephip C-ABI (include/ephip.h).
"""
from . import _lib
from ._lib import EphipError, PREC_BF16X3, PREC_FP32
from . import model

__all__ = ["_lib", "model", "EphipError", "PREC_FP32", "PREC_BF16X3", "core", "eph", "bte"]


def __getattr__(name):
    # core/eph import torch lazily so that host-only users (fixtures, the model generator) do not pay for it
    if name in ("core", "eph", "bte"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
