"""Autotuner for the Fourier-contraction kernels.

The reference hard-codes its crossovers ("tuned on one GPU", src/wannier_to_bloch_batched.jl:397-406).
Here each contraction kind has several kernel variants (LDS staging depth, single/double buffering,
and for the pre-packed split-BF16 kernel the workgroup tile shape and operand prefetch);
``tune_fourier`` times every variant on the caller's live shape with HIP events on the launch stream
and installs the fastest through ``ephip_ft_set_variant``. Results are plain dicts, so a caller can
log or cache them.
"""
from __future__ import annotations

from . import _lib as L

KIND_FP32_WIDE, KIND_FP32_NARROW, KIND_FP64, KIND_PACKED_BF16, KIND_FUSED_ROT, KIND_FUSED_PROJ = 0, 1, 2, 3, 4, 5
# the staged (GridOpt) kR->kq kernel: stage 3 and the k+q-window rotation on the vector ALUs or on MFMAs
KIND_GRIDOPT_KQ = 6
GRIDOPT_KQ_VARIANTS = (0, 1)
GRIDOPT_KQ_VARIANT_NAMES = {0: "VALU stage 3 + DPP window sums", 1: "MFMA stage 3 + window rotation"}
KIND_KQ_ORDER = 7
KQ_ORDER_VARIANTS = (0, 1)
KQ_ORDER_VARIANT_NAMES = {0: "q1 fastest, unit by unit", 1: "XCD-grouped units"}
KIND_KQ_FOLD = 8
KQ_FOLD_VARIANTS = (0, 1)
KQ_FOLD_VARIANT_NAMES = {0: "folded stage-2/3 sums on symmetric boxes", 1: "general stage-2/3 sums"}
# host-side choice (no library kind): the staged kR->kq stage 1 in each workgroup, or by the pre-pass
KQ_STAGE1_VARIANTS = (0, 1)
KQ_STAGE1_VARIANT_NAMES = {0: "stage 1 per (unit, q1) workgroup", 1: "stage-1 pre-pass into a workspace"}
VARIANTS = (0, 1, 2)
VARIANT_NAMES = {0: "single-buffered KC16", 1: "double-buffered KC16", 2: "double-buffered KC8"}
PACKED_VARIANTS = (0, 1, 2, 3, 4, 5, 6, 7)
PACKED_VARIANT_NAMES = {0: "128x128", 1: "128x128 + operand prefetch", 2: "192x64", 3: "192x64 XCD-grouped",
                        4: "192x64 XCD-grouped, mid-sequence LDS writes", 5: "as 4, pipelined", 6: "as 3, pipelined",
                        7: "192x128 eight waves, operand through LDS"}
FUSED_VARIANTS = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12)
PROJ_VARIANTS = (0, 1)
# candidates the outer-k loop times by default: one representative per distinct design (variants that
# measured within noise of, or clearly behind, these stay selectable by set_variant)
FUSED_TUNE = (0, 5, 8, 10)
PROJ_TUNE = (0,)


def get_variant(kind: int) -> int:
    return L.load().ephip_ft_get_variant(kind)


def set_variant(kind: int, variant: int) -> None:
    L.call("ephip_ft_set_variant", kind, variant)


def tune_fourier(kind: int, run, reps: int = 3, variants=VARIANTS) -> dict:
    """Time ``run()`` (which must launch the contraction of this kind) under each variant; keep the
    fastest. Returns {variant: median ms, "best": variant}."""
    import torch
    res = {}
    for v in variants:
        set_variant(kind, v)
        run()  # warm-up (code object load)
        times = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            run()
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        res[v] = sorted(times)[len(times) // 2]
    best = min((v for v in variants), key=lambda v: res[v])
    set_variant(kind, best)
    res["best"] = best
    return res
