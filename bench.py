"""Benchmark: g(k,q) pairs/s of the Wannier->Bloch e-ph hot path on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): seeded synthetic model with
nw = 16 Wannier functions, 4 atoms (nmodes = 12), H(R)/D(R) on a 9^3 R-box, g(Re, Rp) on 7^3 x 7^3
boxes stored complex64 (2.89 GB), fine k-grid 48^3, q-grid 24^3, window E_F +- 0.3 eV.

A step is one tile of ``--qtile`` q-points through the outer-q hot path: RR->Rq for the tile, then for
every q and every k-batch of the window-surviving k: Fourier over Re on the matrix cores, k+q state
lookup, windowed rotations and |g|^2. Setup (model upload, fine-grid states, phonons) is outside the
timed region; inputs are resident in HBM when timing starts.

value = |K_surv| * q-points done / time, where K_surv are the k with at least one band in the window
by the exact filter (the same numerator for every path).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "electronphonon.jl_amd"))

METRIC = ("g(k,q) k·q pairs/s at |g|²≤1e-4 rel err vs FP64; plain vs "
          "mixed+batched+surrogate")
PEAK = {"fp32": 157.3, "bf16x3": 2500.0}  # TFLOP/s, MI355X dense MFMA (fp32, bf16)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_epmat(model_mod, m, workers=16):
    """Build the complex64 "ph" layout e-ph tensor column by column in parallel threads."""
    nr_ep = len(m.irvec_ep)
    nd = m.nw * m.nw * m.nmodes * len(m.irvec_el)
    ep = np.empty((nr_ep, nd), dtype=np.complex64)

    def col(ip):
        ep[ip] = model_mod.eph_column(m.seed, m.nw, m.nmodes, m.irvec_el, m.irvec_ep[ip], ip, m.g_decay,
                                      np.complex64)

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(col, range(nr_ep)))
    return ep


def cpu_baseline(O, m, st, epmat_c64, iq, nk_sample, kgrid):
    """Time the FP64 oracle (C restatement of the reference's CPU outer-q path, OpenMP) on one q:
    RR->Rq for the q and the Rq->kq pairs of the first nk_sample surviving k. Returns pairs/s
    extrapolated to one full q (RR->Rq once + |K_surv| pairs)."""
    import torch
    ksurv = st.ksurv.cpu().numpy()
    nks = len(ksurv)
    ik = ksurv[:nk_sample]
    xk = st.el.xk.cpu().numpy()
    xq = st.ph.xq[iq].cpu().numpy()
    ikq = O.hash_xk(xk[ik] + xq, kgrid)
    need = np.unique(np.concatenate([ik, ikq]))
    U = np.zeros((xk.shape[0], m.nw * m.nw), np.complex128)
    U[need] = st.el.U[torch.from_numpy(need).to(st.el.U.device)].cpu().numpy().reshape(len(need), -1)
    bmin = st.el.filt.band_min.cpu().numpy()
    bmax = st.el.filt.band_max.cpu().numpy()
    uph = st.ph.u[iq].cpu().numpy().astype(np.complex128).ravel()
    ep128 = epmat_c64.astype(np.complex128)
    t0 = time.perf_counter()
    eRpq = O.eph_RR_to_Rq(m.nw, m.nmodes, ep128, m.irvec_el, m.irvec_ep, xq, uph)
    t1 = time.perf_counter()
    O.outer_q_block(m.nw, m.nmodes, eRpq, m.irvec_el, xk, U, bmin, bmax, ik, ikq, st.nbw)
    t2 = time.perf_counter()
    per_k = (t2 - t1) / len(ik)
    t_q = (t1 - t0) + per_k * nks
    return {"value": nks / t_q, "unit": "pairs/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"FP64 C oracle (restatement of the reference CPU outer-q path, OpenMP): 1 q "
                      f"(RR->Rq {t1 - t0:.2f} s) + Rq->kq for {len(ik)} of {nks} surviving k "
                      f"({t2 - t1:.2f} s), extrapolated to one full q",
            "seconds": t2 - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qtile", type=int, default=8)
    ap.add_argument("--kbatch", type=int, default=16384)
    ap.add_argument("--precision", choices=["fp32", "bf16x3"], default="fp32")
    ap.add_argument("--cpu-sample", type=int, default=3000, help="k-points in the CPU baseline slice (0 = skip)")
    ap.add_argument("--config", default="cfg2")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import torch

    from ephip import _lib as L
    from ephip import core, eph
    from ephip import model as M

    t_setup = time.perf_counter()
    backend = core.gpu_backend(0)
    m, grids = M.make_config_model(args.config, with_g=False)
    log(f"[bench] generating e-ph tensor ({args.config})")
    epmat = make_epmat(M, m)
    log(f"[bench] e-ph tensor {epmat.nbytes / 1e9:.2f} GB generated in {time.perf_counter() - t_setup:.1f} s")
    epmat_dev = core.WannierObject(m.irvec_ep, torch.from_numpy(epmat).to(backend.torch_device), check=False)
    w = grids["window"]
    st = eph.setup_outer_q(m, grids["kgrid"], grids["qgrid"], (-w, w), backend, epmat_dev=epmat_dev)
    core.synchronize(backend)
    nks = int(st.ksurv.shape[0])
    nk_total = int(np.prod(grids["kgrid"]))
    nq_total = int(np.prod(grids["qgrid"]))
    log(f"[bench] setup {time.perf_counter() - t_setup:.1f} s: |K_surv| = {nks} of {nk_total}, nbw = {st.nbw}")

    prec = L.PREC_FP32 if args.precision == "fp32" else L.PREC_BF16X3
    loop = eph.OuterQLoop(st, nk_batch=args.kbatch, nq_tile=args.qtile, precision=prec)
    ft_events = []
    loop.ft_timer = ft_events
    for s in range(args.warmup):
        loop.run_qtile((s * args.qtile) % nq_total, args.qtile)
    torch.cuda.synchronize()
    ft_events.clear()
    iq_base = args.warmup * args.qtile
    t0 = time.perf_counter()
    for s in range(args.steps):
        loop.run_qtile((iq_base + s * args.qtile) % (nq_total - args.qtile + 1), args.qtile)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pairs = nks * args.qtile * args.steps
    value = pairs / elapsed
    # dominant kernel: the Fourier over Re (Rq->kq), timed with HIP events on its own stream
    durs = [a.elapsed_time(b) * 1e-3 for a, b, _ in ft_events]
    ncols = [n for _, _, n in ft_events]
    nd = m.nw * m.nw * m.nmodes
    nr_el = len(m.irvec_el)
    flops = [8.0 * nd * nr_el * n for n in ncols]
    achieved = sum(flops) / sum(durs) / 1e12
    avg_ms = 1e3 * sum(durs) / len(durs)
    log(f"[bench] {args.steps} steps in {elapsed:.3f} s; FT kernel avg {avg_ms:.3f} ms "
        f"({len(durs)} launches), {achieved:.1f} TFLOP/s")

    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        log("[bench] timing the CPU oracle slice")
        cpu = cpu_baseline(O, m, st, epmat, 0, args.cpu_sample, grids["kgrid"])
        log(f"[bench] cpu baseline {cpu['value']:.3e} pairs/s ({cpu['seconds']:.1f} s)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded model generator, no checkpoint)",
        "config": {"workload": f"{args.config}: nw {m.nw}, nmodes {m.nmodes}, H/D 9^3, g 7^3 x 7^3 c64, "
                               f"k {grids['kgrid']}, q {grids['qgrid']}, window +-{w} eV",
                   "path": "outer-q plain (separate Fourier / rotate kernels)",
                   "K_surv": nks, "pairs_per_step": nks * args.qtile, "q_per_step": args.qtile,
                   "k_batch": args.kbatch, "pre_filter_pairs": nk_total * nq_total},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[args.precision], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[args.precision], "traffic": None,
                     "kernel": "ft_mfma_kernel<float,128> (Rq->kq Fourier over Re)" if args.precision == "fp32"
                     else "ft_bf16x3_kernel<128> (Rq->kq Fourier over Re)",
                     "avg_launch_ms": avg_ms, "flop_per_launch_avg": sum(flops) / len(flops)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
