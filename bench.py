"""Benchmark: g(k,q) pairs/s of the Wannier->Bloch e-ph hot path on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): seeded synthetic model with
nw = 16 Wannier functions, 4 atoms (nmodes = 12), H(R)/D(R) on a 9^3 R-box, g(Re, Rp) on 7^3 x 7^3
boxes stored complex64 (2.89 GB), fine k-grid 48^3, q-grid 24^3, window E_F +- 0.3 eV.

Two paths are timed (``--path both``, the default):
* plain outer-q, FP32 (the reference's shape: separate Fourier and rotation kernels): a step is one
  tile of ``--qtile`` q-points over all window-surviving k;
* fast outer-k (the value): k-side window projection, kept-q compaction, batched Fourier over Rp and
  the k+q/phonon rotations; a step is one batch of ``--kbatch-k`` surviving k against every q, in
  split-BF16 (default) or FP32 Fourier precision; the Fourier kernel variant is tuned on the live batch
  first (``--no-tune`` to skip).
Setup (model upload, fine-grid states, phonons, operand packing) is outside the timed region; inputs
are resident in HBM when timing starts.

value = pairs done / time with pairs = |K_surv| x |Q| per q (plain) or per k (fast), K_surv the k with
at least one band in the window by the exact filter (the same numerator for every path).
``--transport`` runs the transport measurement instead (one JSON line of its own).
"""
from __future__ import annotations

import argparse
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
# HBM bytes per launch of the two Fourier kernels of the fast path on the cfg2 workload, keyed by
# (precision, stage, k batch, kernel variant), from rocprofv3 FETCH_SIZE (x2.00) and WRITE_SIZE (x1.00)
# passes calibrated on a 256 MiB copy (profiles/cfg2_fast_gridopt_hbm_counters.md,
# profiles/cfg2_fast_bf16_mfma_hbm_counters.md). Counters cannot be read
# inside the timed run; a variant or batch without a measurement reports traffic null.
TRAFFIC_BYTES_CFG2 = {
    ("fp32", "kR->kq", 1152, "gridopt"): (3376891 * 2 + 638069) * 1024,  # staged kR->kq + rotations + g2 (VALU)
    ("fp32", "kR->kq", 1152, "gridopt_mfma"): (3536000 * 2 + 819800) * 1024,  # the same, stage 3 on MFMAs
    ("fp32", "RR->kR", 1152, "gridopt"): (204717 * 2 + 274026) * 1024,   # staged RR->kR projection launch (MFMA)
    ("fp32", "kR->kq", 768, "gridopt"): (2186567 * 2 + 414734) * 1024,
    ("fp32", "RR->kR", 768, "gridopt"): (204184 * 2 + 230276) * 1024,
    ("bf16x3", "kR->kq", 384, 10): (2998082 * 2 + 204444) * 1024,   # fused kR->kq + rotations, 8 waves
    ("bf16x3", "RR->kR", 384, "proj0"): (1827474 * 2 + 200604) * 1024,  # RR->kR fused with the projection
}
# kernel names by (stage, variant) for the split-BF16 fast path (tuner kinds 3 and 4)
def _staged(v) -> bool:
    """Variant label of a staged (GridOpt) contraction."""
    return isinstance(v, str) and v.startswith("gridopt")


KERNEL_NAMES = {
    ("RR->kR", "gridopt"): "gridopt_stage1_kernel + gridopt_kR_project_mfma_kernel: staged (GridOpt) RR->kR over "
                           "Re, stages 2/3 with the k-side window projection on FP32 MFMAs",
    ("kR->kq", "gridopt"): "gridopt_kq_rotate_kernel: staged (GridOpt) kR->kq over the q grid with the k+q/phonon "
                           "rotations and g2, FP32 VALU",
    ("kR->kq", "gridopt_mfma"): "gridopt_kq_rotate_mfma_kernel: staged (GridOpt) kR->kq over the q grid, stages 1-2 "
                                "on the VALU, stage 3 and the k+q-window rotation as one FP32 MFMA contraction, "
                                "phonon rotation and g2",
    ("RR->kR", "proj0"): "ft_bf16p8_proj_kernel: eight-wave 192x128 packed split-BF16 RR->kR Fourier over Re "
                         "fused with the k-side window projection",
    ("RR->kR", "proj1"): "ft_bf16p8_proj_kernel<deep>: eight-wave 192x128 packed split-BF16 RR->kR Fourier over "
                         "Re fused with the k-side window projection, loads two stages ahead",
    ("RR->kR", 7): "ft_bf16p8_kernel: eight-wave 192x128 packed split-BF16 RR->kR Fourier over Re",
    ("kR->kq", 5): "ft_bf16p8_rot_kernel: eight-wave packed split-BF16 kR->kq Fourier over Rp fused with the "
                   "k+q/phonon rotations and g2",
    ("kR->kq", 8): "ft_bf16p8_rot_kernel: eight-wave packed split-BF16 kR->kq Fourier over Rp fused with the "
                   "k+q/phonon rotations and g2, units dealt to XCDs in groups of 8",
    ("kR->kq", 10): "ft_bf16p8_rot_kernel<prefetched epilogue>: eight-wave packed split-BF16 kR->kq Fourier over "
                    "Rp fused with the k+q/phonon rotations and g2, units dealt to XCDs in groups of 8",
    ("kR->kq", 12): "ft_bf16p8_rot_kernel<prefetched epilogue, deep>: eight-wave packed split-BF16 kR->kq Fourier "
                    "over Rp fused with the k+q/phonon rotations and g2, loads two stages ahead",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(O, m, st, epmat_c64, nk_sample, kgrid, min_seconds, direct_seconds=3.0):
    """Time the reference's CPU path as its own benchmarks run it: the outer-q loop in GridOpt mode
    (run_eph_over_q_and_k with fourier_mode = "gridopt", benchmark/bench_bte_gpu.jl:58-60), restated in
    FP64 C with OpenMP (oracle/ephor.c ephor_outer_q_gridopt; not the reference itself). Whole q slices in
    grid order from q = 0 (the order the reference's loop takes, which gives its GridOpt caches their
    reuse), each RR->Rq + the Rq->kq pairs of every surviving k, until min_seconds of CPU work; pairs/s on
    the metric's numerator (K_surv per q). Beside it, the direct-sum restatement (ephor_outer_q_block, the
    "normal" Fourier mode) on q slices spread over the grid for direct_seconds, as an extra field."""
    ksurv = st.ksurv.cpu().numpy()
    nks = len(ksurv)
    ik = ksurv if nk_sample < 0 else ksurv[:nk_sample]
    xk = st.el.xk.cpu().numpy()
    xq_all = st.ph.xq.cpu().numpy()
    nq = xq_all.shape[0]
    bmin = st.el.filt.band_min.cpu().numpy()
    bmax = st.el.filt.band_max.cpu().numpy()
    U = st.el.U.cpu().numpy().reshape(xk.shape[0], -1).astype(np.complex128)
    uph_all = st.ph.u.cpu().numpy().astype(np.complex128).reshape(nq, -1)
    om_all = st.ph.omega.cpu().numpy().astype(np.float64)
    ep128 = epmat_c64.astype(np.complex128)
    # GridOpt mode: blocks of consecutive q
    t_go, q_go, pairs_go, blk = 0.0, 0, 0, 4
    while q_go == 0 or (t_go < min_seconds and q_go < nq):
        iqs = np.arange(q_go, min(q_go + blk, nq), dtype=np.int64)
        t0 = time.perf_counter()
        n, _, _, _ = O.outer_q_gridopt(m.nw, m.nmodes, ep128, m.irvec_el, m.irvec_ep, xq_all, uph_all, om_all, iqs,
                                       xk, kgrid, U, bmin, bmax, ik, st.nbw)
        t_go += time.perf_counter() - t0
        q_go += len(iqs)
        pairs_go += n
    value = q_go * nks / t_go if len(ik) == nks else q_go * len(ik) / t_go
    # direct-sum ("normal" Fourier mode) restatement, q spread over the grid
    t_rr = t_kq = 0.0
    qs = []
    while not qs or t_rr + t_kq < direct_seconds:
        iq = (len(qs) * 4999) % nq
        if iq in qs:
            break
        qs.append(iq)
        xq = xq_all[iq]
        ikq = O.hash_xk(xk[ik] + xq, kgrid)
        t0 = time.perf_counter()
        eRpq = O.eph_RR_to_Rq(m.nw, m.nmodes, ep128, m.irvec_el, m.irvec_ep, xq, uph_all[iq])
        t1 = time.perf_counter()
        O.outer_q_block(m.nw, m.nmodes, eRpq, m.irvec_el, xk, U, bmin, bmax, ik, ikq, st.nbw)
        t2 = time.perf_counter()
        t_rr += t1 - t0
        t_kq += t2 - t1
    direct = len(qs) * len(ik) / (t_rr + t_kq)
    return {"value": value, "unit": "pairs/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"FP64 C restatement (not the reference) of the reference CPU outer-q loop in GridOpt mode "
                      f"(fourier_mode=\"gridopt\", OpenMP, {O.num_threads()} threads): q 0..{q_go - 1} in grid "
                      f"order, each RR->Rq + Rq->kq for {len(ik)} of {nks} surviving k ({pairs_go} pairs with a "
                      f"window band at k+q computed); {t_go:.2f} s; measured",
            "seconds": t_go,
            "direct_sum": {"value": direct, "unit": "pairs/s", "seconds": t_rr + t_kq,
                           "sample": f"direct phase sums (fourier_mode=\"normal\"): {len(qs)} q slice(s) "
                                     f"{qs[:4]}{'...' if len(qs) > 4 else ''}, RR->Rq {t_rr:.2f} s, "
                                     f"Rq->kq {t_kq:.2f} s"}}


def oracle_check(m, st, epmat, grids, w, payload, nk_sample=32, nq_sample=2):
    """Check what was timed against the FP64 oracle (test infrastructure, run after the timed region):
    the window filter of the whole fine grid bit for bit (hence K_surv, the metric's numerator), and the
    gauge-invariant |g|^2 of sampled (k, q) pairs of the LAST timed batch (its payload buffers), averaged
    over degenerate bands: RMS (norm-wise) relative error (the metric's 1e-4 bound), global and worst-block
    max|d| / max|ref|."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import eph_checks as C
    from ephip import model as M
    t0 = time.perf_counter()
    f = st.el.filt
    E_o = C.full_grid_states(m, M.kpoints_grid(grids["kgrid"]))
    r = C.filter_compare(E_o, f.keep.cpu().numpy(), f.band_min.cpu().numpy(), f.band_max.cpu().numpy(),
                         st.el.E.cpu().numpy(), -w, w)
    out = {"filter": r, "filter_bit_exact": r["keep_equal"] and r["band_min_equal"] and r["band_max_equal"]}
    if payload is not None:
        nq = int(np.prod(grids["qgrid"]))
        qsel = [int(q) for q in np.random.default_rng(7).choice(nq, nq_sample, replace=False)]
        got = C.sample_units_g2(payload, qsel, nk_sample, torch, seed=8)
        m.epmat_ph = epmat
        g = C.g2_parity(m, epmat, grids["kgrid"], grids["qgrid"], got, f.band_min.cpu().numpy(),
                        f.band_max.cpu().numpy(), st.nbw)
        m.epmat_ph = None
        out.update(g2_rel_err=g["rel_err"], g2_pairs=g["pairs"], g2_max_rel_err=g["max_rel_err"],
                   g2_max_block_rel_err=g["max_block_rel_err"],
                   g2_sample=f"{nk_sample} k of the last timed batch x q {qsel}")
    out["within_bound"] = bool(out["filter_bit_exact"] and out.get("g2_rel_err", 0.0) <= 1e-4)
    out["seconds"] = time.perf_counter() - t0
    log(f"[bench] oracle check: filter bit-exact {out['filter_bit_exact']} (K_surv {r['K_surv_oracle']}), "
        f"|g|^2 rel err {out.get('g2_rel_err')} over {out.get('g2_pairs')} pairs ({out['seconds']:.1f} s)")
    return out


def transport_bench(args):
    """Electron transport at the workload's grids on the time-reversal-symmetric variant of the model
    (real H(R), Hermitian real g): one full pass of the outer-k e-ph loop feeding the Boltzmann
    calculator, velocities, SERTA and iterative BTE, in FP32 and split-BF16 Fourier precision.
    Reports the conductivities, their relative difference and the timings (one JSON line)."""
    import torch

    from ephip import _lib as L
    from ephip import bte, core
    from ephip import model as M
    t0 = time.perf_counter()
    backend = core.gpu_backend(0)
    m, grids = M.make_config_model(args.config, time_reversal=True)
    log(f"[transport] time-reversal model {args.config} built in {time.perf_counter() - t0:.1f} s")
    w = grids["window"]
    epmat_dev = core.WannierObject(m.irvec_ep, torch.from_numpy(m.epmat_ph).to(backend.torch_device), check=False)
    from ephip import eph
    st = eph.setup_outer_q(m, grids["kgrid"], grids["qgrid"], (-w, w), backend, epmat_dev=epmat_dev)
    T = 0.025852  # 300 K
    out = {}
    for name, prec in (("fp32", L.PREC_FP32), ("bf16x3", L.PREC_BF16X3)):
        params = bte.ElectronTransportParams(Tlist=[T], nlist=[1.0], volume=1.0, mulist=[0.0])
        run = bte.run_transport(m, grids["kgrid"], grids["qgrid"], (-w, w), params, backend, precision=prec,
                                smearing=("Gaussian", args.transport_eta), occupation_method=5,
                                nk_batch=args.kbatch_k, max_iter=args.transport_iter, rtol=1e-8, setup=st)
        r = run.result
        out[name] = {"sigma_serta": r.sigma_serta[0].tolist(), "sigma": r.sigma[0].tolist(),
                     "niter": r.niter[0], "converged": r.converged[0], "seconds": run.seconds,
                     "n_states": run.calc.bs.n,
                     "s_in_rows": int(sum(b.nrows for b in run.calc.blocks))}
        log(f"[transport] {name}: sigma_serta diag {np.diag(r.sigma_serta[0])}, sigma diag {np.diag(r.sigma[0])}, "
            f"{r.niter[0]} iterations, times {run.seconds}")
        del run
        torch.cuda.empty_cache()
    rel = lambda a, b: float(np.linalg.norm(np.array(a) - np.array(b)) / np.linalg.norm(np.array(b)))
    res = {"metric": "transport: conductivity at 300 K, 10 meV Gaussian, method 5 (split-BF16 vs FP32 Fourier)",
           "config": {"workload": f"{args.config} time-reversal variant, k {grids['kgrid']}, q {grids['qgrid']}, "
                                  f"window +-{w} eV", "eta_eV": args.transport_eta, "T_eV": T},
           "rel_diff_sigma_serta": rel(out["bf16x3"]["sigma_serta"], out["fp32"]["sigma_serta"]),
           "rel_diff_sigma": rel(out["bf16x3"]["sigma"], out["fp32"]["sigma"]),
           "detail": out}
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qtile", type=int, default=8)
    ap.add_argument("--kbatch", type=int, default=16384, help="k batch of the plain outer-q path")
    ap.add_argument("--kbatch-k", type=int, default=1152, help="k batch (one step) of the fast outer-k path")
    ap.add_argument("--path", choices=["fast", "plain", "both"], default="both")
    ap.add_argument("--precision", choices=["fp32", "bf16x3"], default="bf16x3",
                    help="Fourier precision of the fast path (split-BF16 meets the 1e-4 |g|^2 bound)")
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="surviving k per q slice of the CPU baseline (-1 = all, 0 = skip the baseline)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="q slices are added to the CPU baseline until this much CPU time is spent")
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--tune", dest="tune", action="store_true", default=True,
                    help="autotune the Fourier kernel variant on the live batch first (default)")
    ap.add_argument("--no-tune", dest="tune", action="store_false")
    ap.add_argument("--pipeline", action="store_true",
                    help="run the k batches on two overlapping streams (RR->kR of the next batch beside kR->kq)")
    ap.add_argument("--no-gridopt", dest="gridopt", action="store_false", default=True,
                    help="direct (MFMA) Fourier contractions instead of the staged GridOpt kernels")
    ap.add_argument("--variant", action="append", default=[], metavar="KIND=V",
                    help="pin a Fourier kernel variant (tuner kind=variant, repeatable); skips the tuner")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 256 MiB device-to-device copy kernel (counter calibration)")
    ap.add_argument("--transport", action="store_true",
                    help="run the transport measurement (conductivity, FP32 vs split-BF16) instead of the e-ph bench")
    ap.add_argument("--transport-eta", type=float, default=0.01, help="Gaussian smearing of the transport run (eV)")
    ap.add_argument("--transport-iter", type=int, default=200)
    ap.add_argument("--surrogate", dest="surrogate", action="store_true", default=None,
                    help="learned pre-filter leg: train the band-energy network on 24^3 + 20000 random k, prune the "
                         "fine grid, check zero misses, time the exact and pruned states (default: on for cfg5)")
    ap.add_argument("--no-surrogate", dest="surrogate", action="store_false")
    ap.add_argument("--bf16-leg", dest="bf16_leg", action="store_true", default=None,
                    help="also time the split-BF16 MFMA fast path (default: on for cfg5)")
    ap.add_argument("--no-bf16-leg", dest="bf16_leg", action="store_false")
    ap.add_argument("--graph", dest="graph", action="store_true", default=True,
                    help="also capture the timed k batches of the staged fast path as one HIP graph and time its "
                         "replay (reported beside the eager steps, not as the value; default on)")
    ap.add_argument("--no-graph", dest="graph", action="store_false")
    ap.add_argument("--no-check", dest="check", action="store_false", default=True,
                    help="skip the oracle check (full-grid filter and sampled |g|^2 of the last timed batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note only, not implemented)"}))
        sys.exit(2)

    if args.transport:
        transport_bench(args)
        return

    import torch

    from ephip import _lib as L
    from ephip import core, eph
    from ephip import model as M
    from ephip import tune as T

    t_setup = time.perf_counter()
    backend = core.gpu_backend(0)
    if args.calibrate_copy:
        # a plain copy of known size: 256 MiB read + 256 MiB written, for scaling HBM byte counters
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=backend.torch_device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    m, grids = M.make_config_model(args.config, with_g=False)
    log(f"[bench] generating e-ph tensor ({args.config})")
    epmat = M.build_epmat(m)
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
    nq_grid = nq_total
    kgrid = grids["kgrid"]

    def kept_pairs(ks):
        """Number of (unit, q) columns the outer-k path computes for the surviving-k positions ks
        (host arithmetic on the filter: q kept iff k+q has an in-window band), and the number of
        (unit, q, k+q window band) slots among them (the rotations' algorithmic extent)."""
        bmin = st.el.filt.band_min.cpu().numpy()
        bmax = st.el.filt.band_max.cpu().numpy()
        ksurv = st.ksurv.cpu().numpy()
        xk = M.kpoints_grid(kgrid)
        xq = M.kpoints_grid(grids["qgrid"])
        tot = slots = 0
        for p in ks:
            k = ksurv[p]
            h = O_hash(xk[k] + xq, kgrid)
            kept_h = h[bmin[h] > 0]
            nbk = int(bmax[k] - bmin[k] + 1)
            tot += len(kept_h) * nbk
            slots += int(np.minimum(bmax[kept_h] - bmin[kept_h] + 1, st.nbw).sum()) * nbk
        return tot, slots

    def O_hash(x, ng):
        c = np.mod(np.rint(x * np.asarray(ng)).astype(np.int64), np.asarray(ng))
        return (c[:, 0] * ng[1] + c[:, 1]) * ng[2] + c[:, 2]

    results = {}
    # ---- plain outer-q path (reference CS2 shape: separate Fourier / rotate kernels, full K_surv per q)
    if args.path in ("plain", "both"):
        loop = eph.OuterQLoop(st, nk_batch=args.kbatch, nq_tile=args.qtile, precision=L.PREC_FP32)
        ev = []
        loop.ft_timer = ev
        for s in range(args.warmup):
            loop.run_qtile((s * args.qtile) % nq_total, args.qtile)
        torch.cuda.synchronize()
        ev.clear()
        steps = args.steps if args.path == "plain" else max(2, args.steps // 4)
        iq_base = args.warmup * args.qtile
        t0 = time.perf_counter()
        for s in range(steps):
            loop.run_qtile((iq_base + s * args.qtile) % (nq_total - args.qtile + 1), args.qtile)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        durs = [a.elapsed_time(b) * 1e-3 for a, b, _ in ev]
        nd = m.nw * m.nw * m.nmodes
        flops = [8.0 * nd * len(m.irvec_el) * n for _, _, n in ev]
        results["plain"] = dict(value=nks * args.qtile * steps / el, ms_per_step=1e3 * el / steps, steps=steps,
                                ft_tflops=sum(flops) / sum(durs) / 1e12, ft_avg_ms=1e3 * sum(durs) / len(durs),
                                ft_flop_per_launch=sum(flops) / len(flops),
                                kernel="ft_mfma_kernel<float,128> (Rq->kq Fourier over Re)")
        log(f"[bench] plain outer-q fp32: {results['plain']['value']:.3e} pairs/s, FT "
            f"{results['plain']['ft_tflops']:.1f} TFLOP/s")
        del loop
    # ---- fast outer-k path (k-side projection, k+q window skip)
    if args.path in ("fast", "both"):
        loop = eph.OuterKLoop(st, nk_batch=args.kbatch_k, precision=prec, gridopt=args.gridopt)
        fused = bool(getattr(loop, "fused", False))
        if args.variant:
            for kv in args.variant:
                kind, v = (int(x) for x in kv.split("="))
                T.set_variant(kind, v)
            results["pinned_variants"] = list(args.variant)
            log(f"[bench] pinned variants {args.variant}")
        elif args.tune:
            tuned = loop.tune()
            log(f"[bench] tuner (ms per k batch by variant): {tuned}")
            results["tuner"] = {str(k): ({str(a): b for a, b in v.items()} if isinstance(v, dict) else v)
                                for k, v in tuned.items()}
        ev = []
        loop.ft_timer = ev
        nks_steps = args.kbatch_k
        warm = [s * nks_steps for s in range(args.warmup)]
        k_base = args.warmup * nks_steps
        timed = [(k_base + s * nks_steps) % max(1, nks - nks_steps) for s in range(args.steps)]
        if args.pipeline:  # RR->kR of batch b+1 on a second stream beside kR->kq of batch b
            loop.run_pipelined(warm, nks_steps)
        else:
            for i0 in warm:
                loop.run_kbatch(i0, nks_steps)
        torch.cuda.synchronize()
        ev.clear()
        t0 = time.perf_counter()
        last_payload = None
        if args.pipeline:
            loop.run_pipelined(timed, nks_steps)
        else:
            for i0 in timed:
                last_payload = loop.run_kbatch(i0, nks_steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        # algorithmic flops of the launches actually made (direct or staged kernels), step by step
        cols, f_rrkr_total, f_krkq_total = 0, 0.0, 0.0
        for i0 in timed:
            kept, slots = kept_pairs(range(i0, min(i0 + nks_steps, nks)))
            fr, fk = loop.kernel_flops(i0, nks_steps, kept, slots)
            cols += kept
            f_rrkr_total += fr
            f_krkq_total += fk
        kin = {"RR->kR": [], "kR->kq": []}
        for a, b, lab, fl in ev:
            kin[lab].append(a.elapsed_time(b) * 1e-3)
        t_rrkr, t_krkq = sum(kin["RR->kR"]), sum(kin["kR->kq"])
        staged_rr = getattr(loop, "gridopt", None) is not None
        staged_kq = getattr(loop, "gridopt_q", None) is not None
        results["fast"] = dict(value=nks_steps * nq_grid * args.steps / el, ms_per_step=1e3 * el / args.steps,
                               steps=args.steps, computed_columns=cols,
                               rrkr_tflops=f_rrkr_total / t_rrkr / 1e12,
                               rrkr_avg_ms=1e3 * t_rrkr / len(kin["RR->kR"]),
                               krkq_tflops=f_krkq_total / t_krkq / 1e12,
                               krkq_avg_ms=1e3 * t_krkq / len(kin["kR->kq"]),
                               rrkr_flop_per_launch=f_rrkr_total / len(kin["RR->kR"]),
                               krkq_flop_per_launch=f_krkq_total / len(kin["kR->kq"]),
                               fused_rotation=fused or staged_kq,
                               fused_projection=bool(getattr(loop, "fused_proj", False)) or staged_rr,
                               variant_rrkr=("gridopt" if staged_rr else
                                             f"proj{T.get_variant(T.KIND_FUSED_PROJ)}"
                                             if getattr(loop, "fused_proj", False) else
                                             T.get_variant(T.KIND_PACKED_BF16)),
                               variant_krkq=(("gridopt_mfma" if T.get_variant(T.KIND_GRIDOPT_KQ) == 1 else "gridopt")
                                             if staged_kq else T.get_variant(T.KIND_FUSED_ROT)))
        if args.graph and staged_rr and staged_kq and not args.pipeline:
            # the same timed batches as one captured HIP graph: launch overhead and host work removed
            try:
                bg = loop.capture(timed, nks_steps)
                bg.replay()
                torch.cuda.synchronize()
                t0g = time.perf_counter()
                bg.replay()
                torch.cuda.synchronize()
                elg = time.perf_counter() - t0g
                results["fast_graph"] = dict(value=nks_steps * nq_grid * len(timed) / elg,
                                             ms_per_step=1e3 * elg / len(timed), steps=len(timed))
                log(f"[bench] fast outer-k as one captured graph: {results['fast_graph']['value']:.3e} pairs/s "
                    f"({results['fast_graph']['ms_per_step']:.2f} ms per batch, eager "
                    f"{1e3 * el / args.steps:.2f} ms)")
                del bg
            except Exception as exc:   # capture refused (not a device fault): report, keep the eager numbers
                results["fast_graph"] = dict(error=str(exc))
                log(f"[bench] graph capture failed: {exc}")
        run_as = ("fp32 staged" if staged_rr and staged_kq else
                  "+".join(sorted({"fp32 staged" if st_ else args.precision for st_ in (staged_rr, staged_kq)})))
        log(f"[bench] fast outer-k ({run_as}): {results['fast']['value']:.3e} pairs/s; RR->kR "
            f"{results['fast']['rrkr_tflops']:.1f} TFLOP/s ({results['fast']['rrkr_avg_ms']:.2f} ms), kR->kq "
            f"{results['fast']['krkq_tflops']:.1f} TFLOP/s ({results['fast']['krkq_avg_ms']:.2f} ms)")

    # ---- split-BF16 MFMA fast path (packed operands, fused projection / rotations), a few steps
    if args.bf16_leg is None:
        args.bf16_leg = args.config == "cfg5"
    if args.bf16_leg and args.path in ("fast", "both"):
        del loop
        torch.cuda.empty_cache()
        lb = eph.OuterKLoop(st, nk_batch=384, precision=L.PREC_BF16X3, gridopt=False)
        if args.tune:
            lb.tune()
        nb_steps = max(2, args.steps // 2)
        for i0 in (0, 384):
            lb.run_kbatch(i0, 384)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in range(nb_steps):
            lb.run_kbatch(((2 + s_) * 384) % max(1, nks - 384), 384)
        torch.cuda.synchronize()
        el_b = time.perf_counter() - t0
        results["fast_bf16"] = dict(value=384 * nq_grid * nb_steps / el_b, ms_per_step=1e3 * el_b / nb_steps,
                                    steps=nb_steps, k_batch=384,
                                    kernels="ft_bf16p8_proj_kernel (RR->kR + projection), ft_bf16p8_rot_kernel "
                                            "(kR->kq + rotations + g2), split-BF16 on MFMA")
        log(f"[bench] fast outer-k split-BF16 MFMA: {results['fast_bf16']['value']:.3e} pairs/s")
        del lb
        torch.cuda.empty_cache()
    # ---- learned pre-filter: the band-energy network ahead of the exact states of the fine grid
    if args.surrogate is None:
        args.surrogate = args.config == "cfg5"
    if args.surrogate:
        from ephip import surrogate as S
        ham = core.WannierObject(m.irvec_h, m.ham_r).to_device(backend, dtype=torch.complex128)
        rng = np.random.default_rng(5)
        xtr = torch.from_numpy(np.vstack([M.kpoints_grid((24, 24, 24)), rng.random((20000, 3))])).to(
            backend.torch_device)
        t0 = time.perf_counter()
        st_tr = eph.compute_electron_states(ham, xtr, None)
        sur = S.BandSurrogate(m.nw, backend.torch_device, nf2=32, hidden=128, seed=1)
        sur.train(xtr, st_tr.E, steps=3000)
        e32, e16 = sur.calibrate(xtr, st_tr.E, window=(-w, w))
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        xf = st.el.xk
        t0 = time.perf_counter()
        ex = eph.compute_electron_states(ham, xf, (-w, w))
        torch.cuda.synchronize()
        t_exact = time.perf_counter() - t0
        t0 = time.perf_counter()
        pr, cand = eph.compute_electron_states_pruned(ham, xf, (-w, w), sur)
        torch.cuda.synchronize()
        t_pruned = time.perf_counter() - t0
        ke, kp = ex.filt.keep.bool(), pr.filt.keep.bool()
        misses = int((ke & ~kp).sum())
        same = bool(torch.equal(ke, kp) and torch.equal(ex.filt.band_min, pr.filt.band_min)
                    and torch.equal(ex.filt.band_max, pr.filt.band_max))
        results["surrogate"] = dict(
            train_seconds=t_train, train_points=int(xtr.shape[0]), max_err_fp32_eV=e32, max_err_bf16_eV=e16,
            window_distance_err_eV=sur.margin_error, margin_eV=sur.margin, candidates=int(cand.numel()),
            grid_points=int(xf.shape[0]), exact_survivors=int(ke.sum()), misses=misses, selection_identical=same,
            states_exact_ms=1e3 * t_exact, states_pruned_ms=1e3 * t_pruned)
        log(f"[bench] surrogate: margin {sur.margin:.3f} eV, candidates {cand.numel()} of {xf.shape[0]}, misses "
            f"{misses}; states exact {1e3 * t_exact:.0f} ms, pruned {1e3 * t_pruned:.0f} ms")
        del ex, pr, ham, st_tr
        torch.cuda.empty_cache()
    # ---- whole-job rates (states of the fine grid + every k batch over the full q grid)
    job = None
    if "fast" in results and "plain" in results:
        job_pairs = nks * nq_total
        t_states = results.get("surrogate", {}).get("states_exact_ms", None)
        t_states = (t_states or 0.0) * 1e-3
        n_batches = -(-nks // args.kbatch_k)
        job = {"pairs": job_pairs, "states_seconds": t_states,
               "fast_fp32_staged": job_pairs / (t_states + n_batches * results["fast"]["ms_per_step"] * 1e-3),
               "plain_fp32": job_pairs / (t_states + job_pairs / results["plain"]["value"])}
        if "fast_bf16" in results:
            t_st = t_states
            if "surrogate" in results and results["surrogate"]["misses"] == 0:
                t_st = results["surrogate"]["states_pruned_ms"] * 1e-3
            job["fast_bf16_batched_surrogate"] = job_pairs / (t_st + -(-nks // 384) *
                                                              results["fast_bf16"]["ms_per_step"] * 1e-3)
            job["ratio_bf16_batched_surrogate_vs_plain"] = job["fast_bf16_batched_surrogate"] / job["plain_fp32"]
        job["ratio_fast_vs_plain"] = job["fast_fp32_staged"] / job["plain_fp32"]
        log(f"[bench] whole job: fast {job['fast_fp32_staged']:.3e}, plain {job['plain_fp32']:.3e} pairs/s"
            + (f", split-BF16 + batched + surrogate {job['fast_bf16_batched_surrogate']:.3e}"
               if "fast_bf16_batched_surrogate" in job else ""))

    cpu = None
    if args.cpu_sample != 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        log("[bench] timing the CPU oracle slice")
        cpu = cpu_baseline(O, m, st, epmat, args.cpu_sample, grids["kgrid"], args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.3e} pairs/s ({cpu['seconds']:.1f} s)")

    check = None
    if args.check:   # the FP64 oracle as the checker of what was timed (after the timed region)
        check = oracle_check(m, st, epmat, grids, w, last_payload if "fast" in results else None)

    if "fast" in results:
        r = results["fast"]
        value, ms, steps = r["value"], r["ms_per_step"], r["steps"]
        # dominant kernel: the longer of the two Fourier launches per step
        bf = args.precision == "bf16x3"
        var = {"RR->kR": r.get("variant_rrkr"), "kR->kq": r.get("variant_krkq")}
        stage = "RR->kR" if r["rrkr_avg_ms"] >= r["krkq_avg_ms"] else "kR->kq"
        if stage == "RR->kR":
            if _staged(var[stage]) or bf:
                name = KERNEL_NAMES.get((stage, var[stage])) or (
                    f"ft_bf16p_kernel (packed split-BF16, variant {var[stage]}), RR->kR Fourier over Re")
            else:
                name = "ft_mfma_kernel<float,128>, RR->kR Fourier over Re"
            dom = (name + " (k batch)", r["rrkr_tflops"], r["rrkr_avg_ms"], r["rrkr_flop_per_launch"], stage)
        else:
            if _staged(var[stage]) or r.get("fused_rotation"):
                name = KERNEL_NAMES.get((stage, var[stage])) or (
                    f"ft_bf16p_rot_kernel (variant {var[stage]}): packed split-BF16 kR->kq Fourier over Rp fused "
                    "with the k+q/phonon rotations and g2")
            else:
                name = "batched Fourier kernel, kR->kq over Rp"
            dom = (name + " (units x kept q)", r["krkq_tflops"], r["krkq_avg_ms"], r["krkq_flop_per_launch"],
                   stage)
        # arithmetic of the two contractions as run: the staged (GridOpt) kernels are FP32 whatever the
        # precision argument
        kinds = {"fp32" if _staged(var[k]) else args.precision for k in var}
        path_dtype = "+".join(sorted(kinds))
        dom_fp32 = _staged(var[dom[4]]) or args.precision == "fp32"
        path = "outer-k, k-side window projection, k+q window skip"
        pairs_step = args.kbatch_k * nq_total
    else:
        r = results["plain"]
        value, ms, steps = r["value"], r["ms_per_step"], r["steps"]
        dom = (r["kernel"], r["ft_tflops"], r["ft_avg_ms"], r["ft_flop_per_launch"], "Rq->kq")
        path = "outer-q plain (separate Fourier / rotate kernels)"
        pairs_step = nks * args.qtile
        dom_fp32 = True
    peak = PEAK["fp32" if ("fast" not in results or dom_fp32) else args.precision]
    # the dominant kernel's time goes to its FP32 vector-ALU work: the staged kR->kq kernel in both forms (in
    # the MFMA form only stage 3 and the k+q-window contraction run on the matrix cores, a quarter of its time)
    valu_dom = "fast" in results and _staged(var[dom[4]]) and dom[4] == "kR->kq"
    valu_only = valu_dom and var[dom[4]] == "gridopt"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": path_dtype if "fast" in results else "fp32",
        "data": "synthetic (seeded model generator, no checkpoint)",
        "config": {"workload": f"{args.config}: nw {m.nw}, nmodes {m.nmodes}, H/D 9^3, g 7^3 x 7^3 c64, "
                               f"k {grids['kgrid']}, q {grids['qgrid']}, window +-{w} eV",
                   "path": path, "K_surv": nks, "pairs_per_step": pairs_step,
                   "pre_filter_pairs": nk_total * nq_total,
                   "plain_outer_q_fp32_pairs_per_s": results.get("plain", {}).get("value")},
        "roofline": {"bound": ("valu" if valu_dom else "mfma"),
                     "achieved": dom[1], "peak": peak, "unit": "TFLOP/s",
                     "frac": dom[1] / peak,
                     "compute": ("FP32 VALU, no MFMA: 157.3 TFLOP/s is MI355X's published FP32 peak, which counts "
                                 "packed FP32 (not used here); the unpacked vector peak is 78.6 TFLOP/s"
                                 if valu_only else
                                 "FP32 MFMA for stage 3 and the k+q-window contraction; stages 1-2, the phonon "
                                 "rotation and g2 on the FP32 VALU (no packed FP32); flops of all stages counted"
                                 if "fast" in results and var[dom[4]] == "gridopt_mfma" else
                                 "MFMA (split-BF16 flops counted once)" if "fast" in results and not dom_fp32
                                 else "MFMA FP32"),
                     "traffic": (TRAFFIC_BYTES_CFG2.get(("fp32" if _staged(var[dom[4]]) else args.precision,
                                                         dom[4], args.kbatch_k, var[dom[4]]))
                                 if ("fast" in results and args.config == "cfg2"
                                     and (dom[4] != "kR->kq" or r.get("fused_rotation"))) else None),
                     "kernel": dom[0], "avg_launch_ms": dom[2], "flop_per_launch": dom[3],
                     # FP32 VALU kernels: the same achieved rate against the unpacked vector peak
                     "frac_unpacked_vector_peak": (dom[1] / 78.6 if valu_dom else None)},
        "cpu_baseline": cpu,
        "check": check,
        "job": job,
        "detail": results,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
