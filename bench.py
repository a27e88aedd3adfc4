#!/usr/bin/env python3
"""Benchmark of the SpecAnn alpha-sweep on one MI355X.

A step is one full sweep of the hot path on the workload's graph: 128 alphas (warm-started
Lanczos per alpha, rounding, deduplication) plus the batched local search of the distinct
candidates and the exact energies. The graph is generated and uploaded (J/c, degrees) before the
timed region, so every timed step starts with its inputs resident in HBM.

Reported (one JSON line): value = seconds per FP32 sweep with the tangent predictor (the diagonal
predictor plus the eigenvector tangent's correction from two deflated MINRES steps seeded with the
backward difference). Beside it, timed once each after warm-up: the FP32 sweep with the reference's
default diagonal predictor and the plain FP64 diagonal-predictor form (SpecAnn-64), giving both the
north star's ratio (FP32 tangent against FP64 diagonal) and the like-for-like one (diagonal in both
precisions) with their best-energy gaps, and the end-to-end solve() time from the host matrix (the
reference's `time` column). The roofline object gives the fused scaled SpMV's algorithmic bandwidth
(8 nnz + 4 (n+1) + 12 n bytes per launch in FP32) against the 8 TB/s HBM peak, timed with HIP events
on the launch stream; cpu_baseline times the reference's CPU path (scipy ARPACK + numpy local search)
on one warm-started alpha of the same graph and scales it to the sweep (labelled as extrapolated).

    python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--workload er22|ba20|er1k] [--pred tangent|diagonal|secant]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "alpha-sweep sec at n=2^22 FP32 vs FP64; energy gap; SpMV % of HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "er22": dict(kind="er", n=1 << 22, deg=40, coupling="pm1", seed=42,
                 text="Erdos-Renyi n=2^22, mean degree 40, J uniform in {-1,+1} (seed 42)"),
    "ba20": dict(kind="ba", n=1 << 20, m=20, coupling="pm1", seed=42,
                 text="signed Barabasi-Albert n=2^20, m=20, J uniform in {-1,+1} (seed 42)"),
    "er1k": dict(kind="er", n=1024, deg=40, coupling="neg", seed=42,
                 text="Erdos-Renyi n=1024, p=40/1023, J=-1 (seed 42)"),
    "c5": dict(kind="ba", n=1 << 25, m=10, coupling="neg", seed=42, native=True,
               text="Barabasi-Albert n=2^25, m=10, J=-1 (seed 42, parallel host generator)"),
    "er22n": dict(kind="er", n=1 << 22, deg=40, coupling="pm1", seed=42, native=True,
                  text="Erdos-Renyi n=2^22, mean degree 40, J uniform in {-1,+1} (seed 42, parallel host generator)"),
    "c3": dict(kind="batch", count=256, lo=10, hi=16, deg=40, m=20, coupling="neg", seed=42,
               text="256 instances packed block-diagonally, n = floor(2^U(10,16)) (seed 42), alternating "
                    "Erdos-Renyi (p = 40/(n-1)) and Barabasi-Albert (m = 20), J = -1"),
}


def make_batch(w):
    """The batch's instances: n_i = floor(2^U(lo, hi)) from seed; even i Erdos-Renyi, odd i Barabasi-Albert."""
    from specann_amd import graphs
    rng = np.random.default_rng(w["seed"])
    sizes = np.floor(2.0 ** rng.uniform(w["lo"], w["hi"], w["count"])).astype(np.int64)
    Js = []
    for i, n in enumerate(sizes):
        n = int(n)
        if i % 2 == 0:
            Js.append(graphs.erdos_renyi(n, w["deg"] / (n - 1), seed=w["seed"] + i, kind=w["coupling"]))
        else:
            Js.append(graphs.barabasi_albert(n, w["m"], seed=w["seed"] + i, kind=w["coupling"]))
    return Js


def run_batch(args, w):
    """The batched sweep (all instances per launch) in FP32, the FP64 form beside it, and a sample of
    the instances swept one at a time (the per-instance path) for comparison."""
    import torch
    from specann_amd.batch import BatchProblem, BatchSweep
    from specann_amd.device import DeviceProblem
    from specann_amd.spectral_annealing import SpectralAnnealing
    t0 = time.time()
    Js = make_batch(w)
    N, nnz = sum(J.shape[0] for J in Js), sum(J.nnz for J in Js)
    log(f"batch {args.workload}: {len(Js)} instances, n total {N}, nnz {nnz}, generated in {time.time() - t0:.1f}s")
    pred = args.pred if args.pred in ("diagonal", "secant", "prev") else "secant"

    tunes = {}

    def sweep(dtype, bp=None, lockstep=False):
        if bp is None:
            bp = BatchProblem(Js, dtype)
            tunes[str(dtype)] = bp.autotune()["choice"]
            log(f"batched SpMV tiles ({dtype}): {tunes[str(dtype)]}")
        torch.cuda.synchronize()
        t = time.perf_counter()
        sw = BatchSweep(random_seed=42, num_alphas=args.alphas, pred_mode=pred, lockstep=lockstep)
        out = sw.run(bp)
        torch.cuda.synchronize()
        return time.perf_counter() - t, out, bp, sw

    sec32, out32, bp32, sw32 = sweep(torch.float32)
    log(f"fp32 batched sweep: {sec32:.3f}s ({out32.steps} batched steps, {out32.matvecs} instance operator "
        f"applications, local search {out32.post_sweep_time:.3f}s, {out32.local_passes} passes, "
        f"{sw32.retries} cold retries)")
    lock32 = None
    if not args.skip_lockstep:
        lock32, outl, _, _ = sweep(torch.float32, bp32, lockstep=True)
        same = all(a.best_energy == b.best_energy for a, b in zip(out32.results, outl.results))
        log(f"fp32 lockstep batched sweep: {lock32:.3f}s ({outl.steps} batched steps; best energies equal: {same})")
    spmv_sec, spmv_bytes = spmv_roofline(bp32.prob)
    achieved = spmv_bytes / spmv_sec / 1e9
    del bp32
    sec64 = out64 = None
    if not args.skip_fp64:
        sec64, out64, _, _ = sweep(torch.float64)
        log(f"fp64 batched sweep: {sec64:.3f}s")
    sample = min(args.c3_sample, len(Js))
    solo = None
    if sample:
        t = time.perf_counter()
        for J in Js[:sample]:
            SpectralAnnealing("bench", "synthetic", random_seed=42, precision=32, verbose=False,
                              num_alphas=args.alphas, pred_mode=pred).run_on_device(DeviceProblem(J, torch.float32))
        torch.cuda.synchronize()
        solo = time.perf_counter() - t
        log(f"{sample} instances swept one at a time: {solo:.3f}s")
    e32 = [r.best_energy for r in out32.results]
    gaps = None
    if out64 is not None:
        e64 = [r.best_energy for r in out64.results]
        gaps = [(a - b) / abs(b) for a, b in zip(e32, e64) if b]
    conv = float(np.mean([r["eigsh_converged"] for res in out32.results for r in res.rows]))
    cpu = None
    if not args.skip_cpu:
        # the CPU path (scipy ARPACK on S_rho, numpy local search) on the instances of up to 2^13 nodes
        # among every 16th one, extrapolated to the batch by stored entries
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import specann_oracle as O
        picks = [i for i in range(0, len(Js), 16) if Js[i].shape[0] <= (1 << 13)]
        t = time.perf_counter()
        for i in picks:
            O.cpu_sweep(Js[i], precision=32, num_alphas=args.alphas, pred_mode="diagonal", seed=42)
        csec = time.perf_counter() - t
        snnz = sum(Js[i].nnz for i in picks)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": csec * nnz / snnz if snnz else None, "unit": "s (extrapolated by stored entries)",
               "cores": 1, "kind": "port",
               "sample": (f"the full {args.alphas}-alpha CPU path (scipy ARPACK eigsh on S_rho, diagonal predictor, "
                          f"numpy local search) on {len(picks)} of the instances (every 16th with n <= 2^13: n = "
                          f"{[int(Js[i].shape[0]) for i in picks]}), measured {csec:.1f} s for {snnz} stored "
                          f"entries; value = that x {nnz} / {snnz} (extrapolated). Single-threaded; {cores} cores "
                          f"visible")}
        log(f"cpu sample: {csec:.3f}s on {len(picks)} instances")
    line = {
        "metric": METRIC, "value": sec32, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": sec32 * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": w["text"], "alphas": args.alphas, "predictor": pred, "instances": len(Js),
                   "n_total": N, "nnz_total": nnz,
                   "layout": ("packed block-diagonal CSR; the autotuner picks plain per-instance tiles (tile size, lanes "
                              "per row), the instance-staged layout (rows per workgroup, lanes per row) or, in FP32, "
                              "the instance-streamed layout (rows per workgroup)"),
                   "tiles": tunes},
        "fp64_sweep_s": sec64, "fp32_vs_fp64_speedup": (sec64 / sec32) if sec64 else None,
        # (E32 - E64) / |E64| per instance: negative = FP32 found the lower energy
        "energy_gap_max": max(abs(x) for x in gaps) if gaps else None,
        "energy_gap_median": float(np.median(np.abs(gaps))) if gaps else None,
        "energy_gap_fp32_worse_max": max([0.0] + gaps) if gaps else None,
        "instances_gap_above_1e-3": int(sum(abs(x) > 1e-3 for x in gaps)) if gaps else None,
        "converged_fraction": conv, "cold_retries": sw32.retries,
        "lockstep_sweep_s": lock32,
        "batched_steps": out32.steps, "operator_applications": out32.matvecs,
        "one_at_a_time_sample": {"instances": sample, "seconds": solo,
                                 "extrapolated_all_s": solo * len(Js) / sample if solo else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": {"instance-staged": "spmv_inst_kernel<float> (instance slices of the gathered vector "
                                                   "staged in LDS)",
                                "instance-streamed": "spmv_istream_kernel (each workgroup's rows of one instance "
                                                     "streamed by column, FP64 row sums in LDS)"}.get(
                                    tunes[str(torch.float32)]["kind"],
                                    "spmv_tile_kernel<float> (fused scaled SpMV over the packed CSR)"),
                     "bytes_per_launch": spmv_bytes, "us_per_launch": spmv_sec * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def make_graph(w):
    """The workload's graph. ER 2^22 and BA 2^20 come from the numpy generators (the instances every
    earlier measurement here was taken on); BA 2^25 from libspecann's parallel host generator (the numpy
    one takes half a minute there)."""
    from specann_amd import graphs
    native = w.get("native", False)
    if w["kind"] == "er":
        f = graphs.erdos_renyi_native if native else graphs.erdos_renyi
        return f(w["n"], w["deg"] / (w["n"] - 1), seed=w["seed"], kind=w["coupling"])
    f = graphs.barabasi_albert_native if native else graphs.barabasi_albert
    return f(w["n"], w["m"], seed=w["seed"], kind=w["coupling"])


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_roofline(prob, reps=30):
    """Average duration of the fused scaled SpMV kernel (the Lanczos step's launch), HIP events."""
    import torch
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    n, nnz = prob.n, prob.nnz
    s = torch.empty(n, dtype=prob.dtype, device=prob.device)
    prob.alpha_scaling(0.5, 1e-12, s)
    g = torch.empty(n, dtype=prob.dtype, device=prob.device)
    prob.fill_normal(7, 0, g)
    y = torch.empty_like(g)
    fn = f"sa_spmv_gather_{prob.sfx}"
    for _ in range(3):
        call(fn, prob.csr_ref, ptr(s), ptr(g), 0.0, None, 1.0, ptr(y), stream())
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        call(fn, prob.csr_ref, ptr(s), ptr(g), 0.0, None, 1.0, ptr(y), stream())
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    esz = 4 if prob.sfx == "f32" else 8
    alg = (4 + esz) * nnz + 4 * (n + 1) + 3 * esz * n
    return sec, alg


def spmv_kernel_name(choice, tname):
    """The SpMV kernel(s) one application of the chosen layout launches."""
    if choice.get("layout") == "blocked":
        return f"spmv_block_kernel<{tname}> + combine (LDS-blocked fused scaled SpMV)"
    if choice.get("layout") == "cstream":
        return "spmv_cstream_kernel + combine (column-streamed fused scaled SpMV)"
    if choice.get("layout") == "rstream":
        return "spmv_cstream_kernel<*, true> (row-owned column-streamed fused scaled SpMV, no combine)"
    if choice.get("layout") == "pblock":
        return "pb_expand_kernel + pb_reduce_kernel (two-pass fused scaled SpMV)"
    if choice.get("npanels"):
        return f"spmv_panel_kernel<{tname}> + combine (XCD-panel fused scaled SpMV)"
    return f"spmv_tile_kernel<{tname}> (fused scaled CSR SpMV)"


def measured_traffic(workload, dtype, layout):
    """HBM bytes per SpMV application measured by the FETCH_SIZE / WRITE_SIZE counter passes committed
    under profiles/ (scaled by a copy kernel), if they were taken on this workload and layout."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"spmv_traffic_{workload}_{dtype}.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    want = {"npanels": layout["npanels"], "tile_nnz": layout["tile_nnz"]}
    if layout.get("layout") == "blocked":
        want = {"layout": "blocked", **{k: layout[k] for k in ("brows", "bcols", "bnr", "bnc")}}
    if layout.get("layout") in ("cstream", "rstream", "pblock"):
        want = {"layout": layout["layout"], **{k: layout[k] for k in ("brows", "bnr")}}
    if rec.get("layout") != want:
        return None
    return float(rec["traffic"])


def copy_calibration(n=1 << 26, reps=10):
    """A plain copy of n floats (sa_scale_copy_f32 with scale 1: 4n bytes read, 4n written), the known-size
    kernel the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch
    from specann_amd._lib import call
    from specann_amd.device import ptr, stream
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    call("sa_scale_copy_f32", n, ptr(x), 1.0, ptr(y), stream())
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        call("sa_scale_copy_f32", n, ptr(x), 1.0, ptr(y), stream())
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps, 8 * n


def cpu_sample(J, prob32, alphas, candidates, alpha_index=64):
    """The reference's CPU path (scipy ARPACK eigsh on S_rho with the diagonal predictor, then the numpy
    local search) timed on one warm-started alpha of the schedule, scaled to the sweep.

    y_last at alpha_index - 1 comes from the GPU (so no cold CPU solve is needed); the CPU then runs the
    diagonal predictor, its residual gate, ARPACK's eigsh(which='SA', tol 1e-6, ncv 33) and the
    local search of the rounded candidate. Extrapolation: that alpha's eigensolve seconds x the sweep's
    alphas + the local search seconds x the sweep's distinct candidates."""
    import torch
    import scipy.sparse.linalg as spla
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import specann_oracle as O
    from specann_amd.eigsh import ScaledOperator, eigsh
    a = O.alpha_values(alphas)
    i = min(max(1, alpha_index), len(a) - 1)
    s_prev = torch.empty(prob32.n, dtype=torch.float32, device=prob32.device)
    prob32.alpha_scaling(a[i - 1], 1e-12, s_prev)
    _, v, _ = eigsh(ScaledOperator(prob32, s_prev, None), which="LA", tol=1e-6, maxiter=4000)
    y_last = v[:, 0].double().cpu().numpy()
    y_last /= np.linalg.norm(y_last)
    Jn, d, c = O.normalize(J, np.float32)
    Jint = O.integer_couplings(J)
    n = J.shape[0]
    t0 = time.perf_counter()
    s, _, g = O.alpha_scaling(d, a[i])
    rho = O.shift_rho(None, g)
    y0 = O.diagonal_predictor(y_last, d, a[i] - a[i - 1])
    counter = [0]

    def mv(x):
        counter[0] += 1
        x = x.astype(np.float32)
        return (np.float32(rho) * x - s * (Jn @ (s * x))).astype(np.float32)

    v1 = O.n_matvec(Jn, s, y0)
    gate = float(np.linalg.norm(v1 - (y0 @ v1) * y0))
    start = y0 if gate <= 1e-3 else y_last
    op = spla.LinearOperator((n, n), matvec=mv, dtype=np.float32)
    try:
        w, V = spla.eigsh(op, k=1, which="SA", tol=1e-6, maxiter=40, v0=start.astype(np.float32), ncv=33)
    except spla.ArpackNoConvergence as e:
        V = e.eigenvectors
    t_eig = time.perf_counter() - t0
    spins = O.round_spins(s.astype(np.float64) * V[:, 0])
    t1 = time.perf_counter()
    res = O.local_search(Jint, spins[None, :], 42)
    t_ls = time.perf_counter() - t1
    return {"eig_s": t_eig, "matvecs": counter[0] + 1, "ls_s": t_ls, "ls_passes": res.total_passes,
            "alpha": a[i], "value": t_eig * alphas + t_ls * candidates}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed sweeps first (first-use allocations and code loading stay out of the timing)")
    ap.add_argument("--workload", default="er22", choices=sorted(WORKLOADS))
    ap.add_argument("--alphas", type=int, default=128)
    ap.add_argument("--pred", default="tangent", choices=["diagonal", "tangent", "secant", "prev"],
                    help="predictor of the timed FP32 sweep (the sweeps beside it use the diagonal one)")
    ap.add_argument("--tangent-steps", type=int, default=2)
    ap.add_argument("--skip-fp32-diagonal", action="store_true",
                    help="skip the FP32 diagonal-predictor sweep (the Lanczos applications the predictor saves)")
    ap.add_argument("--ncv", type=int, default=None, help="Lanczos basis size (default: the reference's min(33, n-1))")
    ap.add_argument("--rows-out", default=None, help="write the FP32 sweep's per-alpha rows (JSON) to this file")
    ap.add_argument("--keep", type=int, default=None,
                    help="Ritz vectors a thick restart keeps (default: min(12, ncv / 2))")
    ap.add_argument("--lanczos-graphs", type=int, default=None, choices=(0, 1),
                    help="replay Lanczos step chunks from HIP graphs (default: only for n <= 2^18)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--reorth", default="selective", choices=("selective", "full"),
                    help="Lanczos reorthogonalisation: selective (default) or full at every step (the reference's)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-e2e", action="store_true", help="skip the end-to-end solve() timing")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV timing loop and a copy kernel of known size (for counter profiling)")
    ap.add_argument("--c3-sample", type=int, default=16,
                    help="batch workload: instances also swept one at a time, for comparison")
    ap.add_argument("--skip-lockstep", action="store_true",
                    help="batch workload: skip the lockstep schedule's sweep (every instance at the same alpha)")
    ap.add_argument("--layout", default=None, metavar="NPANELS,TILE",
                    help="fixed SpMV layout instead of the autotuner's choice, e.g. 8,1024, blocked (LDS-blocked "
                         "layout; blocked:BROWS,BCOLS for a given geometry), cstream (column-streamed), rstream "
                         "(column-streamed, one workgroup per row block) or pblock (two-pass), "
                         "each also as NAME-balanced (row blocks balanced by entries)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import torch
    from specann_amd.device import DeviceProblem, require_gpu
    from specann_amd.spectral_annealing import SpectralAnnealing
    require_gpu()
    w = WORKLOADS[args.workload]
    if w["kind"] == "batch":
        return run_batch(args, w)
    t0 = time.time()
    J = make_graph(w)
    log(f"graph {args.workload}: n={J.shape[0]} nnz={J.nnz} generated in {time.time() - t0:.1f}s")
    prob32 = DeviceProblem(J, torch.float32)
    if args.layout and args.layout.split("-")[0] in ("cstream", "rstream", "pblock"):
        name, bal = args.layout.split("-")[0], args.layout.endswith("-balanced")
        build = {"cstream": prob32.build_cstream, "rstream": prob32.build_rstream, "pblock": prob32.build_pblock}[name]
        if not build(balanced=bal):
            raise SystemExit(f"the {args.layout} layout cannot hold this matrix")
        tune32 = {"choice": {"layout": name, "npanels": 8 if name == "cstream" else 0,
                             "tile_nnz": prob32.tile_nnz, "seconds": None,
                             **{k: v for k, v in prob32.block.items() if k not in ("layout", "chunks", "column_blocks")}},
                  "tried": []}
    elif args.layout and args.layout.startswith("blocked"):
        geom = [int(v) for v in args.layout.split(":")[1].split(",")] if ":" in args.layout else [None, None]
        if not prob32.build_blocks(brows=geom[0], bcols=geom[1]):
            raise SystemExit("the LDS-blocked layout cannot hold this matrix")
        tune32 = {"choice": {"layout": "blocked", "npanels": 8, "tile_nnz": prob32.tile_nnz, "seconds": None,
                             **{k: v for k, v in prob32.block.items() if k != "entries"}}, "tried": []}
    elif args.layout:
        npanels, tile = (int(v) for v in args.layout.split(","))
        prob32.set_tiling(tile)
        if npanels:
            prob32.build_panels(npanels)
        tune32 = {"choice": {"npanels": npanels, "tile_nnz": tile, "seconds": None}, "tried": []}
    else:
        tune32 = prob32.autotune()
    torch.cuda.synchronize()
    log(f"uploaded FP32 problem; SpMV layout: {tune32['choice']} of {tune32['tried']}")

    spmv_sec, spmv_bytes = spmv_roofline(prob32)
    achieved = spmv_bytes / spmv_sec / 1e9
    log(f"spmv f32: {spmv_sec * 1e6:.1f} us/launch, {achieved:.0f} GB/s algorithmic")
    if args.spmv_only:
        copy_sec, copy_bytes = copy_calibration()
        print(json.dumps({"spmv_us": spmv_sec * 1e6, "spmv_gbs": achieved, "spmv_alg_bytes": spmv_bytes,
                          "copy_us": copy_sec * 1e6, "copy_bytes": copy_bytes}))
        return 0

    def solver(precision, pred="diagonal"):
        return SpectralAnnealing("bench", "synthetic", random_seed=42, precision=precision, verbose=False,
                                 num_alphas=args.alphas, pred_mode=pred, tangent_steps=args.tangent_steps,
                                 eigsh_ncv=args.ncv, reorth=args.reorth, eigsh_keep=args.keep,
                                 lanczos_graphs=args.lanczos_graphs)

    # the FP32 eigenpairs of the (untimed) first warm-up sweep: true residuals now, vectors kept for the
    # angle to the FP64 sweep's eigenvectors at the same alphas
    pairs = {"res32": {}, "vec32": {}, "vec64": {}}

    def keep32(ai, alpha, prob, lz, s, scale, eo):
        pairs["res32"][ai] = SpectralAnnealing._true_residual(prob, lz, s, eo.lambda1, 1.0 / abs(scale))
        pairs["vec32"][ai] = lz.Y * scale

    for i in range(args.warmup):
        sw = solver(32, args.pred)
        if i == 0 and not args.skip_fp64:
            sw.eigvec_hook = keep32
        out = sw.run_on_device(prob32)
        log(f"warmup {i}: {out.sweep_time + out.post_sweep_time:.2f}s")
    torch.cuda.synchronize()
    times, outs = [], []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = solver(32, args.pred).run_on_device(prob32)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        outs.append(out)
        log(f"step {i}: {times[-1]:.3f}s (sweep {out.sweep_time:.3f}s, local search {out.post_sweep_time:.3f}s, "
            f"{out.matvecs} Lanczos + {out.tangent_matvecs} tangent operator applications, {out.num_candidates} "
            f"candidates, best {out.best_energy})")
    if args.rows_out:
        keys = ("alpha_index", "alpha", "eigsh_iterations", "pred_residual", "pred_accepted", "lanczos_restarts",
                "lambda1", "residual_estimate", "tangent_iterations", "candidate", "eigsh_retry")
        with open(args.rows_out, "w") as f:
            json.dump([{k: r.get(k) for k in keys} for r in outs[-1].rows], f)
    sec32 = float(np.mean(times))
    e32 = outs[-1].best_energy
    matvecs = outs[-1].matvecs + outs[-1].tangent_matvecs
    conv = float(np.mean([bool(r['eigsh_converged']) for r in outs[-1].rows]))

    sec32d = e32d = apps32d = None
    if args.pred != "diagonal" and not args.skip_fp32_diagonal:
        torch.cuda.synchronize()
        t = time.perf_counter()
        outd = solver(32, "diagonal").run_on_device(prob32)
        torch.cuda.synchronize()
        sec32d, e32d, apps32d = time.perf_counter() - t, outd.best_energy, outd.matvecs
        log(f"fp32 diagonal sweep: {sec32d:.3f}s, best {e32d}, {apps32d} operator applications")

    cores = len(os.sched_getaffinity(0))
    cpu = None
    if not args.skip_cpu and J.shape[0] <= (1 << 14):
        # small graphs: the whole CPU path of the sweep (scipy ARPACK on S_rho, numpy local search)
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import specann_oracle as O
        t = time.perf_counter()
        cres = O.cpu_sweep(J, precision=32, num_alphas=args.alphas, pred_mode="diagonal", seed=42)
        csec = time.perf_counter() - t
        cpu = {"value": csec, "unit": "s", "cores": 1, "kind": "port",
               "sample": (f"the full {args.alphas}-alpha sweep on the CPU path (scipy ARPACK eigsh on the S_rho "
                          f"operator, diagonal predictor, numpy local search; {cores} cores visible, "
                          f"scipy's CSR product and the local search single-threaded), best energy "
                          f"{cres.best_energy}")}
        log(f"cpu sweep: {csec:.3f}s, best {cres.best_energy}")
    elif not args.skip_cpu:
        k = outs[-1].num_candidates or 1
        c = cpu_sample(J, prob32, args.alphas, k)
        cpu = {"value": c["value"], "unit": "s (extrapolated from a measured alpha)", "cores": 1, "kind": "port",
               "sample": (f"one warm-started alpha ({c['alpha']:.4f}, mid-schedule) of the CPU path on the same "
                          f"graph, measured: diagonal predictor + gate + scipy ARPACK eigsh(S_rho, which='SA', "
                          f"tol 1e-6, ncv 33) {c['eig_s']:.1f} s over {c['matvecs']} operator applications, numpy "
                          f"local search of its candidate {c['ls_s']:.1f} s ({c['ls_passes']} passes); value = "
                          f"{c['eig_s']:.1f} s x {args.alphas} alphas + {c['ls_s']:.1f} s x {k} distinct "
                          f"candidates of the GPU sweep (extrapolated). scipy's CSR product and the local search "
                          f"run single-threaded; {cores} cores visible, OMP_NUM_THREADS="
                          f"{os.environ.get('OMP_NUM_THREADS')}")}
        log(f"cpu sample: {c}")

    e2e = None
    if not args.skip_e2e:
        # the reference's `time` column: SpectralAnnealing.solve(J) from the host matrix (upload,
        # normalisation, SpMV layout and autotune, sweep, local search), result files in a scratch dir
        import tempfile
        here = os.getcwd()
        with tempfile.TemporaryDirectory() as tmp:
            os.chdir(tmp)
            try:
                torch.cuda.synchronize()
                t = time.perf_counter()
                e_e2e, _, _ = solver(32, args.pred).solve(J)
                torch.cuda.synchronize()
                e2e = time.perf_counter() - t
            finally:
                os.chdir(here)
        log(f"end-to-end solve(): {e2e:.3f}s, best {e_e2e}")

    sec64 = e64 = apps64 = None
    if not args.skip_fp64:
        del prob32
        torch.cuda.empty_cache()
        prob64 = DeviceProblem(J, torch.float64)
        tune64 = prob64.autotune()
        log(f"FP64 autotuned SpMV: {tune64['choice']}")
        # untimed FP64 warm-up (code loading, first-use allocations), as the FP32 sweeps had
        w64 = SpectralAnnealing("bench", "synthetic", random_seed=42, precision=64, verbose=False,
                                num_alphas=min(16, args.alphas), reorth=args.reorth)
        w64.run_on_device(prob64)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s64 = solver(64)
        if pairs["vec32"]:
            s64.eigvec_hook = lambda ai, alpha, prob, lz, s, scale, eo: pairs["vec64"].__setitem__(ai, lz.Y * scale)
        out64 = s64.run_on_device(prob64)
        torch.cuda.synchronize()
        sec64 = time.perf_counter() - t
        e64 = out64.best_energy
        apps64 = out64.matvecs
        log(f"fp64 sweep: {sec64:.3f}s, best {e64}, {out64.matvecs} operator applications")
        del prob64

    def gap(a, b):
        return abs(a - b) / abs(b) if (a is not None and b) else None

    eigenpairs = None
    if pairs["vec32"] and pairs["vec64"]:
        # angle between the FP32 and FP64 unit eigenvectors of each alpha: 2 asin(|u -+ v| / 2)
        angles = {}
        for ai, v64 in pairs["vec64"].items():
            if ai in pairs["vec32"]:
                u = pairs["vec32"][ai].double()
                d = min(torch.linalg.vector_norm(u - v64).item(), torch.linalg.vector_norm(u + v64).item())
                angles[ai] = 2.0 * math.asin(min(1.0, d / 2.0))
        res = np.array(list(pairs["res32"].values()))
        ang = np.array(list(angles.values()))
        worst = max(angles, key=angles.get) if angles else None
        eigenpairs = {"alphas_compared": len(angles), "fp32_true_residual_max": float(res.max()),
                      "fp32_true_residual_median": float(np.median(res)),
                      "fp32_residual_tol": 1e-6, "angle_to_fp64_max": float(ang.max()) if ang.size else None,
                      "angle_to_fp64_median": float(np.median(ang)) if ang.size else None,
                      "alpha_index_of_max_angle": worst,
                      "angles_above_1e-3": int((ang > 1e-3).sum()),
                      "note": "FP32 tangent sweep (untimed warm-up) against the FP64 diagonal sweep, same alphas; "
                              "residual = ||N y - lambda y|| / ||y|| of the FP32 eigenpair, N y taken by the FP32 "
                              "SpMV (its rounding, ~1e-7 of ||y||, comes on top of the Lanczos stopping test's "
                              "tolerance fp32_residual_tol)"}
        pairs.clear()

    line = {
        "metric": METRIC, "value": sec32, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec32 * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": w["text"], "alphas": args.alphas, "predictor": args.pred,
                   "local_search": "batched deterministic", "precision": "FP32 vectors, FP64 Lanczos scalars",
                   "spmv_layout": tune32["choice"]},
        # the north star's comparison: FP32 with this run's predictor against plain FP64 (diagonal)
        "fp64_diagonal_sweep_s": sec64,
        f"fp32_{args.pred}_vs_fp64_diagonal_speedup": (sec64 / sec32) if sec64 else None,
        f"energy_gap_fp32_{args.pred}_vs_fp64_diagonal": gap(e32, e64),
        # like for like: the same (diagonal) predictor in both precisions, both warmed up
        "fp32_diagonal_sweep_s": sec32d,
        "fp32_vs_fp64_speedup_diagonal": (sec64 / sec32d) if (sec64 and sec32d) else None,
        "energy_gap_diagonal_fp32_vs_fp64": gap(e32d, e64),
        "fp64_operator_applications": apps64, "fp32_diagonal_operator_applications": apps32d,
        "best_energy_fp32": e32, "best_energy_fp32_diagonal": e32d, "best_energy_fp64_diagonal": e64,
        "lanczos_applications_saved_vs_diagonal": (apps32d - matvecs) if apps32d is not None else None,
        "operator_applications": matvecs, "lanczos_applications": outs[-1].matvecs,
        "tangent_applications": outs[-1].tangent_matvecs, "converged_fraction": conv,
        "solve_end_to_end_s": e2e,
        "reorth": args.reorth, "orth_full_steps": outs[-1].orth_full_steps,
        "orth_selective_steps": outs[-1].orth_selective_steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, "f32", tune32["choice"]),
                     "kernel": spmv_kernel_name(tune32["choice"], "float"),
                     "bytes_per_launch": spmv_bytes, "us_per_launch": spmv_sec * 1e6},
        "cpu_baseline": cpu,
        "eigenpairs": eigenpairs,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
