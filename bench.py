#!/usr/bin/env python3
"""HODLR solves/sec on one MI355X: mixed precision + FP64 refinement vs plain FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 5] [--no-cpu]

Workload (default, --config 5): N = 2^20 rows, 10 levels of 1024-row leaves,
rank 128 with decaying spectrum U = Q_U diag(10^(-k/32)), V = Q_V, leaf blocks
randn/32 + 4 I, 32 right-hand sides uniform [0,1), seed 1245667, generated on
the host and resident in HBM before timing.

One step = one complete solve of the system for all its right-hand sides:
  * mixed (the reported value): FP32 factor+solve (MFMA f32), then FP64
    residuals and FP32 correction solves until max_j ||b_j - A x_j|| / ||b_j||
    <= 1e-12, host checks included;
  * fp64 (reported beside it): one fused FP64 factor+solve pass, the
    reference's bfs_solve; its residual is checked after the timed region.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fastsolver_amd import gen, hodlr  # noqa: E402

METRIC = "HODLR solves/sec at rel. residual ≤1e-12: mixed+refine vs FP64, 1 MI355X"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}
TOL = 1e-12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    hodlr.device_synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    hodlr.device_synchronize()
    return (time.perf_counter() - t0) / steps, out


def measured_traffic(cat, cfg):
    """HBM bytes per launch of `cat` on workload cfg<cfg>, from the calibrated
    rocprofv3 PMC summaries (lines 'TRAFFIC <category> cfg<k> <bytes>')."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for f in sorted(os.listdir(pdir)):
        if not f.endswith((".md", ".txt")):
            continue
        for line in open(os.path.join(pdir, f)):
            w = line.split()
            if len(w) == 4 and w[0] == "TRAFFIC" and w[1] == cat and w[2] == f"cfg{cfg}":
                return float(w[3]), "profiles/" + f
    return None, None


def roofline_entry(prof, cat, precision):
    d = prof[cat]
    if d["launches"] == 0 or d["ms"] <= 0:
        return None
    t = d["ms"] / 1e3
    peak_f = PEAK["fp64"] if precision == "fp64" else PEAK["fp32"]
    ai = d["flops"] / d["bytes"] if d["bytes"] > 0 else float("inf")
    ridge = peak_f / PEAK["hbm"]
    if ai >= ridge:
        ach = d["flops"] / t / 1e12
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak_f / 1e12, "unit": "TFLOP/s",
                "frac": round(ach * 1e12 / peak_f, 4), "traffic": None, "kernel": cat,
                "launches": d["launches"], "avg_launch_ms": round(d["ms"] / d["launches"], 4),
                "flops_per_launch": d["flops"] / d["launches"], "bytes_per_launch": d["bytes"] / d["launches"]}
    ach = d["bytes"] / t / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(ach * 1e9 / PEAK["hbm"], 4), "traffic": None, "kernel": cat, "launches": d["launches"],
            "avg_launch_ms": round(d["ms"] / d["launches"], 4), "flops_per_launch": d["flops"] / d["launches"],
            "bytes_per_launch": d["bytes"] / d["launches"]}


def cpu_baseline(spec, X, V, K, threads):
    """The oracle (reference task bodies over MKL) on the host cores: one full
    solve of the same workload, tasks of a level spread over OpenMP threads."""
    import oracle
    Xc = X.copy(order="F")
    Kc = K.copy()
    t0 = time.perf_counter()
    rc = oracle.hodlr_solve(spec.as_dict(), Xc, V, Kc, order=1, nthreads=threads)
    t = time.perf_counter() - t0
    del Kc
    return rc, t, Xc[:, :spec.nrhs].copy(order="F"), oracle.lapack_name()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--maxiter", type=int, default=20)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--tune", type=int, default=2, help="autotuner rounds before timing (0 = off)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        return 2

    spec = gen.config(args.config)[0]
    nl, md, xc, vc, kp = gen.tree_info(spec)
    log(f"[bench] config {args.config}: N={spec.n} levels={md} leaves={nl} rank={spec.ranks} c={spec.nrhs}")
    t0 = time.perf_counter()
    X, V, K = gen.generate(spec)
    log(f"[bench] generated inputs on the host in {time.perf_counter() - t0:.1f}s "
        f"({(X.nbytes + V.nbytes + K.nbytes) / 1e9:.1f} GB)")
    t0 = time.perf_counter()
    ds = hodlr.DeviceSystem.from_host(spec, X, V, K)
    log(f"[bench] uploaded in {time.perf_counter() - t0:.1f}s on {hodlr.device_name()}")

    # ------------------------------------------------------------ plain FP64
    p64 = hodlr.Plan([ds], hodlr.PREC_FP64)
    info64 = p64.info()
    if args.tune:
        big = p64.autotune(args.tune)
        log(f"[bench] fp64 autotune: {big} launches on 128x128 tiles")
    if args.graph:
        p64.use_graph(True)
    t64, _ = timed_steps(p64.factor_solve, args.steps, args.warmup)
    rc, li, ni = p64.status()
    relres64, _ = p64.residual()
    x64 = ds.solution()
    log(f"[bench] fp64: {t64 * 1e3:.2f} ms/solve, relres {relres64:.2e}, status {rc}")
    p64.profile(None)
    p64.factor_solve()
    prof64 = p64.profile_read()
    p64.profile([])
    p64.close()
    del p64

    # ------------------------------------------------------------ mixed + refine
    p32 = hodlr.Plan([ds], hodlr.PREC_FP32)
    ws32 = p32.info().workspace_bytes
    if args.tune:
        big = p32.autotune(args.tune)
        log(f"[bench] mixed autotune: {big} launches on 128x128 tiles")
    if args.graph:
        p32.use_graph(True)
    rep_box = {}

    def mixed():
        rep_box["r"] = p32.solve(TOL, args.maxiter)
        return rep_box["r"]

    # warm up, then find the dominant kernel of one profiled (untimed) step
    for _ in range(args.warmup):
        mixed()
    p32.profile(None)
    mixed()
    prof_all = p32.profile_read()
    p32.profile([])
    p32.profile_reset()
    dom = max(prof_all, key=lambda k: prof_all[k]["ms"])
    # timed steps: only the dominant kernel's launches carry event pairs
    p32.profile([dom])
    tmix, rep = timed_steps(mixed, args.steps, 0)
    prof_dom = p32.profile_read()
    p32.profile([])
    xmix = ds.solution()
    log(f"[bench] mixed: {tmix * 1e3:.2f} ms/solve, {rep.iterations} refinement sweeps, "
        f"relres {rep.relres:.2e} (first {rep.relres_first:.2e}); dominant kernel {dom}")
    roof = roofline_entry(prof_dom, dom, "fp32" if dom not in ("node_gemm", "node_lu", "node_trsm", "matvec",
                                                                "reduce_sum", "norm", "solve_node") else "fp64")
    if roof is not None:
        roof["traffic"], roof["traffic_source"] = measured_traffic(dom, args.config)
    value = 1.0 / tmix
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(tmix * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (host Philox4x32-10, seed 1245667)",
        "config": {"workload": f"cfg{args.config}: N={spec.n}, {md} levels, {nl} leaves of {spec.n // nl}, "
                               f"rank {spec.ranks[0] if len(spec.ranks) == 1 else spec.ranks}, {spec.nrhs} rhs, "
                               + ("decay 10^(-k/%g)" % spec.decay if spec.uv_mode == gen.UV_DECAY else "uniform U,V"),
                   "n": spec.n, "leaves": nl, "rank": spec.ranks, "nrhs": spec.nrhs, "tol": TOL},
        "mixed": {"ms": round(tmix * 1e3, 3), "iterations": rep.iterations, "relres": rep.relres,
                  "relres_first": rep.relres_first},
        "fp64": {"ms": round(t64 * 1e3, 3), "solves_per_s": round(1.0 / t64, 3), "relres": relres64,
                 "gflops": round(info64.flops_factor / t64 / 1e9, 1)},
        "speedup_mixed_vs_fp64": round(t64 / tmix, 3),
        "flops_factor": info64.flops_factor,
        "launches_fp64_pass": info64.launches_factor,
        "workspace_gb": {"fp64": round(info64.workspace_bytes / 1e9, 2), "mixed": round(ws32 / 1e9, 2)},
        "roofline": roof,
        "profile_fp64_ms": {k: round(v["ms"], 3) for k, v in prof64.items() if v["launches"]},
        "profile_mixed_ms": {k: round(v["ms"], 3) for k, v in prof_all.items() if v["launches"]},
        "cpu_baseline": None,
    }
    p32.close()
    del p32, ds

    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"[bench] cpu baseline: oracle on {threads} host threads ...")
        rc, tcpu, xref, lapack = cpu_baseline(spec, X, V, K, threads)
        e64 = float(np.linalg.norm(x64 - xref) / np.linalg.norm(xref))
        emix = float(np.linalg.norm(xmix - xref) / np.linalg.norm(xref))
        out["fp64"]["err_vs_oracle"] = e64
        out["mixed"]["err_vs_oracle"] = emix
        out["cpu_baseline"] = {"value": round(1.0 / tcpu, 5), "unit": "solves/s", "cores": threads,
                               "kind": "port", "sample": f"one full solve of the same workload (status {rc}), "
                               f"reference task bodies over {os.path.basename(lapack)}, level tasks over "
                               f"{threads} OpenMP threads with sequential BLAS",
                               "seconds": round(tcpu, 3)}
        out["fp64_vs_cpu"] = round(tcpu / t64, 2)
        log(f"[bench] cpu: {tcpu:.2f} s/solve; fp64 err {e64:.2e}, mixed err {emix:.2e}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
