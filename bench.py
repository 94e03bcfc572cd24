#!/usr/bin/env python3
"""HODLR solves/sec on one MI355X: mixed precision + FP64 refinement vs plain FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 5] [--no-cpu]

--config k picks workload family k of fastsolver_amd.gen.config:
  1  N = 16384, 64 leaves of 256, rank 30, 1 rhs (plain FP64 family)
  2  N = 131072, 256 leaves of 512, rank 64, 4 rhs (plain FP64 family)
  3  N = 262144, 256 leaves of 1024, rank 100 decaying, 8 rhs (FP32 then BF16 + refinement)
  4  512 uneven systems in one plan, one step = all 512 solves (FP32 + refinement)
  5  N = 2^20, 1024 leaves of 1024, rank 128 decaying 10^(-k/32), 32 rhs (BF16 + refinement; default)
Inputs: leaf blocks randn*s + 4 I, seed 1245667, generated on the host and
resident in HBM before timing.

One step = one complete solve of every system for all its right-hand sides:
  * mixed: the family's low-precision factor+solve (MFMA), then FP64 residuals
    and correction solves until max_j ||b_j - A x_j|| / ||b_j|| <= 1e-12; the
    convergence check runs on the device and the whole solve (pass, residuals,
    sweeps) replays one captured hipGraph, with one host synchronisation;
  * fp64: one fused FP64 factor+solve pass (the reference's bfs_solve), graph
    replay; its residual is checked after the timed region.
`value` is the family's headline path (mixed for 3-5, fp64 for the plain-FP64
families 1-2); every other path is reported beside it.  The timed steps record
no events.  The per-category profile comes from a separate untimed eager step;
the dominant category's launches are then timed with HIP events around each
launch over a few eager solves.  Prints one JSON line on stdout; progress goes
to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fastsolver_amd import gen, hodlr  # noqa: E402

METRIC = "HODLR solves/sec at rel. residual ≤1e-12: mixed+refine vs FP64, 1 MI355X"
# MI355X published dense peaks (FLOP/s) and HBM3E bandwidth (B/s)
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}
TOL = 1e-12
PNAME = {hodlr.PREC_FP64: "fp64", hodlr.PREC_FP32: "fp32", hodlr.PREC_BF16: "bf16"}
# headline path per family (None: plain FP64) and the mixed paths measured beside it
FAMILY = {1: (None, [hodlr.PREC_FP32]), 2: (None, [hodlr.PREC_FP32]),
          3: (hodlr.PREC_BF16, [hodlr.PREC_FP32]), 4: (hodlr.PREC_FP32, [hodlr.PREC_BF16]),
          5: (hodlr.PREC_BF16, [hodlr.PREC_FP32])}
CPU_SHARE = 16   # host threads one GPU's share of the box provides


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


def measured_traffic(cat, cfg, prec):
    """HBM bytes per launch of `cat` on workload cfg<cfg> in precision prec, from
    the calibrated rocprofv3 PMC summaries (lines 'TRAFFIC <cat> cfg<k> <prec> <bytes>')."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for f in sorted(os.listdir(pdir)):
        if not f.endswith((".md", ".txt")):
            continue
        for line in open(os.path.join(pdir, f)):
            w = line.split()
            if len(w) == 5 and w[0] == "TRAFFIC" and w[1] == cat and w[2] == f"cfg{cfg}" and w[3] == prec:
                return float(w[4]), "profiles/" + f
    return None, None


def category_math(cat, prec):
    """Arithmetic type of a profile category in a plan of precision `prec`."""
    wt = "fp64" if prec == hodlr.PREC_FP64 else "fp32"
    mt = PNAME[prec]
    if cat in ("leaf_copy", "copy", "norm", "reduce_sum"):
        return None
    if cat in ("lu_panel", "lu_rowblock", "lu_misc"):
        return wt
    if cat in ("node_gemm", "node_lu", "node_trsm", "solve_node") or cat.startswith("matvec"):
        return "fp64"
    return mt


def roofline_entry(d, cat, prec):
    """Bound, achieved rate and fraction of one category (algorithmic work over
    measured time)."""
    if d["launches"] == 0 or d["ms"] <= 0:
        return None
    t = d["ms"] / 1e3
    m = category_math(cat, prec)
    peak_f = PEAK[m] if m else None
    ai = d["flops"] / d["bytes"] if d["bytes"] > 0 else float("inf")
    e = {"kernel": cat, "math": m, "launches": d["launches"], "ms": round(d["ms"], 4),
         "avg_launch_ms": round(d["ms"] / d["launches"], 4), "flops_per_launch": d["flops"] / d["launches"],
         "bytes_per_launch": d["bytes"] / d["launches"], "ai_flop_per_byte": round(ai, 2)}
    floor_s = max(d["flops"] / peak_f if peak_f else 0.0, d["bytes"] / PEAK["hbm"])
    e["floor_ms"] = round(floor_s * 1e3, 4)
    if peak_f and ai >= peak_f / PEAK["hbm"]:
        ach = d["flops"] / t
        e.update(bound="mfma", achieved=round(ach / 1e12, 3), peak=peak_f / 1e12, unit="TFLOP/s",
                 frac=round(ach / peak_f, 4))
    else:
        ach = d["bytes"] / t
        e.update(bound="hbm", achieved=round(ach / 1e9, 1), peak=PEAK["hbm"] / 1e9, unit="GB/s",
                 frac=round(ach / PEAK["hbm"], 4))
    return e


def per_pass(prof, prec, step_ms):
    """Per-category roofline of one profiled step and the total fraction:
    sum over categories of max(F/peak, B/BW) divided by the measured step time."""
    rows = {}
    floor = 0.0
    for cat, d in prof.items():
        e = roofline_entry(d, cat, prec)
        if e is None:
            continue
        rows[cat] = {k: e[k] for k in ("bound", "math", "ms", "launches", "achieved", "unit", "frac", "floor_ms")}
        floor += e["floor_ms"]
    prof_ms = sum(d["ms"] for d in prof.values())
    return rows, {"floor_ms": round(floor, 3), "profiled_ms": round(prof_ms, 3), "step_ms": round(step_ms, 3),
                  "frac": round(floor / step_ms, 4) if step_ms > 0 else None}


def hodlr_flops(sp):
    """Algorithmic flops of one fused factor+solve of system `sp`."""
    nl, md, xc, vc, kp = gen.tree_info(sp)
    r = list(sp.ranks) if len(sp.ranks) > 1 else [sp.ranks[0]] * md
    m = sp.n / nl
    rs = [0.0]
    for x in r:
        rs.append(rs[-1] + x)
    f = nl * ((2.0 / 3.0) * m ** 3 + 2.0 * m * m * (sp.nrhs + rs[md]))
    for l in range(md):
        f += 2.0 * sp.n * r[l] * (sp.nrhs + rs[l + 1]) + 2.0 * sp.n * r[l] * (sp.nrhs + rs[l])
        f += 2 ** l * ((8.0 / 3.0) * r[l] ** 3 + 6.0 * r[l] ** 2 * (sp.nrhs + rs[l]))
    return f


def workload(cfg, specs):
    sp = specs[0]
    if len(specs) > 1:
        return {"workload": f"cfg{cfg}: {len(specs)} independent systems, N in [{min(s.n for s in specs)}, "
                            f"{max(s.n for s in specs)}] (sum {sum(s.n for s in specs)}), leaves <= {sp.threshold}, "
                            f"per-level ranks in [{min(min(s.ranks) for s in specs)}, "
                            f"{max(max(s.ranks) for s in specs)}], 1..{max(s.nrhs for s in specs)} rhs",
                "systems": len(specs), "sum_n": sum(s.n for s in specs), "tol": TOL}
    nl, md, xc, vc, kp = gen.tree_info(sp)
    return {"workload": f"cfg{cfg}: N={sp.n}, {md} levels, {nl} leaves of {sp.n // nl}, "
                        f"rank {sp.ranks[0] if len(sp.ranks) == 1 else sp.ranks}, {sp.nrhs} rhs, "
                        + ("decay 10^(-k/%g)" % sp.decay if sp.uv_mode == gen.UV_DECAY else "uniform U,V"),
            "n": sp.n, "leaves": nl, "rank": sp.ranks, "nrhs": sp.nrhs, "tol": TOL}


# ------------------------------------------------------------------ CPU baseline
def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_sample_specs(cfg, specs, nmax):
    """The bounded CPU sample: the same family at N <= nmax rows (same leaf
    size, rank and nrhs), or every 32nd system of a batch."""
    if len(specs) > 1:
        return [specs[i] for i in range(0, len(specs), 32)]
    sp = specs[0]
    scale = 1
    while sp.n // scale > nmax:
        scale *= 2
    return gen.config(cfg, scale=scale) if scale > 1 else [sp]


def time_variant(oracle, samples, variant, threads, runs):
    """Median wall time of `runs` solves (after one warm-up) of the sample."""
    hosts = [gen.generate(s) for s in samples]
    times = []
    for it in range(runs + 1):
        copies = [(X.copy(order="F"), V, K.copy()) for (X, V, K) in hosts]
        t0 = time.perf_counter()
        for s, (X, V, K) in zip(samples, copies):
            if variant == "tasks":
                rc = oracle.hodlr_solve(s.as_dict(), X, V, K, order=1, nthreads=threads)
            elif variant == "blas":
                oracle.set_blas_threads(threads)
                rc = oracle.hodlr_solve(s.as_dict(), X, V, K, order=0)
            else:
                oracle.set_blas_threads(1)
                rc = oracle.hodlr_solve(s.as_dict(), X, V, K, order=1, nthreads=1)
            assert rc == 0, rc
        t = time.perf_counter() - t0
        if it > 0:
            times.append(t)
    oracle.set_blas_threads(threads)
    return statistics.median(times), times


def cpu_baseline(cfg, specs, runs):
    import oracle
    threads = min(CPU_SHARE, os.cpu_count() or 1)
    f_full = sum(hodlr_flops(s) for s in specs)
    out = {"variants": {}}
    plan = {"tasks": (threads, 1 << 17), "blas": (threads, 1 << 17), "1core": (1, 1 << 14)}
    best = None
    for variant, (nt, nmax) in plan.items():
        samples = cpu_sample_specs(cfg, specs, nmax)
        f_s = sum(hodlr_flops(s) for s in samples)
        med, times = time_variant(oracle, samples, variant, nt, runs)
        t_full = med * f_full / f_s
        out["variants"][variant] = {"threads": nt, "sample": (f"{len(samples)} of {len(specs)} systems"
                                                              if len(specs) > 1 else f"N={samples[0].n}"),
                                    "median_s": round(med, 4), "runs_s": [round(t, 4) for t in times],
                                    "gflops": round(f_s / med / 1e9, 1), "full_s": round(t_full, 3)}
        log(f"[bench] cpu {variant}: {nt} threads, sample median {med:.3f} s "
            f"({f_s / med / 1e9:.1f} GFLOP/s) -> {t_full:.2f} s per full step")
        if best is None or t_full < out["variants"][best]["full_s"]:
            best = variant
    out["best"] = best
    out["lapack"] = oracle.lapack_name()
    out["lapack_version"] = oracle.lapack_version()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle error")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU runs per variant (median)")
    ap.add_argument("--maxiter", type=int, default=20)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--tune", type=int, default=2, help="autotuner rounds before timing (0 = off)")
    ap.add_argument("--tiling", type=int, default=-1,
                    help="with --tune 0: force every tunable launch to its smaller (0) or larger (1) choice")
    ap.add_argument("--paths", default="", help="comma list of mixed paths to run (fp32,bf16); default: family's")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        return 2

    headline, beside = FAMILY[args.config]
    mixed_paths = ([headline] if headline is not None else []) + beside
    if args.paths:
        mixed_paths = [{"fp32": hodlr.PREC_FP32, "bf16": hodlr.PREC_BF16}[p] for p in args.paths.split(",")]
    specs = gen.config(args.config)
    batch = len(specs) > 1
    # host copies are kept for the systems the oracle error is taken on
    sample = set(range(0, len(specs), 32)) if batch else {0}
    t0 = time.perf_counter()
    dss, hosts, nbytes = [], {}, 0
    for i, sp in enumerate(specs):
        X, V, K = gen.generate(sp)
        nbytes += X.nbytes + V.nbytes + K.nbytes
        dss.append(hodlr.DeviceSystem.from_host(sp, X, V, K))
        if i in sample and not args.no_cpu:
            hosts[i] = (X, V, K)
        del X, V, K
    spec = specs[0]
    nl, md, xc, vc, kp = gen.tree_info(spec)
    if batch:
        log(f"[bench] config {args.config}: {len(specs)} systems, N in [{min(s.n for s in specs)}, "
            f"{max(s.n for s in specs)}], sum N {sum(s.n for s in specs)}")
    else:
        log(f"[bench] config {args.config}: N={spec.n} levels={md} leaves={nl} rank={spec.ranks} c={spec.nrhs}")
    log(f"[bench] generated and uploaded {nbytes / 1e9:.1f} GB in {time.perf_counter() - t0:.1f}s "
        f"on {hodlr.device_name()}")

    def prepare(plan):
        if not args.tune and args.tiling >= 0:
            plan.force_tiling(args.tiling)
        big = plan.autotune(args.tune) if args.tune else 0
        if args.graph:
            plan.use_graph(True)
        return big

    # ------------------------------------------------------------ plain FP64
    p64 = hodlr.Plan(dss, hodlr.PREC_FP64)
    info64 = p64.info()
    big = prepare(p64)
    t64, _ = timed_steps(p64.factor_solve, args.steps, args.warmup)
    rc, li, ni = p64.status()
    relres64, _ = p64.residual()
    x64 = {i: dss[i].solution() for i in sample}
    log(f"[bench] fp64: {t64 * 1e3:.2f} ms/solve, relres {relres64:.2e}, status {rc}, "
        f"{big} launches on the larger tiling")
    p64.use_graph(False)
    p64.profile(None)
    p64.factor_solve()
    prof64 = p64.profile_read()
    p64.profile([])
    p64.close()
    del p64
    res = {"fp64": {"ms": round(t64 * 1e3, 3), "solves_per_s": round(1.0 / t64, 3), "relres": relres64,
                    "gflops": round(info64.flops_factor / t64 / 1e9, 1),
                    "profile_ms": {k: round(v["ms"], 3) for k, v in prof64.items() if v["launches"]}}}
    res["fp64"]["passes"], res["fp64"]["total"] = per_pass(prof64, hodlr.PREC_FP64, t64 * 1e3)
    sols = {"fp64": x64}

    # ------------------------------------------------------------ mixed + refine
    dom_roof = None
    for prec in mixed_paths:
        name = PNAME[prec]
        plan = hodlr.Plan(dss, prec)
        ws = plan.info().workspace_bytes
        big = prepare(plan)
        box = {}

        def mixed():
            box["r"] = plan.solve(TOL, args.maxiter)
            return box["r"]

        for _ in range(args.warmup):
            mixed()
        # one untimed, fully profiled step (eager launches, every category timed)
        plan.profile(None)
        mixed()
        prof = plan.profile_read()
        plan.profile([])
        plan.profile_reset()
        dom = max(prof, key=lambda k: prof[k]["ms"])
        # timed steps: the shipped path (graph replay when on), no events
        mixed()   # (re-)capture outside the timed region
        tmix, rep = timed_steps(mixed, args.steps, 0)
        sols[name] = {i: dss[i].solution() for i in sample}
        log(f"[bench] {name}: {tmix * 1e3:.2f} ms/solve, {rep.iterations} refinement sweeps, relres "
            f"{rep.relres:.2e} (first {rep.relres_first:.2e}); dominant category {dom}; "
            f"{big} launches on the larger tiling")
        r = {"ms": round(tmix * 1e3, 3), "solves_per_s": round(1.0 / tmix, 3), "iterations": rep.iterations,
             "relres": rep.relres, "relres_first": rep.relres_first, "workspace_gb": round(ws / 1e9, 2),
             "dominant": dom, "profile_ms": {k: round(v["ms"], 3) for k, v in prof.items() if v["launches"]}}
        r["passes"], r["total"] = per_pass(prof, prec, tmix * 1e3)
        res[name] = r
        if prec == headline:
            # the dominant category's launches, each bracketed by HIP events on
            # the plan's stream, over eager solves after the timed region
            nr = max(1, min(args.steps, 5))
            plan.profile([dom])
            plan.profile_reset()
            for _ in range(nr):
                mixed()
            prof_dom = plan.profile_read()
            plan.profile([])
            dom_roof = roofline_entry(prof_dom[dom], dom, prec)
            if dom_roof is not None:
                dom_roof["timing"] = (f"HIP events around each launch on the plan's stream over {nr} eager solves "
                                      "after the timed (graph-replayed) steps")
                dom_roof["traffic"], dom_roof["traffic_source"] = measured_traffic(dom, args.config, name)
        plan.close()
        del plan

    if headline is None:
        value_ms, dtype = res["fp64"]["ms"], "f64"
        e = prof64.get(max(prof64, key=lambda k: prof64[k]["ms"]))
        dcat = max(prof64, key=lambda k: prof64[k]["ms"])
        dom_roof = roofline_entry(e, dcat, hodlr.PREC_FP64)
        if dom_roof is not None:
            dom_roof["timing"] = "HIP events around each launch of an untimed profiled pass"
            dom_roof["traffic"], dom_roof["traffic_source"] = measured_traffic(dcat, args.config, "fp64")
    else:
        value_ms = res[PNAME[headline]]["ms"]
        dtype = PNAME[headline] + "+f64"
    out = {
        "metric": METRIC, "value": round(1e3 / value_ms, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": value_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic (host Philox4x32-10, seed 1245667)",
        "config": workload(args.config, specs),
        "headline": PNAME[headline] if headline is not None else "fp64",
        "paths": res,
        "speedup_mixed_vs_fp64": {PNAME[p]: round(res["fp64"]["ms"] / res[PNAME[p]]["ms"], 3) for p in mixed_paths},
        "flops_factor": info64.flops_factor,
        "launches_fp64_pass": info64.launches_factor,
        "roofline": dom_roof,
        "cpu_baseline": None,
    }
    del dss

    if not args.no_cpu:
        import oracle
        threads = min(CPU_SHARE, os.cpu_count() or 1)
        # the oracle's solution of the benchmarked systems (for err_vs_oracle)
        errs = {k: 0.0 for k in sols}
        t_or = 0.0
        for i in sorted(sample):
            Xc, V, Kc = hosts[i][0].copy(order="F"), hosts[i][1], hosts[i][2].copy()
            t1 = time.perf_counter()
            rc = oracle.hodlr_solve(specs[i].as_dict(), Xc, V, Kc, order=1, nthreads=threads)
            t_or += time.perf_counter() - t1
            assert rc == 0, rc
            xref = Xc[:, :specs[i].nrhs]
            for k in sols:
                errs[k] = max(errs[k], float(np.linalg.norm(sols[k][i] - xref) / np.linalg.norm(xref)))
            del Xc, Kc
        for k in sols:
            res[k]["err_vs_oracle"] = errs[k]
        hosts.clear()
        log(f"[bench] oracle solve of the benchmarked systems: {t_or:.2f} s; errors {errs}")
        cb = cpu_baseline(args.config, specs, args.cpu_runs)
        best = cb["variants"][cb["best"]]
        hi = host_info()
        out["cpu_baseline"] = {
            "value": round(1.0 / best["full_s"], 5), "unit": "solves/s", "cores": best["threads"], "kind": "port",
            "sample": (f"fastest of three variants ({cb['best']}): the oracle (the reference's task bodies over "
                       f"{os.path.basename(cb['lapack'])}) on {best['sample']} of the same family, median of "
                       f"{args.cpu_runs} runs after 1 warm-up, scaled to the benchmarked step by algorithmic flops"),
            "seconds": best["full_s"], "variants": cb["variants"], "lapack": cb["lapack"],
            "lapack_version": cb["lapack_version"], "host": hi, "cpu_share_threads": threads,
            "oracle_full_step_s": round(t_or, 3)}
        out["fp64_vs_cpu"] = round(best["full_s"] / (res["fp64"]["ms"] / 1e3), 2)
        out["value_vs_cpu"] = round(best["full_s"] / (value_ms / 1e3), 2)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
