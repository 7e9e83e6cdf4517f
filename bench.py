"""Benchmark of the MI355X FMM force evaluation (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--config plummer4m|plummer1m|batched256|learned8m]

A step is one full force evaluation of the device-resident particle set:
bbox, the reference tree rule (depth loop, Morton keys, sort, levels,
lists), P2M, M2M, M2L, L2L, L2P and P2P, i.e. nbfmm_eval_accel_dev.  Inputs
are in HBM before the timed region; value = N * K / elapsed (particle-force
evaluations per second).

The metric is quoted at a relative force error <= 1e-6 against the FP64
direct sum.  Before timing, every precision mode is tuned over the expansion
order p, the P2P cut, the leaf capacity and the class-order tolerance
(nbody-fmm_amd/tune.py) on one sample of 4096 targets; the fastest setting
meeting the gate there is timed, and its error and the gate are then judged
on a second, disjoint sample of 4096 targets (held out from the choice).

The roofline record comes from extra evaluations with the phases run one
after another (nbfmm_set_timing mode 2): each phase, P2P included, is timed
alone with HIP events on the stream its kernels run on, and set against its
algorithmic work (flops for M2L and P2P, bytes for the others).

The default run (plummer4m) also carries sub-records for the other
workloads: the 1M plain FP64 configuration with 5 RK4 steps, the batch of 256
systems with 20 RK4 steps in both precisions, and the 8M clustered mix with
the learned order selection.

The learned order selection is timed beside the fixed orders (mode
"mixed_learned"): its model is trained in the run, on the GPU, from
FMM-vs-direct labels of another Plummer sphere of the same size (seed 11).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbody-fmm_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import nbfmm  # noqa: E402
import tune  # noqa: E402
import workloads  # noqa: E402

# MI355X published peaks: TFLOP/s (vector and dense MFMA alike) and HBM bytes/s
PEAK = {"fp64": 78.6, "fp32": 157.3}
HBM = 8.0e12
METRIC = "particle-force evals/s at rel. force err <=1e-6 vs FP64 direct; mixed vs plain"
CONFIGS = {
    # N = 4,194,304 Plummer (a = 1, M = 1, equal masses), leaf capacity 64: the accuracy-gated metric's workload
    # (tols: class_order_tol values tried; 0 = every M2L pair at p)
    "plummer4m": dict(n=1 << 22, leaf_capacity=64, leaves=(32, 64), orders=(10, 11, 12), cuts=(256, 512, 1024, 2048),
                      tols=(0.0, 0.05)),
    # N = 1,048,576 Plummer, p = 10, leaf capacity 64 (the plain FP64 configuration)
    "plummer1m": dict(n=1 << 20, leaf_capacity=64, leaves=(32, 64), orders=(10, 11, 12), cuts=(256, 512, 1024, 2048),
                      tols=(0.0, 0.05)),
    # 256 independent Plummer spheres, sizes log-uniform in [1e3, 6.4e4] (~3.5M in all), a uniform in [0.5, 2]
    "batched256": dict(n=None, leaf_capacity=32, leaves=(32,), orders=(12, 13, 14), cuts=(2048, 4096, 8192)),
    # N = 8,388,608: 32 Plummer spheres (centres uniform in [0,100)^3, a log-uniform in [0.1, 2], sphere mass
    # proportional to a, equal particle counts); learned order selection against the fixed orders
    "learned8m": dict(n=1 << 23, leaf_capacity=32, leaves=(32,), orders=(10, 11, 12, 13, 14), cuts=(1024,)),
}
PREC_NAMES = {0: "fp64", 1: "mixed"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync(ctx):
    nbfmm.hip().hipStreamSynchronize(ctx.stream)


def cpu_baseline(budget_s=20.0):
    """The reference's own FMM::update (oracle/_ref: its sources compiled -O2,
    one thread, its hardcoded L = 2 and leaf threshold 20) on a Plummer
    sample sized to the time budget."""
    import orc
    if not orc.ref_available():
        return None
    n = 8192
    while True:
        xyz, m = workloads.plummer(n, seed=42)
        t0 = time.perf_counter()
        orc.ref_fmm(xyz, m)
        dt = time.perf_counter() - t0
        # FMM cost is about linear in N: double the sample while the next one
        # stays in the budget, so the last one takes budget/2 .. budget
        if dt * 2 > budget_s or n >= (1 << 20):
            break
        n *= 2
    out = {
        "value": n / dt, "unit": "evals/s", "cores": 1, "kind": "reference",
        "sample": f"reference FMM::update (its L=2, leaf threshold 20; sources built -O2, 1 thread) on a "
                  f"Plummer sample of N={n}, one call, {dt:.2f} s",
    }
    # the same path at the headline's accuracy: the oracle's exact FMM (the C
    # restatement, -O2, 1 thread) at p = 11, leaf capacity 32, cut 1024 --
    # what a CPU needs for 1e-6, which the reference's L = 2 cannot reach
    n2 = 16384
    while True:
        xyz, m = workloads.plummer(n2, seed=42)
        t0 = time.perf_counter()
        orc.fmm_exact(xyz, m, 11, 32, G=1.0, center_mode=0, p2p_cut=1024, class_order_tol=0.05)
        dt2 = time.perf_counter() - t0
        if dt2 * 2 > budget_s or n2 >= (1 << 20):
            break
        n2 *= 2
    out["port_at_headline_settings"] = {
        "value": n2 / dt2, "unit": "evals/s", "cores": 1, "kind": "port",
        "sample": f"the oracle's exact FMM (C restatement, -O2, 1 thread) at p=11, leaf capacity 32, P2P cut "
                  f"1024, class-order tol 0.05 on a Plummer sample of N={n2}, one call, {dt2:.2f} s"}
    log(f"[bench] CPU: reference L=2 {n / dt:.3e} evals/s (N={n}); oracle p=11 {n2 / dt2:.3e} evals/s (N={n2})")
    return out


def timed(ctx, n, dx, dm, da, steps, warmup):
    for _ in range(warmup):
        ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
    sync(ctx)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
    sync(ctx)
    wall = time.perf_counter() - t0
    return wall, ctx.stats()


def rk4_direct_vs_graph(ctx, n, xyz, vel, dm, da, nsteps, dt=1e-4, off=None):
    """nsteps classic RK4 steps from (xyz, vel), run directly and replayed from
    a captured graph (nbfmm_set_graph): each run takes one untimed step first
    (the graph run's first step runs directly and captures the graph), then
    nsteps timed steps in one call.  Returns ms per step of both, the graph
    statistics and the relative difference of the two final states."""
    out = {}
    states = {}
    for graph in (0, 1):
        cx, cv = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(vel)

        def steps(k):
            if off is None:
                ctx.step_rk4_dev(n, cx.ptr, cv.ptr, da.ptr, dm.ptr, dt, k, 1)
            else:
                ctx.batch_step_rk4_dev(off, cx.ptr, cv.ptr, da.ptr, dm.ptr, dt, k, 1)
        ctx.set_graph(graph)
        try:
            before = ctx.graph_stats()
            steps(1)
            sync(ctx)
            t0 = time.perf_counter()
            steps(nsteps)
            sync(ctx)
            ms = (time.perf_counter() - t0) * 1e3 / nsteps
            after = ctx.graph_stats()
        finally:
            ctx.set_graph(0)
        states[graph] = (cx.download(ctx.stream), cv.download(ctx.stream))
        cx.free()
        cv.free()
        out["graph_ms_per_step" if graph else "direct_ms_per_step"] = ms
        if graph:
            out["graph_stats"] = {k: after[k] - before[k] for k in after}
    (xd, vd), (xg, vg) = states[0], states[1]
    out["graph_vs_direct_rel_diff"] = {
        "displacement": float(np.linalg.norm(xg - xd) / np.linalg.norm(xd - xyz)),
        "velocity_change": float(np.linalg.norm(vg - vd) / np.linalg.norm(vd - vel))}
    out["graph_speedup"] = out["direct_ms_per_step"] / out["graph_ms_per_step"]
    out["steps"] = nsteps
    return out


MS_FIELDS = ("ms_tree", "ms_lists", "ms_p2m", "ms_m2m", "ms_m2l", "ms_l2l", "ms_l2p", "ms_p2p", "ms_total",
             "ms_m2l_kernel", "ms_p2p_kernel")


TIMING_MODE = 1  # 2 with --timing isolated: every evaluation of the run


def isolated(ctx, n, dx, dm, da, reps=3):
    """Per-phase times with the phases one after another (timing mode 2),
    averaged over `reps` evaluations after one warm-up; returns (ms, stats)."""
    ctx.set_timing(2)
    try:
        ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
        rows = []
        for _ in range(reps):
            ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
            rows.append(ctx.stats().as_dict())
    finally:
        ctx.set_timing(TIMING_MODE)
    ms = {k: float(np.mean([r[k] for r in rows])) for k in MS_FIELDS}
    return ms, rows[-1]


def phase_table(st, ms, prec, order, t_eval_ms):
    """Per-phase roofline: algorithmic flops F and bytes B per evaluation,
    t_roof = max(F / peak, B / HBM), frac = t_roof / (phase time alone)."""
    w = 8 if prec == "fp64" else 4
    K = (order + 1) ** 2
    N = st["_n"]
    k = st["depth_sorts"]
    peak = PEAK["fp64" if prec == "fp64" else "fp32"] * 1e12
    p2p_alone = ms["ms_p2p_kernel"]
    rows = [
        # (phase, bound, flops, bytes, ms alone)
        ("tree", "hbm", 0.0, N * (120 + 32 * (k - 1)), ms["ms_tree"]),
        ("lists", "latency", 0.0, 0.0, ms["ms_lists"]),
        ("up-pass planning + P2M", "hbm", 0.0, st["n_p2m_terms"] * 32 + st["n_p2m_cells"] * K * w, ms["ms_p2m"]),
        ("M2M", "hbm", 0.0, st["n_m2m_cells"] * 9 * K * w, ms["ms_m2m"]),
        ("M2L", "mfma", st["flop_m2l"], (st["n_m2l_sources"] + st["n_m2l_targets"]) * K * w, ms["ms_m2l"]),
        ("L2L", "hbm", 0.0, st["n_l2l_cells"] * 9 * K * w, ms["ms_l2l"]),
        ("L2P", "hbm", 0.0, N * 56 + st["n_l2p_cells"] * K * w, ms["ms_l2p"]),
        ("P2P", "valu", st["flop_p2p"], N * (32 + 48), p2p_alone),
        ("finish", "hbm", 0.0, N * (24 + 24 + 24 + 4), max(ms["ms_p2p"] - p2p_alone, 0.0)),
    ]
    out, roof_sum, alone_sum = [], 0.0, 0.0
    for name, bound, F, B, t in rows:
        t_roof = max(F / peak, B / HBM) * 1e3
        roof_sum += t_roof
        alone_sum += t
        out.append({"phase": name, "bound": bound, "flop": F, "bytes": B, "ms_alone": round(t, 4),
                    "ms_roofline": round(t_roof, 4), "frac": (t_roof / t) if t > 0 and t_roof > 0 else None})
    return out, {"sum_ms_roofline": roof_sum, "sum_ms_alone": alone_sum, "t_eval_ms": t_eval_ms,
                 "frac_of_eval": roof_sum / t_eval_ms, "frac_of_phases_alone": roof_sum / alone_sum}


def direct_ref(ctx, n, dx, dm, tgt):
    return tune.direct_reference(ctx, n, dx, dm, tgt)


# ---------------------------------------------------------------------------------------------
def bench_single(args, cfg, name, precs, learned=True, cpu=True):
    """Tunes and times one Plummer sphere; returns the result record."""
    n = cfg["n"]
    orders = (args.order,) if args.order else cfg["orders"]
    cuts = (args.p2p_cut,) if args.p2p_cut is not None else cfg["cuts"]
    leaves = (args.leaf,) if args.leaf else cfg["leaves"]
    tols = (args.class_order_tol,) if args.class_order_tol is not None else cfg.get("tols", (0.0,))
    log(f"[bench] config={name} N={n} leaves={leaves} orders={orders} cuts={cuts} tols={tols}")
    xyz, m = workloads.plummer(n, seed=42)
    ctx = nbfmm.Context(order=orders[0], leaf_capacity=cfg["leaf_capacity"], G=1.0, eps2=0.0, p2p_cut=cuts[0])
    ctx.set_timing(TIMING_MODE)
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
    tgt = workloads.sample_targets(n, args.targets)
    hold = workloads.sample_targets_disjoint(n, args.targets, tgt)
    ref = direct_ref(ctx, n, dx, dm, tgt)
    ref_hold = direct_ref(ctx, n, dx, dm, hold)
    log("[bench] direct references done (tuning sample and held-out sample)")
    results = {}
    for prec in precs:
        ctx.set_params(precision=prec)
        best, table = tune.tune(ctx, n, dx, dm, da, tgt, ref, orders, cuts, leaves=leaves, gate=args.gate,
                                precision=prec, log=log, tols=tols)
        if best is None:
            best = min(table, key=lambda r: r["err"]["l2"])
            log(f"[bench] no setting meets the gate for {PREC_NAMES[prec]}; timing the most accurate one")
        ctx.set_params(order=best["order"], p2p_cut=best["p2p_cut"], leaf_capacity=best["leaf_capacity"],
                       precision=prec, class_order_tol=best["class_order_tol"])
        tiles, tile_table = (ctx.tiles(), [])
        if not args.no_tile_tuning:
            tiles, tile_table = tune.tune_tiles(ctx, n, dx, dm, da, prec, log=log)
        wall, st = timed(ctx, n, dx, dm, da, args.steps, args.warmup)
        acc = da.download(ctx.stream)
        err = tune.rel_errors(acc[hold], ref_hold)
        err_tune = tune.rel_errors(acc[tgt], ref)
        ms_step = wall * 1e3 / args.steps
        ms_iso, st_iso = isolated(ctx, n, dx, dm, da)
        st_iso["_n"] = n
        res = {"value": n * args.steps / wall, "ms_per_step": ms_step, "order": best["order"],
               "p2p_cut": best["p2p_cut"], "leaf_capacity": best["leaf_capacity"],
               "class_order_tol": best["class_order_tol"], "err": err, "err_tuning_sample": err_tune,
               "meets_gate": err["l2"] <= args.gate, "stats": st.as_dict(), "isolated_ms": ms_iso,
               "isolated_stats": st_iso, "tiles": tiles, "tile_tuning": tile_table,
               "tuning": [{"leaf_capacity": r["leaf_capacity"], "order": r["order"], "p2p_cut": r["p2p_cut"],
                           "class_order_tol": r["class_order_tol"], "ms": round(r["ms"], 3), "l2": r["err"]["l2"]}
                          for r in table]}
        results[PREC_NAMES[prec]] = res
        log(f"[bench] {name} {PREC_NAMES[prec]}: leaf={best['leaf_capacity']} p={best['order']} "
            f"cut={best['p2p_cut']} tol={best['class_order_tol']} {ms_step:.2f} ms/step; held-out L2 err "
            f"{err['l2']:.3e} (p99 {err['p99']:.2e}, max {err['max']:.2e}); phases alone: "
            + " ".join(f"{k[3:]}={v:.2f}" for k, v in ms_iso.items()))
    if learned and "mixed" in results and not args.no_learned:
        lr = learned_mode(ctx, n, dx, dm, da, hold, ref_hold, results["mixed"], args)
        if lr is not None:
            results["mixed_learned"] = lr
    if args.calib_copy:
        nb = args.calib_copy << 20
        src, dst = nbfmm.DeviceArray(nbytes=nb, shape=(nb // 8,)), nbfmm.DeviceArray(nbytes=nb, shape=(nb // 8,))
        if nbfmm.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3) != 0:
            raise RuntimeError("calibration copy failed")
        log(f"[bench] calibration copy of {args.calib_copy} MiB done")
    ctx.close()
    return results, len(hold)


def roofline_record(head_name, head, config):
    """The dominant kernel's roofline object (P2P or the M2L contraction, the
    longer of the two when each runs alone) plus the per-phase table."""
    st, ms = head["isolated_stats"], head["isolated_ms"]
    prec = "fp64" if head_name == "fp64" else "mixed"
    unit_prec = "fp64" if prec == "fp64" else "fp32"
    phases, overall = phase_table(st, ms, prec, head["order"], head["ms_per_step"])
    if ms["ms_p2p_kernel"] >= ms["ms_m2l_kernel"]:
        kern, bound, avg, F = ("k_p2p_tasks + k_p2p_wide (P2P, vector ALUs)", "valu", ms["ms_p2p_kernel"],
                               st["flop_p2p"])
        traffic, tsrc = read_traffic("k_p2p_tasks", config, prec, head)
    else:
        kern, bound, avg, F = ("k_m2l_cls / k_m2l_bf3 (M2L contraction, MFMA)", "mfma", ms["ms_m2l_kernel"],
                               st["flop_m2l"])
        traffic, tsrc = read_traffic("k_m2l_cls", config, prec, head)
    ach = F / (avg * 1e-3) / 1e12
    peak = PEAK[unit_prec]
    return {"kernel": kern, "bound": bound, "avg_ms": avg,
            "avg_ms_note": "kernel-alone time per evaluation: HIP events on the kernel's stream around its "
                           "launches, in an evaluation whose phases run one after another (timing mode 2)",
            "flop_per_launch": F, "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic_source": tsrc,
            "traffic": traffic, "phases": phases, "overall": overall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="plummer4m", choices=sorted(CONFIGS))
    ap.add_argument("--order", type=int, default=None, help="fix p instead of tuning")
    ap.add_argument("--p2p-cut", type=int, default=None, help="fix the P2P cut instead of tuning")
    ap.add_argument("--leaf", type=int, default=None, help="fix the leaf capacity instead of tuning")
    ap.add_argument("--class-order-tol", type=float, default=None,
                    help="fix the distance-dependent M2L order tolerance instead of tuning (0 = off)")
    ap.add_argument("--precision", default="all", choices=["all", "fp64", "mixed"])
    ap.add_argument("--timing", default="normal", choices=["normal", "isolated"],
                    help="isolated: every evaluation runs its phases one after another (for profiles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned order selection")
    ap.add_argument("--no-subrecords", action="store_true", help="only the headline workload")
    ap.add_argument("--no-tile-tuning", action="store_true", help="keep the built-in tile shapes")
    ap.add_argument("--targets", type=int, default=4096)
    ap.add_argument("--gate", type=float, default=1e-6)
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="after timing, copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding over several GPUs is designed, not built).", file=sys.stderr)
        return 2
    precs = [0, 1] if args.precision == "all" else [0 if args.precision == "fp64" else 1]
    cfg = CONFIGS[args.config]
    global TIMING_MODE
    TIMING_MODE = 2 if args.timing == "isolated" else 1
    if args.config == "batched256":
        rec = bench_batched(args, cfg, precs)
        print(json.dumps(line_from_sub(args, rec)))
        return 0
    if args.config == "learned8m":
        rec = bench_learned(args, cfg)
        print(json.dumps(line_from_sub(args, rec)))
        return 0
    results, ntarg = bench_single(args, cfg, args.config, precs)
    mixed_ok = [k for k in ("mixed", "mixed_learned") if k in results and results[k]["meets_gate"]]
    head_name = (min(mixed_ok, key=lambda k: results[k]["ms_per_step"]) if mixed_ok
                 else ("mixed" if "mixed" in results else "fp64"))
    head = results[head_name]
    st = head["stats"]
    roof = roofline_record(head_name, head, args.config)
    subs = {}
    if args.config == "plummer4m" and not args.no_subrecords and args.precision == "all":
        subs["plummer1m_fp64"] = bench_1m(args)
        subs["batched256"] = bench_batched(args, CONFIGS["batched256"], [0, 1], quick=True)
        subs["learned8m"] = bench_learned(args, CONFIGS["learned8m"])
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    err = head["err"]
    out = {
        "metric": METRIC,
        "value": head["value"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "f32/f64", "data": "synthetic",
        "error_note": (f"held-out sample of {ntarg} targets: relative L2 error {err['l2']:.3e} (the gated "
                       f"figure); per-target relative error p99 {err['p99']:.2e}, max {err['max']:.2e}"),
        "config": {
            "workload": f"Plummer sphere a=1 M=1 equal masses, N={cfg['n']}, {head_name}, leaf capacity "
                        f"{head['leaf_capacity']}, p={head['order']}, P2P cut {head['p2p_cut']}, M2L class-order tol "
                        f"{head.get('class_order_tol', 0.0)}, autotuned tiles",
            "N": cfg["n"], "precision": head_name, "order": head["order"], "p2p_cut": head["p2p_cut"],
            "class_order_tol": head.get("class_order_tol", 0.0),
            "leaf_capacity": head["leaf_capacity"],
            "error_gate_l2": args.gate, "error_targets": ntarg,
            "error_sample": "held out: disjoint from the 4096 targets the tuner chose the setting on",
            "rel_force_err_vs_direct": err, "rel_force_err_tuning_sample": head["err_tuning_sample"],
            "meets_gate": head["meets_gate"],
            "phases_ms": {k: st[k] for k in MS_FIELDS},
            "phases_ms_alone": head["isolated_ms"],
            "counts": {k: st[k] for k in ("depth", "n_cells", "n_leaves", "n_il", "n_m2l", "n_m2l_targets",
                                           "n_m2l_sources", "n_p2p_pairs", "n_p2m_cells", "n_m2m_cells",
                                           "n_l2l_cells", "n_l2p_cells")},
            "tiles": head.get("tiles"),
            "modes": {k: {kk: v[kk] for kk in ("value", "ms_per_step", "order", "p2p_cut", "leaf_capacity",
                                               "class_order_tol", "err", "meets_gate", "isolated_ms", "tiles",
                                               "tuning", "tile_tuning", "learned")
                          if kk in v}
                      for k, v in results.items()},
            "mixed_over_plain": (results["mixed"]["value"] / results["fp64"]["value"]
                                 if "mixed" in results and "fp64" in results else None),
            "subrecords": subs,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def line_from_sub(args, rec):
    """A bench line for --config batched256 / learned8m from its sub-record."""
    return {"metric": METRIC, "value": rec["value"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": rec.get("dtype", "f32/f64"), "data": "synthetic",
            "config": rec, "roofline": None, "cpu_baseline": None}


def bench_1m(args):
    """The plain FP64 configuration: N = 1,048,576 Plummer, p = 10, leaf
    capacity 64 (cut 1024): one evaluation timed, its error on 4096 sampled
    targets, and 5 classic RK4 steps dt = 1e-4."""
    n = 1 << 20
    xyz, m, vel = workloads.plummer(n, seed=42, with_velocity=True)
    ctx = nbfmm.Context(order=10, leaf_capacity=64, G=1.0, eps2=0.0, p2p_cut=1024, precision=0)
    try:
        dx, dm, dv = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m), nbfmm.DeviceArray(vel)
        da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
        tgt = workloads.sample_targets(n, args.targets)
        ref = direct_ref(ctx, n, dx, dm, tgt)
        wall, st = timed(ctx, n, dx, dm, da, args.steps, args.warmup)
        err = tune.rel_errors(da.download(ctx.stream)[tgt], ref)
        rec = ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, 1, integrator=2, every=1)  # warm-up step
        sync(ctx)
        t0 = time.perf_counter()
        rec = ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, 5, integrator=2, every=5)
        rk4_ms = (time.perf_counter() - t0) * 1e3 / 5
        dE = float(abs(rec[-1, 2] - rec[0, 2]) / abs(rec[0, 2]))
        graph = rk4_direct_vs_graph(ctx, n, xyz, vel, dm, da, 5)
    finally:
        ctx.close()
    # small systems, where launches and host round trips set the step time
    small = {}
    for ns in (500, 16384, 131072):
        sx, sm, sv = workloads.plummer(ns, seed=7, with_velocity=True)
        sctx = nbfmm.Context(order=8, leaf_capacity=16 if ns == 500 else 32, G=1.0, eps2=0.0, p2p_cut=512,
                             precision=0)
        try:
            sdx, sdm = nbfmm.DeviceArray(sx), nbfmm.DeviceArray(sm)
            sda = nbfmm.DeviceArray(nbytes=ns * 24, shape=(ns, 3))
            small[str(ns)] = rk4_direct_vs_graph(sctx, ns, sx, sv, sdm, sda, 20)
        finally:
            sctx.close()
        g = small[str(ns)]
        log(f"[bench] RK4 N={ns} FP64 p=8: direct {g['direct_ms_per_step']:.2f} ms/step, graph "
            f"{g['graph_ms_per_step']:.2f} ms/step ({g['graph_speedup']:.2f}x) {g['graph_stats']}")
    r = {"workload": "Plummer N=1,048,576, plain FP64, p=10, leaf capacity 64, P2P cut 1024",
         "value": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps, "dtype": "f64",
         "err": err, "depth": st.depth, "rk4_classic_ms_per_step": rk4_ms,
         "rk4_energy_drift_5_steps": dE,
         "rk4_note": "5 classic RK4 steps dt=1e-4 (4 evaluations each) with the energy monitor before and after "
                     "(two extra evaluations with potentials inside the timed span)",
         "rk4_graph": graph,
         "rk4_graph_small_systems": small}
    log(f"[bench] plummer1m fp64: {r['ms_per_step']:.2f} ms/eval, L2 err {err['l2']:.3e}, RK4 {rk4_ms:.1f} ms/step, "
        f"|dE/E| {dE:.1e}; RK4 direct {graph['direct_ms_per_step']:.2f} ms/step, graph "
        f"{graph['graph_ms_per_step']:.2f} ms/step {graph['graph_stats']}, rel diff {graph['graph_vs_direct_rel_diff']}")
    return r


def bench_batched(args, cfg, precs, quick=False):
    """The batched workload: all systems in one evaluation (nbfmm_batch_eval_accel_dev).
    Error: per-system relative L2 error on 16 sampled targets per system against
    that system's FP64 direct sum; the gate applies to the worst system.
    Then 20 classic RK4 steps dt = 1e-4 of the whole batch."""
    systems = workloads.batched_systems(256, seed=11, with_velocity=True)
    sizes = np.array([len(s[1]) for s in systems], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    xyz = np.concatenate([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    vel = np.concatenate([s[2] for s in systems])
    n = int(off[-1])
    log(f"[bench] batched: {len(sizes)} systems, N={n} (sizes {sizes.min()}..{sizes.max()})")
    ctx = nbfmm.Context(order=cfg["orders"][0], leaf_capacity=cfg["leaf_capacity"], G=1.0, eps2=0.0,
                        p2p_cut=cfg["cuts"][0], precision=1)
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
    rng = np.random.default_rng(4096)
    tg = [np.sort(rng.choice(int(sizes[k]), 16, replace=False)) for k in range(len(sizes))]
    refs = []
    for k in range(len(sizes)):
        sx, sm = nbfmm.DeviceArray(xyz[off[k]:off[k + 1]]), nbfmm.DeviceArray(m[off[k]:off[k + 1]])
        refs.append(direct_ref(ctx, int(sizes[k]), sx, sm, tg[k]))
        sx.free()
        sm.free()

    def errs():
        a = da.download(ctx.stream)
        return [float(np.linalg.norm(a[off[k] + tg[k]] - refs[k]) / np.linalg.norm(refs[k])) for k in range(len(sizes))]

    results = {}
    orders = (args.order,) if (args.order and not quick) else cfg["orders"]
    cuts = (args.p2p_cut,) if (args.p2p_cut is not None and not quick) else ((4096,) if quick else cfg["cuts"])
    for prec in precs:
        best, table = None, []
        for p in orders:
            for cut in cuts:
                ctx.set_params(order=p, p2p_cut=cut, precision=prec)
                ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
                ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
                ms, e = ctx.stats().ms_total, errs()
                table.append({"order": p, "p2p_cut": cut, "ms": round(ms, 3), "max_l2": max(e)})
                log(f"[bench] batched {PREC_NAMES[prec]} p={p} cut={cut}: {ms:.2f} ms, worst system {max(e):.3e}")
                if max(e) <= args.gate and (best is None or ms < best["ms"]):
                    best = table[-1]
        if best is None:
            best = min(table, key=lambda r: r["max_l2"])
        ctx.set_params(order=best["order"], p2p_cut=best["p2p_cut"], precision=prec)
        tiles, tile_table = ctx.tiles(), []
        if not args.no_tile_tuning:
            tiles, tile_table = tune.tune_tiles(ctx, n, dx, dm, da, prec, log=log, reps=3,
                                                run=lambda: ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr))
        for _ in range(args.warmup):
            ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
        sync(ctx)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
        sync(ctx)
        wall = time.perf_counter() - t0
        e = errs()
        # 20 classic RK4 steps of the whole batch, dt = 1e-4, on a copy of the state
        vx, cx = nbfmm.DeviceArray(vel), nbfmm.DeviceArray(xyz)
        sync(ctx)
        t1 = time.perf_counter()
        ctx.batch_step_rk4_dev(off, cx.ptr, vx.ptr, da.ptr, dm.ptr, 1e-4, 20, 1)
        sync(ctx)
        rk4_ms = (time.perf_counter() - t1) * 1e3 / 20
        ok = bool(np.all(np.isfinite(cx.download(ctx.stream))))
        cx.free()
        vx.free()
        # the same 20 steps replayed from a captured graph (one persistent step graph)
        graph = rk4_direct_vs_graph(ctx, n, xyz, vel, dm, da, 20, off=off)
        log(f"[bench] batched {PREC_NAMES[prec]} RK4: direct {graph['direct_ms_per_step']:.2f} ms/step, graph "
            f"{graph['graph_ms_per_step']:.2f} ms/step {graph['graph_stats']}, rel diff "
            f"{graph['graph_vs_direct_rel_diff']}")
        results[PREC_NAMES[prec]] = {"value": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps,
                                     "order": best["order"], "p2p_cut": best["p2p_cut"],
                                     "worst_system_l2": max(e), "mean_system_l2": float(np.mean(e)),
                                     "meets_gate": max(e) <= args.gate, "rk4_ms_per_step": rk4_ms,
                                     "rk4_steps": 20, "rk4_finite": ok, "rk4_graph": graph,
                                     "tuning": table, "tiles": tiles,
                                     "tile_tuning": tile_table}
        log(f"[bench] batched {PREC_NAMES[prec]}: {results[PREC_NAMES[prec]]['ms_per_step']:.2f} ms/eval, "
            f"worst system {max(e):.3e}, RK4 {rk4_ms:.1f} ms/step")
    ctx.close()
    if "mixed" in results and (results["mixed"]["meets_gate"] or "fp64" not in results
                               or not results["fp64"]["meets_gate"]):
        head_name = "mixed"
    else:
        head_name = "fp64"
    head = results[head_name]
    return {"workload": f"256 independent Plummer spheres (sizes log-uniform 1e3..6.4e4, a in [0.5, 2]), N={n}, "
                        f"one batched evaluation, {head_name}, p={head['order']}, cut {head['p2p_cut']}",
            "N": n, "systems": len(sizes), "error_gate_worst_system_l2": args.gate, "precision": head_name,
            "value": head["value"], "ms_per_step": head["ms_per_step"],
            "dtype": "f32/f64" if head_name == "mixed" else "f64",
            "modes": results,
            "mixed_over_plain": (results["mixed"]["value"] / results["fp64"]["value"]
                                 if "mixed" in results and "fp64" in results else None)}


def bench_learned(args, cfg):
    """The learned-selection workload: the lowest fixed order meeting the gate
    (checked on sampled targets) against the learned per-cell orders, whose
    model is trained in the run on two other mixes of 2^20 particles; fixed
    order 14 is the fallback of the checked evaluation.  Mixed precision, cubic
    cells."""
    n = cfg["n"]
    cut = args.p2p_cut if args.p2p_cut is not None else cfg["cuts"][0]
    xyz, m = workloads.plummer_mix(n, n_spheres=32, seed=7)
    # the mix's bounding box is not a cube (about 1.3 : 1 : 1): cubic cells
    # keep every interaction class admissible for M2L
    ctx = nbfmm.Context(order=14, leaf_capacity=cfg["leaf_capacity"], G=1.0, eps2=0.0, p2p_cut=cut, precision=1,
                        cubic_box=1)
    t0 = time.time()
    ctx.order_model_init(seed=7)
    rows = 0
    for seed in (201, 202):
        xt, mt = workloads.plummer_mix(1 << 20, n_spheres=32, seed=seed)
        tx, tm = nbfmm.DeviceArray(xt), nbfmm.DeviceArray(mt)
        rows += ctx.order_dataset_dev(len(mt), tx.ptr, tm.ptr, stride=2, seed=seed, tol=0.5 * args.gate)
        tx.free()
        tm.free()
    loss = ctx.order_train(epochs=100, batch=256, lr=2e-3, seed=3)
    t_train = time.time() - t0
    log(f"[bench] learned8m: {rows} labelled cells, loss {loss[0]:.3f} -> {loss[-1]:.3f}, {t_train:.1f} s")
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))

    def run(order, select):
        ctx.set_params(order=order, order_select=select, order_min=4, order_conf=0.9)
        err, fb = ctx.eval_accel_checked_dev(n, dx.ptr, dm.ptr, da.ptr, ntgt=args.targets, seed=9, gate=args.gate,
                                             fallback_order=14)
        wall, st = timed(ctx, n, dx, dm, da, args.steps, args.warmup)
        return {"order": order, "learned": bool(select), "value": n * args.steps / wall,
                "ms_per_step": wall * 1e3 / args.steps, "checked_err_l2": err, "fell_back": fb,
                "meets_gate": err <= args.gate and not fb, "mean_order": st.mean_order, "flop_m2l": st.flop_m2l}

    fixed = None
    table = []
    for p in cfg["orders"]:
        r = run(p, 0)
        table.append(r)
        log(f"[bench] learned8m fixed p={p}: {r['ms_per_step']:.2f} ms, err {r['checked_err_l2']:.3e}, "
            f"fell back {r['fell_back']}")
        if r["meets_gate"]:
            fixed = r
            break
    learned = []
    for pmax in sorted({fixed["order"] if fixed else 14, 14}):
        r = run(pmax, 1)
        learned.append(r)
        log(f"[bench] learned8m learned max order {pmax}: {r['ms_per_step']:.2f} ms, err {r['checked_err_l2']:.3e}, "
            f"fell back {r['fell_back']}, mean order {r['mean_order']:.2f}, M2L {r['flop_m2l']:.3e} flop")
    ctx.close()
    ok = [r for r in learned if r["meets_gate"]]
    best_l = min(ok, key=lambda r: r["ms_per_step"]) if ok else None
    cands = [r for r in [fixed, best_l] if r is not None]
    head = min(cands, key=lambda r: r["ms_per_step"]) if cands else table[-1]
    return {"workload": f"32 Plummer spheres (centres in [0,100)^3, a log-uniform 0.1..2), N={n}, mixed, "
                        f"cut {cut}, cubic box; learned per-cell orders vs the lowest fixed order meeting the gate",
            "N": n, "error_gate_l2": args.gate, "error_targets": args.targets,
            "value": head["value"], "ms_per_step": head["ms_per_step"], "head": head,
            "learned_over_fixed": (best_l["value"] / fixed["value"] if best_l and fixed else None),
            "fixed_scan": table, "learned": learned, "training_rows": rows, "training_s": round(t_train, 1),
            "loss_first_last": [float(loss[0]), float(loss[-1])]}


def learned_mode(ctx, n, dx, dm, da, tgt, ref, mixed, args):
    """Trains the order model on another Plummer sphere of the same size (GPU
    kernels of the library), then times the learned selection with the
    highest order at the tuned mixed order and one above; returns the fastest
    that meets the gate (or the most accurate).  Errors on the held-out sample."""
    t0 = time.time()
    ctx.set_params(order=14, order_select=0, p2p_cut=mixed["p2p_cut"], leaf_capacity=mixed["leaf_capacity"],
                   precision=1, class_order_tol=mixed.get("class_order_tol", 0.0))
    ctx.order_model_init(seed=7)
    ctx.order_dataset_put()
    xt, mt = workloads.plummer(n, seed=11)
    tx, tm = nbfmm.DeviceArray(xt), nbfmm.DeviceArray(mt)
    rows = ctx.order_dataset_dev(n, tx.ptr, tm.ptr, stride=8, seed=11, tol=0.5 * args.gate)
    tx.free()
    tm.free()
    loss = ctx.order_train(epochs=60, batch=256, lr=2e-3, seed=3)
    t_train = time.time() - t0
    log(f"[bench] learned: {rows} labelled cells, loss {loss[0]:.3f} -> {loss[-1]:.3f}, {t_train:.1f} s")
    runs = []
    for pmax in sorted({mixed["order"], min(14, mixed["order"] + 1)}):
        ctx.set_params(order=pmax, order_select=1, order_min=4, order_conf=0.9)
        wall, st = timed(ctx, n, dx, dm, da, args.steps, args.warmup)
        err = tune.rel_errors(da.download(ctx.stream)[tgt], ref)
        runs.append({"value": n * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps, "order": pmax,
                     "tiles": ctx.tiles(),
                     "p2p_cut": mixed["p2p_cut"], "leaf_capacity": mixed["leaf_capacity"],
                     "class_order_tol": mixed.get("class_order_tol", 0.0), "err": err,
                     "meets_gate": err["l2"] <= args.gate, "stats": st.as_dict(), "tuning": [],
                     "learned": {"max_order": pmax, "order_conf": 0.9, "mean_order": st.mean_order,
                                 "flop_m2l": st.flop_m2l, "training_rows": rows, "training_s": round(t_train, 2),
                                 "loss_first_last": [float(loss[0]), float(loss[-1])]}})
        log(f"[bench] mixed_learned max order {pmax}: {runs[-1]['ms_per_step']:.2f} ms/step, mean order "
            f"{st.mean_order:.2f}, L2 err {err['l2']:.3e}")
    ok = [r for r in runs if r["meets_gate"]]
    best = min(ok, key=lambda r: r["ms_per_step"]) if ok else min(runs, key=lambda r: r["err"]["l2"])
    ctx.set_params(order=best["order"], order_select=1)
    best["isolated_ms"], best["isolated_stats"] = isolated(ctx, n, dx, dm, da)
    best["isolated_stats"]["_n"] = n
    ctx.set_params(order_select=0)
    return best


def read_traffic(kernel, config, prec, head):
    """HBM bytes per evaluation of `kernel` (with its companion kernels, as the
    summary states) from the committed counter summary
    profiles/<kernel>_<config>_traffic.md, measured at the head's settings:
    the line 'traffic_bytes[<prec> p=<order> cut=<cut> leaf=<leaf> tol=<tol>]: X'.
    Returns (bytes or None when no summary was measured at these settings, source)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}_traffic.md")
    key = (f"traffic_bytes[{prec} p={head['order']} cut={head['p2p_cut']} leaf={head['leaf_capacity']} "
           f"tol={head['class_order_tol']:g}]:")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith(key):
                return float(line.split(":", 1)[1]), f"profiles/{kernel}_{config}_traffic.md, {key[:-1]}"
    return None, f"no rocprofv3 FETCH_SIZE/WRITE_SIZE summary at these settings ({key[:-1]})"


if __name__ == "__main__":
    sys.exit(main())
