"""Benchmark of the MI355X FMM force evaluation (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config plain1m|plain4m]

A step is one full force evaluation of the device-resident particle set:
bbox, the reference tree rule (depth loop, Morton keys, sort, levels,
lists), P2M, M2L, L2P and P2P, i.e. nbfmm_eval_accel_dev.  Inputs are in
HBM before the timed region; value = N * K / elapsed (particle-force
evaluations per second).  Accuracy is measured outside the timed region
against the FP64 direct sum (compensated) on 4096 sampled targets.
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
import workloads  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
CONFIGS = {
    # configs[1] of the benchmark definition: plain FP64, N=1,048,576 Plummer, leaf capacity 64
    "plain1m": dict(n=1 << 20, order=10, leaf_capacity=64, precision=0, p2p_cut=1024),
    "plain4m": dict(n=1 << 22, order=10, leaf_capacity=64, precision=0, p2p_cut=1024),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(budget_s=20.0):
    """The reference's own FMM::update (oracle/_ref, -O2, 1 thread, its
    hardcoded L=2 / leaf threshold 20) on a Plummer sample sized to the budget."""
    import orc
    if not orc.ref_available():
        return None
    n = 8192
    xyz, m = workloads.plummer(n, seed=42)
    t0 = time.perf_counter()
    orc.ref_fmm(xyz, m)
    dt = time.perf_counter() - t0
    # grow the sample while it stays inside the budget (FMM cost ~ linear in N)
    while dt * 4 < budget_s / 2 and n < (1 << 18):
        n *= 4
        xyz, m = workloads.plummer(n, seed=42)
        t0 = time.perf_counter()
        orc.ref_fmm(xyz, m)
        dt = time.perf_counter() - t0
    return {
        "value": n / dt, "unit": "evals/s", "cores": 1, "kind": "reference",
        "sample": f"reference FMM::update (L=2, leaf threshold 20 as hardcoded, -O2 build of its sources), "
                  f"Plummer N={n}, one call, {dt:.2f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="plain1m", choices=sorted(CONFIGS))
    ap.add_argument("--order", type=int, default=None)
    ap.add_argument("--p2p-cut", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--targets", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design in DESIGN.md, not built).", file=sys.stderr)
        return 2
    cfg = dict(CONFIGS[args.config])
    if args.order:
        cfg["order"] = args.order
    if args.p2p_cut is not None:
        cfg["p2p_cut"] = args.p2p_cut
    n = cfg["n"]
    log(f"[bench] config={args.config} {cfg}")
    xyz, m = workloads.plummer(n, seed=42)
    ctx = nbfmm.Context(order=cfg["order"], leaf_capacity=cfg["leaf_capacity"], G=1.0, eps2=0.0,
                        precision=cfg["precision"], p2p_cut=cfg["p2p_cut"])
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
    for w in range(args.warmup):
        ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
        log(f"[bench] warmup {w} done: {ctx.stats().ms_total:.2f} ms")
    st = ctx.stats()
    # timed region: K evaluations, per-kernel events on the library's stream
    timer = nbfmm.Timer(ctx.stream)
    nbfmm.hip().hipStreamSynchronize(ctx.stream)
    m2l_ms, p2p_ms = [], []
    t0 = time.perf_counter()
    timer.start()
    for k in range(args.steps):
        ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
        s = ctx.stats()
        m2l_ms.append(s.ms_m2l_kernel)
        p2p_ms.append(s.ms_p2p_kernel)
    gpu_ms = timer.stop()
    wall = time.perf_counter() - t0
    nbfmm.hip().hipStreamSynchronize(ctx.stream)
    ms_per_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step (events {gpu_ms / args.steps:.2f})")
    acc = da.download(ctx.stream)
    # accuracy vs the FP64 direct sum on sampled targets (outside the timed region)
    tgt = workloads.sample_targets(n, args.targets)
    dt_ = nbfmm.DeviceArray(tgt)
    dd = nbfmm.DeviceArray(nbytes=len(tgt) * 24, shape=(len(tgt), 3))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, len(tgt), dt_.ptr, 1, dd.ptr)
    ref = dd.download(ctx.stream)
    diff = acc[tgt] - ref
    e_i = np.linalg.norm(diff, axis=1) / np.linalg.norm(ref, axis=1)
    err = {"l2": float(np.linalg.norm(diff) / np.linalg.norm(ref)), "max": float(e_i.max()),
           "p99": float(np.percentile(e_i, 99))}
    log(f"[bench] error vs direct: {err}")
    s = ctx.stats()
    m2l_avg = float(np.mean(m2l_ms))
    p2p_avg = float(np.mean(p2p_ms))
    # dominant kernel's roofline
    if m2l_avg >= p2p_avg:
        ach = s.flop_m2l / (m2l_avg * 1e-3) / 1e12
        roof = {"kernel": "k_m2l_mfma (FP64 MFMA)", "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "traffic": None,
                "avg_ms": m2l_avg, "flop_per_launch": s.flop_m2l}
    else:
        ach = s.flop_p2p / (p2p_avg * 1e-3) / 1e12
        roof = {"kernel": "k_p2p (FP64 VALU)", "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "traffic": None,
                "avg_ms": p2p_avg, "flop_per_launch": s.flop_p2p}
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "particle-force evals/s at rel. force err <=1e-6 vs FP64 direct; mixed vs plain",
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Plummer sphere a=1 M=1, N={n}, p={cfg['order']}, leaf capacity "
                               f"{cfg['leaf_capacity']}, {'plain FP64' if cfg['precision'] == 0 else 'mixed'}",
                   "N": n, "order": cfg["order"], "leaf_capacity": cfg["leaf_capacity"],
                   "p2p_cut": cfg["p2p_cut"], "precision": "fp64" if cfg["precision"] == 0 else "mixed",
                   "rel_force_err_vs_direct": err, "error_targets": len(tgt),
                   "meets_1e-6": err["l2"] <= 1e-6,
                   "phases_ms": {k: getattr(s, k) for k in ("ms_tree", "ms_lists", "ms_p2m", "ms_m2l", "ms_l2p",
                                                             "ms_p2p", "ms_total")},
                   "counts": {k: getattr(s, k) for k in ("depth", "n_cells", "n_leaves", "n_il", "n_m2l",
                                                          "n_m2l_targets", "n_m2l_sources", "n_p2p_pairs")}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
