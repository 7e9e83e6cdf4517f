"""Accuracy-gated choice of the expansion order, the P2P cut, the leaf capacity and
the distance-dependent M2L order tolerance (class_order_tol), then of the tile
shapes and launch split of the P2P and M2L kernels (tune_tiles).

For a particle set resident on the device, every (order, p2p_cut, leaf capacity, tol) candidate
is evaluated (one warm-up evaluation, then one timed with the library's
events), its relative force error is measured against the FP64 direct sum
(compensated, eps2 = 0) on sampled targets, and the fastest candidate whose
global L2 relative error meets the gate is returned.

Error measure: e = sqrt(sum |a - a_dir|^2 / sum |a_dir|^2) over the sampled
targets; max and p99 of the per-target |a - a_dir| / |a_dir| are reported
beside it.
"""
import numpy as np

import nbfmm


def direct_reference(ctx, n, d_xyz, d_m, tgt):
    """FP64 direct-sum accelerations (compensated) of the targets `tgt`."""
    d_tgt = nbfmm.DeviceArray(np.ascontiguousarray(tgt, dtype=np.int64))
    d_ref = nbfmm.DeviceArray(nbytes=len(tgt) * 24, shape=(len(tgt), 3))
    ctx.direct_accel_dev(n, d_xyz.ptr, d_m.ptr, len(tgt), d_tgt.ptr, 1, d_ref.ptr)
    ref = d_ref.download(ctx.stream)
    d_tgt.free()
    d_ref.free()
    return ref


def rel_errors(acc_t, ref):
    diff = acc_t - ref
    e_i = np.linalg.norm(diff, axis=1) / np.linalg.norm(ref, axis=1)
    return {"l2": float(np.linalg.norm(diff) / np.linalg.norm(ref)), "max": float(e_i.max()),
            "p99": float(np.percentile(e_i, 99))}


def tune(ctx, n, d_xyz, d_m, d_acc, tgt, ref, orders, cuts, leaves=(None,), gate=1e-6, precision=0, log=None,
         tols=(0.0,)):
    """Returns (best, table): best = dict(order, p2p_cut, leaf_capacity, class_order_tol, ms, err) or None."""
    table = []
    best = None
    for leaf in leaves:
        if leaf is not None:
            ctx.set_params(leaf_capacity=leaf)
        for p in orders:
            for tol in tols:
                for cut in cuts:
                    ctx.set_params(order=p, p2p_cut=cut, precision=precision, class_order_tol=tol)
                    ctx.eval_accel_dev(n, d_xyz.ptr, d_m.ptr, d_acc.ptr)  # warm-up at this setting
                    ts = []
                    for _ in range(3):  # settings within ~0.2 ms of each other: a median, not one sample
                        ctx.eval_accel_dev(n, d_xyz.ptr, d_m.ptr, d_acc.ptr)
                        ts.append(ctx.stats().ms_total)
                    ms = float(np.median(ts))
                    acc = d_acc.download(ctx.stream)
                    err = rel_errors(acc[tgt], ref)
                    row = {"order": p, "p2p_cut": cut, "leaf_capacity": int(ctx.params.leaf_capacity),
                           "class_order_tol": tol, "ms": ms, "err": err}
                    table.append(row)
                    if log:
                        log(f"[tune] precision={precision} leaf={row['leaf_capacity']} p={p} tol={tol} cut={cut}: "
                            f"{ms:.2f} ms, L2 err {err['l2']:.3e}")
                    if err["l2"] <= gate and (best is None or ms < best["ms"]):
                        best = row
    ctx.set_params(class_order_tol=0.0)
    return best, table


# Tile fields searched per precision, in this order, with the values tried
# (the built-in default is among them)
TILE_GRID = {
    1: [("p2p_group_f32", (8, 12, 16, 24)), ("p2p_narrow_max_f32", (16, 32, 64)),
        ("p2p_early_share_f32", (0.3, 0.5, 0.7)), ("p2p_early_launches_f32", (8, 16, 24)), ("m2l_slices", (1, 2)),
        ("m2l_item_pairs", (512, 1024, 2048)), ("p2p_queue", (0, 1, 2)), ("p2p_rest_launches", (1, 4, 6, 8)),
        ("m2l_mode", (1, 2))],
    0: [("p2p_group_f64", (4, 8, 12)), ("p2p_narrow_max_f64", (16, 32, 64)),
        ("p2p_early_share_f64", (0.1, 0.2, 0.4)), ("p2p_early_launches_f64", (1, 4, 16)), ("m2l_slices", (1, 2)),
        ("m2l_item_pairs", (512, 1024, 2048)), ("p2p_queue", (0, 1, 2)), ("p2p_rest_launches", (1, 4, 6))],
}


def _median_ms(ctx, run, reps):
    run()
    ms = []
    for _ in range(reps):
        run()
        ms.append(ctx.stats().ms_total)
    return float(np.median(ms))


def tune_tiles(ctx, n, d_xyz, d_m, d_acc, precision, reps=5, log=None, grid=None, margin=0.01, run=None):
    """Coordinate search over the tile fields of TILE_GRID[precision] at the
    context's parameters (`run` performs one evaluation; by default
    nbfmm_eval_accel_dev of the given particles): each field in turn takes the value of lowest median
    evaluation time (over `reps` evaluations) with the others fixed; a value
    replaces the current one only when faster by more than `margin`.  Tile
    shapes change speed, not results (beyond the rounding of the near-field
    atomics).  Returns (tiles dict, table of every trial)."""
    grid = grid if grid is not None else TILE_GRID[precision]
    if run is None:  # one evaluation of the particle set (the batch evaluator for batches)
        def run():
            ctx.eval_accel_dev(n, d_xyz.ptr, d_m.ptr, d_acc.ptr)
    best = ctx.tiles()
    best_ms = _median_ms(ctx, run, reps)
    table = [{"tiles": dict(best), "ms": round(best_ms, 3)}]
    for field, values in grid:
        cur = best[field]
        for v in values:
            if v == cur:
                continue
            ctx.set_tiles(**{field: v})
            ms = _median_ms(ctx, run, reps)
            table.append({"tiles": ctx.tiles(), "ms": round(ms, 3)})
            if log:
                log(f"[tiles] precision={precision} {field}={v}: {ms:.2f} ms (best {best_ms:.2f})")
            if ms < best_ms * (1.0 - margin):
                best, best_ms = ctx.tiles(), ms
            ctx.set_tiles(**best)
    ctx.set_tiles(**best)
    return best, table
