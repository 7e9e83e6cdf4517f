"""Accuracy-gated choice of the expansion order, the P2P cut, the leaf capacity and
the distance-dependent M2L order tolerance (class_order_tol).

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
                    ctx.eval_accel_dev(n, d_xyz.ptr, d_m.ptr, d_acc.ptr)
                    ctx.eval_accel_dev(n, d_xyz.ptr, d_m.ptr, d_acc.ptr)
                    ms = ctx.stats().ms_total
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
