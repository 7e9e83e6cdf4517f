"""The target-major mixed-precision M2L (tile field m2l_mode 2) against the
class-major contraction with per-pair partials (m2l_mode 1, the default).

Both evaluate the same M2L sums with the same bf16x3 products; they differ
in the order of the additions (the target-major kernel adds each class's
FP32 contraction to per-target FP64 sums on chip, the class-major one sums
the FP32 pair partials in FP64 afterwards) and in using E_S T(d) E_S for the operator of the offset
class S d (a sign change S of the axes), which the oracle's exact M2L
satisfies exactly (tests/test_oracle_ops.py::test_m2l_reflection_symmetry).
So accelerations and potentials agree to FP32 rounding of the far field,
the error against the FP64 evaluation is unchanged, and the flop count is
the same.  Cases: fixed orders (one operator width), class orders (several
widths in one launch), operators sliced over workgroups (p = 13, 14, and
m2l_slices 2), an elongated box (non-cubic cells), a batch of uneven
systems, and the potentials."""
import numpy as np
import pytest

import nbfmm
import workloads

pytestmark = pytest.mark.gpu

TOL = 5e-7  # relative L2 difference of the two modes (FP32 rounding of the far field in another order)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _both(ctx, run):
    out = {}
    default = ctx.tiles()
    try:
        for mode in (1, 2):
            ctx.set_tiles(m2l_mode=mode)
            out[mode] = (run(), ctx.stats())
    finally:
        ctx.set_tiles(**default)
    return out


@pytest.mark.parametrize("order,tol,tiles", [(8, 0.0, {}), (11, 0.05, {}), (11, 0.05, {"m2l_slices": 2}),
                                             (13, 0.05, {}), (14, 0.0, {}), (6, 0.05, {})])
def test_target_major_equals_class_major(order, tol, tiles):
    xyz, m = workloads.plummer(120000, seed=61)
    ctx = nbfmm.Context(order=order, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=256, precision=1,
                        class_order_tol=tol)
    try:
        if tiles:
            ctx.set_tiles(**tiles)
        out = _both(ctx, lambda: ctx.eval_accel(xyz, m))
        (a1, s1), (a2, s2) = out[1], out[2]
        assert s1.n_m2l > 0 and s1.n_m2l == s2.n_m2l
        e = _rel(a2, a1)
        ctx.set_params(precision=0)
        a64 = ctx.eval_accel(xyz, m)
        e1, e2 = _rel(a1, a64), _rel(a2, a64)
        print(f"[m2l_tgt] p={order} tol={tol} {tiles}: modes differ by {e:.2e}; against FP64 "
              f"class-major {e1:.3e}, target-major {e2:.3e}; M2L flops {s1.flop_m2l:.4e} / {s2.flop_m2l:.4e}, "
              f"mean order {s1.mean_order:.3f} / {s2.mean_order:.3f}", flush=True)
        assert e < TOL, e
        assert e2 <= 1.05 * e1 + 1e-8, (e1, e2)
        assert abs(s2.flop_m2l - s1.flop_m2l) <= 1e-9 * s1.flop_m2l
        assert abs(s2.mean_order - s1.mean_order) <= 1e-9 * order
    finally:
        ctx.close()


def test_target_major_elongated_box_and_potentials():
    xyz, m = workloads.plummer(80000, seed=62)
    xyz = xyz * np.array([4.0, 1.0, 0.5])
    n = len(m)
    ctx = nbfmm.Context(order=10, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=256, precision=1,
                        class_order_tol=0.05)
    try:
        dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
        da, dp = nbfmm.DeviceArray(np.zeros((n, 3))), nbfmm.DeviceArray(np.zeros(n))

        def run():
            ctx.eval_accel_pot_dev(n, dx.ptr, dm.ptr, da.ptr, dp.ptr)
            return da.download(ctx.stream), dp.download(ctx.stream)
        out = _both(ctx, run)
        ea = _rel(out[2][0][0], out[1][0][0])
        ep = _rel(out[2][0][1], out[1][0][1])
        print(f"[m2l_tgt] elongated box: accelerations differ by {ea:.2e}, potentials by {ep:.2e}", flush=True)
        assert ea < TOL and ep < TOL, (ea, ep)
    finally:
        ctx.close()


def test_target_major_batch():
    sizes = [5000, 300, 21000, 64, 9000, 40000]
    xs = [workloads.plummer(s, seed=70 + i)[0] * (1.0 + 0.3 * i) + 50.0 * i for i, s in enumerate(sizes)]
    ms = [workloads.plummer(s, seed=70 + i)[1] for i, s in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ctx = nbfmm.Context(order=11, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=256, precision=1,
                        class_order_tol=0.05)
    try:
        bx, bm = nbfmm.DeviceArray(np.concatenate(xs)), nbfmm.DeviceArray(np.concatenate(ms))
        ba = nbfmm.DeviceArray(np.zeros((int(off[-1]), 3)))

        def run():
            ctx.batch_eval_accel_dev(off, bx.ptr, bm.ptr, ba.ptr)
            return ba.download(ctx.stream)
        out = _both(ctx, run)
        a1, a2 = out[1][0], out[2][0]
        for i in range(len(sizes)):
            sl = slice(int(off[i]), int(off[i + 1]))
            e = _rel(a2[sl], a1[sl])
            print(f"[m2l_tgt] batch system {i} ({sizes[i]} particles): modes differ by {e:.2e}", flush=True)
            assert e < TOL, (i, e)
    finally:
        ctx.close()


def test_m2l_mode_validated():
    ctx = nbfmm.Context(order=6, leaf_capacity=32, precision=1)
    try:
        assert ctx.tiles()["m2l_mode"] == 1
        with pytest.raises(nbfmm.NbfmmError):
            ctx.set_tiles(m2l_mode=3)
        ctx.set_tiles(m2l_mode=2)
        assert ctx.tiles()["m2l_mode"] == 2
        ctx.set_tiles(m2l_mode=0)
        assert ctx.tiles()["m2l_mode"] == 1
    finally:
        ctx.close()
