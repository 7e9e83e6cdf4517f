"""GPU tests of integration steps replayed from a captured HIP graph.

With graphs on, every evaluation of a step is planned on the device (no size
is read back to the host: launches and buffers are sized by capacities from
the last direct evaluation, kernels read the actual sizes from device
memory), and the whole step -- four evaluations and the updates for RK4 -- is
captured once and replayed.  A replayed step must equal the direct step to
rounding (the near-field sums meet in FP64 atomics, whose order varies); a
step whose sizes outgrow the capacities must leave the state untouched and
run again directly.  The reference advances one step per call of
RungeKutta4::integrate (runge_kutta_4.cpp:10-93)."""
import numpy as np
import pytest

import nbfmm
import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = nbfmm.Context(order=8, leaf_capacity=32, G=1.0, eps2=0.0)
    yield c
    c.close()


def _steps(ctx, graph, kind, x, v, m, dt, calls, off=None):
    """Runs `calls` (a list of step counts, one API call each) from (x, v);
    returns the final state and the graph statistics of the run."""
    ctx.set_graph(graph)
    n = len(m)
    dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (x, v, np.zeros((n, 3)), m))
    before = ctx.graph_stats()
    try:
        for k in calls:
            if off is not None:
                if kind == 0:
                    ctx.batch_step_se_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, k)
                else:
                    ctx.batch_step_rk4_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, k, 0 if kind == 1 else 1)
            elif kind == 0:
                ctx.step_se_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, k)
            else:
                ctx.step_rk4_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, k, 0 if kind == 1 else 1)
        xs, vs = dx.download(ctx.stream), dv.download(ctx.stream)
    finally:
        ctx.set_graph(0)
    after = ctx.graph_stats()
    return xs, vs, {k: after[k] - before[k] for k in after}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_graph_steps_equal_direct_steps(ctx, precision, kind):
    # 20k Plummer sphere; mixed precision with distance-dependent class orders
    # (per-order M2L launches, the device schedule of every order)
    x, m = workloads.plummer(20000, seed=31)
    v = np.random.default_rng(5).normal(0, 0.3, x.shape)
    tol = 0.05 if precision == 1 else 0.0
    ctx.set_params(order=8 if precision == 0 else 10, leaf_capacity=32, p2p_cut=512, G=1.0, eps2=0.0,
                   precision=precision, class_order_tol=tol)
    print(f"[graph] precision={precision} kind={kind}", flush=True)
    dt = 1e-3
    calls = [1, 3, 2]  # the graph is kept across calls with the same arrays (one step per call too)
    xd, vd, sd = _steps(ctx, 0, kind, x, v, m, dt, calls)
    xg, vg, sg = _steps(ctx, 1, kind, x, v, m, dt, calls)
    ex, ev = _rel(xg - x, xd - x), _rel(vg - v, vd - v)
    print(f"[graph] displacement rel diff {ex:.2e}, velocity change rel diff {ev:.2e}, stats {sg}")
    assert sd == {"captures": 0, "launches": 0, "redone": 0, "direct": 0}
    assert sg["captures"] == 1 and sg["direct"] == 1 and sg["launches"] == 5 and sg["redone"] == 0, sg
    assert ex < 1e-12 and ev < 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_step_overflow_runs_again(ctx, precision):
    # G = 0: the particles move ballistically and exactly alike in both runs,
    # while the tree changes from step to step: half of a uniform cube
    # contracts to a tenth of its size (deeper tree), then passes through the
    # centre and spreads again.  A replay planned on the previous step's sizes
    # overflows (another depth), leaves the state alone and runs again directly.
    x, m = workloads.uniform_cube(6000, seed=8)
    v = np.zeros_like(x)
    dt = 0.1
    v[:3000] = -0.9 * (x[:3000] - 0.5) / dt
    ctx.set_params(order=6, leaf_capacity=16, p2p_cut=256, G=0.0, eps2=0.0, precision=precision,
                   class_order_tol=0.0)
    print(f"[graph-overflow] precision={precision}", flush=True)
    xd, vd, _ = _steps(ctx, 0, 0, x, v, m, dt, [5])
    xg, vg, sg = _steps(ctx, 1, 0, x, v, m, dt, [5])
    print(f"[graph-overflow] stats {sg}")
    assert sg["redone"] >= 1, sg
    assert sg["launches"] + sg["direct"] - sg["redone"] == 5, sg  # every step replayed or run directly
    assert np.array_equal(xg, xd) and np.array_equal(vg, vd)


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_batch_steps_equal_direct(ctx, precision):
    # a batch of uneven systems: placement, forest and every phase in the graph
    rng = np.random.default_rng(12)
    sizes = [300, 2500, 6000, 900, 40]
    xs, ms = [], []
    for k, n in enumerate(sizes):
        xk, mk = workloads.plummer(n, seed=200 + k, a=rng.uniform(0.5, 2.0))
        xs.append(xk + rng.uniform(-20, 20, 3))
        ms.append(mk)
    x, m = np.concatenate(xs), np.concatenate(ms)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    v = rng.normal(0, 0.2, x.shape)
    ctx.set_params(order=8, leaf_capacity=16, p2p_cut=256, G=1.0, eps2=0.0, precision=precision,
                   class_order_tol=0.0)
    print(f"[graph-batch] precision={precision}", flush=True)
    xd, vd, _ = _steps(ctx, 0, 2, x, v, m, 1e-3, [4], off=off)
    xg, vg, sg = _steps(ctx, 1, 2, x, v, m, 1e-3, [4], off=off)
    ex, ev = _rel(xg - x, xd - x), _rel(vg - v, vd - v)
    print(f"[graph-batch] displacement rel diff {ex:.2e}, velocity change rel diff {ev:.2e}, stats {sg}")
    assert sg["captures"] == 1 and sg["launches"] == 3 and sg["redone"] == 0, sg
    assert ex < 1e-12 and ev < 1e-10


def test_graph_run_with_monitor(ctx):
    # the multi-step driver with the energy monitor between replayed steps
    x, m = workloads.plummer(8000, seed=44)
    v = np.random.default_rng(9).normal(0, 0.3, x.shape)
    ctx.set_params(order=8, leaf_capacity=32, p2p_cut=512, G=1.0, eps2=0.0, precision=0, class_order_tol=0.0)
    n = len(m)
    out = []
    for graph in (0, 1):
        ctx.set_graph(graph)
        dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (x, v, np.zeros((n, 3)), m))
        try:
            diag = ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-3, 6, integrator=2, every=2)
        finally:
            ctx.set_graph(0)
        out.append((dx.download(ctx.stream), dv.download(ctx.stream), diag))
    (xd, vd, dd), (xg, vg, dg) = out
    assert _rel(xg - x, xd - x) < 1e-12 and _rel(vg - v, vd - v) < 1e-10
    assert np.allclose(np.asarray(dg), np.asarray(dd), rtol=1e-12, atol=1e-14)


def test_graph_rejects_learned_orders(ctx):
    ctx.set_params(order=10, order_select=1, order_min=4, class_order_tol=0.0)
    try:
        with pytest.raises(nbfmm.NbfmmError):
            ctx.set_graph(1)
    finally:
        ctx.set_params(order_select=0)
        ctx.set_graph(0)
