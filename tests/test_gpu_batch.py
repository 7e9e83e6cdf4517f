"""GPU tests of the batched systems (many independent systems in one
evaluation).  No system may act on another: with every interaction-list pair
evaluated directly the batch equals each system's own direct sum, even for
systems that overlap in space; with the FMM each system keeps the accuracy
of its own evaluation; the batched integrators equal the per-system ones."""
import numpy as np
import pytest

import nbfmm
import orc
import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _systems(sizes, seed):
    rng = np.random.default_rng(seed)
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = workloads.plummer(int(n), seed=seed * 100 + k, a=rng.uniform(0.5, 2.0))
        # systems 0 and 1 overlap; the others are spread out
        xs.append(x + (0.0 if k < 2 else rng.uniform(-50, 50, 3)))
        ms.append(m * rng.uniform(0.5, 2.0))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.concatenate(xs), np.concatenate(ms), off


@pytest.fixture(scope="module")
def ctx():
    c = nbfmm.Context(order=10, leaf_capacity=16, G=1.0, eps2=0.0)
    yield c
    c.close()


def _direct(ctx, x, m):
    n = len(m)
    dx, dm, da = nbfmm.DeviceArray(x), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, n, None, 1, da.ptr)
    return da.download(ctx.stream)


def test_batch_direct_equals_per_system(ctx):
    x, m, off = _systems([700, 1500, 300, 2600, 900, 1200, 450], seed=3)
    n = len(m)
    dx, dm, da = nbfmm.DeviceArray(x), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.set_params(order=4, leaf_capacity=16, p2p_cut=1 << 50)
    ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
    a = da.download(ctx.stream)
    for k in range(len(off) - 1):
        s = slice(off[k], off[k + 1])
        ref = _direct(ctx, x[s], m[s])
        e = _rel(a[s], ref)
        # placement rounds positions at 1e-16 of the grid (close pairs see ~1e-12);
        # a leak between the overlapping systems 0 and 1 would be of order 1
        assert e < 1e-10, (k, e)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_fmm_accuracy(ctx, precision):
    sizes = [3000, 800, 5000, 1200, 2500, 600, 4000, 1500, 2000, 3500, 900, 1800]
    x, m, off = _systems(sizes, seed=5)
    n = len(m)
    dx, dm, da = nbfmm.DeviceArray(x), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.set_params(order=10, leaf_capacity=16, p2p_cut=256, precision=precision)
    try:
        ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
        a = da.download(ctx.stream)
        worst = 0.0
        for k in range(len(off) - 1):
            s = slice(off[k], off[k + 1])
            ref = _direct(ctx, x[s], m[s])
            e = _rel(a[s], ref)
            # the same system evaluated by itself
            e1 = _rel(ctx.eval_accel(x[s], m[s]), ref)
            worst = max(worst, e)
            assert e < 1e-5 and e < 20 * e1 + 1e-7, (k, e, e1)
        print(f"[batch] precision={precision} worst per-system rel L2 error {worst:.2e}")
    finally:
        ctx.set_params(precision=0)


def test_batch_integrators_match_per_system(ctx):
    x, m, off = _systems([600, 1100, 400], seed=9)
    rng = np.random.default_rng(1)
    v = rng.normal(0, 0.1, x.shape)
    n = len(m)
    ctx.set_params(order=4, leaf_capacity=16, p2p_cut=1 << 50)
    for kind in ("se", "rk4"):
        dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (x, v, np.zeros((n, 3)), m))
        if kind == "se":
            ctx.batch_step_se_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-3, 3)
        else:
            ctx.batch_step_rk4_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-3, 3, 1)
        xb, vb = dx.download(ctx.stream), dv.download(ctx.stream)
        for k in range(len(off) - 1):
            s = slice(off[k], off[k + 1])
            nk = int(off[k + 1] - off[k])
            ex, ev, ea, em = (nbfmm.DeviceArray(a) for a in (x[s], v[s], np.zeros((nk, 3)), m[s]))
            if kind == "se":
                ctx.step_se_dev(nk, ex.ptr, ev.ptr, ea.ptr, em.ptr, 1e-3, 3)
            else:
                ctx.step_rk4_dev(nk, ex.ptr, ev.ptr, ea.ptr, em.ptr, 1e-3, 3, 1)
            assert _rel(xb[s], ex.download(ctx.stream)) < 1e-13, (kind, k)
            assert _rel(vb[s] - v[s], ev.download(ctx.stream) - v[s]) < 1e-10, (kind, k)


def test_batch_errors(ctx):
    x, m, off = _systems([500, 500], seed=2)
    dx, dm, da = nbfmm.DeviceArray(x), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((len(m), 3)))
    ctx.set_params(order=4, leaf_capacity=16, p2p_cut=256)
    for bad in ([1, 500, 1000], [0, 500, 500, 1000], [0]):
        with pytest.raises(nbfmm.NbfmmError):
            ctx.batch_eval_accel_dev(np.array(bad), dx.ptr, dm.ptr, da.ptr)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_small_systems(ctx, precision):
    # every system fits in one leaf (1 to 15 particles, leaf capacity 16): the
    # tree is still refined one level below the system cells, and each system
    # gets its own direct sum (single particles: zero acceleration)
    rng = np.random.default_rng(21)
    sizes = [5, 1, 15, 9, 2, 12, 7, 15, 3, 11]
    x, m, off = _systems([20] * len(sizes), seed=13)
    xs, ms = [], []
    for k, nk in enumerate(sizes):
        s = slice(off[k], off[k] + nk)
        xs.append(x[s] * rng.uniform(0.5, 3.0))
        ms.append(m[s])
    x, m = np.concatenate(xs), np.concatenate(ms)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = len(m)
    dx, dm, da = nbfmm.DeviceArray(x), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.set_params(order=8, leaf_capacity=16, p2p_cut=256, precision=precision)
    try:
        print(f"[batch-small] precision={precision} systems={len(sizes)} n={n}", flush=True)
        ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
        a = da.download(ctx.stream)
        tol = 1e-12 if precision == 0 else 1e-6
        for k in range(len(sizes)):
            s = slice(off[k], off[k + 1])
            if sizes[k] == 1:
                assert np.all(a[s] == 0.0), k
                continue
            ref = orc.direct_targets(x[s], m[s], np.arange(sizes[k]), G=1.0, eps2=0.0)
            assert _rel(a[s], ref) < tol, (k, _rel(a[s], ref))
    finally:
        ctx.set_params(precision=0)
