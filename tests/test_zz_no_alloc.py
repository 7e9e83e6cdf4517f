"""The allocation-free direct sum (nbfmm_direct_accel_noalloc_dev) against
nbfmm_direct_accel_dev, through the C ABI: bit for bit in both modes, on two
calls in a row (the second smaller, so stale partial sums of the first would
show), and sizes beyond the fixed scratch rejected before any GPU work."""
import numpy as np
import pytest

import nbfmm
import workloads

pytestmark = pytest.mark.gpu


def _run(fn, n, dx, dm, tgt, mode):
    dt = nbfmm.DeviceArray(np.ascontiguousarray(tgt, dtype=np.int64))
    da = nbfmm.DeviceArray(nbytes=len(tgt) * 24, shape=(len(tgt), 3))
    fn(n, dx.ptr, dm.ptr, len(tgt), dt.ptr, mode, da.ptr)
    return da


def test_noalloc_direct_equals_direct():
    ctx = nbfmm.Context(order=4, leaf_capacity=16, G=1.0, eps2=0.0)
    try:
        for mode in (1, 0):
            for n, nt, seed in ((200000, 5000, 1), (30000, 700, 2)):
                xyz, m = workloads.plummer(n, seed=seed)
                dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
                tgt = workloads.sample_targets(n, nt, seed=seed + 10)
                print(f"[noalloc] mode={mode} n={n} ntgt={nt}", flush=True)
                a = _run(ctx.direct_accel_noalloc_dev, n, dx, dm, tgt, mode).download(ctx.stream)
                b = _run(ctx.direct_accel_dev, n, dx, dm, tgt, mode).download(ctx.stream)
                assert np.all(np.isfinite(a))
                assert np.array_equal(a, b), (mode, n, float(np.abs(a - b).max()))
    finally:
        ctx.close()


def test_noalloc_direct_rejects_too_many_targets():
    ctx = nbfmm.Context(order=4, leaf_capacity=16, G=1.0, eps2=0.0)
    try:
        n = 16
        dx, dm = nbfmm.DeviceArray(np.zeros((n, 3))), nbfmm.DeviceArray(np.ones(n))
        nt = nbfmm.NOALLOC_MAX_TARGETS + 1
        # the size check comes before any kernel: the (unread) target and
        # output buffers are a small stand-in
        dt = nbfmm.DeviceArray(np.zeros(8, np.int64))
        da = nbfmm.DeviceArray(np.zeros(8))
        with pytest.raises(nbfmm.NbfmmError) as e:
            ctx.direct_accel_noalloc_dev(n, dx.ptr, dm.ptr, nt, dt.ptr, 1, da.ptr)
        assert e.value.status == 1 and "at most" in str(e.value)
    finally:
        ctx.close()
