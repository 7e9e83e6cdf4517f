"""Tile shapes (nbfmm_set_tiles) change speed, not results: every P2P target
group, narrow/wide threshold, first-share split, M2L slice count and the
persistent work-queue P2P gives the
default evaluation to rounding (the near-field sums meet in FP64 atomics in a
varying order), in both precisions; invalid tiles are rejected."""
import numpy as np
import pytest

import nbfmm
import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    c = nbfmm.Context(order=8, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=512)
    yield c
    c.close()


CASES = {
    1: [dict(p2p_group_f32=g) for g in (8, 12, 24)] + [dict(p2p_narrow_max_f32=1), dict(p2p_narrow_max_f32=32)]
       + [dict(p2p_early_share_f32=0.0, p2p_early_launches_f32=3), dict(p2p_early_share_f32=1.0),
          dict(m2l_slices=2), dict(m2l_item_pairs=128), dict(m2l_item_pairs=16384), dict(p2p_rest_launches=3),
          dict(p2p_rest_launches=8, p2p_early_share_f32=0.0)]
       + [dict(p2p_queue=1), dict(p2p_queue=2, p2p_queue_waves0=4, p2p_queue_chunk=1),
          dict(p2p_queue=2, p2p_queue_waves0=1, p2p_group_f32=24, p2p_narrow_max_f32=1),
          dict(p2p_queue=2, p2p_queue_waves0=16, p2p_queue_chunk=64, p2p_early_share_f32=0.0)],
    0: [dict(p2p_group_f64=g) for g in (4, 12)] + [dict(p2p_narrow_max_f64=1), dict(p2p_narrow_max_f64=16)]
       + [dict(p2p_early_share_f64=0.7, p2p_early_launches_f64=8), dict(m2l_slices=2), dict(m2l_item_pairs=256),
          dict(p2p_rest_launches=5)]
       + [dict(p2p_queue=1, p2p_queue_chunk=3), dict(p2p_queue=2, p2p_queue_waves0=2),
          dict(p2p_queue=2, p2p_early_share_f64=1.0, p2p_narrow_max_f64=32)],
}


@pytest.mark.parametrize("precision", [0, 1])
def test_tiles_same_result(ctx, precision):
    xyz, m = workloads.plummer(60000, seed=29)
    ctx.set_params(precision=precision, order=10 if precision == 0 else 8)
    default = ctx.tiles()
    try:
        ref = ctx.eval_accel(xyz, m)
        for kw in CASES[precision]:
            ctx.set_tiles(**kw)
            a = ctx.eval_accel(xyz, m)
            e = _rel(a, ref)
            print(f"[tiles] precision={precision} {kw}: rel diff {e:.2e}", flush=True)
            assert e < 1e-13 if precision == 0 else e < 1e-6, (kw, e)
            ctx.set_tiles(**default)
    finally:
        ctx.set_tiles(**default)
        ctx.set_params(precision=0)


@pytest.mark.parametrize("precision", [0, 1])
def test_queue_potentials_and_batch(ctx, precision):
    """The persistent work-queue P2P gives the per-task launches' potentials
    and accelerations, alone and over a batch of uneven systems."""
    ctx.set_params(precision=precision, order=10 if precision == 0 else 8)
    default = ctx.tiles()
    tol = 1e-13 if precision == 0 else 1e-6
    try:
        xyz, m = workloads.plummer(40000, seed=31)
        n = len(m)
        dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
        da, dp = nbfmm.DeviceArray(np.zeros((n, 3))), nbfmm.DeviceArray(np.zeros(n))
        out = []
        for q in (0, 2):
            ctx.set_tiles(p2p_queue=q)
            ctx.eval_accel_pot_dev(n, dx.ptr, dm.ptr, da.ptr, dp.ptr)
            out.append((da.download(ctx.stream), dp.download(ctx.stream)))
        for k in range(2):
            e = _rel(out[1][k], out[0][k])
            print(f"[queue] precision={precision} {'accel' if k == 0 else 'phi'}: rel diff {e:.2e}", flush=True)
            assert e < tol, (k, e)
        sizes = [5000, 300, 21000, 64, 9000, 1500]
        xs = [workloads.plummer(s, seed=40 + i)[0] + 30.0 * i for i, s in enumerate(sizes)]
        ms = [workloads.plummer(s, seed=40 + i)[1] for i, s in enumerate(sizes)]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        bx, bm = nbfmm.DeviceArray(np.concatenate(xs)), nbfmm.DeviceArray(np.concatenate(ms))
        ba = nbfmm.DeviceArray(np.zeros((int(off[-1]), 3)))
        res = []
        for q in (0, 2):
            ctx.set_tiles(p2p_queue=q)
            ctx.batch_eval_accel_dev(off, bx.ptr, bm.ptr, ba.ptr)
            res.append(ba.download(ctx.stream))
        e = _rel(res[1], res[0])
        print(f"[queue] precision={precision} batch of {len(sizes)}: rel diff {e:.2e}", flush=True)
        assert e < tol, e
    finally:
        ctx.set_tiles(**default)
        ctx.set_params(precision=0)


def test_tiles_rejected(ctx):
    default = ctx.tiles()
    for kw in (dict(p2p_group_f32=10), dict(p2p_group_f64=16), dict(p2p_narrow_max_f32=65),
               dict(p2p_early_share_f64=1.5), dict(p2p_early_launches_f32=65), dict(m2l_slices=3), dict(m2l_item_pairs=100),
               dict(p2p_rest_launches=9), dict(p2p_queue=3), dict(p2p_queue_waves0=17), dict(p2p_queue_chunk=65), dict(p2p_queue_waves0=-1)):
        with pytest.raises(nbfmm.NbfmmError) as e:
            ctx.set_tiles(**kw)
        assert e.value.status == 1
        assert ctx.tiles() == default
    # zeros (and a negative share) select the built-in defaults
    ctx.set_tiles(p2p_group_f32=0, p2p_early_share_f32=-1.0, m2l_slices=0, p2p_queue_waves0=0, p2p_queue_chunk=0,
                  m2l_item_pairs=0, p2p_rest_launches=0)
    assert ctx.tiles() == default
