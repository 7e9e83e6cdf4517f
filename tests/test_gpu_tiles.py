"""Tile shapes (nbfmm_set_tiles) change speed, not results: every P2P target
group, narrow/wide threshold, first-share split and M2L slice count gives the
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
    1: [dict(p2p_group_f32=g) for g in (8, 12, 24)] + [dict(p2p_narrow_max_f32=1), dict(p2p_narrow_max_f32=256)]
       + [dict(p2p_early_share_f32=0.0, p2p_early_launches_f32=3), dict(p2p_early_share_f32=1.0),
          dict(m2l_slices=2)],
    0: [dict(p2p_group_f64=g) for g in (4, 12)] + [dict(p2p_narrow_max_f64=1), dict(p2p_narrow_max_f64=256)]
       + [dict(p2p_early_share_f64=0.7, p2p_early_launches_f64=8), dict(m2l_slices=2)],
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


def test_tiles_rejected(ctx):
    default = ctx.tiles()
    for kw in (dict(p2p_group_f32=10), dict(p2p_group_f64=16), dict(p2p_narrow_max_f32=300),
               dict(p2p_early_share_f64=1.5), dict(p2p_early_launches_f32=65), dict(m2l_slices=3)):
        with pytest.raises(nbfmm.NbfmmError) as e:
            ctx.set_tiles(**kw)
        assert e.value.status == 1
        assert ctx.tiles() == default
    # zeros (and a negative share) select the built-in defaults
    ctx.set_tiles(p2p_group_f32=0, p2p_early_share_f32=-1.0, m2l_slices=0)
    assert ctx.tiles() == default
