"""The benchmark configurations run at their specification on the GPU, and
the reference's own FMM outputs compared with the GPU's.

  config 1  N = 500 uniform in [0,1)^3, masses U[0.5,1.5], seed 1234, p = 8,
            leaf capacity 16, SyplecticEuler dt = 1e-3 for 10 steps; checked
            against the oracle's trajectory and against NaiveUpdate
  config 2  N = 1,048,576 Plummer, plain FP64, p = 10, leaf capacity 64: tree
            bit-exact against the oracle's reference rule, one evaluation
            against the FP64 direct sum, 5 RK4 steps dt = 1e-4 with the
            energy and momentum monitor
  config 4  256 Plummer spheres of 1e3..6.4e4 particles in one batch, mixed
            precision: every system within 1e-6 of its own direct sum, then
            20 RK4 steps dt = 1e-4 with each system's energy drift bounded
  config 5  N = 8,388,608 clustered mix of 32 Plummer spheres, learned order
            selection through the checked evaluation, gated at 1e-6
  golden    the reference's FMM::update (its L = 2, leaf threshold 20, G =
            6.6743e-11) on t100 / u500 / u2000 (tests/golden/ref_fmm.npz,
            written by the reference build): the GPU's error against the FP64
            direct sum is no larger than the reference's own, and the
            reference's test bound |a_fmm - a_naive| <= 1e-2
            (tests/test_fmm_vs_naive_3d.cpp:195-199) holds.
"""
import os

import numpy as np
import pytest

import nbfmm
import orc
import workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _say(*a):
    print(*a, flush=True)


def _direct(ctx, n, dx, dm, tgt):
    """FP64 compensated direct sum (eps2 = 0) at the targets, on the GPU."""
    dt_ = nbfmm.DeviceArray(np.ascontiguousarray(tgt, dtype=np.int64))
    dr = nbfmm.DeviceArray(nbytes=len(tgt) * 24, shape=(len(tgt), 3))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, len(tgt), dt_.ptr, 1, dr.ptr)
    r = dr.download(ctx.stream)
    dt_.free()
    dr.free()
    return r


# ---- config 1 -----------------------------------------------------------------------
def test_config1_se_500():
    xyz, m = workloads.uniform_cube(500, seed=1234)
    n = len(m)
    vel, acc0 = np.zeros((n, 3)), np.zeros((n, 3))
    p, thr, dt, steps = 8, 16, 1e-3, 10
    ctx = nbfmm.Context(order=p, leaf_capacity=thr, G=1.0, eps2=0.0, p2p_cut=0)
    try:
        _say(f"[cfg1] N={n} p={p} leaf={thr} SE dt={dt} x {steps}")
        # the first force evaluation against the direct sum (NaiveUpdate without softening)
        a0 = ctx.eval_accel(xyz, m)
        e0 = _rel(a0, orc.direct_targets(xyz, m, np.arange(n), G=1.0, eps2=0.0))
        _say(f"[cfg1] first evaluation vs direct sum (eps2=0): rel L2 {e0:.3e}")
        assert e0 < 2e-5
        dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (xyz, vel, acc0, m))
        ctx.step_se_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, steps)
        gx, gv = dx.download(ctx.stream), dv.download(ctx.stream)
    finally:
        ctx.close()
    # the oracle's exact FMM on the same tree rule, same steps
    cfg = orc.UpdCfg(orc.UPD_FMM_EXACT, p, thr, 1.0, 0.0, 0, 0)
    ox, ov, oa = xyz, vel, acc0
    for _ in range(steps):
        ox, ov, oa = orc.step("se", ox, ov, oa, m, cfg, dt)
    ex = _rel(gx - xyz, ox - xyz)
    _say(f"[cfg1] displacement vs oracle FMM trajectory: rel {ex:.2e}; positions rel {_rel(gx, ox):.2e}")
    assert _rel(gx, ox) <= 1e-13
    assert ex < 1e-10
    # against NaiveUpdate (naive_update.cpp:9-35): without softening the
    # trajectories differ by the FMM's truncation (6.5e-6 in the force),
    # amplified by the close pairs of unit masses at rest (accelerations up to
    # ~1e4, so dt = 1e-3 resolves them poorly); with the reference's eps2 =
    # 1e-4 close pairs differ by O(eps2 / r^2) as well
    for eps2 in (0.0, 1e-4):
        ncfg = orc.UpdCfg(orc.UPD_NAIVE, p, thr, 1.0, eps2, 0, 0)
        nx, nv, na = xyz, vel, acc0
        for _ in range(steps):
            nx, nv, na = orc.step("se", nx, nv, na, m, ncfg, dt)
        e = _rel(gx - xyz, nx - xyz)
        _say(f"[cfg1] displacement vs NaiveUpdate trajectory (eps2={eps2}): rel {e:.2e}")
        if eps2 == 0.0:
            assert e < 1e-2


# ---- config 2 -----------------------------------------------------------------------
@pytest.fixture(scope="module")
def plummer1m():
    return workloads.plummer(1 << 20, seed=42, with_velocity=True)


def test_config2_tree_bitexact_1m(plummer1m):
    xyz, m, _ = plummer1m
    ctx = nbfmm.Context(order=10, leaf_capacity=64, G=1.0, eps2=0.0, p2p_cut=1024)
    try:
        _say("[cfg2] tree of 1M Plummer, leaf 64: GPU")
        _, gt = ctx.eval_tree(xyz, m, with_il=True)
    finally:
        ctx.close()
    _say(f"[cfg2] GPU depth {gt['depth']}, {len(gt['key'])} cells, {len(gt['il'])} IL entries; oracle tree")
    ot = orc.tree(xyz, 64)
    assert gt["depth"] == ot["depth"]
    for k in ("level_start", "key", "pbeg", "pend", "parent", "perm", "nbr_off", "nbr", "il_off", "il"):
        assert np.array_equal(gt[k], ot[k]), k


def test_config2_eval_and_rk4_1m(plummer1m):
    xyz, m, vel = plummer1m
    n = len(m)
    tgt = workloads.sample_targets(n, 4096, seed=2)
    ctx = nbfmm.Context(order=10, leaf_capacity=64, G=1.0, eps2=0.0, p2p_cut=1024)
    try:
        dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
        da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
        ref = _direct(ctx, n, dx, dm, tgt)
        errs = {}
        for cut in (0, 1024):  # 0 = every interaction-list pair by M2L, as the reference
            ctx.set_params(p2p_cut=cut)
            ctx.eval_accel_dev(n, dx.ptr, dm.ptr, da.ptr)
            a = da.download(ctx.stream)[tgt]
            e_i = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
            errs[cut] = _rel(a, ref)
            _say(f"[cfg2] FP64 p=10 cut={cut}: rel L2 {errs[cut]:.3e}, p99 {np.percentile(e_i, 99):.2e}, "
                 f"max {e_i.max():.2e}")
        assert errs[0] < 5e-6 and errs[1024] < 5e-6
        # 5 classic RK4 steps with the monitor after each
        dv = nbfmm.DeviceArray(vel)
        rec = ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, 5, integrator=2, every=1)
        x5 = dx.download(ctx.stream)
    finally:
        ctx.close()
    E, P = rec[:, 2], rec[:, 3:6]
    dE = float(np.abs(E - E[0]).max() / abs(E[0]))
    pscale = float(np.sum(m * np.linalg.norm(vel, axis=1)))
    dP = float(np.abs(P - P[0]).max() / pscale)
    _say(f"[cfg2] RK4 dt=1e-4 x5: |dE/E| {dE:.2e}, |dP| / sum m|v| {dP:.2e}, max displacement "
         f"{np.abs(x5 - xyz).max():.2e}")
    assert np.all(np.isfinite(x5))
    assert dE < 1e-6
    assert dP < 1e-7


def _graph_vs_direct(ctx, n, xyz, vel, m, nsteps, off=None):
    """nsteps classic RK4 steps dt = 1e-4 run directly and with graphs on (the
    first step direct and captured, the rest replayed); the two final states."""
    out = []
    for graph in (0, 1):
        dx, dv, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(vel), nbfmm.DeviceArray(m)
        da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
        ctx.set_graph(graph)
        try:
            before = ctx.graph_stats()
            if off is None:
                ctx.step_rk4_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, nsteps, 1)
            else:
                ctx.batch_step_rk4_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, nsteps, 1)
            after = ctx.graph_stats()
        finally:
            ctx.set_graph(0)
        out.append((dx.download(ctx.stream), dv.download(ctx.stream), {k: after[k] - before[k] for k in after}))
    (xd, vd, _), (xg, vg, st) = out
    return _rel(xg - xyz, xd - xyz), _rel(vg - vel, vd - vel), st


def test_config2_graph_rk4_1m(plummer1m):
    # the captured RK4 step equals the direct one on the configuration's input
    xyz, m, vel = plummer1m
    n = len(m)
    ctx = nbfmm.Context(order=10, leaf_capacity=64, G=1.0, eps2=0.0, p2p_cut=1024)
    try:
        _say("[cfg2] 4 classic RK4 steps, direct and replayed from a graph")
        ex, ev, st = _graph_vs_direct(ctx, n, xyz, vel, m, 4)
    finally:
        ctx.close()
    _say(f"[cfg2] graph vs direct: displacement {ex:.2e}, velocity change {ev:.2e}, stats {st}")
    assert st["captures"] == 1 and st["launches"] == 3 and st["redone"] == 0, st
    assert ex < 1e-12 and ev < 1e-10


# ---- config 4 -----------------------------------------------------------------------
def test_config4_graph_rk4():
    # the batch of 256 systems (mixed precision, p = 13, cut 4096): replayed
    # RK4 steps equal the direct ones
    systems = workloads.batched_systems(256, seed=11, with_velocity=True)
    sizes = np.array([len(s[1]) for s in systems], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    xyz = np.concatenate([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    vel = np.concatenate([s[2] for s in systems])
    n = int(off[-1])
    ctx = nbfmm.Context(order=13, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=4096, precision=1)
    try:
        _say(f"[cfg4] {len(sizes)} systems, N={n}: 6 classic RK4 steps, direct and replayed from a graph")
        ex, ev, st = _graph_vs_direct(ctx, n, xyz, vel, m, 6, off=off)
    finally:
        ctx.close()
    _say(f"[cfg4] graph vs direct: displacement {ex:.2e}, velocity change {ev:.2e}, stats {st}")
    assert st["captures"] == 1 and st["launches"] == 5 and st["redone"] == 0, st
    assert ex < 1e-12 and ev < 1e-10


def test_config4_batched256():
    systems = workloads.batched_systems(256, seed=11, with_velocity=True)
    sizes = np.array([len(s[1]) for s in systems], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    xyz = np.concatenate([s[0] for s in systems])
    m = np.concatenate([s[1] for s in systems])
    vel = np.concatenate([s[2] for s in systems])
    n = int(off[-1])
    nsys = len(sizes)
    ctx = nbfmm.Context(order=13, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=4096, precision=1)
    try:
        dx, dm, dv = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m), nbfmm.DeviceArray(vel)
        da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
        _say(f"[cfg4] {nsys} systems, N={n}: batched evaluation, mixed p=13 cut 4096")
        ctx.batch_eval_accel_dev(off, dx.ptr, dm.ptr, da.ptr)
        a = da.download(ctx.stream)
        rng = np.random.default_rng(77)
        errs = []
        sub = nbfmm.Context(order=13, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=4096)
        energy0 = []
        try:
            for k in range(nsys):
                s = slice(off[k], off[k + 1])
                nk = int(sizes[k])
                sx, sm, sv = nbfmm.DeviceArray(xyz[s]), nbfmm.DeviceArray(m[s]), nbfmm.DeviceArray(vel[s])
                tg = np.sort(rng.choice(nk, min(nk, 64), replace=False))
                ref = _direct(sub, nk, sx, sm, tg)
                errs.append(_rel(a[s][tg], ref))
                energy0.append(_energy(sub, nk, sx, sv, sm))
            worst = max(errs)
            _say(f"[cfg4] worst system rel L2 {worst:.3e} (64 sampled targets each), mean {np.mean(errs):.3e}")
            assert worst <= 1e-6
            # 20 classic RK4 steps of the whole batch
            _say("[cfg4] 20 RK4 steps dt=1e-4")
            ctx.batch_step_rk4_dev(off, dx.ptr, dv.ptr, da.ptr, dm.ptr, 1e-4, 20, 1)
            x1, v1 = dx.download(ctx.stream), dv.download(ctx.stream)
            assert np.all(np.isfinite(x1)) and np.all(np.isfinite(v1))
            drift = []
            for k in range(nsys):
                s = slice(off[k], off[k + 1])
                nk = int(sizes[k])
                sx, sm, sv = nbfmm.DeviceArray(x1[s]), nbfmm.DeviceArray(m[s]), nbfmm.DeviceArray(v1[s])
                e1 = _energy(sub, nk, sx, sv, sm)
                drift.append(abs(e1 - energy0[k]) / abs(energy0[k]))
            _say(f"[cfg4] energy drift per system: max {max(drift):.2e}, median {np.median(drift):.2e}")
            assert max(drift) < 1e-5
        finally:
            sub.close()
    finally:
        ctx.close()


def _energy(ctx, n, dx, dv, dm):
    """Total energy of one system (FMM potential at the context's settings, FP64)."""
    da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
    dp = nbfmm.DeviceArray(nbytes=n * 8, shape=(n,))
    ctx.eval_accel_pot_dev(n, dx.ptr, dm.ptr, da.ptr, dp.ptr)
    d = ctx.diagnostics_dev(n, dx.ptr, dv.ptr, dm.ptr, dp.ptr)
    return d["total"]


# ---- config 5 -----------------------------------------------------------------------
def test_config5_learned_8m():
    n = 1 << 23
    ctx = nbfmm.Context(order=11, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=1024, precision=1, cubic_box=1)
    try:
        _say("[cfg5] training the order model on a 2^18 clustered mix")
        ctx.order_model_init(seed=7)
        xt, mt = workloads.plummer_mix(1 << 18, n_spheres=32, seed=201)
        tx, tm = nbfmm.DeviceArray(xt), nbfmm.DeviceArray(mt)
        rows = ctx.order_dataset_dev(len(mt), tx.ptr, tm.ptr, stride=2, seed=201, tol=0.5e-6)
        tx.free()
        tm.free()
        loss = ctx.order_train(epochs=40, batch=256, lr=2e-3, seed=3)
        _say(f"[cfg5] {rows} labelled cells, loss {loss[0]:.3f} -> {loss[-1]:.3f}; 8M mix")
        xyz, m = workloads.plummer_mix(n, n_spheres=32, seed=7)
        dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
        da = nbfmm.DeviceArray(nbytes=n * 24, shape=(n, 3))
        ctx.set_params(order_select=nbfmm.NBFMM_ORDER_LEARNED, order_min=4, order_conf=0.9)
        err, fb = ctx.eval_accel_checked_dev(n, dx.ptr, dm.ptr, da.ptr, ntgt=4096, seed=9, gate=1e-6,
                                             fallback_order=14)
        st = ctx.stats()
        _say(f"[cfg5] checked learned evaluation: err {err:.3e}, fell back {fb}, mean order {st.mean_order:.2f}")
        assert err <= 1e-6
        # the check's error is the error of what is left in d_acc: recheck on
        # other targets
        tgt = workloads.sample_targets(n, 2048, seed=5)
        ref = _direct(ctx, n, dx, dm, tgt)
        e2 = _rel(da.download(ctx.stream)[tgt], ref)
        _say(f"[cfg5] error on 2048 other targets {e2:.3e}")
        assert e2 <= 2e-6
    finally:
        ctx.close()


# ---- the reference's own outputs ------------------------------------------------------
@pytest.mark.parametrize("case", ["t100", "u500", "u2000"])
def test_reference_golden_inputs(case):
    g = np.load(os.path.join(GOLDEN, "ref_fmm.npz"))
    G = float(g["G"])
    xyz, m = g[f"{case}_xyz"], g[f"{case}_m"]
    n = len(m)
    ref_fmm, ref_naive = g[f"{case}_fmm"], g[f"{case}_naive"]
    direct = orc.direct_targets(xyz, m, np.arange(n), G=G, eps2=0.0)
    e_ref = _rel(ref_fmm, direct)
    ctx = nbfmm.Context(order=2, leaf_capacity=20, G=G, eps2=1e-4, p2p_cut=0)
    try:
        # NaiveUpdate: the reference's own output, bit for bit
        naive = ctx.direct_accel(xyz, m)
        assert np.array_equal(naive, ref_naive)
        out = {}
        for p in (2, 8):
            ctx.set_params(order=p)
            out[p] = ctx.eval_accel(xyz, m)
    finally:
        ctx.close()
    e2, e8 = _rel(out[2], direct), _rel(out[8], direct)
    _say(f"[golden] {case}: reference FMM (L=2, thr 20) rel L2 vs direct {e_ref:.3e}; GPU p=2 {e2:.3e}, "
         f"p=8 {e8:.3e}; GPU p=2 vs reference FMM {_rel(out[2], ref_fmm):.3e}")
    assert e2 <= e_ref
    assert e8 <= e_ref
    # the reference's test bound (absolute, per component, against NaiveUpdate)
    assert np.abs(out[2] - ref_naive).max() <= 1e-2
