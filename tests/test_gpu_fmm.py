"""GPU parity tests: libnbfmm.so (through its C ABI) against the CPU oracle.

Bars: bit-exact for the tree (keys, leaves, order, lists) and for the
NaiveUpdate formula; FP64 FMM within 1e-12 relative (L2 over all particles)
of the oracle's exact FMM on the same tree, lists and expansion order.
"""
import numpy as np
import pytest

import nbfmm
import orc
import workloads

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    c = nbfmm.Context(order=4, leaf_capacity=16, G=1.0, eps2=0.0)
    yield c
    c.close()


def _check_tree(gt, ot, n):
    assert gt["depth"] == ot["depth"]
    assert np.array_equal(gt["level_start"], ot["level_start"])
    assert np.array_equal(gt["key"], ot["key"])
    assert np.array_equal(gt["pbeg"], ot["pbeg"])
    assert np.array_equal(gt["pend"], ot["pend"])
    assert np.array_equal(gt["parent"], ot["parent"])
    assert np.array_equal(gt["perm"], ot["perm"])
    assert np.array_equal(gt["nbr_off"], ot["nbr_off"])
    assert np.array_equal(gt["nbr"], ot["nbr"])
    if "il" in gt:
        assert np.array_equal(gt["il_off"], ot["il_off"])
        assert np.array_equal(gt["il"], ot["il"])


@pytest.mark.parametrize("case", ["uniform2k", "plummer50k", "tiny", "clustered", "duplicates"])
def test_tree_bitexact(ctx, case):
    if case == "duplicates":
        # 40 particles at one point: the depth loop runs to the deepest level
        # (16), so every level's cells come from one set of launches
        xyz, m = workloads.uniform_cube(3000, seed=4)
        xyz[:40] = xyz[0]
        thr = 16
    elif case == "uniform2k":
        xyz, m = workloads.uniform_cube(2000, seed=3)
        thr = 16
    elif case == "plummer50k":
        xyz, m = workloads.plummer(50000, seed=5)
        thr = 64
    elif case == "tiny":
        xyz, m = workloads.uniform_cube(7, seed=9)
        thr = 20
    else:
        xyz, m = workloads.plummer_mix(20000, n_spheres=6, seed=2)
        thr = 20
    ctx.set_params(leaf_capacity=thr, order=4, p2p_cut=0)
    print(f"[tree] case={case} n={len(m)} thr={thr}", flush=True)
    _, gt = ctx.eval_tree(xyz, m, with_il=True)
    ot = orc.tree(xyz, thr)
    _check_tree(gt, ot, len(m))


def test_tree_depth_guess():
    """Each evaluation first sorts at the previous tree's depth and replays the
    reference's depth loop on every level's fullest cell.  A sequence of inputs
    whose depths go up and down takes every branch (confirmed, shallower:
    sorted again, deeper: the loop goes on); every tree is the oracle's, with
    the reference loop's iteration count, and a repeated input sorts once."""
    c = nbfmm.Context(order=4, leaf_capacity=16, G=1.0, eps2=0.0, p2p_cut=0)
    dup, dm = workloads.uniform_cube(3000, seed=4)
    dup[:40] = dup[0]
    seq = [("plummer50k", *workloads.plummer(50000, seed=5), 64),
           ("plummer50k again", *workloads.plummer(50000, seed=5), 64),
           ("uniform2k", *workloads.uniform_cube(2000, seed=3), 16),
           ("duplicates", dup, dm, 16),
           ("duplicates again", dup, dm, 16),
           ("clustered", *workloads.plummer_mix(20000, n_spheres=6, seed=2), 20),
           ("tiny", *workloads.uniform_cube(7, seed=9), 20),
           ("plummer50k thr 8", *workloads.plummer(50000, seed=5), 8)]
    try:
        prev = None
        for name, xyz, m, thr in seq:
            c.set_params(leaf_capacity=thr)
            _, gt = c.eval_tree(xyz, m)
            st = c.stats()
            ot = orc.tree(xyz, thr)
            print(f"[guess] {name}: depth {gt['depth']} (previous {prev}), iterations {st.depth_iterations}, "
                  f"sorts {st.depth_sorts}", flush=True)
            _check_tree(gt, ot, len(m))
            assert st.depth_iterations == len(ot["trace"])
            if prev == gt["depth"]:
                assert st.depth_sorts == 1
            else:
                assert st.depth_sorts >= 2 or prev is None
            prev = gt["depth"]
    finally:
        c.close()


def test_direct_naive_bitwise(ctx):
    xyz, m = workloads.uniform_cube(1500, seed=11)
    ctx.set_params(G=6.6743e-11, eps2=1e-4)
    a = ctx.direct_accel(xyz, m)
    b = orc.direct_naive(xyz, m, G=6.6743e-11, eps2=1e-4)
    ctx.set_params(G=1.0, eps2=0.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("p,cut", [(2, 0), (4, 0), (8, 0), (10, 256), (12, 1024), (13, 512), (14, 0), (6, 1 << 40)])
def test_fmm_parity_exact(ctx, p, cut):
    xyz, m = workloads.uniform_cube(3000, seed=p)
    ctx.set_params(order=p, leaf_capacity=16, p2p_cut=cut, G=1.0, eps2=0.0)
    print(f"[fmm] p={p} cut={cut}", flush=True)
    a = ctx.eval_accel(xyz, m)
    b = orc.fmm_exact(xyz, m, p, 16, G=1.0, center_mode=0, p2p_cut=cut)
    assert _rel(a, b) < 1e-12


@pytest.mark.parametrize("cut", [0, 256])
def test_fmm_parity_elongated(ctx, cut):
    # a 2.5 : 1.25 : 1 bounding box: the nearest classes along the short axes
    # are not admissible for M2L and go to P2P (same rule in the oracle)
    rng = np.random.default_rng(17)
    xyz = rng.random((6000, 3)) * np.array([2.5, 1.25, 1.0])
    m = rng.uniform(0.5, 1.5, 6000)
    ctx.set_params(order=8, leaf_capacity=16, p2p_cut=cut, G=1.0, eps2=0.0)
    a = ctx.eval_accel(xyz, m)
    b = orc.fmm_exact(xyz, m, 8, 16, G=1.0, center_mode=0, p2p_cut=cut)
    assert _rel(a, b) < 1e-12


def test_elongated_box_converges(ctx):
    # with the inadmissible classes evaluated directly the error keeps falling with p
    rng = np.random.default_rng(18)
    xyz = rng.random((8000, 3)) * np.array([1.6, 1.0, 1.0])
    m = rng.uniform(0.5, 1.5, 8000)
    ctx.set_params(G=1.0, eps2=0.0)
    d = ctx.direct_accel(xyz, m)
    err = {}
    for p in (4, 8, 12):
        ctx.set_params(order=p, leaf_capacity=16, p2p_cut=0)
        err[p] = _rel(ctx.eval_accel(xyz, m), d)
    print(f"[elongated] errors {err}")
    assert err[12] < 0.1 * err[8] < 0.01 * err[4]


def test_fmm_parity_elongated_large_cells(ctx):
    # cells of about 1100 particles at level 2: pairs of the inadmissible
    # classes with n_t n_s >= 2^20 stay M2L, smaller ones go to P2P (same cap
    # in the oracle)
    rng = np.random.default_rng(21)
    xyz = rng.random((70000, 3)) * np.array([2.5, 1.25, 1.0])
    m = rng.uniform(0.5, 1.5, 70000)
    ctx.set_params(order=6, leaf_capacity=1300, p2p_cut=0, G=1.0, eps2=0.0)
    a = ctx.eval_accel(xyz, m)
    b = orc.fmm_exact(xyz, m, 6, 1300, G=1.0, center_mode=0, p2p_cut=0)
    ctx.set_params(leaf_capacity=16)
    assert _rel(a, b) < 1e-12


@pytest.mark.parametrize("case", ["elongated", "plummer"])
def test_tree_bitexact_cubic(ctx, case):
    if case == "elongated":
        rng = np.random.default_rng(22)
        xyz = rng.random((9000, 3)) * np.array([4.0, 1.0, 0.5]) + np.array([-3.0, 7.0, 0.25])
        m = np.ones(9000)
    else:
        xyz, m = workloads.plummer(30000, seed=23)
    ctx.set_params(leaf_capacity=16, order=4, p2p_cut=0, cubic_box=1)
    try:
        _, gt = ctx.eval_tree(xyz, m, with_il=True)
    finally:
        ctx.set_params(cubic_box=0)
    ot = orc.tree(xyz, 16, cubic=True)
    _check_tree(gt, ot, len(m))


@pytest.mark.parametrize("p,cut", [(8, 0), (10, 512)])
def test_fmm_parity_cubic(ctx, p, cut):
    rng = np.random.default_rng(24)
    xyz = rng.random((6000, 3)) * np.array([4.0, 1.0, 1.0])
    m = rng.uniform(0.5, 1.5, 6000)
    ctx.set_params(order=p, leaf_capacity=16, p2p_cut=cut, G=1.0, eps2=0.0, cubic_box=1)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(cubic_box=0)
    b = orc.fmm_exact(xyz, m, p, 16, G=1.0, center_mode=0, p2p_cut=cut, cubic=True)
    assert _rel(a, b) < 1e-12


def test_cubic_box_converges_on_long_box(ctx):
    # a 4 : 1 : 1 box on its bounding cube: every class is admissible and the
    # error falls geometrically with p
    rng = np.random.default_rng(25)
    xyz = rng.random((8000, 3)) * np.array([4.0, 1.0, 1.0])
    m = rng.uniform(0.5, 1.5, 8000)
    ctx.set_params(G=1.0, eps2=0.0)
    d = ctx.direct_accel(xyz, m)
    err = {}
    try:
        for p in (4, 8, 12):
            ctx.set_params(order=p, leaf_capacity=16, p2p_cut=0, cubic_box=1)
            err[p] = _rel(ctx.eval_accel(xyz, m), d)
    finally:
        ctx.set_params(cubic_box=0)
    print(f"[cubic] errors {err}")
    assert err[12] < 5e-6 and err[12] < 0.05 * err[8] < 0.0025 * err[4]


@pytest.mark.parametrize("p,cut,tol", [(8, 0, 0.1), (12, 256, 0.05), (6, 0, 1.0)])
def test_fmm_parity_class_orders(ctx, p, cut, tol):
    # distance-dependent M2L orders: the same class table and per-pair
    # truncation as the oracle
    xyz, m = workloads.uniform_cube(3000, seed=30 + p)
    ctx.set_params(order=p, leaf_capacity=16, p2p_cut=cut, G=1.0, eps2=0.0, class_order_tol=tol)
    try:
        a = ctx.eval_accel(xyz, m)
        st = ctx.stats()
    finally:
        ctx.set_params(class_order_tol=0.0)
    b = orc.fmm_exact(xyz, m, p, 16, G=1.0, center_mode=0, p2p_cut=cut, class_order_tol=tol)
    assert _rel(a, b) < 1e-12
    assert st.mean_order < p  # some pairs run below p


@pytest.mark.parametrize("tol", [0.05, 0.3])
def test_mixed_parity_class_orders(ctx, tol):
    xyz, m = workloads.plummer(20000, seed=31)
    ctx.set_params(order=11, leaf_capacity=32, p2p_cut=1024, G=1.0, eps2=0.0, precision=1, class_order_tol=tol)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(class_order_tol=0.0, precision=0)
    b = orc.fmm_exact(xyz, m, 11, 32, G=1.0, center_mode=0, p2p_cut=1024, class_order_tol=tol)
    assert _rel(a, b) < 2e-6


def test_class_orders_bad_params(ctx):
    with pytest.raises(nbfmm.NbfmmError):
        ctx.set_params(order=3, class_order_tol=0.1)
    with pytest.raises(nbfmm.NbfmmError):
        ctx.set_params(class_order_tol=-0.5)
    ctx.set_params(order=4, class_order_tol=0.0)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["duplicates", "leaf500"])
def test_fmm_parity_big_leaves(ctx, case, precision):
    # P2M of one wavefront per cell over leaves of many particles, and P2P
    # pairs at zero distance: 40 coincident particles (the depth loop runs to
    # level 16, the deepest leaf holds all 40), and leaves of 415 to 511
    # particles (30k uniform, leaf capacity 5000: depth 2)
    if case == "duplicates":
        xyz, m = workloads.uniform_cube(3000, seed=4)
        xyz[:40] = xyz[0]
        p, thr, cut = 8, 16, 0
    else:
        xyz, m = workloads.uniform_cube(30000, seed=23)
        p, thr, cut = 8, 5000, 0
    ctx.set_params(order=p, leaf_capacity=thr, p2p_cut=cut, G=1.0, eps2=0.0, precision=precision)
    print(f"[big-leaves] case={case} precision={precision}", flush=True)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(precision=0)
    b = orc.fmm_exact(xyz, m, p, thr, G=1.0, center_mode=0, p2p_cut=cut)
    err = _rel(a, b)
    print(f"[big-leaves] rel L2 vs oracle {err:.3e}")
    assert err < (1e-12 if precision == 0 else 2e-6)
    if case == "duplicates":
        # coincident particles exert nothing on each other: the 40 share one
        # acceleration, that of the other 2960 particles
        assert np.allclose(a[:40], a[0], rtol=1e-12 if precision == 0 else 1e-6, atol=0)


@pytest.mark.parametrize("precision", [0, 1])
def test_isolated_timing_same_result(ctx, precision):
    # timing mode 2 runs the phases one after another (P2P alone at the end):
    # the same result to rounding (the near-field atomics may add in another
    # order), and every phase timed
    xyz, m = workloads.plummer(30000, seed=19)
    ctx.set_params(order=8, leaf_capacity=32, p2p_cut=512, G=1.0, eps2=0.0, precision=precision)
    try:
        a = ctx.eval_accel(xyz, m)
        ctx.set_timing(2)
        b = ctx.eval_accel(xyz, m)
        st = ctx.stats()
    finally:
        ctx.set_timing(1)
        ctx.set_params(precision=0)
    assert _rel(b, a) < 1e-13
    assert st.isolated == 1
    for k in ("ms_tree", "ms_lists", "ms_p2m", "ms_m2l", "ms_l2p", "ms_p2p", "ms_p2p_kernel"):
        assert getattr(st, k) > 0.0, k
    assert st.ms_m2m >= 0.0 and st.ms_l2l >= 0.0
    assert st.n_p2m_cells > 0 and st.n_l2p_cells > 0
    print(f"[isolated] precision={precision} " + " ".join(f"{k}={getattr(st, k):.3f}" for k in
                                                            ("ms_tree", "ms_m2l", "ms_p2p_kernel", "ms_total")))


def test_fmm_parity_plummer(ctx):
    xyz, m = workloads.plummer(40000, seed=1)
    ctx.set_params(order=8, leaf_capacity=64, p2p_cut=512, G=1.0, eps2=0.0)
    a = ctx.eval_accel(xyz, m)
    b = orc.fmm_exact(xyz, m, 8, 64, G=1.0, center_mode=0, p2p_cut=512)
    assert _rel(a, b) < 1e-12


def test_all_p2p_equals_direct(ctx):
    # every interaction-list pair evaluated directly: the FMM is the direct sum
    xyz, m = workloads.plummer(20000, seed=8)
    ctx.set_params(order=4, leaf_capacity=32, p2p_cut=1 << 50, G=1.0, eps2=0.0)
    a = ctx.eval_accel(xyz, m)
    n = len(m)
    dx, dm, da = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, n, None, 1, da.ptr)
    d = da.download(ctx.stream)
    assert _rel(a, d) < 1e-13


def test_fmm_accuracy_vs_direct(ctx):
    xyz, m = workloads.plummer(100000, seed=2)
    n = len(m)
    tgt = workloads.sample_targets(n, 2048)
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    dt_ = nbfmm.DeviceArray(tgt)
    dd = nbfmm.DeviceArray(np.zeros((len(tgt), 3)))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, len(tgt), dt_.ptr, 1, dd.ptr)
    ref = dd.download(ctx.stream)
    # the GPU direct oracle agrees with the CPU oracle on a few targets
    few = tgt[:32]
    cpu = orc.direct_targets(xyz, m, few, G=1.0)
    assert _rel(ref[:32], cpu) < 1e-14
    errs = []
    for p in (4, 8, 12):
        ctx.set_params(order=p, leaf_capacity=64, p2p_cut=1024)
        a = ctx.eval_accel(xyz, m)
        errs.append(_rel(a[tgt], ref))
    print("errors vs direct", errs)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 5e-7


def test_se_rk4_steps(ctx):
    xyz, m, vel = workloads.plummer(3000, seed=4, with_velocity=True)
    n = len(m)
    acc0 = np.zeros((n, 3))
    p, thr, cut, dt = 6, 16, 256, 1e-3
    ctx.set_params(order=p, leaf_capacity=thr, p2p_cut=cut, G=1.0, eps2=0.0)
    cfg = orc.UpdCfg(orc.UPD_FMM_EXACT, p, thr, 1.0, 0.0, 0, cut)
    for kind, mode in (("se", 0), ("rk4", 0), ("rk4", 1)):
        dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (xyz, vel, acc0, m))
        if kind == "se":
            ctx.step_se_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, 2)
        else:
            ctx.step_rk4_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, 2, mode)
        gx, gv = dx.download(ctx.stream), dv.download(ctx.stream)
        ox, ov, oa = xyz, vel, acc0
        for _ in range(2):
            ox, ov, oa = orc.step(kind, ox, ov, oa, m, cfg, dt, rk4_mode=mode)
        assert _rel(gx, ox) < 1e-14, kind
        assert _rel(gv - vel, ov - vel) < 1e-10, kind


def test_bad_inputs(ctx):
    xyz, m = workloads.uniform_cube(100, seed=1)
    xyz[5, 1] = np.nan
    with pytest.raises(nbfmm.NbfmmError) as e:
        ctx.eval_accel(xyz, m)
    assert e.value.status == 2
    with pytest.raises(nbfmm.NbfmmError):
        ctx.set_params(order=0)
    ctx.params.order = 4
    one = ctx.eval_accel(np.zeros((1, 3)), np.ones(1))
    assert np.all(one == 0)
    same = ctx.eval_accel(np.zeros((50, 3)), np.ones(50))
    assert np.all(same == 0)


# ---- mixed precision (FP32 P2P with compensated sums, FP32 MFMA M2L, FP32 P2M/L2P) ----
# Tolerance against the FP64 oracle on the same tree, lists, p and cut: 2e-6
# relative L2 (FP32 rounding of ~1e-7 per term over the expansion and pair
# sums); a wrong index, sign or layout gives errors of order 1.
@pytest.mark.parametrize("p,cut", [(4, 0), (8, 256), (10, 0), (12, 1024), (13, 256), (14, 0)])
def test_mixed_parity(ctx, p, cut):
    xyz, m = workloads.uniform_cube(3000, seed=p + 7)
    ctx.set_params(order=p, leaf_capacity=16, p2p_cut=cut, G=1.0, eps2=0.0, precision=1)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(precision=0)
    b = orc.fmm_exact(xyz, m, p, 16, G=1.0, center_mode=0, p2p_cut=cut)
    err = _rel(a, b)
    print(f"[mixed] p={p} cut={cut} rel L2 vs FP64 oracle {err:.3e}")
    assert err < 2e-6


def test_mixed_fp64_levels(ctx):
    # mixed precision with the coarse M2M/L2L levels (or all of them) in FP64:
    # the same FP64 oracle bar; the all-FP64 translations come closer to it
    xyz, m = workloads.plummer(40000, seed=37)
    b = orc.fmm_exact(xyz, m, 10, 32, G=1.0, center_mode=0, p2p_cut=256)
    err = {}
    try:
        for k in (0, 3, 99):
            ctx.set_params(order=10, leaf_capacity=32, p2p_cut=256, G=1.0, eps2=0.0, precision=1, mixed_fp64_levels=k)
            err[k] = _rel(ctx.eval_accel(xyz, m), b)
    finally:
        ctx.set_params(precision=0, mixed_fp64_levels=0)
    print(f"[mixed-fp64-levels] rel L2 vs FP64 oracle {err}")
    assert all(e < 2e-6 for e in err.values())
    assert err[99] <= err[0] * 1.05


def test_mixed_all_p2p_vs_direct(ctx):
    # everything direct: the FP32 compensated P2P against the FP64 direct sum,
    # within the metric's 1e-6 (FP32 offsets from the task anchor bound the
    # error of the closest pairs)
    xyz, m = workloads.plummer(20000, seed=8)
    n = len(m)
    ctx.set_params(order=4, leaf_capacity=32, p2p_cut=1 << 50, G=1.0, eps2=0.0, precision=1)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(precision=0)
    dx, dm, da = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m), nbfmm.DeviceArray(np.zeros((n, 3)))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, n, None, 1, da.ptr)
    d = da.download(ctx.stream)
    err = _rel(a, d)
    print(f"[mixed] all-P2P rel L2 vs direct {err:.3e}")
    assert err < 1e-6


def test_mixed_accuracy_vs_direct(ctx):
    xyz, m = workloads.plummer(100000, seed=2)
    n = len(m)
    tgt = workloads.sample_targets(n, 2048)
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    dt_ = nbfmm.DeviceArray(tgt)
    dd = nbfmm.DeviceArray(np.zeros((len(tgt), 3)))
    ctx.direct_accel_dev(n, dx.ptr, dm.ptr, len(tgt), dt_.ptr, 1, dd.ptr)
    ref = dd.download(ctx.stream)
    ctx.set_params(order=12, leaf_capacity=64, p2p_cut=1024, precision=1)
    try:
        a = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(precision=0)
    err = _rel(a[tgt], ref)
    print(f"[mixed] p=12 rel L2 vs direct {err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_m2l_in_passes(ctx, precision):
    # the translation done in many passes over the class-sorted pairs (a small
    # partial buffer) gives the same result as one pass (mixed precision: the
    # class-major contraction, m2l_mode 1, which has the passes)
    xyz, m = workloads.uniform_cube(3000, seed=8)
    ctx.set_params(order=8, leaf_capacity=16, p2p_cut=0, G=1.0, eps2=0.0, precision=precision, m2l_pass_pairs=5000)
    ctx.set_tiles(m2l_mode=1)
    try:
        a = ctx.eval_accel(xyz, m)
        npass = -(-ctx.stats().n_m2l // 5000)
    finally:
        ctx.set_params(precision=0, m2l_pass_pairs=0)
        ctx.set_tiles(m2l_mode=0)
    b = orc.fmm_exact(xyz, m, 8, 16, G=1.0, center_mode=0, p2p_cut=0)
    print(f"[passes] precision={precision} passes={npass} rel L2 {_rel(a, b):.3e}")
    assert npass > 10
    assert _rel(a, b) < (1e-12 if precision == 0 else 2e-6)


def test_operator_cache_follows_geometry_order_and_precision():
    # the operators are rebuilt only when the leaf geometry, the order or the
    # precision changed since the last build: a context that goes through
    # boxes of other shapes, orders and precisions gives what a fresh context
    # gives for each evaluation
    xa, ma = workloads.plummer(30000, seed=81)
    xb = xa * np.array([3.0, 1.0, 0.4])
    runs = [(xa, 1, 9), (xb, 1, 9), (xa, 1, 9), (xa, 0, 9), (xa, 1, 9), (xa, 1, 7), (xb, 0, 7), (xa, 1, 9)]
    c = nbfmm.Context(order=9, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=256, class_order_tol=0.05)
    try:
        for k, (x, prec, p) in enumerate(runs):
            c.set_params(precision=prec, order=p)
            a = c.eval_accel(x, ma)
            f = nbfmm.Context(order=p, leaf_capacity=32, G=1.0, eps2=0.0, p2p_cut=256, class_order_tol=0.05,
                              precision=prec)
            try:
                b = f.eval_accel(x, ma)
            finally:
                f.close()
            e = _rel(a, b)
            print(f"[ops cache] run {k} precision={prec} p={p}: rel diff to a fresh context {e:.2e}", flush=True)
            assert e < (1e-13 if prec == 0 else 1e-7), (k, e)
    finally:
        c.close()
