"""ctypes binding of the CPU oracle (liborc.so) and of the reference build
(_ref/libnbref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libnbref.so")

_vp, _i64, _dbl, _u32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


_orc = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise RuntimeError(f"{ORC_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(ORC_PATH)
        L.orc_morton_encode3.restype = ctypes.c_uint64
        L.orc_morton_encode3.argtypes = [_u32, _u32, _u32]
        L.orc_morton_decode3.argtypes = [ctypes.c_uint64, _vp, _vp, _vp]
        L.orc_tree_build.restype = _vp
        L.orc_tree_build.argtypes = [_i64, _vp, _u32]
        L.orc_tree_build_box.restype = _vp
        L.orc_tree_build_box.argtypes = [_i64, _vp, _u32, ctypes.c_int]
        L.orc_tree_free.argtypes = [_vp]
        L.orc_tree_depth.restype = ctypes.c_int
        L.orc_tree_depth.argtypes = [_vp]
        for f in ("orc_tree_ncells", "orc_tree_nbr_total", "orc_tree_il_total"):
            getattr(L, f).restype = _i64
            getattr(L, f).argtypes = [_vp]
        L.orc_tree_depth_trace.restype = ctypes.c_int
        L.orc_tree_depth_trace.argtypes = [_vp, _vp, _vp]
        for f in ("orc_tree_geometry",):
            getattr(L, f).argtypes = [_vp, _vp, _vp, _vp]
        L.orc_tree_levels.argtypes = [_vp, _vp]
        L.orc_tree_cells.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.orc_tree_perm.argtypes = [_vp, _vp]
        L.orc_tree_particle_keys.argtypes = [_vp, _vp]
        L.orc_tree_lists.argtypes = [_vp, _vp, _vp, _vp, _vp]
        for f in ("orc_ref_powerset", "orc_ref_irregular", "orc_exact_regular", "orc_exact_singular"):
            getattr(L, f).argtypes = [_u32, _vp, _vp]
        for f in ("orc_ref_p2m_add", "orc_exact_p2m_add"):
            getattr(L, f).argtypes = [_u32, _vp, _dbl, _vp]
        for f in ("orc_ref_m2m", "orc_ref_m2l", "orc_ref_l2l", "orc_ref_l2p", "orc_exact_m2m", "orc_exact_m2l",
                  "orc_exact_l2l"):
            getattr(L, f).argtypes = [_u32, _vp, _vp, _vp]
        for f in ("orc_exact_l2p", "orc_exact_m2p"):
            getattr(L, f).argtypes = [_u32, _vp, _vp, _vp, _vp]
        L.orc_fmm_refcompat.argtypes = [_i64, _vp, _vp, _u32, _u32, _dbl, _vp]
        L.orc_fmm_exact.argtypes = [_i64, _vp, _vp, _u32, _u32, _dbl, ctypes.c_int, _i64, _vp]
        L.orc_fmm_exact_box.argtypes = [_i64, _vp, _vp, _u32, _u32, _dbl, ctypes.c_int, _i64, ctypes.c_int, _vp]
        L.orc_fmm_exact_opt.argtypes = [_i64, _vp, _vp, _u32, _u32, _dbl, ctypes.c_int, _i64, ctypes.c_int, _dbl,
                                        _vp]
        L.orc_class_orders.argtypes = [_vp, ctypes.c_int, _dbl, _vp]
        L.orc_class_orders.restype = None
        L.orc_direct_naive.argtypes = [_i64, _vp, _vp, _dbl, _dbl, _vp]
        L.orc_direct_targets.argtypes = [_i64, _vp, _vp, _dbl, _dbl, _i64, _vp, _vp]
        L.orc_step_se.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, _dbl]
        L.orc_step_rk4.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, _dbl, ctypes.c_int]
        _orc = L
    return _orc


class UpdCfg(ctypes.Structure):
    _fields_ = [("updater", ctypes.c_int), ("order", ctypes.c_uint32), ("thr", ctypes.c_uint32),
                ("G", ctypes.c_double), ("eps2", ctypes.c_double), ("center_mode", ctypes.c_int),
                ("p2p_cut", ctypes.c_int64)]


UPD_NAIVE, UPD_FMM_REF, UPD_FMM_EXACT = 0, 1, 2


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def morton_encode(x, y, z):
    return orc().orc_morton_encode3(x, y, z)


def morton_decode(k):
    c = [ctypes.c_uint32() for _ in range(3)]
    orc().orc_morton_decode3(k, *(ctypes.byref(v) for v in c))
    return tuple(v.value for v in c)


def class_orders(ext, p, tol):
    """M2L order of each of the 343 offset classes (index (dz+3)*49 + (dy+3)*7 + (dx+3))
    for a box of extents `ext` at order p and class_order_tol `tol`."""
    q = np.zeros(343, np.int32)
    e = _f64(np.asarray(ext, dtype=np.float64))
    orc().orc_class_orders(_p(e), int(p), float(tol), _p(q))
    return q


def tree(xyz, thr, cubic=False):
    """The reference tree rule as numpy arrays (cells level 0 first); with
    cubic=True on the bounding cube centred on the reference's box."""
    xyz = _f64(xyz)
    n = len(xyz)
    L = orc()
    h = L.orc_tree_build_box(n, _p(xyz), thr, int(cubic))
    if not h:
        raise ValueError("oracle tree build rejected the input")
    try:
        D = L.orc_tree_depth(h)
        nc = L.orc_tree_ncells(h)
        t = {"depth": D}
        d = np.zeros(32, np.int32)
        mc = np.zeros(32, np.uint32)
        k = L.orc_tree_depth_trace(h, _p(d), _p(mc))
        t["trace"] = list(zip(d[:k].tolist(), mc[:k].tolist()))
        lo, hi, ls = np.zeros(3), np.zeros(3), np.zeros(3)
        L.orc_tree_geometry(h, _p(lo), _p(hi), _p(ls))
        t.update(lo=lo, hi=hi, leaf_size=ls)
        t["level_start"] = np.zeros(D + 2, np.int64)
        L.orc_tree_levels(h, _p(t["level_start"]))
        t["key"] = np.zeros(nc, np.uint64)
        t["pbeg"], t["pend"], t["parent"] = (np.zeros(nc, np.int64) for _ in range(3))
        L.orc_tree_cells(h, _p(t["key"]), _p(t["pbeg"]), _p(t["pend"]), _p(t["parent"]))
        t["perm"] = np.zeros(n, np.int64)
        L.orc_tree_perm(h, _p(t["perm"]))
        t["pkey"] = np.zeros(n, np.uint64)
        L.orc_tree_particle_keys(h, _p(t["pkey"]))
        nn, ni = L.orc_tree_nbr_total(h), L.orc_tree_il_total(h)
        t["nbr_off"], t["il_off"] = np.zeros(nc + 1, np.int64), np.zeros(nc + 1, np.int64)
        t["nbr"], t["il"] = np.zeros(max(nn, 1), np.int64), np.zeros(max(ni, 1), np.int64)
        L.orc_tree_lists(h, _p(t["nbr_off"]), _p(t["nbr"]), _p(t["il_off"]), _p(t["il"]))
        t["nbr"], t["il"] = t["nbr"][:nn], t["il"][:ni]
        return t
    finally:
        L.orc_tree_free(h)


def fmm_refcompat(xyz, m, L=2, thr=20, G=1.0):
    xyz, m = _f64(xyz), _f64(m)
    acc = np.zeros_like(xyz)
    rc = orc().orc_fmm_refcompat(len(m), _p(xyz), _p(m), L, thr, G, _p(acc))
    if rc:
        raise ValueError("oracle refcompat FMM failed")
    return acc


def fmm_exact(xyz, m, p, thr, G=1.0, center_mode=0, p2p_cut=0, cubic=False, class_order_tol=0.0):
    xyz, m = _f64(xyz), _f64(m)
    acc = np.zeros_like(xyz)
    rc = orc().orc_fmm_exact_opt(len(m), _p(xyz), _p(m), p, thr, G, center_mode, p2p_cut, int(cubic),
                                 float(class_order_tol), _p(acc))
    if rc:
        raise ValueError("oracle exact FMM failed")
    return acc


def direct_naive(xyz, m, G=1.0, eps2=1e-4):
    xyz, m = _f64(xyz), _f64(m)
    acc = np.zeros_like(xyz)
    orc().orc_direct_naive(len(m), _p(xyz), _p(m), G, eps2, _p(acc))
    return acc


def direct_targets(xyz, m, tgt, G=1.0, eps2=0.0):
    xyz, m = _f64(xyz), _f64(m)
    tgt = np.ascontiguousarray(tgt, dtype=np.int64)
    acc = np.zeros((len(tgt), 3))
    orc().orc_direct_targets(len(m), _p(xyz), _p(m), G, eps2, len(tgt), _p(tgt), _p(acc))
    return acc


def step(kind, xyz, vel, acc, m, cfg, dt, rk4_mode=0):
    """One integrator step on copies; returns (xyz, vel, acc)."""
    x, v, a, m = _f64(xyz).copy(), _f64(vel).copy(), _f64(acc).copy(), _f64(m)
    if kind == "se":
        rc = orc().orc_step_se(len(m), _p(x), _p(v), _p(a), _p(m), ctypes.byref(cfg), dt)
    else:
        rc = orc().orc_step_rk4(len(m), _p(x), _p(v), _p(a), _p(m), ctypes.byref(cfg), dt, rk4_mode)
    if rc:
        raise ValueError("oracle step failed")
    return x, v, a


# ---- the reference's own build (oracle/_ref) -----------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_PATH} is not built; run `make -C oracle ref`")
        L = ctypes.CDLL(REF_PATH)
        L.ref_grav_const.restype = _dbl
        for f in ("ref_fmm_update", "ref_naive_update"):
            getattr(L, f).argtypes = [_i64, _vp, _vp, _vp]
        L.ref_integrate.argtypes = [_i64, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dbl, ctypes.c_int,
                                    ctypes.c_int]
        L.ref_tree_build.restype = ctypes.c_int
        L.ref_tree_build.argtypes = [_i64, _vp, _vp]
        L.ref_tree_level_count.restype = _i64
        L.ref_tree_level_count.argtypes = [ctypes.c_int]
        L.ref_tree_level_keys.argtypes = [ctypes.c_int, _vp]
        L.ref_tree_leaf_particles.restype = _i64
        L.ref_tree_leaf_particles.argtypes = [ctypes.c_uint64, _vp]
        L.ref_tree_lists.argtypes = [ctypes.c_int, ctypes.c_uint64, _vp, _vp, _vp, _vp]
        for f in ("ref_powerset", "ref_irregular"):
            getattr(L, f).argtypes = [ctypes.c_uint, _vp, _vp]
        L.ref_p2m.argtypes = [ctypes.c_uint, _i64, _vp, _vp, _vp]
        for f in ("ref_m2m", "ref_m2l", "ref_l2l", "ref_l2p"):
            getattr(L, f).argtypes = [ctypes.c_uint, _vp, _vp, _vp]
        _ref = L
    return _ref


def ref_fmm(xyz, m):
    xyz, m = _f64(xyz), _f64(m)
    acc = np.zeros_like(xyz)
    ref().ref_fmm_update(len(m), _p(xyz), _p(m), _p(acc))
    return acc


def ref_naive(xyz, m):
    xyz, m = _f64(xyz), _f64(m)
    acc = np.zeros_like(xyz)
    ref().ref_naive_update(len(m), _p(xyz), _p(m), _p(acc))
    return acc


def ref_integrate(xyz, vel, m, kind, updater, dt, nsteps, preupdate=True):
    """Reference integrator steps; returns (xyz, vel, carried acceleration)."""
    x, v, m = _f64(xyz).copy(), _f64(vel).copy(), _f64(m)
    a = np.zeros_like(x)
    ref().ref_integrate(len(m), _p(x), _p(v), _p(a), _p(m), kind, updater, dt, nsteps, int(preupdate))
    return x, v, a


def ref_tree(xyz, m):
    """The reference Tree: depth, per-level sorted keys, leaf particle lists and
    per-cell list keys (in list order)."""
    xyz, m = _f64(xyz), _f64(m)
    L = ref()
    D = L.ref_tree_build(len(m), _p(xyz), _p(m))
    try:
        levels = []
        for d in range(D + 1):
            k = np.zeros(L.ref_tree_level_count(d), np.uint64)
            L.ref_tree_level_keys(d, _p(k))
            levels.append(k)
        leaves = {}
        for k in levels[D]:
            cnt = L.ref_tree_leaf_particles(int(k), None)
            idx = np.zeros(cnt, np.int64)
            L.ref_tree_leaf_particles(int(k), _p(idx))
            leaves[int(k)] = idx
        lists = {}
        for d in range(D + 1):
            for k in levels[d]:
                nn, ni = ctypes.c_int64(), ctypes.c_int64()
                L.ref_tree_lists(d, int(k), ctypes.byref(nn), ctypes.byref(ni), None, None)
                nb, il = np.zeros(nn.value, np.uint64), np.zeros(ni.value, np.uint64)
                L.ref_tree_lists(d, int(k), ctypes.byref(nn), ctypes.byref(ni), _p(nb), _p(il))
                lists[(d, int(k))] = (nb, il)
        return {"depth": D, "levels": levels, "leaves": leaves, "lists": lists}
    finally:
        L.ref_tree_free()
