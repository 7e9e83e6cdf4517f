"""ctypes binding of libnbfmm.so (the C ABI in include/nbfmm.h).

Used by the tests and bench.py.  Host arrays are numpy float64; device
buffers are raw HIP allocations made through libamdhip64 (no torch import,
which keeps GPU runs short).  Importing this module does not touch the GPU.
The library must have been built (nbody-fmm_amd/build.py); a missing
library raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbfmm.so")

NBFMM_PLAIN_FP64 = 0
NBFMM_MIXED = 1
NBFMM_ORDER_FIXED = 0
NBFMM_ORDER_LEARNED = 1
MLP_FEATURES = 16
MLP_ORDERS = 11
MLP_ORDER_LO = 4
MLP_PARAMS = 5963
MLP_WEIGHTS = 5995


class Params(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_uint32),
        ("leaf_capacity", ctypes.c_uint32),
        ("G", ctypes.c_double),
        ("eps2", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("m2l_pass_pairs", ctypes.c_int32),
        ("p2p_cut", ctypes.c_int64),
        ("order_select", ctypes.c_int32),
        ("order_min", ctypes.c_int32),
        ("order_conf", ctypes.c_double),
        ("cubic_box", ctypes.c_int32),
        ("class_order_tol", ctypes.c_double),
        ("mixed_fp64_levels", ctypes.c_int32),
    ]


class Tiles(ctypes.Structure):
    _fields_ = [
        ("p2p_group_f32", ctypes.c_int32),
        ("p2p_group_f64", ctypes.c_int32),
        ("p2p_narrow_max_f32", ctypes.c_int32),
        ("p2p_narrow_max_f64", ctypes.c_int32),
        ("p2p_early_share_f32", ctypes.c_double),
        ("p2p_early_share_f64", ctypes.c_double),
        ("p2p_early_launches_f32", ctypes.c_int32),
        ("p2p_early_launches_f64", ctypes.c_int32),
        ("m2l_slices", ctypes.c_int32),
        ("p2p_queue", ctypes.c_int32),
        ("p2p_queue_waves0", ctypes.c_int32),
        ("p2p_queue_chunk", ctypes.c_int32),
        ("m2l_item_pairs", ctypes.c_int32),
        ("p2p_rest_launches", ctypes.c_int32),
        ("m2l_mode", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [
        ("depth", ctypes.c_int32),
        ("depth_iterations", ctypes.c_int32),
        ("n_cells", ctypes.c_int64),
        ("n_leaves", ctypes.c_int64),
        ("n_il", ctypes.c_int64),
        ("n_m2l", ctypes.c_int64),
        ("n_m2l_targets", ctypes.c_int64),
        ("n_m2l_sources", ctypes.c_int64),
        ("n_p2p_pairs", ctypes.c_int64),
        ("n_p2m_terms", ctypes.c_int64),
        ("n_l2p_terms", ctypes.c_int64),
        ("ms_tree", ctypes.c_double),
        ("ms_lists", ctypes.c_double),
        ("ms_p2m", ctypes.c_double),
        ("ms_m2l", ctypes.c_double),
        ("ms_l2p", ctypes.c_double),
        ("ms_p2p", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_m2l_kernel", ctypes.c_double),
        ("ms_p2p_kernel", ctypes.c_double),
        ("flop_m2l", ctypes.c_double),
        ("flop_p2p", ctypes.c_double),
        ("mean_order", ctypes.c_double),
        ("fallback", ctypes.c_int32),
        ("isolated", ctypes.c_int32),
        ("n_p2m_cells", ctypes.c_int64),
        ("n_m2m_cells", ctypes.c_int64),
        ("n_l2l_cells", ctypes.c_int64),
        ("n_l2p_cells", ctypes.c_int64),
        ("ms_m2m", ctypes.c_double),
        ("ms_l2l", ctypes.c_double),
        ("depth_sorts", ctypes.c_int32),
        ("reserved1", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


EXPORTS = [
    "nbfmm_default_params", "nbfmm_create", "nbfmm_destroy", "nbfmm_set_params", "nbfmm_stream",
    "nbfmm_eval_accel", "nbfmm_eval_accel_dev", "nbfmm_direct_accel", "nbfmm_direct_accel_dev",
    "nbfmm_step_se_dev", "nbfmm_step_rk4_dev", "nbfmm_get_stats", "nbfmm_set_timing",
    "nbfmm_keep_lists", "nbfmm_tree_sizes", "nbfmm_tree_get", "nbfmm_last_error",
    "nbfmm_eval_accel_pot_dev", "nbfmm_diagnostics_dev", "nbfmm_run_dev",
    "nbfmm_order_model_init", "nbfmm_order_model_set", "nbfmm_order_model_get", "nbfmm_order_dataset_dev",
    "nbfmm_order_dataset_get", "nbfmm_order_dataset_put", "nbfmm_order_train", "nbfmm_order_grad",
    "nbfmm_order_predict", "nbfmm_eval_accel_checked_dev",
    "nbfmm_batch_eval_accel_dev", "nbfmm_batch_step_se_dev", "nbfmm_batch_step_rk4_dev",
    "nbfmm_direct_accel_noalloc_dev", "nbfmm_set_tiles", "nbfmm_get_tiles",
    "nbfmm_set_graph", "nbfmm_graph_stats",
]
NOALLOC_SCRATCH_BYTES = 48 << 20
NOALLOC_MAX_TARGETS = 1 << 20

DIAG_NAMES = ["kinetic", "potential", "total", "px", "py", "pz", "lx", "ly", "lz", "mass"]

_lib = None


def lib():
    """Loads libnbfmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run nbody-fmm_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.nbfmm_last_error.restype = ctypes.c_char_p
        L.nbfmm_stream.restype = vp
        L.nbfmm_stream.argtypes = [vp]
        for f in ("nbfmm_eval_accel", "nbfmm_eval_accel_dev", "nbfmm_direct_accel"):
            getattr(L, f).argtypes = [vp, i64, vp, vp, vp]
        L.nbfmm_direct_accel_dev.argtypes = [vp, i64, vp, vp, i64, vp, ctypes.c_int, vp]
        L.nbfmm_direct_accel_noalloc_dev.argtypes = [vp, i64, vp, vp, i64, vp, ctypes.c_int, vp]
        L.nbfmm_step_se_dev.argtypes = [vp, i64, vp, vp, vp, vp, dbl, ctypes.c_int]
        L.nbfmm_step_rk4_dev.argtypes = [vp, i64, vp, vp, vp, vp, dbl, ctypes.c_int, ctypes.c_int]
        L.nbfmm_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.nbfmm_destroy.argtypes = [vp]
        L.nbfmm_set_params.argtypes = [vp, ctypes.POINTER(Params)]
        L.nbfmm_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.nbfmm_set_tiles.argtypes = [vp, ctypes.POINTER(Tiles)]
        L.nbfmm_get_tiles.argtypes = [vp, ctypes.POINTER(Tiles)]
        L.nbfmm_set_timing.argtypes = [vp, ctypes.c_int]
        L.nbfmm_keep_lists.argtypes = [vp, ctypes.c_int]
        L.nbfmm_tree_sizes.argtypes = [vp, vp, vp, vp, vp]
        L.nbfmm_tree_get.argtypes = [vp] + [vp] * 10
        L.nbfmm_eval_accel_pot_dev.argtypes = [vp, i64, vp, vp, vp, vp]
        L.nbfmm_diagnostics_dev.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        L.nbfmm_run_dev.argtypes = [vp, i64, vp, vp, vp, vp, dbl, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        u64, i32 = ctypes.c_uint64, ctypes.c_int32
        L.nbfmm_order_model_init.argtypes = [vp, u64]
        L.nbfmm_order_model_set.argtypes = [vp, vp, i64]
        L.nbfmm_order_model_get.argtypes = [vp, vp, i64]
        L.nbfmm_order_dataset_dev.argtypes = [vp, i64, vp, vp, i64, u64, dbl, vp]
        L.nbfmm_order_dataset_get.argtypes = [vp, vp, vp, i64, vp]
        L.nbfmm_order_dataset_put.argtypes = [vp, i64, vp, vp]
        L.nbfmm_order_train.argtypes = [vp, i32, i32, dbl, u64, vp]
        L.nbfmm_order_grad.argtypes = [vp, i32, vp, vp, vp, vp]
        L.nbfmm_order_predict.argtypes = [vp, i64, vp, vp]
        L.nbfmm_eval_accel_checked_dev.argtypes = [vp, i64, vp, vp, vp, i64, u64, dbl, i32, vp, vp]
        L.nbfmm_batch_eval_accel_dev.argtypes = [vp, i32, vp, vp, vp, vp]
        L.nbfmm_batch_step_se_dev.argtypes = [vp, i32, vp, vp, vp, vp, vp, dbl, ctypes.c_int]
        L.nbfmm_batch_step_rk4_dev.argtypes = [vp, i32, vp, vp, vp, vp, vp, dbl, ctypes.c_int, ctypes.c_int]
        L.nbfmm_set_graph.argtypes = [vp, ctypes.c_int]
        L.nbfmm_graph_stats.argtypes = [vp, vp]
        _lib = L
    return _lib


class NbfmmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"nbfmm status {status}: {msg}")
        self.status = status


def _check(rc):
    if rc != 0:
        raise NbfmmError(rc, lib().nbfmm_last_error().decode())


def default_params(**kw):
    p = Params()
    lib().nbfmm_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One nbfmm context (device buffers + HIP stream) on the current device."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        h = ctypes.c_void_p()
        _check(lib().nbfmm_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nbfmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        """Updates the named parameters; on failure the old ones stay in force."""
        old = {k: getattr(self.params, k) for k in kw}
        for k, v in kw.items():
            setattr(self.params, k, v)
        try:
            _check(lib().nbfmm_set_params(self.h, ctypes.byref(self.params)))
        except NbfmmError:
            for k, v in old.items():
                setattr(self.params, k, v)
            raise

    @property
    def stream(self):
        return lib().nbfmm_stream(self.h)

    def eval_accel(self, xyz, m):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        acc = np.empty_like(xyz)
        _check(lib().nbfmm_eval_accel(self.h, len(m), _ptr(xyz), _ptr(m), _ptr(acc)))
        return acc

    def eval_accel_dev(self, n, d_xyz, d_m, d_acc):
        _check(lib().nbfmm_eval_accel_dev(self.h, n, d_xyz, d_m, d_acc))

    def direct_accel(self, xyz, m):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        acc = np.empty_like(xyz)
        _check(lib().nbfmm_direct_accel(self.h, len(m), _ptr(xyz), _ptr(m), _ptr(acc)))
        return acc

    def direct_accel_dev(self, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc):
        _check(lib().nbfmm_direct_accel_dev(self.h, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc))

    def direct_accel_noalloc_dev(self, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc):
        """As direct_accel_dev, allocating no device memory (fixed per-GPU scratch)."""
        _check(lib().nbfmm_direct_accel_noalloc_dev(self.h, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc))

    def step_se_dev(self, n, dx, dv, da, dm, dt, nsteps):
        _check(lib().nbfmm_step_se_dev(self.h, n, dx, dv, da, dm, dt, nsteps))

    def step_rk4_dev(self, n, dx, dv, da, dm, dt, nsteps, mode):
        _check(lib().nbfmm_step_rk4_dev(self.h, n, dx, dv, da, dm, dt, nsteps, mode))

    def eval_accel_pot_dev(self, n, d_xyz, d_m, d_acc, d_phi):
        _check(lib().nbfmm_eval_accel_pot_dev(self.h, n, d_xyz, d_m, d_acc, d_phi))

    def diagnostics_dev(self, n, d_xyz, d_vel, d_m, d_phi=None):
        """Dict of the 10 sums (DIAG_NAMES) of a device-resident state."""
        out = np.zeros(10)
        _check(lib().nbfmm_diagnostics_dev(self.h, n, d_xyz, d_vel, d_m, d_phi, _ptr(out)))
        return dict(zip(DIAG_NAMES, out.tolist()))

    def run_dev(self, n, dx, dv, da, dm, dt, nsteps, integrator=0, every=1, monitor=True):
        """Advances nsteps on the device; returns the diagnostics records
        (array of shape (1 + nsteps // every, 10), columns DIAG_NAMES) or None."""
        rec = np.zeros((1 + nsteps // every, 10)) if monitor else None
        _check(lib().nbfmm_run_dev(self.h, n, dx, dv, da, dm, dt, nsteps, integrator, every,
                                   _ptr(rec) if monitor else None))
        return rec

    # ---- learned order selection ----
    def order_model_init(self, seed=0):
        _check(lib().nbfmm_order_model_init(self.h, seed))

    def order_model_set(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().nbfmm_order_model_set(self.h, _ptr(w), len(w)))

    def order_model_get(self):
        w = np.zeros(MLP_WEIGHTS, np.float32)
        _check(lib().nbfmm_order_model_get(self.h, _ptr(w), len(w)))
        return w

    def order_dataset_dev(self, n, d_xyz, d_m, stride, seed=0, tol=1e-6):
        """Labels the evaluation cells of one particle set; returns the rows added."""
        added = ctypes.c_int64()
        _check(lib().nbfmm_order_dataset_dev(self.h, n, d_xyz, d_m, stride, seed, tol, ctypes.byref(added)))
        return added.value

    def order_dataset(self):
        """(features [n, 16] float32, labels [n] int8 as order - 4) of the training set."""
        n = ctypes.c_int64()
        _check(lib().nbfmm_order_dataset_get(self.h, None, None, 0, ctypes.byref(n)))
        f = np.zeros((n.value, MLP_FEATURES), np.float32)
        y = np.zeros(n.value, np.int8)
        if n.value:
            _check(lib().nbfmm_order_dataset_get(self.h, _ptr(f), _ptr(y), n.value, None))
        return f, y

    def order_dataset_put(self, feat=None, label=None):
        """Appends rows; with no arguments clears the training set."""
        if feat is None:
            _check(lib().nbfmm_order_dataset_put(self.h, 0, None, None))
            return
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        label = np.ascontiguousarray(label, dtype=np.int8)
        _check(lib().nbfmm_order_dataset_put(self.h, len(label), _ptr(feat), _ptr(label)))

    def order_train(self, epochs, batch=1024, lr=3e-3, seed=0):
        loss = np.zeros(max(epochs, 1), np.float32)
        _check(lib().nbfmm_order_train(self.h, epochs, batch, lr, seed, _ptr(loss)))
        return loss[:epochs]

    def order_grad(self, feat, label):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        label = np.ascontiguousarray(label, dtype=np.int8)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        _check(lib().nbfmm_order_grad(self.h, len(label), _ptr(feat), _ptr(label), _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def order_predict(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        prob = np.zeros((len(feat), MLP_ORDERS), np.float32)
        _check(lib().nbfmm_order_predict(self.h, len(feat), _ptr(feat), _ptr(prob)))
        return prob

    def eval_accel_checked_dev(self, n, d_xyz, d_m, d_acc, ntgt=4096, seed=0, gate=1e-6, fallback_order=14):
        """Evaluation plus sampled direct-sum check; returns (error, fell_back)."""
        err = ctypes.c_double()
        fb = ctypes.c_int32()
        _check(lib().nbfmm_eval_accel_checked_dev(self.h, n, d_xyz, d_m, d_acc, ntgt, seed, gate, fallback_order,
                                                  ctypes.byref(err), ctypes.byref(fb)))
        return err.value, bool(fb.value)

    # ---- batched systems ----
    def batch_eval_accel_dev(self, offsets, d_xyz, d_m, d_acc):
        """offsets: int64 array of nsys + 1 (system k = particles [offsets[k], offsets[k+1]))."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        _check(lib().nbfmm_batch_eval_accel_dev(self.h, len(off) - 1, _ptr(off), d_xyz, d_m, d_acc))

    def batch_step_se_dev(self, offsets, dx, dv, da, dm, dt, nsteps):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        _check(lib().nbfmm_batch_step_se_dev(self.h, len(off) - 1, _ptr(off), dx, dv, da, dm, dt, nsteps))

    def batch_step_rk4_dev(self, offsets, dx, dv, da, dm, dt, nsteps, mode):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        _check(lib().nbfmm_batch_step_rk4_dev(self.h, len(off) - 1, _ptr(off), dx, dv, da, dm, dt, nsteps, mode))

    def set_graph(self, on):
        """Integration steps captured once as a HIP graph and replayed (1), or run directly (0)."""
        _check(lib().nbfmm_set_graph(self.h, int(on)))

    def graph_stats(self):
        """{captures, launches, redone, direct}: graphs captured, steps replayed, replays whose sizes
        overflowed (run again directly), steps run directly."""
        out = np.zeros(4, dtype=np.int64)
        _check(lib().nbfmm_graph_stats(self.h, _ptr(out)))
        return dict(zip(("captures", "launches", "redone", "direct"), (int(v) for v in out)))

    def stats(self):
        s = Stats()
        _check(lib().nbfmm_get_stats(self.h, ctypes.byref(s)))
        return s

    def tiles(self):
        """The tile shapes in force (dict of nbfmm_tiles fields, defaults resolved)."""
        t = Tiles()
        _check(lib().nbfmm_get_tiles(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tiles(self, **kw):
        """Changes the named tile fields (the others keep their current values)."""
        t = Tiles()
        _check(lib().nbfmm_get_tiles(self.h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().nbfmm_set_tiles(self.h, ctypes.byref(t)))

    def set_timing(self, mode):
        """0 off, 1 on, 2 on with every phase (P2P included) run alone."""
        _check(lib().nbfmm_set_timing(self.h, int(mode)))

    def keep_lists(self, on):
        _check(lib().nbfmm_keep_lists(self.h, int(on)))

    def tree(self, with_il=False):
        """Tree of the last evaluation as numpy arrays (cells level 0 first)."""
        depth = ctypes.c_int32()
        nc, nn, ni = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().nbfmm_tree_sizes(self.h, ctypes.byref(depth), ctypes.byref(nc), ctypes.byref(nn),
                                      ctypes.byref(ni)))
        D, nc, nn, ni = depth.value, nc.value, nn.value, ni.value
        t = {
            "depth": D,
            "level_start": np.zeros(D + 2, np.int64),
            "key": np.zeros(nc, np.uint64),
            "pbeg": np.zeros(nc, np.int32),
            "pend": np.zeros(nc, np.int32),
            "parent": np.zeros(nc, np.int32),
            "nbr_off": np.zeros(nc + 1, np.int32),
            "nbr": np.zeros(max(nn, 1), np.int32),
        }
        n = int(self._last_n)
        t["perm"] = np.zeros(n, np.int32)
        il_off = il = None
        if with_il:
            il_off = np.zeros(nc + 1, np.int32)
            il = np.zeros(max(ni, 1), np.int32)
        _check(lib().nbfmm_tree_get(
            self.h, _ptr(t["level_start"]), _ptr(t["key"]), _ptr(t["pbeg"]), _ptr(t["pend"]), _ptr(t["parent"]),
            _ptr(t["perm"]), _ptr(t["nbr_off"]), _ptr(t["nbr"]),
            _ptr(il_off) if with_il else None, _ptr(il) if with_il else None))
        t["nbr"] = t["nbr"][:nn]
        if with_il:
            t["il_off"], t["il"] = il_off, il[:ni]
        return t

    def eval_tree(self, xyz, m, with_il=False):
        """Evaluates and returns (acc, tree)."""
        self.keep_lists(with_il)
        acc = self.eval_accel(xyz, m)
        self._last_n = len(m)
        return acc, self.tree(with_il)


# ---- minimal HIP runtime access (device memory for the *_dev entry points) ----
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return _hip


def device_count():
    try:
        n = ctypes.c_int(0)
        if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
            return 0
        return n.value
    except OSError:
        return 0


class DeviceArray:
    """A raw device allocation holding a copy of a numpy array."""

    def __init__(self, arr=None, nbytes=None, dtype=np.float64, shape=None):
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            nbytes, dtype, shape = arr.nbytes, arr.dtype, arr.shape
        self.nbytes, self.dtype, self.shape = int(nbytes), np.dtype(dtype), shape
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {rc}")
        self.ptr = p
        if arr is not None:
            self.upload(arr)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, _ptr(arr), self.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def download(self, stream=None):
        """Copies to host; `stream` (e.g. Context.stream) is synchronised first,
        since work ordered on a non-blocking stream is not seen by hipMemcpy."""
        if stream is not None:
            rc = hip().hipStreamSynchronize(stream)
            if rc != 0:
                raise RuntimeError(f"hipStreamSynchronize failed: {rc}")
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(_ptr(out), self.ptr, self.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """HIP events on a given stream (the stream the kernels run on)."""

    def __init__(self, stream):
        self.stream = stream
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        hip().hipEventCreate(ctypes.byref(self.a))
        hip().hipEventCreate(ctypes.byref(self.b))

    def start(self):
        hip().hipEventRecord(self.a, self.stream)

    def stop(self):
        hip().hipEventRecord(self.b, self.stream)
        hip().hipEventSynchronize(self.b)
        ms = ctypes.c_float()
        hip().hipEventElapsedTime(ctypes.byref(ms), self.a, self.b)
        return ms.value
