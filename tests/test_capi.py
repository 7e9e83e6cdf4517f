"""CPU: libnbfmm.so loads, exports every function include/nbfmm.h declares,
and rejects bad parameters before it touches the GPU."""
import ctypes
import os
import re

import nbfmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(nbfmm_\w+)\s*\(", txt, re.M))
    return names


def test_exports_every_declared_symbol():
    lib = nbfmm.lib()
    declared = _declared()
    assert len(declared) >= 17
    for name in declared:
        assert hasattr(lib, name), name
    assert declared >= set(nbfmm.EXPORTS)


def test_default_params_are_the_reference_settings():
    p = nbfmm.default_params()
    # tree.cpp:94 (L = 2), tree.cpp:32 (threshold 20), system_update_method.hpp:17,
    # naive_update.cpp:27 (1e-4)
    assert (p.order, p.leaf_capacity, p.G, p.eps2, p.precision, p.p2p_cut) == (2, 20, 6.6743e-11, 1e-4, 0, 0)


def test_bad_parameters_rejected_without_gpu():
    lib = nbfmm.lib()
    h = ctypes.c_void_p()
    for kw in (dict(order=0), dict(order=15), dict(leaf_capacity=0), dict(p2p_cut=-1), dict(eps2=-1.0),
               dict(precision=7)):
        p = nbfmm.default_params(**kw)
        assert lib.nbfmm_create(ctypes.byref(p), ctypes.byref(h)) == 1
        assert lib.nbfmm_last_error()
    assert lib.nbfmm_create(None, ctypes.byref(h)) == 1
