"""CPU: the oracle's tree rule against the reference's golden vectors, the
reference build itself, closed forms and brute force."""
import ctypes
import json
import os

import numpy as np
import pytest

import orc
import workloads

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- Morton keys (morton_code.cpp:9-36) ----------------------------------------
def _spread_ref(c):
    out = 0
    for k in range(21):
        out |= ((c >> k) & 1) << (3 * k)
    return out


def test_morton_closed_form_and_roundtrip():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 16, 3))
        k = orc.morton_encode(x, y, z)
        assert k == _spread_ref(x) | (_spread_ref(y) << 1) | (_spread_ref(z) << 2)
        assert orc.morton_decode(k) == (x, y, z)
    assert orc.morton_encode(1, 0, 0) == 1 and orc.morton_encode(0, 1, 0) == 2 and orc.morton_encode(0, 0, 1) == 4
    assert orc.morton_encode(2, 0, 0) == 8


# ---- tests/test_tree.cpp:42-99 golden (glibc srand(42)/rand()) --------------------
def _glibc_particles(n, bound, seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    RAND_MAX = 2147483647.0
    pts = []
    for _ in range(n):
        x = (libc.rand() * 2.0 * bound) / RAND_MAX - bound
        y = (libc.rand() * 2.0 * bound) / RAND_MAX - bound
        z = (libc.rand() * 2.0 * bound) / RAND_MAX - bound
        pts.append((x, y, z))
    return np.array(pts)


# leaf key -> particle indices in leaf order, as asserted by test_tree.cpp:64-98
TEST_TREE_EXPECTED = {0: [1, 4, 7, 14], 1: [11], 3: [2, 13], 4: [0, 3], 5: [5], 6: [10, 12], 7: [6, 8, 9]}


def _leaves(t):
    D = t["depth"]
    b, e = t["level_start"][D], t["level_start"][D + 1]
    return {int(t["key"][c]): [int(t["perm"][s]) for s in range(t["pbeg"][c], t["pend"][c])] for c in range(b, e)}


def test_tree_golden_fixture_matches_reference_test():
    with open(os.path.join(GOLDEN, "test_tree_15.json")) as f:
        g = json.load(f)
    xyz = np.array(g["positions"])
    t = orc.tree(xyz, 20)
    assert t["depth"] == 1
    leaves = _leaves(t)
    assert leaves == {int(k): v for k, v in g["leaves"].items()}
    assert leaves == TEST_TREE_EXPECTED
    assert 2 not in leaves


def test_tree_golden_glibc_rand():
    xyz = _glibc_particles(15, 100.0, 42)
    with open(os.path.join(GOLDEN, "test_tree_15.json")) as f:
        g = json.load(f)
    assert np.array_equal(xyz, np.array(g["positions"]))
    assert _leaves(orc.tree(xyz, 20)) == TEST_TREE_EXPECTED


# ---- depth loop closed form (tree.cpp:32-69) ---------------------------------------
def test_depth_loop_rule():
    xyz, _ = workloads.plummer(30000, seed=3)
    for thr in (16, 20, 64):
        t = orc.tree(xyz, thr)
        maxcnt, depth = len(xyz), 0
        for d, mc in t["trace"]:
            nl = (maxcnt - 1) // thr + 1
            depth = min(16, depth + nl.bit_length() // 3 + 1)
            assert d == depth
            # recompute the max occupancy independently
            ls = (t["hi"] - t["lo"]) / float(1 << d)
            q = np.minimum(((xyz - t["lo"]) / ls).astype(np.int64), (1 << d) - 1)
            _, counts = np.unique(q, axis=0, return_counts=True)
            assert counts.max() == mc
            maxcnt = mc
        assert t["depth"] == depth
        assert depth == 16 or maxcnt <= thr


# ---- lists (node_i.cpp:23-99) against brute force ------------------------------------
def test_lists_brute_force():
    xyz, _ = workloads.plummer(4000, seed=12)
    t = orc.tree(xyz, 8)
    for lev in range(t["depth"] + 1):
        b, e = t["level_start"][lev], t["level_start"][lev + 1]
        coords = {c: np.array(orc.morton_decode(int(t["key"][c]))) for c in range(b, e)}
        for c in range(b, min(e, b + 300)):
            cc = coords[c]
            nb, il = [], []
            for s in range(b, e):
                d = coords[s] - cc
                par = (coords[s] >> 1) - (cc >> 1)
                if np.abs(par).max() > 1:
                    continue  # outside the parent's neighbourhood
                if np.abs(d).max() <= 1:
                    if s != c:
                        nb.append(s)
                else:
                    il.append(s)
            got_nb = set(t["nbr"][t["nbr_off"][c]:t["nbr_off"][c + 1]].tolist())
            got_il = set(t["il"][t["il_off"][c]:t["il_off"][c + 1]].tolist())
            assert got_nb == set(nb) and got_il == set(il)


# ---- against the reference build ---------------------------------------------------------
@pytest.mark.parametrize("case", ["uniform1000", "uniform6000", "plummer3000", "clustered", "test_tree"])
def test_tree_vs_reference_build(ref_lib, case):
    if case == "uniform1000":
        xyz, m = workloads.uniform_cube(1000, seed=1)
    elif case == "uniform6000":
        xyz, m = workloads.uniform_cube(6000, seed=2)
    elif case == "plummer3000":
        xyz, m = workloads.plummer(3000, seed=3)
    elif case == "clustered":
        xyz, m = workloads.plummer_mix(3000, n_spheres=8, seed=4, box=10.0)
    else:
        xyz = _glibc_particles(15, 100.0, 42)
        m = np.full(15, 10.0)
    r = orc.ref_tree(xyz, m)
    t = orc.tree(xyz, 20)
    D = t["depth"]
    assert r["depth"] == D
    assert D <= 9, "reference keys are 32-bit; parity holds for depth <= 9"
    for lev in range(D + 1):
        b, e = t["level_start"][lev], t["level_start"][lev + 1]
        assert np.array_equal(r["levels"][lev], t["key"][b:e])
    assert {k: v.tolist() for k, v in r["leaves"].items()} == _leaves(t)
    for lev in range(D + 1):
        b, e = t["level_start"][lev], t["level_start"][lev + 1]
        for c in range(b, e):
            rn, ri = r["lists"][(lev, int(t["key"][c]))]
            assert np.array_equal(rn, t["key"][t["nbr"][t["nbr_off"][c]:t["nbr_off"][c + 1]]])
            assert np.array_equal(ri, t["key"][t["il"][t["il_off"][c]:t["il_off"][c + 1]]])


def test_degenerate_inputs():
    # one particle, coincident particles, zero extent in one axis
    t = orc.tree(np.zeros((1, 3)), 20)
    assert t["depth"] == 1 and len(t["key"]) == 2
    t = orc.tree(np.zeros((40, 3)), 20)
    assert t["depth"] == 16
    xyz, _ = workloads.uniform_cube(500, seed=5)
    xyz[:, 2] = 0.25
    t = orc.tree(xyz, 20)
    assert all(orc.morton_decode(int(k))[2] == 0 for k in t["key"])
    bad = xyz.copy()
    bad[3, 0] = np.inf
    with pytest.raises(ValueError):
        orc.tree(bad, 20)
