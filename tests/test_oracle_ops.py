"""CPU: the oracle's operators, FMM drivers, direct sums and integrators,
pinned against the reference's golden vectors (tests/golden, made by the
reference's own build), the reference build itself where present, and the
mathematics (closed forms, the addition theorem, exact identities,
convergence to the direct sum)."""
import math
import os

import numpy as np
import pytest
from scipy.special import lpmv

import orc
import workloads

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = lambda a: a.ctypes.data_as(__import__("ctypes").c_void_p)  # noqa: E731


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---- reference-compatible operators: golden vectors, bit-exact ------------------------
@pytest.mark.parametrize("L", [2, 5, 8])
def test_refcompat_ops_golden(L):
    g = np.load(os.path.join(GOLDEN, "ref_ops.npz"))
    k = lambda s: g[f"L{L}_{s}"]  # noqa: E731
    O = orc.orc()
    ps = np.zeros((L + 1) ** 2 + 4)
    O.orc_ref_powerset(L, P(k("r").copy()), P(ps))
    assert np.array_equal(ps[:(L + 1) ** 2], k("powerset")[:(L + 1) ** 2])
    O.orc_ref_irregular(L, P(k("r").copy()), P(ps))
    assert np.array_equal(ps[:(L + 1) ** 2], k("irregular")[:(L + 1) ** 2])
    M = np.zeros(2 * L * L + 4 * L + 1)
    for r, q in zip(k("rel"), k("q")):
        O.orc_ref_p2m_add(L, P(r.copy()), float(q), P(M))
    assert np.array_equal(M, k("p2m"))
    out = np.zeros_like(M)
    O.orc_ref_m2m(L, P(M), P(k("d").copy()), P(out))
    assert np.array_equal(out, k("m2m"))
    loc = np.zeros(2 * (L + 1) ** 2)
    O.orc_ref_m2l(L, P(M), P(k("D").copy()), P(loc))
    assert np.array_equal(loc, k("m2l"))
    l2 = np.zeros_like(loc)
    O.orc_ref_l2l(L, P(loc), P(k("d").copy()), P(l2))
    assert np.array_equal(l2, k("l2l"))
    gr = np.zeros(3)
    O.orc_ref_l2p(L, P(loc), P(k("d").copy()), P(gr))
    assert np.array_equal(gr, k("l2p"))


@pytest.mark.parametrize("L", [1, 3, 6, 10, 12])
def test_refcompat_ops_vs_reference_build(ref_lib, L):
    rng = np.random.default_rng(L)
    O, R = orc.orc(), ref_lib
    r = rng.normal(size=3)
    a, b = np.zeros((L + 1) ** 2 + 4), np.zeros((L + 1) ** 2 + 4)
    O.orc_ref_powerset(L, P(r), P(a))
    R.ref_powerset(L, P(r), P(b))
    assert np.array_equal(a[:(L + 1) ** 2], b[:(L + 1) ** 2])
    rel = rng.normal(size=(6, 3)) * 0.3
    q = rng.uniform(0.5, 1.5, 6)
    Ma, Mb = np.zeros(2 * L * L + 4 * L + 1), np.zeros(2 * L * L + 4 * L + 1)
    for i in range(6):
        O.orc_ref_p2m_add(L, P(rel[i].copy()), float(q[i]), P(Ma))
    R.ref_p2m(L, 6, P(rel), P(q), P(Mb))
    assert np.array_equal(Ma, Mb)
    D = np.array([2.5, -1.0, 1.5])
    la, lb = np.zeros(2 * (L + 1) ** 2), np.zeros(2 * (L + 1) ** 2)
    O.orc_ref_m2l(L, P(Ma), P(D), P(la))
    R.ref_m2l(L, P(Ma), P(D), P(lb))
    assert np.array_equal(la, lb)


# ---- reference FMM::update and NaiveUpdate: golden vectors, bit-exact ------------------
@pytest.mark.parametrize("case", ["t100", "u500", "u2000"])
def test_refcompat_fmm_and_naive_golden(case):
    g = np.load(os.path.join(GOLDEN, "ref_fmm.npz"))
    xyz, m, G = g[f"{case}_xyz"], g[f"{case}_m"], float(g["G"])
    assert np.array_equal(orc.fmm_refcompat(xyz, m, 2, 20, G=G), g[f"{case}_fmm"])
    assert np.array_equal(orc.direct_naive(xyz, m, G=G, eps2=1e-4), g[f"{case}_naive"])


def test_refcompat_fmm_vs_reference_build(ref_lib):
    xyz, m = workloads.plummer(6000, seed=6)
    G = ref_lib.ref_grav_const()
    assert np.array_equal(orc.fmm_refcompat(xyz, m, 2, 20, G=G), orc.ref_fmm(xyz, m))


def test_reference_far_field_does_not_converge():
    # The reference's conventions (multipole_set3D.cpp:45-118, local_set3D.cpp:51-210)
    # do not reproduce the direct sum: its error stays at several percent.  The
    # exact operators on the same tree converge.
    xyz, m = workloads.uniform_cube(3000, seed=1)
    d = orc.direct_targets(xyz, m, np.arange(3000), G=1.0)
    e_ref = _rel(orc.fmm_refcompat(xyz, m, 2, 20, G=1.0), d)
    e_ex = _rel(orc.fmm_exact(xyz, m, 2, 20, G=1.0, center_mode=1), d)
    assert e_ref > 0.02 and e_ex < 0.02


# ---- exact harmonics: closed forms and the addition theorem --------------------------
def _tri(n, m):
    return n * (n + 1) // 2 + m


def test_exact_harmonics_closed_form():
    p = 10
    rng = np.random.default_rng(3)
    O = orc.orc()
    for _ in range(5):
        r = rng.normal(size=3)
        R = np.zeros(2 * (p + 1) * (p + 2) // 2)
        S = np.zeros_like(R)
        O.orc_exact_regular(p, P(r), P(R))
        O.orc_exact_singular(p, P(r), P(S))
        rr = np.linalg.norm(r)
        ct, ph = r[2] / rr, math.atan2(r[1], r[0])
        for n in range(p + 1):
            for m in range(n + 1):
                Pnm = lpmv(m, n, ct)  # includes the Condon-Shortley phase
                Rw = rr ** n / math.factorial(n + m) * Pnm * complex(math.cos(m * ph), math.sin(m * ph))
                Sw = math.factorial(n - m) / rr ** (n + 1) * Pnm * complex(math.cos(m * ph), math.sin(m * ph))
                i = _tri(n, m)
                assert abs(complex(R[2 * i], R[2 * i + 1]) - Rw) <= 1e-12 * max(1.0, abs(Rw))
                assert abs(complex(S[2 * i], S[2 * i + 1]) - Sw) <= 1e-12 * max(1.0, abs(Sw))


def test_exact_addition_theorem():
    p = 24
    O = orc.orc()
    x = np.array([1.3, -0.7, 2.1])
    y = np.array([0.2, 0.1, -0.25])
    R = np.zeros(2 * (p + 1) * (p + 2) // 2)
    S = np.zeros_like(R)
    O.orc_exact_regular(p, P(y), P(R))
    O.orc_exact_singular(p, P(x), P(S))
    tot = 0.0
    for n in range(p + 1):
        for m in range(-n, n + 1):
            i = _tri(n, abs(m))
            r = complex(R[2 * i], R[2 * i + 1])
            s = complex(S[2 * i], S[2 * i + 1])
            if m < 0:
                r = (-1) ** m * r.conjugate()
                s = (-1) ** m * s.conjugate()
            tot += (r.conjugate() * s).real
    assert abs(tot - 1.0 / np.linalg.norm(x - y)) < 1e-14


def test_m2l_reflection_symmetry():
    """The identity the target-major M2L kernel (csrc/m2l.hip, k_m2l_tgt)
    rests on: for a sign change S of the axes, M2L(S d) = E_S M2L(d) E_S with
    E_S diagonal +-1 per coefficient -- (-1)^m times -1 for an imaginary part
    under x -> -x, -1 for an imaginary part under y -> -y, (-1)^(n+m) under
    z -> -z.  Checked on the oracle's exact M2L for every sign change of a
    generic offset; a wrong sign rule (e.g. (-1)^n for z) fails it."""
    p = 11
    O = orc.orc()
    T = (p + 1) * (p + 2) // 2
    rng = np.random.default_rng(5)

    def signs(sx, sy, sz, zrule=lambda n, m: (n + m) & 1):
        e = np.ones(2 * T)
        for n in range(p + 1):
            for m in range(n + 1):
                for part in (0, 1):
                    bit = 0
                    if sx:
                        bit ^= (m + part) & 1
                    if sy:
                        bit ^= part
                    if sz:
                        bit ^= zrule(n, m)
                    e[2 * _tri(n, m) + part] = -1.0 if bit else 1.0
        return e

    M = rng.normal(size=2 * T)
    for n in range(p + 1):
        M[2 * _tri(n, 0) + 1] = 0.0
    D = np.array([2.0, -1.0, 3.0])
    L0 = np.zeros(2 * T)
    O.orc_exact_m2l(p, P(M), P(D), P(L0))
    for s in range(1, 8):
        sx, sy, sz = s & 1, (s >> 1) & 1, (s >> 2) & 1
        e = signs(sx, sy, sz)
        DS = np.ascontiguousarray(D * np.array([-1.0 if sx else 1.0, -1.0 if sy else 1.0, -1.0 if sz else 1.0]))
        Ms = np.ascontiguousarray(e * M)
        L1 = np.zeros(2 * T)
        O.orc_exact_m2l(p, P(Ms), P(DS), P(L1))
        assert np.max(np.abs(e * L1 - L0)) <= 1e-13 * np.max(np.abs(L0)), s
        if sz:
            bad = signs(sx, sy, sz, zrule=lambda n, m: n & 1)
            L2 = np.zeros(2 * T)
            O.orc_exact_m2l(p, P(np.ascontiguousarray(bad * M)), P(DS), P(L2))
            assert np.max(np.abs(bad * L2 - L0)) > 1e-3 * np.max(np.abs(L0)), s


def test_exact_m2m_identity_and_convergence():
    O = orc.orc()
    rng = np.random.default_rng(4)
    src = rng.uniform(-0.5, 0.5, (15, 3))
    q = rng.uniform(0.5, 1.5, 15)
    x = np.array([3.1, -2.2, 1.7])
    d = x - src
    gd = -(q[:, None] * d / np.linalg.norm(d, axis=1)[:, None] ** 3).sum(0)
    prev = None
    for p in (4, 8, 12):
        ts = (p + 1) * (p + 2) // 2
        c1, c2 = np.zeros(3), np.array([0.1, -0.2, 0.15])
        M1, M2 = np.zeros(2 * ts), np.zeros(2 * ts)
        for s, w in zip(src, q):
            O.orc_exact_p2m_add(p, P((s - c1).copy()), float(w), P(M1))
            O.orc_exact_p2m_add(p, P((s - c2).copy()), float(w), P(M2))
        Mt = np.zeros(2 * ts)
        O.orc_exact_m2m(p, P(M1), P((c1 - c2).copy()), P(Mt))
        assert np.abs(Mt - M2).max() <= 1e-13 * np.abs(M2).max()  # exact identity
        assert abs(M1[0] - q.sum()) < 1e-13  # monopole = total mass
        ct = np.array([3.0, -2.0, 1.5])
        Lc, g = np.zeros(2 * ts), np.zeros(3)
        O.orc_exact_m2l(p, P(M1), P((ct - c1).copy()), P(Lc))
        O.orc_exact_l2p(p, P(Lc), P((x - ct).copy()), None, P(g))
        err = np.linalg.norm(g - gd) / np.linalg.norm(gd)
        gm = np.zeros(3)
        O.orc_exact_m2p(p, P(M1), P((x - c1).copy()), None, P(gm))
        assert np.linalg.norm(gm - gd) / np.linalg.norm(gd) < 1e-3
        if prev is not None:
            assert err < prev / 5
        prev = err
    assert prev < 1e-7


def test_exact_l2l_composition_and_finite_difference():
    O = orc.orc()
    p = 8
    ts = (p + 1) * (p + 2) // 2
    rng = np.random.default_rng(5)
    M = np.zeros(2 * ts)
    for s in rng.uniform(-0.3, 0.3, (6, 3)):
        O.orc_exact_p2m_add(p, P(s.copy()), 1.0, P(M))
    Lc = np.zeros(2 * ts)
    O.orc_exact_m2l(p, P(M), P(np.array([4.0, 1.0, -2.0])), P(Lc))
    d1, d2 = np.array([0.1, -0.05, 0.2]), np.array([-0.07, 0.12, 0.03])
    a, b, c = np.zeros(2 * ts), np.zeros(2 * ts), np.zeros(2 * ts)
    O.orc_exact_l2l(p, P(Lc), P(d1), P(a))
    O.orc_exact_l2l(p, P(a), P(d2), P(b))
    O.orc_exact_l2l(p, P(Lc), P((d1 + d2).copy()), P(c))
    assert np.abs(b - c).max() <= 1e-13 * np.abs(c).max()
    t = np.array([0.05, 0.1, -0.08])
    g = np.zeros(3)
    O.orc_exact_l2p(p, P(Lc), P(t), None, P(g))
    h = 1e-5
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fp, fm = __import__("ctypes").c_double(), __import__("ctypes").c_double()
        O.orc_exact_l2p(p, P(Lc), P(t + e), __import__("ctypes").byref(fp), P(np.zeros(3)))
        O.orc_exact_l2p(p, P(Lc), P(t - e), __import__("ctypes").byref(fm), P(np.zeros(3)))
        assert abs((fp.value - fm.value) / (2 * h) - g[k]) < 1e-8 * np.abs(g).max()


# ---- exact FMM: convergence and the all-direct limit -------------------------------------
def test_exact_fmm_converges_to_direct():
    xyz, m = workloads.uniform_cube(3000, seed=1)
    d = orc.direct_targets(xyz, m, np.arange(3000), G=1.0)
    errs = [_rel(orc.fmm_exact(xyz, m, p, 16, G=1.0), d) for p in (2, 6, 10)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5
    # every interaction-list pair evaluated directly: exactly the direct sum
    full = orc.fmm_exact(xyz, m, 4, 16, G=1.0, p2p_cut=1 << 50)
    assert _rel(full, d) < 1e-13
    # substitution removes truncation error
    assert _rel(orc.fmm_exact(xyz, m, 6, 16, G=1.0, p2p_cut=256), d) < errs[1]


def test_exact_fmm_converges_on_elongated_box():
    # a 1.6 : 1 : 1 box: the nearest classes along the short axes converge at
    # a rate above 1 and are evaluated directly; the error keeps falling with p
    rng = np.random.default_rng(18)
    xyz = rng.random((2000, 3)) * np.array([1.6, 1.0, 1.0])
    m = rng.uniform(0.5, 1.5, 2000)
    d = orc.direct_targets(xyz, m, np.arange(2000), G=1.0)
    errs = [_rel(orc.fmm_exact(xyz, m, p, 16, G=1.0), d) for p in (4, 8, 12)]
    assert errs[2] < 0.1 * errs[1] < 0.01 * errs[0]


def test_cubic_tree_equals_reference_tree_on_a_cube():
    # when the bounding box already is a cube the cubic option changes nothing
    rng = np.random.default_rng(26)
    xyz = rng.random((3000, 3)) * 2.0 - 1.0
    xyz[:8] = np.array([[x, y, z] for x in (-1.0, 1.0) for y in (-1.0, 1.0) for z in (-1.0, 1.0)])
    a, b = orc.tree(xyz, 16), orc.tree(xyz, 16, cubic=True)
    for k in ("level_start", "key", "pbeg", "pend", "parent", "perm", "nbr_off", "nbr", "il_off", "il", "lo", "hi"):
        assert np.array_equal(a[k], b[k]), k


def test_cubic_tree_geometry():
    # the box is the bounding cube centred on the bounding box and contains it
    rng = np.random.default_rng(27)
    xyz = rng.random((4000, 3)) * np.array([5.0, 1.0, 0.25]) + np.array([1.0, -2.0, 3.0])
    t = orc.tree(xyz, 16, cubic=True)
    lo, hi = xyz.min(0), xyz.max(0)
    ext = (hi - lo).max()
    assert np.allclose(t["hi"] - t["lo"], ext, rtol=1e-15, atol=0)
    assert np.allclose(0.5 * (t["lo"] + t["hi"]), 0.5 * (lo + hi), rtol=0, atol=1e-15)
    assert np.all(t["lo"] <= lo) and np.all(t["hi"] >= hi)
    # every particle lies in the leaf its key names
    leaf = np.array(t["leaf_size"])
    assert np.allclose(leaf, ext / (1 << t["depth"]), rtol=1e-15, atol=0)


def test_exact_fmm_converges_on_long_box_cubic():
    # a 4 : 1 : 1 box: on its bounding cube every class converges like a cube's
    rng = np.random.default_rng(28)
    xyz = rng.random((2000, 3)) * np.array([4.0, 1.0, 1.0])
    m = rng.uniform(0.5, 1.5, 2000)
    d = orc.direct_targets(xyz, m, np.arange(2000), G=1.0)
    errs = [_rel(orc.fmm_exact(xyz, m, p, 16, G=1.0, cubic=True), d) for p in (4, 8, 12)]
    assert errs[2] < 3e-6 and errs[2] < 0.05 * errs[1] < 0.0025 * errs[0]


def _class_of(dx, dy, dz):
    return (dz + 3) * 49 + (dy + 3) * 7 + (dx + 3)


def test_class_orders_table_shape():
    # cubes: the face classes keep p, orders stay in [4, p], fall with |d|
    # (never rise), classes inside the near field keep p; tol 0 keeps p
    for p in (6, 11, 14):
        assert np.all(orc.class_orders([1, 1, 1], p, 0.0) == p)
        for tol in (0.01, 0.1, 1.0):
            q = orc.class_orders([1, 1, 1], p, tol)
            assert q[_class_of(2, 0, 0)] == p and q[_class_of(0, 0, -2)] == p
            assert q[_class_of(1, 1, 1)] == p and q[_class_of(0, 0, 0)] == p
            assert q.min() >= min(4, p) and q.max() <= p
            far = [(dx, dy, dz) for dx in range(-3, 4) for dy in range(-3, 4) for dz in range(-3, 4)
                   if max(abs(dx), abs(dy), abs(dz)) >= 2]
            by_d = sorted(far, key=lambda o: o[0] ** 2 + o[1] ** 2 + o[2] ** 2)
            qs = [q[_class_of(*o)] for o in by_d]
            assert all(a >= b for a, b in zip(qs, qs[1:]))
            # symmetric under sign changes and axis permutations
            assert q[_class_of(3, 1, 2)] == q[_class_of(-2, 3, -1)]
        # a larger tolerance never raises an order
        assert np.all(orc.class_orders([1, 1, 1], p, 1.0) <= orc.class_orders([1, 1, 1], p, 0.1))
        assert orc.class_orders([1, 1, 1], p, 1.0)[_class_of(3, 3, 3)] < p


def test_class_orders_truncation_matches_model():
    # per-class M2L error of uniform cells at order q against the direct field
    # falls per degree by at most 0.55 * rate (the rule's model), rate = sqrt(3) / |d|
    rng = np.random.default_rng(40)
    P = 12

    def tsz(p):
        return (p + 1) * (p + 2) // 2
    L = orc.orc()
    for cl in ((2, 2, 2), (3, 0, 0), (3, 3, 3)):
        e = np.zeros(P + 1)
        for _ in range(6):
            src = rng.random((16, 3)) - 0.5 + np.array(cl, float)
            tgt = rng.random((16, 3)) - 0.5
            d = tgt[:, None, :] - src[None, :, :]
            f = -(d / ((d ** 2).sum(-1) ** 1.5)[..., None]).sum(1)
            M = np.zeros(2 * tsz(P))
            for j in range(16):
                r = np.ascontiguousarray(src[j] - np.array(cl, float))
                L.orc_exact_p2m_add(P, orc._p(r), 1.0, orc._p(M))
            D = np.ascontiguousarray(-np.array(cl, float))
            for q in (6, 9):
                Lc = np.zeros(2 * tsz(P))
                L.orc_exact_m2l(q, orc._p(M), orc._p(D), orc._p(Lc))
                for i in range(16):
                    g = np.zeros(3)
                    L.orc_exact_l2p(q, orc._p(Lc), orc._p(np.ascontiguousarray(tgt[i])), None, orc._p(g))
                    e[q] += ((g - f[i]) ** 2).sum()
        rate = np.sqrt(3.0) / np.linalg.norm(cl)
        decay = (np.sqrt(e[9]) / np.sqrt(e[6])) ** (1.0 / 3.0)
        assert decay < 0.55 * rate * 1.15, (cl, decay, rate)


def test_exact_fmm_class_orders():
    # tol 0 is the fixed-order FMM bit for bit; small tol stays within the
    # model's quadrature sum of the ~150 classes each within tol of the face
    # class's error; lower tol never gives a larger error here
    xyz, m = workloads.uniform_cube(3000, seed=41)
    d = orc.direct_targets(xyz, m, np.arange(3000), G=1.0)
    a0 = orc.fmm_exact(xyz, m, 8, 16, G=1.0)
    assert np.array_equal(orc.fmm_exact(xyz, m, 8, 16, G=1.0, class_order_tol=0.0), a0)
    e0 = _rel(a0, d)
    errs = [_rel(orc.fmm_exact(xyz, m, 8, 16, G=1.0, class_order_tol=t), d) for t in (0.02, 0.1, 0.5)]
    assert errs[0] < 1.3 * e0
    assert errs[0] <= errs[1] <= errs[2]
    assert errs[1] < np.sqrt(1 + 150 * 0.1 ** 2) * 1.3 * e0


def test_direct_naive_two_body():
    xyz = np.array([[-10.0, 0, 0], [10.0, 0, 0]])
    m = np.array([1e9, 1e9])
    G = 6.6743e-11
    a = orc.direct_naive(xyz, m, G=G, eps2=1e-4)
    want = G * 1e9 / ((400.0 + 1e-4) * 20.0) * 20.0
    assert abs(a[0, 0] - want) < 1e-15 * want and abs(a[1, 0] + want) < 1e-15 * want
    assert np.all(a[:, 1:] == 0)


# ---- integrators (symplectic_euler.cpp:8-15, runge_kutta_4.cpp:10-93) ----------------------
def test_se_golden():
    g = np.load(os.path.join(GOLDEN, "ref_se.npz"))
    G = 6.6743e-11
    for upd, name in ((orc.UPD_FMM_REF, "fmm"), (orc.UPD_NAIVE, "naive")):
        cfg = orc.UpdCfg(upd, 2, 20, G, 1e-4, 0, 0)
        x, v = g["x0"], g["v0"]
        a = orc.direct_naive(x, g["m"], G=G, eps2=1e-4) if upd == orc.UPD_NAIVE else orc.fmm_refcompat(x, g["m"], 2, 20, G)
        for _ in range(10):
            x, v, a = orc.step("se", x, v, a, g["m"], cfg, float(g["dt"]))
        assert np.array_equal(x, g[f"{name}_x"]) and np.array_equal(v, g[f"{name}_v"])
        assert np.array_equal(a, g[f"{name}_a"])


def test_rk4_reference_semantics_golden():
    g = np.load(os.path.join(GOLDEN, "ref_rk4.npz"))
    G = 6.6743e-11
    cfg = orc.UpdCfg(orc.UPD_NAIVE, 2, 20, G, 1e-4, 0, 0)
    x, v = g["x0"], g["v0"]
    a = orc.direct_naive(x, g["m"], G=G, eps2=1e-4)  # the update() before integrating
    a0 = a.copy()
    for _ in range(2):
        x, v, a = orc.step("rk4", x, v, a, g["m"], cfg, float(g["dt"]), rk4_mode=0)
    assert np.array_equal(x, g["naive_x"]) and np.array_equal(v, g["naive_v"])
    assert np.array_equal(a, g["naive_a"]) and np.array_equal(a, a0)  # carried acceleration unchanged


def test_se_closed_form_one_step():
    xyz = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    v = np.array([[0.0, 0.1, 0], [0.0, -0.1, 0]])
    m = np.array([1.0, 2.0])
    cfg = orc.UpdCfg(orc.UPD_NAIVE, 2, 20, 1.0, 0.0, 0, 0)
    x1, v1, a1 = orc.step("se", xyz, v, np.zeros((2, 3)), m, cfg, 0.01)
    assert np.allclose(a1[0], [2.0, 0, 0]) and np.allclose(a1[1], [-1.0, 0, 0])
    assert np.allclose(v1, v + 0.01 * a1) and np.allclose(x1, xyz + 0.01 * v1)


def test_rk4_classic_order():
    # classic RK4 on a circular two-body orbit: error shrinks ~16x when dt halves
    m = np.array([1.0, 0.0])
    xyz = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    v = np.array([[0.0, 0, 0], [0.0, 1.0, 0]])
    cfg = orc.UpdCfg(orc.UPD_NAIVE, 2, 20, 1.0, 0.0, 0, 0)
    T = 0.5
    errs = []
    for nsteps in (20, 40):
        x, vv, a = xyz, v, np.zeros((2, 3))
        for _ in range(nsteps):
            x, vv, a = orc.step("rk4", x, vv, a, m, cfg, T / nsteps, rk4_mode=1)
        exact = np.array([math.cos(T), math.sin(T), 0.0])
        errs.append(np.linalg.norm(x[1] - x[0] - exact))
    assert errs[0] / errs[1] > 12
