"""GPU tests of the potential, the diagnostics and the multi-step driver.

The potential phi_i = -G sum_{j != i} m_j / r_ij is checked against a
brute-force numpy sum (exact in the all-direct limit, converging with p
otherwise); the diagnostics against numpy sums of the same state; the driver
against the plain integrator calls (the monitor must not change the
trajectory) and against energy conservation of classic RK4.
"""
import numpy as np
import pytest

import nbfmm
import workloads

pytestmark = pytest.mark.gpu


def _direct_phi(xyz, m, G=1.0):
    d = xyz[:, None, :] - xyz[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    return -G * (m[None, :] / r).sum(1)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    c = nbfmm.Context(order=8, leaf_capacity=16, G=1.0, eps2=0.0)
    yield c
    c.close()


def _eval_pot(ctx, xyz, m):
    n = len(m)
    dx, dm = nbfmm.DeviceArray(xyz), nbfmm.DeviceArray(m)
    da, dp = nbfmm.DeviceArray(np.zeros((n, 3))), nbfmm.DeviceArray(np.zeros(n))
    ctx.eval_accel_pot_dev(n, dx.ptr, dm.ptr, da.ptr, dp.ptr)
    return da.download(ctx.stream), dp.download(ctx.stream)


@pytest.mark.parametrize("precision,p,cut,tol", [(0, 4, 1 << 50, 1e-12), (0, 12, 256, 1e-7), (1, 12, 256, 2e-6)])
def test_potential_vs_direct(ctx, precision, p, cut, tol):
    xyz, m = workloads.plummer(3000, seed=12)
    ctx.set_params(order=p, leaf_capacity=16, p2p_cut=cut, precision=precision)
    try:
        acc, phi = _eval_pot(ctx, xyz, m)
        acc_only = ctx.eval_accel(xyz, m)
    finally:
        ctx.set_params(precision=0)
    ref = _direct_phi(xyz, m)
    err = _rel(phi, ref)
    print(f"[pot] precision={precision} p={p} cut={cut} rel L2 {err:.3e}")
    assert err < tol
    # the potential path leaves the accelerations as they are (equal up to the
    # order in which the near-field atomics of different tasks land)
    assert _rel(acc, acc_only) < 1e-14


def test_diagnostics_sums(ctx):
    rng = np.random.default_rng(3)
    n = 5000
    x, v = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    m, phi = rng.uniform(0.5, 1.5, n), rng.normal(size=n)
    dx, dv, dm, dp = (nbfmm.DeviceArray(a) for a in (x, v, m, phi))
    d = ctx.diagnostics_dev(n, dx.ptr, dv.ptr, dm.ptr, dp.ptr)
    ref = {
        "kinetic": 0.5 * (m * (v * v).sum(1)).sum(), "potential": 0.5 * (m * phi).sum(),
        "px": (m * v[:, 0]).sum(), "py": (m * v[:, 1]).sum(), "pz": (m * v[:, 2]).sum(),
        "mass": m.sum(),
    }
    L = (m[:, None] * np.cross(x, v)).sum(0)
    ref.update(lx=L[0], ly=L[1], lz=L[2])
    ref["total"] = ref["kinetic"] + ref["potential"]
    for k, val in ref.items():
        assert abs(d[k] - val) <= 1e-11 * max(1.0, abs(val)), k
    # without potentials the potential term is 0
    assert ctx.diagnostics_dev(n, dx.ptr, dv.ptr, dm.ptr)["potential"] == 0.0


def test_run_driver_matches_integrator_and_conserves_energy(ctx):
    xyz, m, vel = workloads.plummer(2000, seed=21, with_velocity=True)
    n = len(m)
    ctx.set_params(order=10, leaf_capacity=16, p2p_cut=256, G=1.0, eps2=0.0, precision=0)
    steps, every, dt = 20, 5, 1e-3
    # classic RK4 with the monitor
    dx, dv, da, dm = (nbfmm.DeviceArray(a) for a in (xyz, vel, np.zeros((n, 3)), m))
    rec = ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, steps, integrator=2, every=every)
    x1, v1 = dx.download(ctx.stream), dv.download(ctx.stream)
    # the same steps without the monitor
    ex, ev, ea, em = (nbfmm.DeviceArray(a) for a in (xyz, vel, np.zeros((n, 3)), m))
    ctx.step_rk4_dev(n, ex.ptr, ev.ptr, ea.ptr, em.ptr, dt, steps, 1)
    # same trajectory (near-field atomics land in any order: rounding-level differences)
    assert _rel(x1, ex.download(ctx.stream)) < 1e-12
    assert _rel(v1 - vel, ev.download(ctx.stream) - vel) < 1e-10
    assert rec.shape == (1 + steps // every, 10)
    e = rec[:, 2]
    drift = np.abs(e - e[0]).max() / abs(e[0])
    dp = np.linalg.norm(rec[:, 3:6] - rec[0, 3:6], axis=1)
    p_scale = (m[:, None] * np.abs(vel)).sum()
    print(f"[run] energy drift {drift:.3e}; momentum change {dp.max():.3e} (scale {p_scale:.3e})")
    # initial kinetic and potential energies follow the state
    assert abs(rec[0, 0] - 0.5 * (m * (vel * vel).sum(1)).sum()) < 1e-12
    assert abs(rec[0, 1] - 0.5 * (m * _direct_phi(xyz, m)).sum()) < 1e-7 * abs(rec[0, 1])
    assert drift < 1e-5
    assert dp.max() < 1e-6 * p_scale
    # symplectic Euler through the driver equals nbfmm_step_se_dev too
    dx, dv, da = (nbfmm.DeviceArray(a) for a in (xyz, vel, np.zeros((n, 3))))
    ctx.run_dev(n, dx.ptr, dv.ptr, da.ptr, dm.ptr, dt, 3, integrator=0, monitor=False)
    ex, ev, ea = (nbfmm.DeviceArray(a) for a in (xyz, vel, np.zeros((n, 3))))
    ctx.step_se_dev(n, ex.ptr, ev.ptr, ea.ptr, dm.ptr, dt, 3)
    assert _rel(dx.download(ctx.stream), ex.download(ctx.stream)) < 1e-12
