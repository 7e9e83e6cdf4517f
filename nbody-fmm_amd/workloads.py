"""Synthetic inputs of the benchmark configurations (numpy, seeded).

  uniform_cube      positions U[0,1)^3, masses U[0.5,1.5]           (N=500 case)
  plummer           Plummer sphere, a=1, M=1, equal masses; Aarseth-Henon-Wielen
                    (1974) sampling with the mass-fraction cut X < 0.999
                    (r < 38.7a), resampled; isotropic velocities by rejection
  plummer_mix       K Plummer spheres, centres U[0,100)^3, a log-uniform [0.1,2],
                    sphere mass proportional to a, equal particle counts
  batched_systems   sizes round(exp(U[ln 1e3, ln 6.4e4])), each a Plummer sphere
                    with a ~ U[0.5, 2]

All use numpy's PCG64 generator with the given seed.
"""
import numpy as np


def uniform_cube(n, seed=1234):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0.0, 1.0, (n, 3))
    m = rng.uniform(0.5, 1.5, n)
    return xyz, m


def _plummer_radius(rng, n, a, cut):
    X = rng.uniform(0.0, 1.0, n)
    bad = (X >= cut) | (X <= 0.0)
    while bad.any():
        X[bad] = rng.uniform(0.0, 1.0, int(bad.sum()))
        bad = (X >= cut) | (X <= 0.0)
    return a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)


def _isotropic(rng, n, r):
    u = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([r * s * np.cos(ph), r * s * np.sin(ph), r * u], axis=1)


def plummer(n, seed=42, a=1.0, total_mass=1.0, cut=0.999, with_velocity=False, G=1.0):
    rng = np.random.default_rng(seed)
    r = _plummer_radius(rng, n, a, cut)
    xyz = _isotropic(rng, n, r)
    m = np.full(n, total_mass / n)
    if not with_velocity:
        return xyz, m
    # speed by rejection from g(q) = q^2 (1 - q^2)^{7/2}, v = q v_esc
    q = np.empty(n)
    todo = np.ones(n, bool)
    while todo.any():
        k = int(todo.sum())
        x1 = rng.uniform(0.0, 1.0, k)
        x2 = rng.uniform(0.0, 0.1, k)
        ok = x2 < x1 * x1 * (1.0 - x1 * x1) ** 3.5
        idx = np.nonzero(todo)[0]
        q[idx[ok]] = x1[ok]
        todo[idx[ok]] = False
    vesc = np.sqrt(2.0 * G * total_mass / a) * (1.0 + (r / a) ** 2) ** -0.25
    vel = _isotropic(rng, n, q * vesc)
    return xyz, m, vel


def plummer_mix(n, n_spheres=32, seed=7, box=100.0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, box, (n_spheres, 3))
    a = np.exp(rng.uniform(np.log(0.1), np.log(2.0), n_spheres))
    counts = np.full(n_spheres, n // n_spheres)
    counts[: n - counts.sum()] += 1
    mass = a / a.sum()
    xs, ms = [], []
    for k in range(n_spheres):
        x, m = plummer(int(counts[k]), seed=seed * 1000 + k, a=a[k], total_mass=mass[k])
        xs.append(x + centres[k])
        ms.append(m)
    return np.concatenate(xs), np.concatenate(ms)


def batched_systems(n_systems=256, seed=11, lo=1e3, hi=6.4e4, with_velocity=False):
    """List of (xyz, m) per system, or (xyz, m, vel) with velocities (the
    positions are the same either way)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_systems))).astype(np.int64)
    a = rng.uniform(0.5, 2.0, n_systems)
    systems = [plummer(int(sizes[k]), seed=seed * 10000 + k, a=a[k], with_velocity=with_velocity)
               for k in range(n_systems)]
    return systems


def sample_targets(n, k, seed=4096):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(k, n), replace=False)).astype(np.int64)


def sample_targets_disjoint(n, k, exclude, seed=8191):
    """k targets drawn without replacement from [0, n) minus `exclude` (a
    held-out sample beside the one a tuner chose its setting on)."""
    rng = np.random.default_rng(seed)
    pool = np.setdiff1d(np.arange(n, dtype=np.int64), np.asarray(exclude, dtype=np.int64), assume_unique=False)
    return np.sort(rng.choice(pool, size=min(k, len(pool)), replace=False)).astype(np.int64)
