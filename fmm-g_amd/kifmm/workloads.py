"""Synthetic inputs of the benchmark configurations (host numpy, seeded).

The reference's tt0 driver draws its own points (tt0.cpp:211-298); these are the
distributions the benchmark configurations name.  Seeds: base 20261202, streams
base + 10*c (positions), +1 (densities), +2 (second density), +3 (check
targets), +4 (finite-difference coordinates) for configuration index c.
"""
import numpy as np

BASE_SEED = 20261202
HI = 1.0 - 2.0 ** -53  # largest double below 1: points stay inside [0,1)^3


def seed(config, stream):
    return BASE_SEED + 10 * config + stream


def uniform(n, s):
    rng = np.random.default_rng(s)
    return np.minimum(rng.random((n, 3)), HI)


def spheres(n, s, nsph=32, rmin=0.02, rmax=0.08, cmin=0.1, cmax=0.9):
    """Points uniform by area on non-overlapping spheres, with outward normals."""
    rng = np.random.default_rng(s)
    ctr, rad = [], []
    while len(ctr) < nsph:
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(cmin, cmax, 3)
        if all(np.linalg.norm(c - cj) > r + rj for cj, rj in zip(ctr, rad)):
            ctr.append(c)
            rad.append(r)
    ctr, rad = np.array(ctr), np.array(rad)
    w = rad ** 2 / np.sum(rad ** 2)
    k = rng.choice(nsph, size=n, p=w)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pos = np.clip(ctr[k] + rad[k][:, None] * d, 0.0, HI)
    return pos, d


def gaussians(n, s, ng=16, smin=0.005, smax=0.05, cmin=0.2, cmax=0.8, frac=0.9):
    """90% from isotropic Gaussians (clipped to the unit cube), 10% uniform."""
    rng = np.random.default_rng(s)
    sig = rng.uniform(smin, smax, ng)
    ctr = rng.uniform(cmin, cmax, (ng, 3))
    from_g = rng.random(n) < frac
    k = rng.integers(0, ng, n)
    pos = rng.random((n, 3))
    m = int(from_g.sum())
    pos[from_g] = ctr[k[from_g]] + sig[k[from_g]][:, None] * rng.standard_normal((m, 3))
    return np.clip(pos, 0.0, HI)


def densities(n, dof, s, kind="uniform"):
    rng = np.random.default_rng(s)
    if kind == "uniform":
        return rng.uniform(-1.0, 1.0, n * dof)
    return rng.standard_normal(n * dof)
