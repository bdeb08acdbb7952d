"""Seeded synthetic suspensions for tests and benchmarks.

``random_packing(N, phi, seed)`` places N unit spheres at volume fraction phi
in a periodic cube with every centre-to-centre distance >= 2 + min_gap.
Random sequential addition saturates near phi = 0.38, so positions are
grown and relaxed instead: random centres, then repeated removal of pair
overlaps while the exclusion diameter is raised in stages to 2 + min_gap.
"""
import math

import numpy as np
from scipy.spatial import cKDTree


def box_length(N, phi, radius=1.0):
    return (N * 4.0 / 3.0 * math.pi * radius ** 3 / phi) ** (1.0 / 3.0)


def _min_image(d, L):
    return d - L * np.rint(d / L)


def _relax(x, L, D, max_iter=4000, slack=1.01):
    """push overlapping pairs apart until all distances >= D (x in [0, L))"""
    for _ in range(max_iter):
        tree = cKDTree(x, boxsize=L)
        pairs = tree.query_pairs(D, output_type="ndarray")
        if len(pairs) == 0:
            return x, True
        i, j = pairs[:, 0], pairs[:, 1]
        d = _min_image(x[j] - x[i], L)
        r = np.sqrt((d * d).sum(axis=1))
        r = np.maximum(r, 1e-9)
        push = 0.5 * (D * slack - r)[:, None] * (d / r[:, None])
        n = x.shape[0]
        disp = np.empty_like(x)
        for c in range(3):
            disp[:, c] = np.bincount(j, weights=push[:, c], minlength=n) - np.bincount(i, weights=push[:, c], minlength=n)
        x = np.mod(x + disp, L)
    return x, False


def random_packing(N, phi, seed=0, min_gap=1e-3, radius=1.0):
    """returns (pos float32[N,3] in [-L/2, L/2), L)"""
    rng = np.random.default_rng(seed)
    L = box_length(N, phi, radius)
    x = rng.uniform(0.0, L, size=(N, 3))
    Dt = 2.0 * radius + min_gap
    stages = 1 if phi < 0.15 else int(4 + 20 * phi)
    for s in range(1, stages + 1):
        D = Dt * (s / stages) ** (1.0 / 3.0) if stages > 1 else Dt
        x, ok = _relax(x, L, D)
    x, ok = _relax(x, L, Dt)
    if not ok:
        raise RuntimeError("packing did not relax at phi=%g" % phi)
    pos = (x - 0.5 * L).astype(np.float32)
    # verify in float32 with the exact exclusion distance
    tree = cKDTree(np.mod(pos.astype(np.float64) + 0.5 * L, L), boxsize=L)
    if len(tree.query_pairs(Dt - 1e-7, output_type="ndarray")):
        x, ok = _relax(np.mod(pos.astype(np.float64) + 0.5 * L, L), L, Dt + 1e-6)
        pos = (x - 0.5 * L).astype(np.float32)
    # keep strictly inside [-L/2, L/2) after the float cast
    pos = np.where(pos >= np.float32(0.5 * L), pos - np.float32(L), pos).astype(np.float32)
    return pos, float(L)


def random_forces(N, seed=0):
    """force U[-1,1]^3 with zero mean, zero torque, as float32[6N]"""
    rng = np.random.default_rng(seed + 7919)
    f = rng.uniform(-1.0, 1.0, size=(N, 3))
    f -= f.mean(axis=0)
    F = np.zeros((N, 6), dtype=np.float32)
    F[:, :3] = f
    return F.reshape(-1)


def pos4(pos):
    """float32[N,3] -> float32[N,4] (w = 0)"""
    out = np.zeros((pos.shape[0], 4), dtype=np.float32)
    out[:, :3] = pos
    return out


# the configurations the benchmarks and parity tests use
CONFIGS = {
    "C1": dict(N=343, phi=0.20, shear_rate=0.0, gmres_tol=1e-6, T=0.0),
    "C2": dict(N=4096, phi=0.30, shear_rate=0.0, gmres_tol=1e-3, T=1.0),
    "C3": dict(N=32768, phi=0.45, shear_rate=1.0, gmres_tol=1e-4, T=0.0),
}


def jittered_lattice(n_side, phi, seed=0, radius=1.0, min_gap=1e-3):
    """n_side^3 spheres on a simple cubic lattice filling a periodic cube at volume
    fraction phi, each coordinate displaced uniformly by up to a quarter of the lattice
    gap less min_gap (two lattice neighbours then approach by at most half the gap, and
    no two spheres overlap).  Returns (pos float32[N,3] in [-L/2, L/2), L)."""
    N = n_side ** 3
    L = box_length(N, phi, radius)
    a = L / n_side
    gap = a - 2.0 * radius
    if gap <= 2.0 * min_gap:
        raise ValueError("phi %g is too dense for a simple cubic lattice" % phi)
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * a - 0.5 * L
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    j = 0.25 * (gap - min_gap)
    x = x + rng.uniform(-j, j, size=x.shape)
    pos = x.astype(np.float32)
    pos = np.where(pos >= np.float32(0.5 * L), pos - np.float32(L), pos)
    pos = np.where(pos < np.float32(-0.5 * L), pos + np.float32(L), pos).astype(np.float32)
    return pos, float(L)


def random_orientations(N, seed=0):
    """unit vectors uniform on the sphere, float32[N, 3]"""
    rng = np.random.default_rng(seed + 104729)
    v = rng.standard_normal((N, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)
