"""Shared seeded inputs for the parity tests."""
import functools

import numpy as np

from fsd_amd.workloads import random_packing, random_forces


@functools.lru_cache(maxsize=None)
def packing(N, phi, seed):
    pos, L = random_packing(N, phi, seed=seed)
    return pos, L


def rel_err(a, b):
    """max-norm error relative to the reference's max-norm"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)


def rel_err2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def forces(N, seed):
    return random_forces(N, seed)


def random_vec(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


def sorted_lists(nneigh, nbrs):
    out = []
    k = 0
    for c in nneigh:
        out.append(tuple(sorted(nbrs[k:k + c])))
        k += c
    return out
