"""An ensemble of independent periodic suspensions of uneven size, stepped
concurrently on one GPU: every replica is its own context (own HIP stream,
workspaces, neighbour list, preconditioner and autotuned launch shapes), and a
pool of host threads drives the replicas' fsd_step calls, so that the host
synchronisations inside one replica's step (GMRES's Hessenberg readback, the
preconditioner's host plan, the Lanczos eigensolves) overlap with the kernels of
the others.  The steps are not captured into one hipGraph and no kernel is
persistent: each step keeps those host synchronisations.

    python -m fsd_amd.ensemble [--replicas 128] [--steps 5] [--threads 16]

prints one JSON line: aggregate particle-steps/s over the timed steps and the
p50/p99 of the per-replica step times.  Replica sizes are uniform in
[200, 2000], volume fractions uniform in [0.1, 0.5], Brownian motion at kT = 1,
random forces, solves to a true residual of 1e-4.
"""
import json
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def make_replicas(R, seed=0, n_range=(200, 2000), phi_range=(0.1, 0.5)):
    """(N, phi, positions, L) of R seeded random packings"""
    from .workloads import random_packing
    rng = np.random.default_rng(seed)
    out = []
    for r in range(R):
        N = int(rng.integers(n_range[0], n_range[1] + 1))
        phi = float(rng.uniform(*phi_range))
        pos, L = random_packing(N, phi, seed=int(rng.integers(1 << 30)))
        out.append((N, phi, pos, L))
    return out


def replica_params(T, dt, seed, i, precond, autotune):
    """the solver settings of replica i"""
    return dict(T=T, dt=dt, seed=seed + i, gmres_tol=1e-4, gmres_true_tol=1e-4, gmres_restart=50, gmres_maxiter=2000,
                guess="extrapolate", gram_schmidt_passes=1, precond=precond, jacobi_iters=8, autotune=autotune)


def run_ensemble(R=128, steps=5, warmup=2, threads=16, T=1.0, dt=1e-3, seed=0, precond="ichol_jacobi",
                 autotune=True, log=print, replicas=None, keep=False):
    """steps the replicas (make_replicas(R, seed) unless given) warmup + steps times with a
    pool of host threads; returns the summary (and, with keep, the replicas with their final
    positions)"""
    import torch
    from .stokes import Stokes, ParticleState
    from .workloads import random_forces
    t0 = time.time()
    reps = make_replicas(R, seed=seed) if replicas is None else replicas
    R = len(reps)
    log("# %d replicas generated in %.1f s (N from %d to %d, %d particles in all)"
        % (R, time.time() - t0, min(r[0] for r in reps), max(r[0] for r in reps), sum(r[0] for r in reps)))
    ctxs = []
    for i, (N, phi, pos, L) in enumerate(reps):   # contexts are created one after the other
        st = Stokes(N, L, **replica_params(T, dt, seed, i, precond, autotune))
        state = ParticleState(pos)
        F = torch.from_numpy(random_forces(N, seed=seed + i)).cuda()
        ctxs.append((st, state, F))
    torch.cuda.synchronize()
    times = [[] for _ in range(R)]
    infos = [None] * R
    lock = threading.Lock()

    def one(i, step, record):
        st, state, F = ctxs[i]
        t = time.perf_counter()
        info = st.integrateStepOne(step, state, force6=F, shear_rate=0.0, sync=False)
        dtm = time.perf_counter() - t
        if record:
            with lock:
                times[i].append(dtm)
                infos[i] = info

    with ThreadPoolExecutor(max_workers=threads) as ex:
        for s in range(warmup):
            list(ex.map(lambda i: one(i, s, False), range(R)))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(warmup, warmup + steps):
            list(ex.map(lambda i: one(i, s, True), range(R)))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    ptot = sum(r[0] for r in reps)
    flat = np.array([x for ts in times for x in ts]) * 1e3
    out = {
        "replicas": R, "steps": steps, "warmup": warmup, "threads": threads, "precond": precond,
        "particles": ptot, "particle_steps_per_s": ptot * steps / (t2 - t1),
        "ensemble_steps_per_s": steps / (t2 - t1),
        "replica_step_ms_p50": float(np.percentile(flat, 50)), "replica_step_ms_p99": float(np.percentile(flat, 99)),
        "converged": int(sum(i["converged"] for i in infos if i)), "gmres_iterations_mean":
            float(np.mean([i["iterations"] for i in infos if i])),
    }
    if keep:
        return out, [(N, phi, pos, L, ctxs[i][1].positions()) for i, (N, phi, pos, L) in enumerate(reps)]
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--replicas", type=int, default=128)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--threads", type=int, default=16)
    a = ap.parse_args(argv)
    out = run_ensemble(R=a.replicas, steps=a.steps, warmup=a.warmup, threads=a.threads,
                       log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
