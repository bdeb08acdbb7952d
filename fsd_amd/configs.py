"""The active-suspension configuration: N = 64^3 = 262144 squirmers at volume
fraction 0.4, each swimming with speed 1 along its own random director
(sqm_B1 = 3/2: the Euler move adds (2/3) B1 times the director,
Integrator.cu:126-228), Brownian motion at kT = 1, saddle solves in FP32 GMRES
with FP64-residual refinement to a true residual of 1e-6, the learned
preconditioner with its IC(0) fallback.

    python -m fsd_amd.configs [--steps 50]

prints one JSON line: steps/s, the FP64 true residual of every step, the fraction
of steps whose preconditioner fell back to IC(0), and the HBM rate of spreading
and contraction (algorithmic bytes over HIP-event times of their launches).
The start is a jittered simple cubic lattice (no overlaps) rather than a random
packing, which at this size takes minutes to relax on the host; Brownian motion
and swimming randomise it within the first steps.
"""
import json
import sys
import time

import numpy as np


def spread_bytes(N, M):
    """one spreading launch: particle records (position 16 B, force and couplet 48 B,
    stencil origin 12 B) in, 11 real grids of M points out"""
    return 76 * N + 44 * M


def contract_bytes(N, M):
    """one contraction launch: 11 real grids in, position and origin (28 B) in, 11 floats out"""
    return 44 * M + 72 * N


def run_active(n_side=64, phi=0.4, steps=50, dt=1e-3, T=1.0, seed=2026, calibration=50, tol=1e-6, skip_after=2,
               log=print):
    import torch
    from .stokes import Stokes, ParticleState
    from .workloads import jittered_lattice, random_orientations
    N = n_side ** 3
    t0 = time.time()
    pos, L = jittered_lattice(n_side, phi, seed=seed)
    ori = random_orientations(N, seed=seed)
    log("# lattice start: N=%d phi=%.2f L=%.3f (%.1f s)" % (N, phi, L, time.time() - t0))
    st = Stokes(N, L, T=T, dt=dt, seed=seed, sqm_B1=1.5, gmres_tol=1e-5, gmres_true_tol=tol, gmres_restart=100,
                gmres_maxiter=3000, refine=True, guess="previous", precond="learned", jacobi_iters=5,
                learned_calibration=calibration, fallback_jacobi=8, autotune=True, learned_skip_after=skip_after)
    pse = st.pse_info()
    M = pse["Nx"] * pse["Ny"] * pse["Nz"]
    log("# PSE grid %dx%dx%d, P=%d" % (pse["Nx"], pse["Ny"], pse["Nz"], pse["P"]))
    state = ParticleState(pos, ori=ori)
    st.profile(True, classes=["spread", "contract"], every=4)
    infos, swim = [], []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s in range(steps):
        before = state.positions().astype(np.float64)
        ts = time.perf_counter()
        info = st.integrateStepOne(s, state, force6=None, shear_rate=0.0)
        ms = 1e3 * (time.perf_counter() - ts)
        d = state.positions().astype(np.float64) - before
        d -= L * np.rint(d / L)
        o = state.ori.view(-1, 3).cpu().numpy().astype(np.float64)
        swim.append(float(np.mean(np.sum(d * o, axis=1))) / dt)   # mean speed along the directors
        infos.append(info)
        log("# step %d: %.0f ms, iterations %d, true residual %.2e, preconditioner %d, fallback %d, Lanczos %d/%d, "
            "speed along director %.4f" % (s, ms, info["iterations"], info["true_residual"], info["precond_used"],
                                           info["fallback"], info["lanczos_ff"], info["lanczos_nf"], swim[-1]))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    rates = {}
    for cls, fb in (("spread", spread_bytes(N, M)), ("contract", contract_bytes(N, M))):
        n, ms = st.profile_get(cls)
        avg = ms / max(n, 1)
        rates[cls] = {"avg_launch_ms": avg, "algorithmic_bytes": fb, "GB_s": fb / (avg * 1e-3) / 1e9 if n else None}
    learned_steps = sum(i["precond_used"] == 1 for i in infos)
    return {
        "config": "active suspension N=%d phi=%.2f, swim speed 1, kT=%g, dt=%g, refined to %g" % (N, phi, T, dt, tol),
        "N": N, "phi": phi, "grid": [pse["Nx"], pse["Ny"], pse["Nz"]], "steps": steps,
        "steps_per_s": steps / (t2 - t1), "ms_per_step": 1e3 * (t2 - t1) / steps,
        "converged_steps": int(sum(i["converged"] for i in infos)),
        "true_residual_max": float(max(i["true_residual"] for i in infos)),
        "gmres_iterations_mean": float(np.mean([i["iterations"] for i in infos])),
        "learned_steps": int(learned_steps),
        "fallback_fraction": float(sum(i["fallback"] != 0 for i in infos)) / steps,
        "fallbacks_by_kind": {k: int(sum(i["fallback"] == c for i in infos))
                              for c, k in ((1, "pivot"), (2, "iterations"), (3, "skipped"))},
        "calibration_steps": int(sum(i["precond_used"] == 0 and i["fallback"] == 0 for i in infos)),
        "mean_speed_along_director": float(np.mean(swim)),
        "hbm": rates, "launch_shapes": st.tuning()[0],
    }


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--n-side", type=int, default=64)
    ap.add_argument("--skip-after", type=int, default=2,
                    help="iteration fallbacks in a row after which the learned factor is skipped until the next "
                         "calibration step (0: never)")
    a = ap.parse_args(argv)
    out = run_active(n_side=a.n_side, steps=a.steps, skip_after=a.skip_after,
                     log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
