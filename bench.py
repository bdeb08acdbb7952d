"""Benchmark of the FSD velocity/integration step on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 32768 unit
spheres at volume fraction 0.45 in a periodic cube, simple shear rate 1,
deterministic random forces U[-1,1]^3 (zero mean), T = 0, xi = 0.5, PSE error
1e-3, saddle solve run until the true relative residual ||b - A x|| / ||b||
<= 1e-4.  A "step" is one full Stokes::integrateStepOne: neighbour list,
preconditioner build, force, right-hand side, GMRES solve, explicit Euler move.

Three arms run on the same packing from the same start (--arms):
  ichol        - the reference's preconditioner: IC(0) with exact triangular solves;
  ichol_jacobi - IC(0) applied by 8 block Jacobi iterations per triangular solve;
  learned      - the learned factor (a message-passing network on the lubrication
                 graph correcting a block LDL^T factor on IC(0)'s pattern: four
                 parallel pivot passes, the pair blocks' fill correction) applied by
                 5 block Jacobi iterations, with the IC(0) fallbacks (non-finite
                 network outputs; iteration count above IC(0)'s at the last
                 calibration step, one IC(0) step every --calibration steps; its
                 IC(0) steps applied by 8 Jacobi iterations) and FP64-residual
                 iterative refinement of the solve; a step that does not calibrate
                 keeps the last RCM ordering while its neighbour list stands
                 (--rcm-every-step: the RCM of every step's graph).
Each context autotunes its mobility and R_FU launch shapes at its first step.
`value` is the learned arm's steps/s (all inputs resident in HBM); every arm and
the learned/IC(0) ratio are in `config.arms`.

Prints one JSON line.  `roofline` is computed for the kernel class that takes
the most time in the learned arm, from its algorithmic bytes and its launch
durations measured with HIP events on the library's stream during the timed
steps.  `cpu_baseline` times the FP64 CPU oracle on a bounded sample of the
same step on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at N=32k phi=0.45: learned vs IChol precond at equal residual 1e-4"
HBM_PEAK_GBS = 8000.0
GUESS_TEXT = {"previous": "previous solution", "extrapolate": "linear extrapolation of the last two solutions",
              "quadratic": "quadratic extrapolation of the last three solutions"}
# kernel-class name -> kernel name in rocprofv3 traces
KERNEL_NAMES = {"spread": "spread_tile2_kernel<8, 6>", "green": "fft_x_green_kernel<9>", "contract": "contract_tile2_kernel<13, 6>",
                "realspace": "realspace_kernel<16>", "rfu": "rfu_apply_kernel<8>", "sweep": "tp_sweep1_kernel",
                "ichol": "ichol_level_kernel", "gmres_vec": "multidot_partial4_kernel", "ljit": "lj_fwd_kernel",
                "jacobi": "jac_fwd_kernel"}


def algorithmic_bytes(cls, N, info, pse, st, kbar=25.5, gs_passes=2):
    """bytes a launch of kernel class `cls` must move at minimum (each input
    read once, each output written once), per the layouts in HBM"""
    M = pse["Nx"] * pse["Ny"] * pse["Nz"]
    Mh = pse["Nx"] * pse["Ny"] * (pse["Nz"] // 2 + 1)
    if cls == "spread":      # particle records in (pos 16 + F/C 48 + origin 12), 11 real grids out
        return 76 * N + 44 * M
    if cls == "contract":    # 11 real grids in, pos + origin in, 11 floats per particle out
        return 44 * M + 28 * N + 44 * N
    if cls == "green":       # 11 complex half spectra in and out + k vectors (with the x transforms at Nx = 72)
        return 2 * 88 * Mh + 16 * Mh
    if cls == "realspace":   # nlist entries + per-particle pos/records/wave/addU in, 11N out
        return 4 * st["nlist_total"] + (16 + 48 + 44 + 24 + 44) * N
    if cls == "rfu":         # per pair: index + 2 float4; per particle: self block, U, x_F, out
        return 36 * st["lub_pairs"] + (144 + 24 + 24 + 24) * N
    if cls == "sweep":       # one triangular sweep, one right-hand side (the saddle preconditioner's
        # d = b_F - b_U): every block of the factor below the diagonal once (6x6 FP32, 144 B), the rows'
        # inputs in (24 B) and results out (24 B)
        return 144 * st["lower_blocks"] + (24 + 24) * N
    if cls == "trsv":        # one saddle preconditioner application (five kernels): both sweeps as above;
        # Z d (b_F, b_U in 48 B, Z 144 B, out 24 B), reorder (48 B), Z^T u and the combination
        # (Z 144 B, u 24 B, b_F and b_S 44 B in, x_F, x_S, x_U 68 B out) per particle
        return 2 * (144 * st["lower_blocks"] + 48 * N) + (48 + 144 + 24 + 48 + 144 + 24 + 44 + 68) * N
    if cls == "ljit":        # one block Jacobi iteration of the learned factor: every pair block W (144 B) and
        # inverse pivot (144 B) once, the fixed right-hand side (t or v_k) and the previous iterate in, the
        # new iterate out (24 B each per particle)
        return 144 * (st["lower_blocks"] + N) + 3 * 24 * N
    if cls == "jacobi":      # one block Jacobi iteration on the IC(0) factor: every folded block Z_P L_PQ
        # (144 B) and its column index (4 B) once, the fixed right-hand side and the previous iterate in,
        # the new iterate out (24 B each per particle), the row offsets (4 B)
        return 148 * st["lower_blocks"] + (3 * 24 + 4) * N
    if cls == "ichol":       # one factorisation: A blocks in, L blocks out (lower and diagonal)
        return 144 * 2 * (st["lower_blocks"] + N)
    if cls == "gmres_vec":   # Gram-Schmidt of a 17N-float vector on k basis vectors, per pass h = V^T w
        k = kbar             # (k + 1 reads) and w -= V h (k + 1 reads, w in and out), then the norm
        # (one read, folded into the second pass of CGS2) and the scaling (in and out); k averaged over
        # the solve's iterations
        return 4 * 17 * N * (gs_passes * (2 * k + 4) + 3)
    return None


def cpu_baseline(pos, L, F, shear_rate, gpu_iters):
    """FP64 oracle on this host's cores (OpenMP): the step's fixed part (neighbour list,
    preconditioner build, right-hand side) timed once plus 3 IC(0)-preconditioned saddle
    applications timed and extrapolated to the IC(0) arm's mean GMRES iterations per step.
    A full FP64 step to the true residual 1e-4 takes over a minute at N = 32768, so the
    sample is bounded to about 20 s of CPU work."""
    from oracle.oracle import Oracle, threads
    t0 = time.perf_counter()
    orc = Oracle(pos, L)
    orc.build_precond()
    b = orc.rhs(F, shear_rate)
    t1 = time.perf_counter()
    x = np.random.default_rng(0).standard_normal(len(b))
    k = 3
    for _ in range(k):
        y = orc.saddle_multiply(x)
        x = orc.saddle_precond(y)
        x /= np.linalg.norm(x)
    t2 = time.perf_counter()
    t_fixed = t1 - t0
    t_iter = (t2 - t1) / k
    step_s = t_fixed + gpu_iters * t_iter
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": threads(), "kind": "port",
            "sample": "FP64 oracle (IC(0) preconditioner), one N=32768 step: neighbour list + preconditioner "
                      "build + rhs timed (%.1f s), 3 preconditioned saddle applications timed (%.2f s each), "
                      "extrapolated to %.1f GMRES iterations/step (the IC(0) arm's mean)" % (t_fixed, t_iter, gpu_iters)}


def run_arm(name, kw, pos, L, F, args, dom=None, log=None):
    """warm-up and timed steps of one arm; returns (stats dict, Stokes, kernel-class totals)"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    from fsd_amd import _lib
    N = args.N
    st = Stokes(N, L, xi=0.5, error=1e-3, gmres_tol=1e-4, gmres_true_tol=1e-4, dt=1e-3, gmres_restart=args.restart,
                guess=args.guess, gram_schmidt_passes=args.gram_schmidt, autotune=not args.no_autotune, **kw)
    state = ParticleState(pos)
    st.profile(True)
    infos = []
    for k in range(args.warmup):
        infos.append(st.integrateStepOne(k, state, force6=F, shear_rate=1.0))
        log("# %s warmup step %d: %s" % (name, k, infos[-1]))
    totals = {c: st.profile_get(c) for c in _lib.PROF_NAMES}
    st.profile_reset()
    log("# %s launch shapes (autotuner): %s" % (name, json.dumps(st.tuning())))
    classes = ["spread", "green", "contract", "realspace", "rfu", "sweep", "ichol", "gmres_vec", "ljit", "jacobi"]
    log("# %s warmup kernel-class times (launches, ms): %s" % (name, json.dumps(totals)))
    # timed steps: HIP events around one launch in eight of every kernel class (every launch with
    # --profile-all); the dominant class is the one with the most sampled time in these steps
    st.profile(True, classes=classes, every=1 if args.profile_all else 8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for k in range(args.steps):
        info = st.integrateStepOne(args.warmup + k, state, force6=F, shear_rate=1.0)
        its.append(info["iterations"])
        infos.append(info)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timed = infos[-args.steps:]
    if dom is None:
        dom = max(classes, key=lambda c: st.profile_get(c)[1])
    log("# %s timed steps: iterations %s, initial residuals %s, preconditioner %s"
        % (name, its, ["%.2e" % i["initial_residual"] for i in timed], [i["precond_used"] for i in timed]))
    out = {"ms_per_step": (t1 - t0) * 1e3 / args.steps, "steps_per_s": args.steps / (t1 - t0),
           "launch_shapes": st.tuning()[0],
           "gmres_iterations_mean": float(np.mean(its)),
           "true_residual_max": float(max(i["true_residual"] for i in timed)),
           "nlist_rebuilds": int(sum(i.get("nlist_rebuilt", 0) for i in timed)),
           "nnz": timed[-1]["nnz"], "lub_pairs": timed[-1]["lub_pairs"], "lower_blocks": timed[-1]["lower_blocks"],
           "levels": timed[-1]["levels"]}
    if kw.get("precond") == "learned":
        out["learned_steps"] = int(sum(i["precond_used"] == 1 for i in timed))
        out["fallback_pivot"] = int(sum(i["fallback"] == 1 for i in timed))
        out["fallback_iterations"] = int(sum(i["fallback"] == 2 for i in timed))
        out["calibration_steps"] = int(sum(i["precond_used"] == 0 and i["fallback"] == 0 for i in timed))
    allt = {c: st.profile_get(c) for c in classes}
    log("# %s timed kernel-class times (sampled launches, ms): %s" % (name, json.dumps(allt)))
    return out, st, dom, timed[-1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=32768)
    ap.add_argument("--phi", type=float, default=0.45)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--restart", type=int, default=100, help="GMRES restart length")
    ap.add_argument("--gram-schmidt", type=int, default=1, choices=[1, 2],
                    help="classical Gram-Schmidt passes per Arnoldi step")
    ap.add_argument("--guess", default="extrapolate", choices=["previous", "extrapolate", "quadratic"],
                    help="initial iterate of each step's solve")
    ap.add_argument("--jacobi", type=int, default=5,
                    help="block Jacobi iterations per triangular solve (learned; at N=32768: 2 -> 19.9 steps/s, "
                         "3 -> 27.1, 5 -> 28.4, 8 -> 26.0)")
    ap.add_argument("--calibration", type=int, default=20,
                    help="learned arm: one IC(0) step every this many steps (its iteration count caps the learned "
                         "solves); 20 puts one such step in the 20 timed steps")
    ap.add_argument("--ichol-jacobi", type=int, default=8,
                    help="block Jacobi iterations per triangular solve of the ichol_jacobi arm and of the "
                         "learned arm's IC(0) steps (calibration and fallback)")
    ap.add_argument("--arms", default="ichol,ichol_jacobi,learned",
                    help="comma-separated arms to run (ichol, ichol_jacobi, learned); the last is `value`")
    ap.add_argument("--profile-all", action="store_true", help="time every kernel class (diagnostics)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--rcm-every-step", action="store_true",
                    help="learned arm: order the particles by the RCM of every step's graph (default: the "
                         "last ordering is kept while the neighbour list it was computed under stands)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MB",
                    help="only run 5 device copies of MB megabytes (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.calib_copy:
        import torch
        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("# copy calibration: 5 copies of %d bytes (read %d + write %d per copy)" % (4 * n, 4 * n, 4 * n))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the FSD step of one suspension is not sharded across GPUs "
              "in this build (see the design notes on sharding)", file=sys.stderr)
        sys.exit(2)

    import torch
    from fsd_amd.workloads import random_packing, random_forces

    N = args.N
    t = time.time()
    pos, L = random_packing(N, args.phi, seed=args.seed)
    F = random_forces(N, seed=args.seed)
    log("# workload generated in %.1f s: N=%d phi=%.2f L=%.3f" % (time.time() - t, N, args.phi, L))
    dF = torch.from_numpy(F).cuda()
    arm_kw = {"ichol": dict(precond="ichol"),
              "ichol_jacobi": dict(precond="ichol_jacobi", jacobi_iters=args.ichol_jacobi),
              "learned": dict(precond="learned", jacobi_iters=args.jacobi, learned_calibration=args.calibration,
                              refine=True, fallback_jacobi=args.ichol_jacobi,
                              learned_keep_order=not args.rcm_every_step)}
    names = [a for a in args.arms.split(",") if a]
    arms, head = {}, names[-1]
    for name in names:
        res, st, dom, last = run_arm(name, arm_kw[name], pos, L, dF, args, log=log)
        arms[name] = res
        if name == head:
            head_st, head_dom, head_last = st, dom, last
        else:
            del st
    pse = head_st.pse_info()
    launches, kms = head_st.profile_get(head_dom)
    avg_ms = kms / max(launches, 1)
    it_mean = arms[head]["gmres_iterations_mean"]
    kbar = 0.5 * (min(it_mean, args.restart) + 1.0)
    bytes_per_launch = algorithmic_bytes(head_dom, N, None, pse, head_last, kbar, args.gram_schmidt)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_launch else None
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    # (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by its reading for a plain copy)
    traffic, kname = None, KERNEL_NAMES.get(head_dom)
    shapes = arms[head]["launch_shapes"]
    if head_dom == "realspace":      # the autotuner's lanes per particle
        kname = "realspace_kernel<%d>" % shapes["real_lanes"]
    elif head_dom == "rfu":
        kname = "rfu_apply_kernel<%d>" % shapes["rfu_lanes"]
    tfile = os.path.join(ROOT, "profiles", "traffic_N32768.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        k = tr.get("kernels", {}).get(kname)
        if k:
            traffic = 1024.0 * (k["fetch_kib"] * tr["fetch_factor"] + k["write_kib"] * tr["write_factor"])
    log("# %s: %.2f ms/step, dominant kernel class %s (%d launches, %.3f ms avg)"
        % (head, arms[head]["ms_per_step"], head_dom, launches, avg_ms))
    cpu = None
    if not args.no_cpu:
        log("# CPU baseline (FP64 oracle) ...")
        try:
            ref_iters = arms.get("ichol", arms[head])["gmres_iterations_mean"]
            cpu = cpu_baseline(pos, L, F, 1.0, ref_iters)
        except Exception as e:  # the baseline is reported, never required
            log("# cpu baseline failed: %r" % (e,))
    ratio = None
    if "ichol" in arms and "learned" in arms:
        ratio = arms["learned"]["steps_per_s"] / arms["ichol"]["steps_per_s"]
    out = {
        "metric": METRIC,
        "value": arms[head]["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": arms[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded packing and forces)",
        "config": {
            "workload": "N=32768 phi=0.45 periodic suspension, shear rate 1, random forces, T=0, xi=0.5, "
                        "PSE error 1e-3, GMRES(%d) with %s Gram-Schmidt to true residual 1e-4, initial guess: %s; "
                        "value = %s arm" % (args.restart, "classical" if args.gram_schmidt == 1 else "twice-classical",
                                            GUESS_TEXT[args.guess], head),
            "N": N, "phi": args.phi, "grid": [pse["Nx"], pse["Ny"], pse["Nz"]], "P": pse["P"],
            "preconditioner": head, "jacobi_iters": args.jacobi, "calibration_every": args.calibration,
            "learned_ordering": "RCM of every step" if args.rcm_every_step else
                                "last RCM kept while its neighbour list stands",
            "gmres_restart": args.restart, "initial_guess": args.guess, "gram_schmidt_passes": args.gram_schmidt,
            "arms": arms, "speedup_learned_vs_ichol": ratio,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kname or head_dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
