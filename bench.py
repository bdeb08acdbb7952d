"""Benchmark of the FSD velocity/integration step on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 32768 unit
spheres at volume fraction 0.45 in a periodic cube, simple shear rate 1,
deterministic random forces U[-1,1]^3 (zero mean), T = 0, xi = 0.5, PSE error
1e-3, IC(0)-preconditioned saddle solve run until the true relative residual
||b - A x|| / ||b|| <= 1e-4.  A "step" is one full Stokes::integrateStepOne:
neighbour list, preconditioner build (prune, RCM, IC(0)), force, right-hand
side, GMRES solve, explicit Euler move.

Prints one JSON line.  `value` is steps/s with all inputs resident in HBM.
`roofline` is computed for the kernel class that takes the most time, from its
algorithmic bytes and its launch durations measured with HIP events on the
library's stream during the timed steps.  `cpu_baseline` times the FP64 CPU
oracle on a bounded sample of the same step on this host.
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
                "ichol": "ichol_level_kernel", "gmres_vec": "multidot_partial4_kernel"}


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
    if cls == "ichol":       # one factorisation: A blocks in, L blocks out (lower and diagonal)
        return 144 * 2 * (st["lower_blocks"] + N)
    if cls == "gmres_vec":   # Gram-Schmidt of a 17N-float vector on k basis vectors, per pass h = V^T w
        k = kbar             # (k + 1 reads) and w -= V h (k + 1 reads, w in and out), then the norm
        # (one read, folded into the second pass of CGS2) and the scaling (in and out); k averaged over
        # the solve's iterations
        return 4 * 17 * N * (gs_passes * (2 * k + 4) + 3)
    return None


def cpu_baseline(pos, L, F, shear_rate, gpu_iters, threads):
    """FP64 oracle on this host: the step's fixed part (neighbour list, preconditioner
    build, right-hand side) plus 3 preconditioned saddle applications, extrapolated to
    the GPU's mean GMRES iterations per step"""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle.oracle import Oracle
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
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "FP64 oracle, one N=32768 step: neighbour list + preconditioner build + rhs timed "
                      "(%.1f s), 3 preconditioned saddle applications timed (%.2f s each), extrapolated "
                      "to %.1f GMRES iterations/step" % (t_fixed, t_iter, gpu_iters)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=32768)
    ap.add_argument("--phi", type=float, default=0.45)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--restart", type=int, default=100, help="GMRES restart length")
    ap.add_argument("--gram-schmidt", type=int, default=1, choices=[1, 2],
                    help="classical Gram-Schmidt passes per Arnoldi step (the library's default is 2)")
    ap.add_argument("--guess", default="extrapolate", choices=["previous", "extrapolate", "quadratic"],
                    help="initial guess of each solve: the previous solution (the reference) or its extrapolation")
    ap.add_argument("--profile-all", action="store_true", help="time every kernel class (diagnostics)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MB",
                    help="only run 5 device copies of MB megabytes (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print("# copy calibration: 5 copies of %d bytes (read %d + write %d per copy)" % (4 * n, 4 * n, 4 * n),
              file=sys.stderr)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the FSD step of one suspension is not sharded across GPUs "
              "in this build (see the design notes on sharding)", file=sys.stderr)
        sys.exit(2)

    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    from fsd_amd.workloads import random_packing, random_forces
    from fsd_amd import _lib

    N = args.N
    t = time.time()
    pos, L = random_packing(N, args.phi, seed=args.seed)
    F = random_forces(N, seed=args.seed)
    print("# workload generated in %.1f s: N=%d phi=%.2f L=%.3f" % (time.time() - t, N, args.phi, L),
          file=sys.stderr, flush=True)
    shear_rate = 1.0
    st = Stokes(N, L, xi=0.5, error=1e-3, gmres_tol=1e-4, gmres_true_tol=1e-4, dt=1e-3, gmres_restart=args.restart,
                guess=args.guess, gram_schmidt_passes=args.gram_schmidt)
    pse = st.pse_info()
    state = ParticleState(pos)
    dF = torch.from_numpy(F).cuda()

    # warm-up, timing every kernel class to find the dominant one
    st.profile(True)
    infos = []
    for k in range(args.warmup):
        infos.append(st.integrateStepOne(k, state, force6=dF, shear_rate=shear_rate))
        print("# warmup step %d: %s" % (k, infos[-1]), file=sys.stderr, flush=True)
    totals = {c: st.profile_get(c) for c in _lib.PROF_NAMES}
    st.profile_reset()
    # single kernels (trsv is the five-kernel preconditioner application, reported in the class times)
    kernel_classes = ["spread", "green", "contract", "realspace", "rfu", "sweep", "ichol", "gmres_vec"]
    dom = max(kernel_classes, key=lambda c: totals[c][1])
    print("# warmup kernel-class times (launches, ms): %s" % json.dumps(totals), file=sys.stderr, flush=True)

    # timed steps: HIP events only around the dominant class, one launch in eight (the events'
    # own cost stays out of the step time; the timed launches give the kernel's average duration)
    st.profile(True, classes=None if args.profile_all else [dom], every=1 if args.profile_all else 8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for k in range(args.steps):
        info = st.integrateStepOne(args.warmup + k, state, force6=dF, shear_rate=shear_rate)
        iters.append(info["iterations"])
        infos.append(info)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    print("# timed steps: iterations %s, initial residuals %s"
          % ([i["iterations"] for i in infos[-args.steps:]],
             ["%.2e" % i["initial_residual"] for i in infos[-args.steps:]]), file=sys.stderr, flush=True)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    launches, kms = st.profile_get(dom)
    last = infos[-1]
    avg_ms = kms / max(launches, 1)
    it_mean = float(np.mean(iters))
    kbar = 0.5 * (min(it_mean, args.restart) + 1.0)
    bytes_per_launch = algorithmic_bytes(dom, N, None, pse, last, kbar, args.gram_schmidt)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_launch else None
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    # (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by its reading for a plain copy)
    traffic, kname = None, KERNEL_NAMES.get(dom)
    tfile = os.path.join(ROOT, "profiles", "traffic_N32768.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        k = tr.get("kernels", {}).get(kname)
        if k:
            traffic = 1024.0 * (k["fetch_kib"] * tr["fetch_factor"] + k["write_kib"] * tr["write_factor"])
    if args.profile_all:
        allt = {c: st.profile_get(c) for c in _lib.PROF_NAMES}
        print("# timed kernel-class times (launches, ms): %s" % json.dumps(allt), file=sys.stderr, flush=True)

    print("# timed: %.2f ms/step, dominant kernel class %s (%d launches, %.3f ms avg)"
          % (ms_per_step, dom, launches, avg_ms), file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        print("# CPU baseline (FP64 oracle) ...", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(pos, L, F, shear_rate, float(np.mean(iters)), int(os.environ.get("OMP_NUM_THREADS", "16")))
        except Exception as e:  # the baseline is reported, never required
            print("# cpu baseline failed: %r" % (e,), file=sys.stderr)

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded packing and forces)",
        "config": {
            "workload": "N=32768 phi=0.45 periodic suspension, shear rate 1, random forces, T=0, xi=0.5, "
                        "PSE error 1e-3, IC(0)-preconditioned GMRES(%d) with %s Gram-Schmidt to true residual 1e-4, "
                        "initial guess: %s" % (args.restart, "classical" if args.gram_schmidt == 1 else "twice-classical",
                                               GUESS_TEXT[args.guess]),
            "N": N, "phi": args.phi, "grid": [pse["Nx"], pse["Ny"], pse["Nz"]], "P": pse["P"],
            "preconditioner": "ichol",
            "gmres_restart": args.restart, "initial_guess": args.guess, "gram_schmidt_passes": args.gram_schmidt,
            "gmres_iterations_mean": float(np.mean(iters)),
            "nlist_rebuilds": int(sum(i.get("nlist_rebuilt", 0) for i in infos[-args.steps:])),
            "true_residual_last": last.get("true_residual"),
            "nnz": last.get("nnz"), "lub_pairs": last.get("lub_pairs"), "levels": last.get("levels"),
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kname or dom,
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
