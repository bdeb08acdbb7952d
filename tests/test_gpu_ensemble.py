"""Replicas stepped concurrently from host threads (fsd_amd/ensemble.py): every
replica's trajectory equals the same replica stepped alone (Brownian noise is keyed
by seed and timestep, so the runs differ only by the order of floating-point
atomics in spreading and by the solver stopping at 1e-4: positions to 1e-3 of the
displacement)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_concurrent_replicas_match_single_runs():
    import torch
    from fsd_amd.ensemble import make_replicas, run_ensemble, replica_params
    from fsd_amd.stokes import Stokes, ParticleState
    from fsd_amd.workloads import random_forces
    reps = make_replicas(6, seed=3, n_range=(200, 800))
    print("replica sizes", [r[0] for r in reps], "volume fractions", [round(r[1], 3) for r in reps], flush=True)
    out, final = run_ensemble(steps=2, warmup=0, threads=3, seed=3, autotune=False, replicas=reps, keep=True,
                              log=lambda m: print(m, flush=True))
    print(out)
    assert out["converged"] == 6
    for i in (0, 5):
        N, phi, pos, L, fin = final[i]
        st = Stokes(N, L, **replica_params(1.0, 1e-3, 3, i, "ichol_jacobi", False))
        state = ParticleState(pos)
        F = torch.from_numpy(random_forces(N, seed=3 + i)).cuda()
        for s in range(2):
            st.integrateStepOne(s, state, force6=F, shear_rate=0.0)
        alone = state.positions()
        img = lambda v: v - L * np.rint(v / L)   # minimum image (a particle may cross the box)
        d = np.abs(img(fin - pos)).max()
        e = np.abs(img(fin - alone)).max()
        print("replica", i, "N", N, "max displacement", d, "max difference from the single run", e)
        assert d > 0 and e < 1e-3 * d
