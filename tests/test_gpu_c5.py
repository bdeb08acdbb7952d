"""The active-suspension configuration at its full size (N = 262144, phi = 0.4,
swim speed 1, kT = 1, solves refined in FP64 to 1e-6, the learned preconditioner
with its IC(0) fallback), a few steps, through size-independent properties:
every step converges to the FP64 true residual 1e-6; the PSE grid stays within
256^3; the particles swim at speed 1 along their directors on average (the
Brownian and hydrodynamic parts average out over 262144 random directors to a few
1e-3 of it)."""
import pytest

pytestmark = pytest.mark.gpu


def test_active_suspension_steps():
    from fsd_amd.configs import run_active
    out = run_active(steps=3, log=lambda m: print(m, flush=True))
    print({k: v for k, v in out.items() if k != "hbm"}, out["hbm"])
    assert max(out["grid"]) <= 256
    assert out["converged_steps"] == 3 and out["true_residual_max"] <= 1e-6
    assert abs(out["mean_speed_along_director"] - 1.0) < 0.05
