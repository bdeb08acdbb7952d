"""The autotuner (fsd_autotune): it times every variant of the mobility and R_FU
kernels on the context's own positions and keeps the fastest; every variant it can
choose computes the same operator.  Each forced variant is compared with the FP64
oracle (1e-5, as the other mobility and R_FU parity tests), and a tuned step with
the one it replaces."""
import numpy as np
import pytest

from tests._cases import packing, rel_err, random_vec, forces

pytestmark = pytest.mark.gpu

# C2's packing: its 40^3 grid has the fused x-transform kernel, so every knob has two choices
N_, PHI_, SEED_ = 4096, 0.30, 12


@pytest.fixture(scope="module")
def c2():
    import torch
    from fsd_amd.stokes import Stokes
    from oracle.oracle import Oracle
    pos, L = packing(N_, PHI_, SEED_)
    print("autotune case N", N_, "phi", PHI_, "seed", SEED_, flush=True)
    p4 = np.zeros((N_, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    st = Stokes(N_, L)
    st.build_nlist(dpos)
    orc = Oracle(pos, L)
    gF = random_vec(11 * N_, 3).astype(np.float32)
    U = random_vec(6 * N_, 4).astype(np.float32)
    return st, orc, dpos, gF, U, orc.mobility(gF, 3), orc.rfu(U)


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def test_autotune_chooses_a_variant(c2):
    st, orc, dpos, gF, U, Mref, Rref = c2
    assert st.pse_info()["Nx"] == 40
    choice, us = st.autotune(dpos)
    print("tuning", choice, "best us", us)
    t, _ = st.tuning()
    assert t["tuned"] == 1 and all(v > 0 for v in us.values())
    assert t["real_lanes"] in (4, 8, 16) and t["rfu_lanes"] in (8, 16)
    assert rel_err(st.mobility(dpos, _t(gF), 3).cpu().numpy(), Mref) < 1e-5
    assert rel_err(st.rfu(dpos, _t(U)).cpu().numpy(), Rref) < 1e-5


@pytest.mark.parametrize("spread,contract,xpass", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)])
def test_wave_variants_match_oracle(c2, spread, contract, xpass):
    st, orc, dpos, gF, U, Mref, Rref = c2
    st.set_tuning(spread=spread, contract=contract, xpass=xpass)
    e = rel_err(st.mobility(dpos, _t(gF), 3).cpu().numpy(), Mref)
    print("spread", spread, "contract", contract, "xpass", xpass, "mobility rel err", e)
    assert e < 1e-5


@pytest.mark.parametrize("lanes", [4, 8, 16])
@pytest.mark.parametrize("block", [128, 256])
def test_real_space_variants_match_oracle(c2, lanes, block):
    st, orc, dpos, gF, U, Mref, Rref = c2
    st.set_tuning(spread=0, contract=0, xpass=0, real_lanes=lanes, real_block=block)
    e = rel_err(st.mobility(dpos, _t(gF), 3).cpu().numpy(), Mref)
    print("real-space lanes", lanes, "block", block, "mobility rel err", e)
    assert e < 1e-5


@pytest.mark.parametrize("lanes", [8, 16])
@pytest.mark.parametrize("block", [128, 256])
def test_rfu_variants_match_oracle(c2, lanes, block):
    st, orc, dpos, gF, U, Mref, Rref = c2
    st.set_tuning(rfu_lanes=lanes, rfu_block=block)
    e = rel_err(st.rfu(dpos, _t(U)).cpu().numpy(), Rref)
    print("R_FU lanes", lanes, "block", block, "rel err", e)
    assert e < 1e-5


def test_tuned_step_equals_default_step():
    """a step with autotune on (tuned at its own positions) gives the velocities of the
    step with the default shapes, to FP32 rounding of the changed reduction orders"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = 343, 0.2, 11
    pos, L = packing(N, phi, seed)
    F = torch.from_numpy(forces(N, 31)).cuda()
    vel = []
    for tune in (False, True):
        st = Stokes(N, L, gmres_tol=1e-6, gmres_true_tol=1e-7, gmres_restart=100, refine=True, autotune=tune)
        state = ParticleState(pos)
        info = st.integrateStepOne(0, state, force6=F, shear_rate=1.0)
        assert info["converged"]
        vel.append(state.velocity.cpu().numpy()[:6 * N].astype(np.float64))
        print("autotune", tune, st.tuning())
    e = rel_err(vel[1], vel[0])
    print("tuned vs default step velocities rel err", e)
    assert e < 1e-5
