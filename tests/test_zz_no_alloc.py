"""The allocation-free entry point fsd_rfu_noalloc against fsd_rfu.

Both run the same kernels on the same neighbour list; only where the pair
records live differs (the library's fixed device array vs context buffers),
so the products must agree bit for bit.  Two calls in a row, the second on a
smaller system (N = 4096, then N = 343), check that the array's leftover
contents from the first call do not leak into the second.  A system above the
array's particle limit is refused with an error.
"""
import numpy as np
import pytest

from tests._cases import packing, random_vec

pytestmark = pytest.mark.gpu


def _ctx(N, phi, seed):
    import torch
    from fsd_amd.stokes import Stokes
    pos, L = packing(N, phi, seed)
    print("noalloc case N", N, "phi", phi, "seed", seed, flush=True)
    st = Stokes(N, L)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    st.build_nlist(dpos)
    return st, dpos


def test_noalloc_rfu_matches_rfu_twice_shrinking():
    import torch
    for (N, phi, seed), useed in (((4096, 0.30, 12), 5), ((343, 0.20, 11), 6)):
        st, dpos = _ctx(N, phi, seed)
        U = torch.from_numpy(random_vec(6 * N, useed).astype(np.float32)).cuda()
        ref = st.rfu(dpos, U).cpu().numpy()
        got = st.rfu_noalloc(dpos, U).cpu().numpy()
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), "max diff %g" % np.abs(got - ref).max()


def test_noalloc_rejects_too_many_particles():
    import torch
    from fsd_amd import _lib
    st, dpos = _ctx(9261, 0.20, 13)
    U = torch.zeros(6 * 9261, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.FsdError, match="scratch limit"):
        st.rfu_noalloc(dpos, U)
