"""Per-particle output files of the reference (Stokes::OutputData,
Stokes.cc:703-931), written with the same names, columns and printf formats
so analysis scripts of the reference read them unchanged:

  raw/stresslet.%010d.txt              tag, n S_xx, n S_xy, n S_xz, n S_yy, n S_yz, n S_zz
  raw/position.%010d.txt               tag, x, y, z, director, unwrapped x, y, z
  raw/velocity.%010d.txt               tag, U, Omega
  raw/interparticle_stresslet.%010d.txt  tag, pair-force stresslet (6 components)
  overlap.txt (appended)               timestep, overlapping pairs, max / mean overlap, min gap

(n = number density).  The reference writes them at timestep 0 before the
first step and after every step whose timestep + 1 is a multiple of the
output period, from the step's neighbour list and positions.  The
interparticle stresslet uses the reference's force law in single precision.
"""
import os

import numpy as np


def _min_image(d, box):
    """HOOMD BoxDim::minImage in float (rint of d / L computed as d * (1 / L), tilts xy, xz, yz)"""
    Lx, Ly, Lz, xy, xz, yz = [np.float32(v) for v in box]
    iLx, iLy, iLz = np.float32(1.0) / Lx, np.float32(1.0) / Ly, np.float32(1.0) / Lz
    d = d.astype(np.float32).copy()
    img = np.rint(d[:, 2] * iLz).astype(np.float32)
    d[:, 0] -= np.float32(xz) * Lz * img
    d[:, 1] -= np.float32(yz) * Lz * img
    d[:, 2] -= Lz * img
    img = np.rint(d[:, 1] * iLy).astype(np.float32)
    d[:, 0] -= np.float32(xy) * Ly * img
    d[:, 1] -= Ly * img
    img = np.rint(d[:, 0] * iLx).astype(np.float32)
    d[:, 0] -= Lx * img
    return d


def write_output(timestep, pos, ori, pos_gb, vel11, nneigh, nbrs, box, ndsr, kappa, k_n, beta, epsq,
                 directory="."):
    """write the five output files of one output step; pos, pos_gb: N x 3 (or
    N x 4), ori: N x 3, vel11: 11N, (nneigh, nbrs): the neighbour list (each
    particle's neighbours in nbrs, consecutively)"""
    pos = np.asarray(pos, dtype=np.float32)[:, :3]
    pos_gb = np.asarray(pos_gb, dtype=np.float32)[:, :3]
    ori = np.asarray(ori, dtype=np.float32).reshape(-1, 3)
    vel11 = np.asarray(vel11, dtype=np.float32)
    N = pos.shape[0]
    Lx, Ly, Lz = box[0], box[1], box[2]
    volume = float(Lx) * float(Ly) * float(Lz)
    nden = np.float32(N / volume)
    F0 = np.float32(1.0 / ndsr)
    Ham = np.float32(F0 * np.float32(beta))
    epsq = np.float32(epsq)
    raw = os.path.join(directory, "raw")
    os.makedirs(raw, exist_ok=True)
    ts = "%010d" % int(timestep)
    S = vel11[6 * N:].reshape(N, 5)
    U = vel11[:6 * N].reshape(N, 6)
    # pair quantities over the neighbour list
    ci = np.repeat(np.arange(N), nneigh)
    R = _min_image(pos[np.asarray(nbrs, dtype=np.int64)] - pos[ci], box)
    dist = np.sqrt((R * R).sum(axis=1, dtype=np.float32)).astype(np.float32)
    gap = (dist - np.float32(2.0)).astype(np.float32)
    ov = gap < 0
    max_overlap = float(-gap[ov].min()) if ov.any() else 0.0
    cnt = float(ov.sum())
    avr = float((-gap[ov]).sum()) / cnt if cnt > 0 else 0.0
    min_gap = 0.0 if ov.any() else float(min(10.0, float(gap.min()) if gap.size else 10.0))
    if cnt > 0:
        cnt /= 2.0
    act = dist <= np.float32(2.0 + 10.0 / kappa)
    rmax = np.float32(2.0 + np.sqrt(float(epsq)))
    g1 = (dist - rmax).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        fm = np.where(g1 >= 0,
                      -F0 * np.exp(-g1 * np.float32(kappa)).astype(np.float32)
                      + Ham / (np.float32(12.0) * (g1 * g1 + epsq)),
                      Ham / (np.float32(12.0) * epsq) - F0 - np.float32(k_n) * np.abs(g1)).astype(np.float32)
        nrm = (R / dist[:, None]).astype(np.float32)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    rep = np.zeros((N, 6), dtype=np.float32)
    for c, (a, b) in enumerate(pairs):
        val = np.where(act, nden * (fm * nrm[:, a] * R[:, b]) / np.float32(2.0), 0).astype(np.float32)
        np.add.at(rep[:, c], ci, val)
    with open(os.path.join(raw, "stresslet.%s.txt" % ts), "w") as f0, \
            open(os.path.join(raw, "position.%s.txt" % ts), "w") as f1, \
            open(os.path.join(raw, "interparticle_stresslet.%s.txt" % ts), "w") as f3, \
            open(os.path.join(raw, "velocity.%s.txt" % ts), "w") as f4:
        for i in range(N):
            sxx, sxy, sxz, syz, syy = [np.float32(nden * S[i, c]) for c in range(5)]
            szz = np.float32(-sxx - syy)
            f0.write("%7i %12.3e %12.3e %12.3e %12.3e %12.3e %12.3e \n" % (i, sxx, sxy, sxz, syy, syz, szz))
            f1.write("%7i %15.6e %15.6e %15.6e %12.3e %12.3e %12.3e %15.6e %15.6e %15.6e \n"
                     % (i, pos[i, 0], pos[i, 1], pos[i, 2], ori[i, 0], ori[i, 1], ori[i, 2],
                        pos_gb[i, 0], pos_gb[i, 1], pos_gb[i, 2]))
            f4.write("%7i %15.6e %15.6e %15.6e %12.3e %12.3e %12.3e \n" % ((i,) + tuple(U[i])))
            f3.write("%7i %12.3e %12.3e %12.3e %12.3e %12.3e %12.3e \n" % ((i,) + tuple(rep[i])))
    with open(os.path.join(directory, "overlap.txt"), "a") as f2:
        f2.write("%9i %8.0f %12.3e %12.3e  %12.3e \n" % (int(timestep), cnt, max_overlap, avr, min_gap))
    if max_overlap > 0.1:
        raise RuntimeError("Large overlap: %12.3e" % max_overlap)
