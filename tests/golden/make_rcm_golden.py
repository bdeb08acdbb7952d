"""Writes tests/golden/rcm_golden.npz: pruned-neighbour graphs (packed
adjacency with self loops, as Precondition_Reorder hands them to genrcm) and
the permutations the reference's GENRCM (source/rcm.cpp, built into
oracle/_ref by oracle/build_ref.sh) returns for them.

Run from the repository root: python tests/golden/make_rcm_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def main():
    from _cases import packing
    from oracle.oracle import Oracle
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_rcm.so"))
    lib.ref_genrcm.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    out = {}
    names = []
    graphs = []
    for N, phi, seed in [(64, 0.1, 1), (343, 0.2, 11), (200, 0.4, 2)]:
        pos, L = packing(N, phi, seed)
        o = Oracle(pos, L)
        pn, phead, plist, _ = o.prune()
        graphs.append(("pack_%d_%g" % (N, phi), phead.copy(), plist.copy()))
    # two chains, a triangle and isolated nodes, with self loops
    adj = {i: [i] for i in range(12)}
    for a, b in [(0, 1), (1, 2), (2, 3), (5, 6), (6, 7), (8, 9), (9, 10), (10, 8)]:
        adj[a].append(b)
        adj[b].append(a)
    phead = np.zeros(13, dtype=np.int32)
    plist = []
    for i in range(12):
        plist += sorted(adj[i])
        phead[i + 1] = len(plist)
    graphs.append(("components", phead, np.array(plist, dtype=np.int32)))
    for name, phead, plist in graphs:
        N = len(phead) - 1
        perm = np.zeros(N, dtype=np.int32)
        lib.ref_genrcm(N, int(phead[N]), phead.ctypes.data, plist.ctypes.data, perm.ctypes.data)
        out[name + "_phead"] = phead
        out[name + "_plist"] = plist
        out[name + "_perm"] = perm
        names.append(name)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "rcm_golden.npz"), **out)
    print("wrote", names)


if __name__ == "__main__":
    main()
