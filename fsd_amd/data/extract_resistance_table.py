"""Extract the tabulated two-sphere lubrication resistance functions into a data file.

The reference ships the table as 23 000 assignment statements inside
``source/Stokes_ResistanceTable.cc`` (distances at :56-1055, values at
:1058-23057).  This script reads those numbers (and nothing else) and writes
``resistance_table.txt``: one line per tabulated centre-to-centre distance,
``r  XA11 XA12 YA11 YA12 YB11 YB12 XC11 XC12 YC11 YC12 XG11 XG12 YG11 YG12
YH11 YH12 XM11 XM12 YM11 YM12 ZM11 ZM12``, with the decimal strings kept
verbatim so that the float values the library loads equal the reference's
(decimal -> double -> float, Stokes_ResistanceTable.cc:34-47).

It only runs where the reference tree is mounted; the committed output is what
the library and the tests read.
"""
import os
import re
import sys

SRC = "/root/reference/source/Stokes_ResistanceTable.cc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "resistance_table.txt")

NAMES = ["XA11", "XA12", "YA11", "YA12", "YB11", "YB12", "XC11", "XC12", "YC11", "YC12",
         "XG11", "XG12", "YG11", "YG12", "YH11", "YH12", "XM11", "XM12", "YM11", "YM12",
         "ZM11", "ZM12"]


def main(src=SRC, out=OUT):
    pat = re.compile(r"h_ResTable_(dist|vals)\.data\[\s*(\d+)\s*\]\s*=\s*([-+0-9.eE]+)\s*;")
    dist = {}
    vals = {}
    with open(src, "r", encoding="utf-8", errors="replace") as f:
        for line in f:
            m = pat.search(line)
            if not m:
                continue
            kind, idx, num = m.group(1), int(m.group(2)), m.group(3)
            (dist if kind == "dist" else vals)[idx] = num
    if sorted(dist) != list(range(1000)) or sorted(vals) != list(range(22000)):
        raise SystemExit("unexpected table layout: %d distances, %d values" % (len(dist), len(vals)))
    with open(out, "w") as f:
        f.write("# two-sphere resistance functions, 1000 log-spaced distances (h_min=1e-4, dlog10=0.004305)\n")
        f.write("# columns: r " + " ".join(NAMES) + "\n")
        for i in range(1000):
            f.write(dist[i] + " " + " ".join(vals[22 * i + k] for k in range(22)) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main(*sys.argv[1:])
