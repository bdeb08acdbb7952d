"""Two-sphere lubrication resistance table (Stokes_ResistanceTable.cc:34-23057).

``load_resistance_table()`` returns ``(dist, vals)`` as float32 arrays of 1000
and 22000 entries in the reference's layout: for each of the 1000 log-spaced
distances r = 2 + h, the 22 functions XA11 XA12 YA11 YA12 YB11 YB12 XC11 XC12
YC11 YC12 XG11 XG12 YG11 YG12 YH11 YH12 XM11 XM12 YM11 YM12 ZM11 ZM12.
Values are parsed decimal -> double -> float, as the reference assigns its
double literals to float arrays.
"""
import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "resistance_table.txt")
TABLE_MIN = 0.0001     # h of the first row (Stokes_ResistanceTable.cc:37)
TABLE_DR = 0.004305    # log10 spacing (Stokes_ResistanceTable.cc:40)
CONTACT_ROW = 232      # r = 2.000997, lubrication contact regularisation (Lubrication.cu:172-186)

_cache = None


def load_resistance_table():
    global _cache
    if _cache is None:
        rows = np.loadtxt(DATA, dtype=np.float64, comments="#")
        if rows.shape != (1000, 23):
            raise ValueError("unexpected resistance table shape %s" % (rows.shape,))
        dist = rows[:, 0].astype(np.float32)
        vals = rows[:, 1:].astype(np.float32).reshape(-1)
        _cache = (np.ascontiguousarray(dist), np.ascontiguousarray(vals))
    return _cache
