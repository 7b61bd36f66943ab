"""Golden values of the reference's CH4 test (tests/run_tests.py:19-28 and
tests/refs/ch4_ref.{rad,cube} of markelgtx/INCA_parallel)."""
import os

import numpy as np

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ch4")
INTEGRAL = 45.013363086638655
INTEGRAL_ERR = 0.01336309
ZERO_HALF = 7.7702702634802296          # printed I(0)/2
TWO_POINTS = (15.540540526960413, 1.2854504822876072)
NELEC = 10


def rad_table():
    rows = [l.split() for l in open(os.path.join(G, "ch4_ref.rad"))
            if l.strip() and not l.strip().startswith("#")]
    return np.array([[float(v) for v in r] for r in rows])


def cube_values():
    lines = open(os.path.join(G, "ch4_ref.cube")).read().splitlines()
    return np.array([float(v) for v in lines[11:]])
