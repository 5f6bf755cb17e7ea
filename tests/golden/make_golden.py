"""Write the small golden fixtures of tests/golden/ from the FP64 oracle.

The reference ships no test vectors, so these fixtures are the oracle's FP64 results
on tiny ring meshes (text, 16 significant digits, Matrix::savetxt layout).  They pin
regressions of the oracle itself (tests/test_golden.py) and give the device tests a
fixed expected output that does not depend on re-running the oracle.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mpflow_amd.phantom import ellipse_phantom, frame_seed, inclusion_phantom  # noqa: E402
from tests.problems import make_problem, oracle_of  # noqa: E402

CASES = {
    # name: (rings, electrodes, width, harmonics, height, port height, p2, quirks)
    "p1_2d": (6, 8, 0.2, 1, 1.0, 1.0, False, True),
    "p1_2p5d": (6, 8, 0.2, 3, 0.4, 0.1, False, True),
    "p2_2d": (4, 8, 0.3, 1, 1.0, 1.0, True, True),
}


def case_problem(name):
    nr, L, w, C, h, hp, p2, q = CASES[name]
    p = make_problem(nr, L, w, C=C, height=h, port_height=hp, p2=p2)
    return p, q


def case_gamma(name, p):
    if name == "p1_2p5d":
        return ellipse_phantom(p["nodes"], p["elements"], frame_seed(2, 99))
    return inclusion_phantom(p["nodes"], p["elements"], frame_seed(1, 99), radius=0.35)


def main():
    for name in CASES:
        p, q = case_problem(name)
        o = oracle_of(p, quirks=q)
        g = case_gamma(name, p)
        U, phis, _ = o.solve(g)
        Jr = o.jacobian(g, phis, quirks=True)
        Jx = o.jacobian(g, phis, quirks=False)
        d = os.path.join(HERE, name)
        os.makedirs(d, exist_ok=True)
        np.savetxt(os.path.join(d, "nodes.txt"), p["nodes"], fmt="%.16g")
        np.savetxt(os.path.join(d, "elements.txt"), p["elements"], fmt="%d")
        np.savetxt(os.path.join(d, "ports.txt"), p["ports"], fmt="%d")
        np.savetxt(os.path.join(d, "gamma.txt"), g, fmt="%.16g")
        np.savetxt(os.path.join(d, "U.txt"), U, fmt="%.16g")
        # unpadded rows (d < D, m < M) and every 4th element column keep the files small
        rows = [dd * o.Mp + mm for dd in range(o.D) for mm in range(o.M)]
        cols = np.arange(0, o.E, 4)
        np.savetxt(os.path.join(d, "J_reference.txt"), Jr[np.ix_(rows, cols)], fmt="%.10g")
        if o.C > 1:
            np.savetxt(os.path.join(d, "J_exact.txt"), Jx[np.ix_(rows, cols)], fmt="%.10g")
        print(name, U.shape, Jr.shape)


if __name__ == "__main__":
    main()
