"""Writes tests/golden/ref_angular.npz: the reference's own AngularData
arrays (AngularData.cpp:9-68: omega, omega_pos, Ox, Oy, Oz, unpadded) for
K = 1..7 and its xE1 (E1.cpp:136-159) at sample points, taken from the
reference's host sources compiled by `make -C oracle ref` (oracle/_ref).

Run in a container that holds the reference sources:
    make -C oracle ref && python tests/golden/make_ref_fixtures.py
The GPU tests and the oracle tests read the fixture, so nothing built from
the reference is loaded where the fixture suffices."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

XE1_POINTS = np.array([0.01, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 5.0, 8.0])


def main():
    if oracle.ref() is None:
        sys.exit("oracle/_ref not built: run `make -C oracle ref` where the reference sources are present")
    out = {}
    for K in range(1, 8):
        r = oracle.ref_angular(K)
        for key in ("omega", "omega_pos", "Ox", "Oy", "Oz"):
            out["K%d_%s" % (K, key)] = r[key]
    out["xE1_x"] = XE1_POINTS
    out["xE1"] = np.array([oracle.ref().ref_xE1(float(x)) for x in XE1_POINTS])
    np.savez_compressed(os.path.join(HERE, "ref_angular.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_angular.npz"))


if __name__ == "__main__":
    main()
