"""Writes the regression fixtures in this directory from the oracle.

    python tests/golden/make_golden.py

The reference ships no test vectors for this path, so these fixtures are not
external ground truth: they freeze the oracle's outputs (the library's own
sin/cos/asin/log) for seeded inputs, element by element and for short runs,
so that any later change of the maps or of the math library that alters a
single bit is caught on the CPU (tests/test_golden.py) and on the GPU
(tests/test_gpu_parity.py::test_golden_fixtures)."""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from elegant_amd import configs  # noqa: E402
from elegant_amd.lattice import Element, Lattice  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = {
    "kquad": [Element("Q", "KQUAD", L=0.3, K1=1.2, N_KICKS=8, DX=1e-4, TILT=0.01)],
    "ksext": [Element("S", "KSEXT", L=0.2, K2=-40.0, K1=0.05, J1=0.02, N_KICKS=4)],
    "koct": [Element("O", "KOCT", L=0.2, K3=1000.0)],
    "csbend": [Element("B", "CSBEND", L=0.8, ANGLE=0.04, K1=0.3, K2=2.0, N_KICKS=8, E1=0.02, E2=0.01,
                       HGAP=0.01, DX=1e-4, DZ=2e-4, FSE=1e-3)],
    "drifts": [Element("D", "DRIF", L=1.6), Element("E", "EDRIFT", L=0.7), Element("D1", "DRIF", L=0.4, ORDER=1)],
    "kquse": [Element("QS", "KQUSE", L=0.25, K1=1.4, K2=-35.0, FSE1=2e-3, FSE2=-5e-3, N_KICKS=8, DX=2e-5,
                      TILT=0.03)],
    "mult": [Element("M3", "MULT", L=0.2, KNL=400.0, ORDER=3, N_SLICES=4, TILT=0.02, DX=1e-4, DZ=1e-3),
             Element("M0", "MULT", L=0.0, KNL=-2e-4, ORDER=0, N_SLICES=1),
             Element("M5", "MULT", L=0.1, KNL=2e7, ORDER=5, N_SLICES=2)],
    "rotate_malign": [Element("R", "ROTATE", TILT=0.3), Element("MA", "MALIGN", DX=1e-4, DXP=2e-5, DZ=1e-3,
                                                                 DP=1e-3, ON_PASS=1),
                      Element("D", "DRIF", L=0.5), Element("R0", "ROTATE", TILT=-0.3, ON_PASS=0),
                      Element("ME", "MALIGN", DE=1e-4, DT=1e-12)],
    "radiation": [Element("Q", "KQUAD", L=0.3, K1=1.2, N_KICKS=8, SYNCH_RAD=1),
                  Element("B", "CSBEND", L=0.8, ANGLE=0.04, K1=0.3, N_KICKS=8, E1=0.02, E2=0.01, SYNCH_RAD=1),
                  Element("M", "MULT", L=0.1, KNL=30.0, ORDER=2, N_SLICES=2, SYNCH_RAD=1)],
}


def rows_for(seed, n=64, scale=3.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 10))
    r[:, 0] = rng.normal(0, 3e-3, n) * scale
    r[:, 1] = rng.normal(0, 3e-4, n) * scale
    r[:, 2] = rng.normal(0, 2e-3, n) * scale
    r[:, 3] = rng.normal(0, 2e-4, n) * scale
    r[:, 4] = rng.normal(0, 1e-3, n)
    r[:, 5] = rng.normal(0, 1e-3, n)
    r[:, 6] = np.arange(1, n + 1)
    r[:8, 0] = np.linspace(-0.03, 0.03, 8)  # some outside a 20 mm aperture
    return r


def lattice_for(name):
    if name == "fodo":
        return configs.fodo_ring()
    if name == "mba":
        return configs.mba_ring(configs.p_central_from_mev(3000.0))
    return Lattice([Element("AP", "MAXAMP", X_MAX=0.02, Y_MAX=0.02)] + CASES[name])


def cases():
    out = []
    for i, name in enumerate(CASES):
        out.append((name, lattice_for(name), rows_for(i), 2, configs.p_central_from_mev(2000.0)))
    out.append(("fodo", lattice_for("fodo"), configs.fodo_beam(128, seed=3), 20, configs.p_central_from_mev(1000.0)))
    out.append(("mba", lattice_for("mba"), configs.da_grid(9, 5, 0.03, 0.012), 3, configs.p_central_from_mev(3000.0)))
    return out


def main():
    for name, lat, rows, turns, Po in cases():
        out, n, _, _ = oracle.do_tracking(rows, lat, turns, Po)
        np.savez(os.path.join(HERE, "%s.npz" % name), rows=rows, out=out, n=n, turns=turns, p_central=Po)
        print(name, n, "of", len(rows))


if __name__ == "__main__":
    main()
