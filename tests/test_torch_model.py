"""The PyTorch restatement used to check the adjoint agrees with the oracle on
surviving particles (CPU only).

Bar: relative deviation below 2e-10 on the FODO ring and the element mix, 1e-8
on one turn of the MBA ring (floor: RMS of each coordinate).  The two evaluate
the bend's exact drift in different but equivalent forms (rho0*(f-1) + f*X
against (rho0+X)*cos(phi)/cos(psi) - rho0, with torch's own sin/cos/asin); the
cancellation in either form costs relative accuracy in x and x' at every bend
sub-step (12800 per MBA turn).  The oracle built with the C library's sin, cos
and asin differs from the oracle built with the library's own by the same
amounts (2.3e-9 in x after one MBA turn), so these bars are the floor of FP64
evaluation of this map, not a modelling difference."""
import math

import numpy as np
import pytest
import torch

from elegant_amd import configs
from elegant_amd.lattice import Element, Lattice
from oracle import oracle
from tests.torch_model import TorchRing


def _compare(lat, rows, turns, Po, tol):
    ref, n, _, _ = oracle.do_tracking(rows, lat, turns, Po)
    assert n == len(rows), "test particles must survive"
    ring = TorchRing(lat, Po)
    z0 = torch.tensor(rows[:, :6])
    out = ring.track(z0, torch.zeros(0, dtype=torch.float64), turns).numpy()
    got = {int(r[6]): r[:6] for r in ref[:n]}
    want = np.array([got[int(i)] for i in rows[:, 6]])
    floor = np.maximum(np.sqrt(np.mean(want ** 2, axis=0)), 1e-12)
    rel = np.abs(out - want) / np.maximum(np.abs(want), floor)
    assert rel.max() < tol, "max rel deviation %.3g (per coord %s)" % (rel.max(), rel.max(axis=0))


def test_fodo_two_turns():
    Po = configs.p_central_from_mev(1000.0)
    rows = configs.fodo_beam(40, seed=3)
    _compare(configs.fodo_ring(), rows, 2, Po, 2e-10)


def test_mba_one_turn_with_rf():
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    rows = configs.mba_beam(20, seed=4, scale=4.0)
    oracle.do_tracking(rows, lat, 1, Po)  # fixes the RF fiducial in the lattice structs
    _compare(lat, rows, 1, Po, 1e-8)


def test_edges_offsets_tilts_steering():
    Po = configs.p_central_from_mev(2000.0)
    lat = Lattice([
        Element("Q", "KQUAD", L=0.3, K1=1.1, N_KICKS=8, DX=1e-4, DY=-2e-4, DZ=3e-4, TILT=0.02, FSE=1e-3,
                HKICK=2e-5, VKICK=-1e-5),
        Element("D", "DRIF", L=0.5),
        Element("S", "KSEXT", L=0.2, K2=-30.0, K1=0.05, J1=0.03, N_KICKS=4, DX=-5e-5),
        Element("B", "CSBEND", L=0.9, ANGLE=0.06, K1=-0.3, K2=3.0, N_KICKS=8, E1=0.02, E2=0.03, HGAP=0.012,
                FINT=0.6),
        Element("E", "EDRIFT", L=0.4),
        Element("QS", "KQUSE", L=0.2, K1=0.4, K2=-12.0, FSE1=1e-3, FSE2=-2e-3, N_KICKS=8, DX=3e-5, DY=1e-5,
                TILT=0.01),
        Element("O", "KOCT", L=0.2, K3=2000.0, TILT=-0.1),
        Element("D2", "DRIF", L=0.3, ORDER=1),
    ])
    rows = configs.fodo_beam(30, seed=5)
    rows[:, :4] *= 2.0
    _compare(lat, rows, 3, Po, 2e-10)
