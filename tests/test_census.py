"""The throughput census (elegant_amd/census.py) against instrumented counts.

The counting build of the checker (oracle/or_count.h, liboracle_count.so)
compiles the same restatement as C++ with counted doubles: every +, -, *, /
and sqrt it executes, and every sin, cos, tan, asin and log call, is tallied
(the generic transport-matrix loops by their non-trivial terms).  The
per-particle, per-pass cost of an element is the difference between 3 and 1
particles and between 2 passes and 1 (removing per-call and first-pass work
such as the RF fiducial).  Every census formula must reproduce it exactly, on
every element of the FODO and MBA rings and on variants of each element type.
"""
import math

import numpy as np
import pytest

from elegant_amd import census, configs
from elegant_amd import structs as S
from elegant_amd.lattice import Element, Lattice
from oracle import oracle

OPS = ("add", "sub", "mul", "div", "sqrt")
PO = configs.p_central_from_mev(3000.0)
ROWS = configs.mba_beam(8, seed=1, scale=1.0)  # every coordinate non-zero


def measured(el):
    tot = {}
    for n in (1, 3):
        for passes in (1, 2):
            c, _, alive, _ = oracle.count_ops(ROWS[:n], Lattice([el]), passes, PO)
            assert alive == n, "the count particles must survive %s" % el.params
            tot[(n, passes)] = (sum(c[k] for k in OPS), c["T"])
    per = [((tot[(3, 2)][i] - tot[(1, 2)][i]) - (tot[(3, 1)][i] - tot[(1, 1)][i])) / 2 for i in (0, 1)]
    return per[0], per[1]


def _check(el):
    got = measured(el)
    want = census.per_unit(el)
    assert got == (float(want[0]), float(want[1])), "%s %s: instrumented %s, census %s" % (
        S.TYPE_NAMES[el.type], el.params, got, want)


def _distinct(lat):
    seen, out = set(), []
    for el in lat.elements:
        key = (el.type, tuple(sorted(el.params.items())))
        if key not in seen and el.type not in (S.T_MARK, S.T_MAXAMP):
            seen.add(key)
            out.append(el)
    return out


@pytest.mark.parametrize("ring", ["fodo", "mba"])
def test_census_matches_instrumented_ring_elements(ring):
    lat = configs.fodo_ring() if ring == "fodo" else configs.mba_ring(PO)
    for el in _distinct(lat):
        _check(el)


@pytest.mark.parametrize("typ,k", [("KQUAD", {"K1": 0.8}), ("KSEXT", {"K2": 20.0}), ("KOCT", {"K3": 800.0}),
                                   ("KQUSE", {"K1": 0.8, "K2": 20.0}), ("KSEXT", {"K2": 20.0, "K1": 0.5})])
def test_census_multipole_variants(typ, k):
    for io in (2, 4, 6):
        for nk in (0, 4, 12):
            for extra in ({}, {"DX": 1e-5, "DZ": 0.01}, {"TILT": 0.01}, {"L": 0.0}):
                p = {"L": 0.25, "INTEGRATION_ORDER": io, "N_KICKS": nk}
                p.update(k)
                p.update(extra)
                _check(Element("E", typ, **p))
    _check(Element("E", "KQUAD", L=0.25, K1=0.8, N_KICKS=8, HKICK=1e-5, VKICK=2e-5))


def test_census_csbend_variants():
    ang = 2 * math.pi / 200
    for io in (2, 4):
        for nk in (1, 8):
            for extra in ({}, {"K1": -0.5}, {"K2": 3.0}, {"K1": 0.3, "EXPANSION_ORDER": 6},
                          {"EDGE1_EFFECTS": 1, "EDGE2_EFFECTS": 0, "E1": 0.01},
                          {"EDGE1_EFFECTS": 1, "EDGE2_EFFECTS": 1, "EDGE_ORDER": 2},
                          {"TILT": 0.01, "K1": 0.2}, {"DX": 1e-5, "FSE": 1e-3}):
                p = dict(L=0.8, ANGLE=ang, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0, INTEGRATION_ORDER=io, N_KICKS=nk)
                p.update(extra)
                _check(Element("B", "CSBEND", **p))


def test_census_other_elements():
    rf = dict(L=0.3, VOLT=2e6, PHASE=170.0, FREQ=6.07e8)
    for el in [Element("R", "RFCA", N_KICKS=0, **rf), Element("R", "RFCA", N_KICKS=5, **rf),
               Element("R", "RFCA", **dict(rf, L=0.0)), Element("R", "RFCA", **dict(rf, VOLT=0.0)),
               Element("D", "DRIF", L=0.3), Element("D", "DRIF", L=1.0), Element("D", "DRIF", L=2.0),
               Element("D", "DRIF", L=0.3, ORDER=1), Element("D", "DRIF", L=1.0, ORDER=1),
               Element("E", "EDRIFT", L=0.3), Element("R", "ROTATE", TILT=0.2),
               Element("E", "ENERGY", CENTRAL_ENERGY=5000.0)]:
        _check(el)
    for p in ({}, {"DX": 1e-5}, {"DZ": 0.01}, {"DT": 1e-12}, {"DP": 1e-4}, {"DE": 1e-4}, {"DP": 1e-4, "DE": 1e-4},
              {"DX": 1e-5, "DZ": 0.01, "DT": 1e-12}):
        _check(Element("M", "MALIGN", **p))
    for L in (0.1, 0.0):
        for o in (0, 1, 2, 3, 5, 8):
            for ns in (1, 3):
                _check(Element("M", "MULT", L=L, KNL=10.0 ** (-5 + 2 * o), ORDER=o, N_SLICES=ns))
    _check(Element("M", "MULT", L=0.1, KNL=300.0, ORDER=3, N_SLICES=2, DX=1e-5, TILT=0.1, FACTOR=2.0))


def test_drift_matrix_loop_executes_more_than_the_census():
    """the reference's generic order-2 loop executes 324 operations per particle
    for a drift; the census counts its 11 non-trivial ones"""
    el = Element("D", "DRIF", L=0.3)
    c1, _, _, _ = oracle.count_ops(ROWS[:1], Lattice([el]), 1, PO)
    c3, _, _, _ = oracle.count_ops(ROWS[:3], Lattice([el]), 1, PO)
    executed = sum(c3["all_" + k] - c1["all_" + k] for k in OPS) / 2
    counted = sum(c3[k] - c1[k] for k in OPS) / 2
    assert executed == 324 and counted == 11


def test_ring_census_per_particle_turn():
    """whole rings: the census total equals the instrumented per-particle cost of
    a turn (second turn minus first, 3 particles minus 1)"""
    for lat in (configs.fodo_ring(), configs.mba_ring(PO)):
        tot = {}
        for n in (1, 3):
            for passes in (1, 2):
                c, _, alive, _ = oracle.count_ops(ROWS[:n], lat.copy(), passes, PO)
                assert alive == n
                tot[(n, passes)] = np.array([sum(c[k] for k in OPS), c["T"]], dtype=float)
        per = ((tot[(3, 2)] - tot[(1, 2)]) - (tot[(3, 1)] - tot[(1, 1)])) / 2
        cen = census.lattice_census(lat)
        assert per[0] == cen["flops"] and per[1] == cen["T"]
