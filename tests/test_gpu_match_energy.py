"""Momentum matching (do_match_energy) and the RF cavity's CHANGE_T and
CHANGE_P0 on the GPU against the oracle, bit for bit, reference momentum
included: plain and fused tracking over several passes with losses before and
between the cavities (so that the mean and row 0 are taken in the reference's
row order), matrix and kick methods, ENERGY's matching modes and fixed targets
on the same line, a beam lost before the cavity, the element and matching entry
points, and the batched path's refusal."""
import ctypes as C
import sys

import numpy as np
import pytest

from elegant_amd import _lib, configs, structs as S, tracking
from elegant_amd.lattice import Element, Lattice
from oracle import oracle
from tests.helpers import assert_rows_identical

pytestmark = pytest.mark.gpu

Po = configs.p_central_from_mev(3000.0)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rf(**kw):
    d = dict(L=0.3, VOLT=2e6, PHASE=170.0, FREQ=5e8, N_KICKS=0)
    d.update(kw)
    return d


def line(rf1, rf2=None, energy=None, aperture=0.02, elliptical=False):
    ell = dict(ELLIPTICAL=1, EXPONENT=4) if elliptical else {}
    els = [Element("AP", "MAXAMP", X_MAX=aperture, Y_MAX=0.75 * aperture, **ell),
           Element("Q1", "KQUAD", L=0.25, K1=0.8, N_KICKS=8),
           Element("D1", "DRIF", L=0.5),
           Element("RF1", "RFCA", **rf1),
           Element("B1", "CSBEND", L=0.8, ANGLE=0.03, K1=-0.3, N_KICKS=8)]
    if energy is not None:
        els.append(Element("E1", "ENERGY", **energy))
    els += [Element("D2", "DRIF", L=0.5)]
    if rf2 is not None:
        els.append(Element("RF2", "RFCA", **rf2))
    els += [Element("Q2", "KQUAD", L=0.25, K1=-0.7, N_KICKS=8), Element("D3", "DRIF", L=0.5)]
    return Lattice(els)


def beam(n, seed, scale=1.0, dp_mean=0.0):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, S.NPROP))
    rows[:, :4] = rng.normal(0, [1e-3, 1e-4, 8e-4, 8e-5], (n, 4)) * scale
    rows[:, 4] = rng.normal(0, 2e-3, n)
    rows[:, 5] = rng.normal(dp_mean, 1e-3, n)
    rows[:, 6] = np.arange(1, n + 1)
    return rows


def _both(rows, make, n_passes, mode):
    oracle.delete_phase_references()
    tracking.delete_phase_references()
    ref, nref, _, pref = oracle.do_tracking(rows, make(), n_passes, Po)
    b = tracking.Beam(rows)
    n, _, pgot = tracking.track(b, make(), n_passes, Po, mode=mode, turns_per_launch=2)
    got = b.download()
    b.free()
    return got, n, pgot, ref, nref, pref


CASES = {
    "change_t matrix": lambda: line(rf(CHANGE_T=1)),
    "change_t kicks": lambda: line(rf(CHANGE_T=1, N_KICKS=4), rf(CHANGE_T=1, N_KICKS=3, PHASE=175.0)),
    "change_p0 matrix": lambda: line(rf(CHANGE_P0=1, PHASE=95.0)),
    "change_p0 and t, fixed ENERGY": lambda: line(rf(CHANGE_P0=1, CHANGE_T=1, N_KICKS=4, PHASE=95.0),
                                                  rf(CHANGE_P0=1, PHASE=100.0),
                                                  energy=dict(CENTRAL_MOMENTUM=Po * 1.001)),
    "change_p0 elliptical": lambda: line(rf(CHANGE_P0=1, CHANGE_T=1, PHASE=95.0), elliptical=True),
    "lock phase kick": lambda: line(rf(LOCK_PHASE=1, N_KICKS=1, PHASE=100.0)),
    "linearize matrix": lambda: line(rf(LINEARIZE=1, PHASE=95.0)),
    "linearize change_t Q": lambda: line(rf(LINEARIZE=1, CHANGE_T=1, N_KICKS=1, PHASE=95.0),
                                         rf(LOCK_PHASE=1, Q=2000.0, PHASE=100.0)),
    "match beamline": lambda: line(rf(), energy=dict(MATCH_BEAMLINE=1)),
    "match particles": lambda: line(rf(CHANGE_T=1), energy=dict(MATCH_PARTICLES=1)),
}


@pytest.mark.parametrize("mode", ["plain", "fused"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_lines_vs_oracle(case, mode, ehlib):
    _say("%s %s" % (case, mode))
    rows = beam(2000, 11, scale=8.0, dp_mean=2e-3)  # losses before and between the cavities
    got, n, pgot, ref, nref, pref = _both(rows, CASES[case], 4, mode)
    assert 0 < nref < len(rows)
    assert_rows_identical(got, ref, n, nref, case)
    assert pgot == pref, (pgot, pref)
    if "p0" in case or "beamline" in case:
        assert pref != Po


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_beam_lost_before_the_cavity(mode, ehlib):
    """no particle reaches the matching cavity: only the ENERGY target moves
    the reference momentum, as in the reference"""
    _say("lost before cavity %s" % mode)
    rows = beam(300, 12)
    rows[:, 0] = 0.5
    make = lambda: line(rf(CHANGE_P0=1, CHANGE_T=1, PHASE=95.0), energy=dict(CENTRAL_MOMENTUM=Po * 1.002))
    got, n, pgot, ref, nref, pref = _both(rows, make, 3, mode)
    assert nref == 0
    assert_rows_identical(got, ref, n, nref, "all lost")
    assert pgot == pref == Po * 1.002


@pytest.mark.parametrize("change_beam", [0, 1])
def test_match_energy_entry_point_after_losses(change_beam, ehlib):
    _say("eh_do_match_energy %d" % change_beam)
    rows = beam(3000, 13, scale=8.0, dp_mean=2e-3)
    lat = line(rf(VOLT=0.0))
    ref, nref, _, _ = oracle.do_tracking(rows, lat, 2, Po)
    mref, pref = oracle.match_energy(ref[:nref], Po, change_beam)
    b = tracking.Beam(rows)
    n, _, _ = tracking.track(b, line(rf(VOLT=0.0)), 2, Po, mode="fused")
    assert n == nref < len(rows)
    pgot = b.match_energy(Po, change_beam)
    got = b.download()
    b.free()
    assert pgot == pref
    assert_rows_identical(got[:n], mref, n, nref, "matched")
    assert_rows_identical(got[n:], ref[nref:], 0, 0, "lost rows")


def test_cavity_entry_point_changes_p_central(ehlib):
    _say("eh_simple_rf_cavity CHANGE_P0/CHANGE_T")
    lat_r = line(rf(CHANGE_P0=1, CHANGE_T=1, N_KICKS=4, PHASE=95.0))
    lat_g = line(rf(CHANGE_P0=1, CHANGE_T=1, N_KICKS=4, PHASE=95.0))
    rows = beam(1500, 14, dp_mean=1e-3)
    rows[:, 4] += 7.3
    ref, nref, pref = oracle.element(rows, lat_r, 3, Po, 1.0, 1.3, with_pc=True)
    b = tracking.Beam(rows)
    pc = C.c_double(Po)
    arr, keep = lat_g.c_beamline()
    n = ehlib.eh_simple_rf_cavity(b.handle, len(rows), C.cast(arr[3].p_elem, C.POINTER(S.EH_RFCA)), C.byref(pc), 1.3)
    assert n >= 0, ehlib.eh_last_error()
    assert pc.value == pref != Po
    assert_rows_identical(b.download(), ref, n, nref, "cavity")
    b.free()


def test_batch_refuses_change_t_and_p0(ehlib):
    _say("batch refusals")
    rows = beam(100, 15)
    for kw in (dict(CHANGE_T=1), dict(CHANGE_P0=1), dict(LOCK_PHASE=1, N_KICKS=1), dict(LINEARIZE=1)):
        lat = line(rf(**kw))
        batch = tracking.Batch([rows, rows])
        with pytest.raises(_lib.EHError, match="CHANGE_T, CHANGE_P0, LINEARIZE and LOCK_PHASE"):
            batch.track([lat, lat.copy()], 1, Po)
        batch.free()
