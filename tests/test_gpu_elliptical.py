"""Elliptical MAXAMP on the GPU against the oracle, bit for bit, lost rows
included: the elimit_amplitudes entry point (src/limit_amplitudes.c:498), the
elliptical in-element check of the kick multipoles and CSBEND
(src/multipole.c:1784-1792) and the sweep after each element
(src/do_tracking.c:2506-2517), in plain, fused and batched tracking."""
import ctypes as C
import sys

import numpy as np
import pytest

from elegant_amd import _lib, configs, structs as S, tracking
from elegant_amd.lattice import Element, Lattice
from oracle import oracle
from tests.helpers import assert_rows_identical

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _beam(n, seed, xs=0.012, ys=0.006):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, S.NPROP))
    rows[:, 0] = rng.uniform(-xs, xs, n)
    rows[:, 1] = rng.normal(0, 4e-4, n)
    rows[:, 2] = rng.uniform(-ys, ys, n)
    rows[:, 3] = rng.normal(0, 2e-4, n)
    rows[:, 5] = rng.normal(0, 1e-3, n)
    rows[:, 6] = np.arange(1, n + 1)
    return rows


@pytest.mark.parametrize("xe,ye,extrap", [(2, 0, 1), (2, 0, 0), (4, 0, 0), (2, 6, 1), (8, 2, 0)])
def test_elimit_entry_point(xe, ye, extrap, ehlib):
    _say("elimit entry point xe=%d ye=%d extrap=%d" % (xe, ye, extrap))
    a, b = 0.01, 0.005
    rows = _beam(4000, seed=xe * 10 + ye)
    rows[:8, 0] = [a, -a, np.nextafter(a, 1), 0, np.inf, np.nan, 0, 0]
    rows[:8, 2] = [0, 0, 0, -b, 0, 0, np.nan, -np.inf]
    ref = np.array(rows, copy=True)
    lib = oracle.load()
    nref = oracle._check(lib, lib.or_elimit_amplitudes_flat(ref.ctypes.data, a, b, len(ref), 2.5, 3000.0, extrap,
                                                            xe, ye))
    assert 0 < nref < len(rows)
    beam = tracking.Beam(rows)
    n = ehlib.eh_elimit_amplitudes(beam.handle, a, b, len(rows), 2.5, 3000.0, extrap, 0, xe, ye)
    assert n >= 0, ehlib.eh_last_error()
    assert_rows_identical(beam.download(), ref, n, nref, "elimit xe=%d ye=%d" % (xe, ye))


def test_elimit_entry_point_refusals(ehlib):
    beam = tracking.Beam(_beam(10, seed=1))
    # EH_ERR_ARG (-2) for exponents elimit_amplitudes rejects, EH_ERR_UNSUPPORTED
    # (-1) for an open side
    assert ehlib.eh_elimit_amplitudes(beam.handle, 0.01, 0.01, 10, 0.0, 1.0, 0, 0, 3, 0) == -2
    assert ehlib.eh_elimit_amplitudes(beam.handle, 0.01, 0.01, 10, 0.0, 1.0, 0, 0, 2, 7) == -2
    assert ehlib.eh_elimit_amplitudes(beam.handle, 0.01, 0.01, 10, 0.0, 1.0, 0, 1, 2, 0) == -1


def elliptical_ring(exponent=2, y_exponent=0):
    """kick multipoles, a gradient bend and both drift kinds behind an elliptical
    MAXAMP"""
    return Lattice([
        Element("AP", "MAXAMP", X_MAX=0.015, Y_MAX=0.008, ELLIPTICAL=1, EXPONENT=exponent, YEXPONENT=y_exponent),
        Element("QF", "KQUAD", L=0.3, K1=1.4, N_KICKS=8, DX=1e-4),
        Element("D1", "DRIF", L=0.8),
        Element("B1", "CSBEND", L=1.0, ANGLE=0.1, K1=-0.3, N_KICKS=8, E1=0.05, E2=0.05),
        Element("D2", "EDRIFT", L=0.4),
        Element("SX", "KSEXT", L=0.2, K2=30.0, N_KICKS=4, TILT=0.02),
        Element("QD", "KQUAD", L=0.3, K1=-1.4, N_KICKS=8),
        Element("D3", "DRIF", L=0.8),
    ])


@pytest.mark.parametrize("exps", [(2, 0), (4, 0), (2, 4)])
@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_elliptical_ring(mode, exps, ehlib):
    _say("elliptical ring %s %s" % (mode, exps))
    from tests.test_gpu_parity import _track_both
    lat = elliptical_ring(*exps)
    rows = _beam(3000, seed=5, xs=0.016, ys=0.009)
    Po = configs.p_central_from_mev(1500.0)
    got, n, ref, nref, stats, eff = _track_both(lat, rows, 4, Po, mode, tpl=2)
    assert 0 < nref < len(rows)
    assert_rows_identical(got, ref, n, nref, "elliptical ring %s %s" % (mode, exps))
    assert stats["particle_element_turns"] == eff


@pytest.mark.parametrize("which", ["KQUAD", "CSBEND"])
def test_element_entry_points_with_elliptical_maxamp(which, ehlib):
    """the element routines given an elliptical MAXAMP: in-element losses by the
    (super-)ellipse, no sweep"""
    _say("element entry point %s elliptical" % which)
    Po = configs.p_central_from_mev(1500.0)
    el = elliptical_ring().elements[1 if which == "KQUAD" else 3]
    lat1 = Lattice([el])
    rows = _beam(3000, seed=8, xs=0.016, ys=0.009)
    for amp in (S.EH_MAXAMP(0.015, 0.008, 1, 2, 0, 0), S.EH_MAXAMP(0.015, 0.008, 1, 4, 2, 0)):
        ref, nref = oracle.element(rows, lat1, 0, Po, 1.0, 2.0, maxamp=amp)
        beam = tracking.Beam(rows)
        arr, keep = lat1.c_beamline()
        if which == "KQUAD":
            n = ehlib.eh_multipole_tracking2(beam.handle, len(rows), arr, 0.0, Po, 1.0, C.byref(amp), -1)
        else:
            n = ehlib.eh_track_through_csbend(beam.handle, len(rows), C.cast(arr[0].p_elem, C.POINTER(S.EH_CSBEND)),
                                              0.0, Po, 1.0, C.byref(amp), -1)
        assert n >= 0, ehlib.eh_last_error()
        assert 0 < nref < len(rows)
        assert_rows_identical(beam.download(), ref, n, nref, "%s exponent %d" % (which, amp.exponent))


def test_batch_elliptical_variants():
    _say("batch elliptical")
    Po = configs.p_central_from_mev(1500.0)
    lats = [elliptical_ring(*e) for e in ((2, 0), (4, 0), (2, 4))]
    beams = [_beam(500 + 70 * v, seed=60 + v, xs=0.016, ys=0.009) for v in range(3)]
    batch = tracking.Batch(beams)
    n, _ = batch.track(lats, 3, Po)
    total = 0
    for v, e in enumerate(((2, 0), (4, 0), (2, 4))):
        ref, nref, _, _ = oracle.do_tracking(beams[v], elliptical_ring(*e), 3, Po)
        total += nref
        assert_rows_identical(batch.download(v), ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total


def test_adjoint_refuses_elliptical_maxamp():
    from elegant_amd import adjoint
    _say("adjoint refuses elliptical")
    lat = elliptical_ring()
    Po = configs.p_central_from_mev(1500.0)
    rows = _beam(64, seed=2, xs=0.002, ys=0.001)
    tr = adjoint.AdjointTracker(lat, adjoint.ParamSpec.strengths(lat), Po, 2)
    with pytest.raises(_lib.EHError, match="elliptical"):
        tr.forward(rows)
