"""RF cavity phase references and PMAXIMUM/FIRST fiducials on the GPU against the
oracle, bit for bit, with the same group table history on both sides: plain and
fused tracking, repeated calls with the table kept and deleted, the PMAXIMUM
selection in the reference's row order after losses."""
import ctypes as C
import sys

import numpy as np
import pytest

from elegant_amd import _lib, configs, structs as S, tracking
from oracle import oracle
from tests.helpers import assert_rows_identical
from tests.test_rfca_phase import Po, RF1, RF2, beam, rf_line

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def fresh_tables():
    oracle.delete_phase_references()
    tracking.delete_phase_references()
    yield
    oracle.delete_phase_references()
    tracking.delete_phase_references()


def _both(rows, lat_ref, lat_gpu, n_passes, mode):
    ref, nref, eff, _ = oracle.do_tracking(rows, lat_ref, n_passes, Po)
    b = tracking.Beam(rows)
    n, stats, _ = tracking.track(b, lat_gpu, n_passes, Po, mode=mode, turns_per_launch=2)
    got = b.download()
    b.free()
    return got, n, ref, nref


def _same_state(lat_ref, lat_gpu, groups):
    for i in (RF1, RF2):
        a, g = lat_ref.structs()[i], lat_gpu.structs()[i]
        assert (a.fiducial_seen, a.phase_fiducial) == (g.fiducial_seen, g.phase_fiducial), i
    for ref in groups:
        assert oracle.get_phase_reference(ref) == tracking.get_phase_reference(ref), ref


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_linked_cavities_over_calls(mode, ehlib):
    _say("linked cavities %s" % mode)
    lat_r, lat_g = rf_line(refs=(7, 7)), rf_line(refs=(7, 7))
    b = beam(2000, 1, scale=8.0)  # losses before and between the cavities
    got, n, ref, nref = _both(b, lat_r, lat_g, 3, mode)
    assert nref < len(b)
    assert_rows_identical(got, ref, n, nref, "call 1")
    _same_state(lat_r, lat_g, [7])
    # cavities forget their fiducials; the groups keep their phase
    lat_r.reset_state()
    lat_g.reset_state()
    b2 = beam(1500, 2, scale=6.0)
    got, n, ref, nref = _both(b2, lat_r, lat_g, 2, mode)
    assert_rows_identical(got, ref, n, nref, "call 2")
    _same_state(lat_r, lat_g, [7])
    assert lat_g.structs()[RF1].fiducial_seen == 0
    # a new step: phase references deleted, the first cavity fiducializes again
    oracle.delete_phase_references()
    tracking.delete_phase_references()
    got, n, ref, nref = _both(b2, lat_r, lat_g, 2, mode)
    assert_rows_identical(got, ref, n, nref, "call 3")
    _same_state(lat_r, lat_g, [7])


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_groups_and_zero_voltage(mode, ehlib):
    """two groups, a cavity without voltage leaving its group to the next"""
    _say("groups and zero voltage %s" % mode)
    lat_r = rf_line(refs=(5, 5), volts=(0.0, 2e6))
    lat_g = rf_line(refs=(5, 5), volts=(0.0, 2e6))
    b = beam(1000, 3, scale=4.0)
    got, n, ref, nref = _both(b, lat_r, lat_g, 3, mode)
    assert_rows_identical(got, ref, n, nref, "zero voltage first")
    _same_state(lat_r, lat_g, [5])
    lat_r, lat_g = rf_line(refs=(2, 3)), rf_line(refs=(2, 3))
    got, n, ref, nref = _both(b, lat_r, lat_g, 3, mode)
    assert_rows_identical(got, ref, n, nref, "two groups")
    _same_state(lat_r, lat_g, [2, 3])


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_pmaximum_after_losses(mode, ehlib):
    """the PMAXIMUM fiducial taken in the reference's row order after losses
    before the cavity (the largest momentum carried by a particle lost first)"""
    _say("pmaximum %s" % mode)
    b = beam(3000, 4, scale=8.0)
    k = int(np.argmax(b[:, 5]))
    b[k, 0] = 0.5  # lost at the start
    lat_r, lat_g = rf_line(modes=(S.FID_PMAX, S.FID_PMAX)), rf_line(modes=(S.FID_PMAX, S.FID_PMAX))
    got, n, ref, nref = _both(b, lat_r, lat_g, 2, mode)
    assert_rows_identical(got, ref, n, nref, "pmaximum")
    _same_state(lat_r, lat_g, [])


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_first_fails_like_the_reference(mode, ehlib):
    _say("first %s" % mode)
    lat = rf_line(modes=(S.FID_FIRST, 0))
    b = tracking.Beam(beam(200, 5))
    with pytest.raises(_lib.EHError, match="FID_MODE_FIRST"):
        tracking.track(b, lat, 1, Po, mode=mode)
    b.free()


def test_element_entry_point_group(ehlib):
    """eh_simple_rf_cavity alone: the second cavity of a group takes the first
    one's phase from the library's table"""
    _say("element entry point group")
    lat_r, lat_g = rf_line(refs=(21, 21)), rf_line(refs=(21, 21))
    rows = beam(800, 6)
    for idx, z in ((RF1, 1.05), (RF2, 2.65)):
        ref, nref = oracle.element(rows, lat_r, idx, Po, z - 0.3, z)
        b = tracking.Beam(rows)
        pc = C.c_double(Po)
        arr, keep = lat_g.c_beamline()
        n = ehlib.eh_simple_rf_cavity(b.handle, len(rows), C.cast(arr[idx].p_elem, C.POINTER(S.EH_RFCA)), C.byref(pc),
                                      z)
        assert n >= 0, ehlib.eh_last_error()
        assert_rows_identical(b.download(), ref, n, nref, "cavity %d" % idx)
        b.free()
    _same_state(lat_r, lat_g, [21])
    assert lat_g.structs()[RF2].fiducial_seen == 0


def test_batch_refuses_phase_reference_and_pmaximum(ehlib):
    _say("batch refusals")
    rows = beam(100, 7)
    for lat in (rf_line(refs=(7, 7)), rf_line(modes=(S.FID_PMAX, 0))):
        batch = tracking.Batch([rows, rows])
        with pytest.raises(_lib.EHError, match="PHASE_REFERENCE|TMEAN and LIGHT"):
            batch.track([lat, lat.copy()], 1, Po)
