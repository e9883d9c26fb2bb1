"""RF cavity phase references and fiducial modes in the oracle
(src/simple_rfca.c:62-146 and :412-466, src/phase_reference.c:28-143), pinned by
equivalences that do not restate the formulas:

* a cavity in a PHASE_REFERENCE group set by an earlier cavity tracks exactly as
  if it had that cavity's fiducial as its own, and takes none itself;
* the group keeps its phase across calls until the phase references are
  deleted, is set only by a cavity with voltage reached by particles;
* PMAXIMUM's fiducial is the TMEAN fiducial of a beam holding only the particle
  it selects (the first whose momentum exceeds every earlier one, row 0 included);
  it fails when row 0 has the largest momentum, and FIRST always fails, as the
  reference's loops do in serial elegant."""
import numpy as np
import pytest

from elegant_amd import configs, structs as S
from elegant_amd.lattice import Element, Lattice
from oracle import oracle

Po = configs.p_central_from_mev(3000.0)
RF1, RF2 = 3, 6


def rf_line(refs=(0, 0), modes=(0, 0), volts=(2e6, 2e6), phases=(170.0, 175.0)):
    rf = configs.mba_rf(Po)
    els = [Element("AP", "MAXAMP", X_MAX=0.02, Y_MAX=0.015),
           Element("Q1", "KQUAD", L=0.25, K1=0.8, N_KICKS=8),
           Element("D1", "DRIF", L=0.5),
           Element("RF1", "RFCA", **dict(rf, VOLT=volts[0], PHASE=phases[0], PHASE_REFERENCE=refs[0])),
           Element("B1", "CSBEND", L=0.8, ANGLE=0.03, K1=-0.3, N_KICKS=8),
           Element("D2", "DRIF", L=0.5),
           Element("RF2", "RFCA", **dict(rf, VOLT=volts[1], PHASE=phases[1], PHASE_REFERENCE=refs[1])),
           Element("Q2", "KQUAD", L=0.25, K1=-0.7, N_KICKS=8),
           Element("D3", "DRIF", L=0.5)]
    lat = Lattice(els)
    lat.set_param(RF1, "fiducial_mode", modes[0])
    lat.set_param(RF2, "fiducial_mode", modes[1])
    return lat


def beam(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, S.NPROP))
    rows[:, :4] = rng.normal(0, [1e-3, 1e-4, 8e-4, 8e-5], (n, 4)) * scale
    rows[:, 4] = rng.normal(0, 2e-3, n)
    rows[:, 5] = rng.normal(0, 1e-3, n)
    rows[:, 6] = np.arange(1, n + 1)
    return rows


@pytest.fixture(autouse=True)
def fresh_table():
    oracle.delete_phase_references()
    yield
    oracle.delete_phase_references()


def test_linked_cavity_tracks_with_the_first_cavitys_fiducial():
    b = beam(400, 1)
    lat = rf_line(refs=(7, 7))
    got, n, _, _ = oracle.do_tracking(b, lat, 3, Po)
    s1, s2 = lat.structs()[RF1], lat.structs()[RF2]
    assert s1.fiducial_seen == 1 and s2.fiducial_seen == 0  # the second takes none
    assert oracle.get_phase_reference(7) == (1, s1.phase_fiducial)
    # the same line without a group, the second cavity given the first one's fiducial
    oracle.delete_phase_references()
    ref_lat = rf_line()
    r1 = ref_lat.structs()[RF2]
    r1.fiducial_seen, r1.phase_fiducial = 1, s1.phase_fiducial
    ref, nref, _, _ = oracle.do_tracking(b, ref_lat, 3, Po)
    assert n == nref and np.array_equal(got, ref)
    assert ref_lat.structs()[RF1].phase_fiducial == s1.phase_fiducial  # the group leaves the first alone
    # without the group the second cavity's own fiducial differs
    oracle.delete_phase_references()
    own = rf_line()
    oracle.do_tracking(b, own, 1, Po)
    assert own.structs()[RF2].phase_fiducial != s1.phase_fiducial


def test_group_persists_across_calls_until_deleted():
    lat = rf_line(refs=(11, 11))
    oracle.do_tracking(beam(300, 2), lat, 1, Po)
    phi = lat.structs()[RF1].phase_fiducial
    lat.reset_state()  # the cavities forget their fiducials, the table does not
    other = beam(300, 3, scale=2.0)
    oracle.do_tracking(other, lat, 1, Po)
    assert lat.structs()[RF1].fiducial_seen == 0  # it tracked with the group's phase
    assert oracle.get_phase_reference(11) == (1, phi)
    oracle.delete_phase_references()
    assert oracle.get_phase_reference(11)[0] == 2  # known, not set
    oracle.do_tracking(other, lat, 1, Po)
    assert lat.structs()[RF1].fiducial_seen == 1 and lat.structs()[RF1].phase_fiducial != phi


def test_group_set_only_with_particles_and_voltage():
    lost = beam(50, 4)
    lost[:, 0] = 0.5  # outside the MAXAMP at the start
    lat = rf_line(refs=(4242, 4242))
    _, n, _, _ = oracle.do_tracking(lost, lat, 2, Po)
    assert n == 0 and oracle.get_phase_reference(4242)[0] == 3 and lat.structs()[RF1].fiducial_seen == 0
    # a cavity without voltage leaves the group to the next one
    lat = rf_line(refs=(9, 9), volts=(0.0, 2e6))
    oracle.do_tracking(beam(200, 5), lat, 1, Po)
    s1, s2 = lat.structs()[RF1], lat.structs()[RF2]
    assert s1.fiducial_seen == 0 and s2.fiducial_seen == 1
    assert oracle.get_phase_reference(9) == (1, s2.phase_fiducial)


def _pmax_expected(rows):
    dp = rows[:, 5]
    best, ibest = dp[0], -1
    for i in range(len(dp)):
        if best < dp[i]:
            best, ibest = dp[i], i
    return ibest


def test_pmaximum_is_the_tmean_of_the_selected_particle():
    b = beam(300, 6)
    b[:, 0] *= 0.1  # nothing lost before the first cavity: row order is upload order
    b[5, 5] = b[9, 5] = b[:, 5].max() + 1e-4  # a tie: the earlier row wins
    assert _pmax_expected(b) == 5
    lat = rf_line(modes=(S.FID_PMAX, S.FID_TMEAN))
    oracle.do_tracking(b, lat, 1, Po)
    phi_p = lat.structs()[RF1].phase_fiducial
    for i, same in ((5, True), (9, False)):
        oracle.delete_phase_references()
        one = rf_line()
        oracle.do_tracking(b[i:i + 1], one, 1, Po)
        assert (one.structs()[RF1].phase_fiducial == phi_p) == same


def test_pmaximum_and_first_failures():
    b = beam(100, 7)
    b[0, 5] = b[:, 5].max() + 1e-3  # row 0 has the largest momentum
    with pytest.raises(RuntimeError, match="No available particle for RF cavity"):
        oracle.do_tracking(b, rf_line(modes=(S.FID_PMAX, 0)), 1, Po)
    with pytest.raises(RuntimeError, match="FID_MODE_FIRST"):
        oracle.do_tracking(beam(100, 8), rf_line(modes=(S.FID_FIRST, 0)), 1, Po)
    # a cavity that already has a fiducial, or a T_REFERENCE, takes none
    lat = rf_line(modes=(S.FID_FIRST, 0))
    lat.set_param(RF1, "tReference", 1e-9)
    oracle.do_tracking(beam(100, 8), lat, 1, Po)
    assert lat.structs()[RF1].fiducial_seen == 1
