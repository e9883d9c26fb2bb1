"""Momentum matching (do_match_energy, src/do_tracking.c:3376-3549) and the RF
cavity's CHANGE_P0 and CHANGE_T (src/simple_rfca.c:488-498, :586-588, :864-865,
:977-982) in the oracle, pinned by what the mathematics fixes rather than by
restating the formulas:

* MATCH_BEAMLINE (change_beam 0): the new reference momentum is the beam's mean
  momentum to within 2 ulp of the exact mean of the rounded products (a Kahan
  sum; the plain sum misses by many ulp on the same beam), every particle keeps
  its momentum, the mean delta becomes 0, and a beam within 1e-14 of the
  reference is left alone, bit for bit;
* MATCH_PARTICLES (change_beam 1): every momentum moves by the same amount, the
  mean momentum becomes the reference, each time of flight is kept;
* CHANGE_P0: the cavity followed by the matching, the reference momentum then
  the survivors' mean momentum;
* Q: the fill factor sqrt(1 - exp(-dt/tau)) on the energy gain, none before the
  fiducial time, the whole gain many fill times after it;
* LOCK_PHASE: the cavity's phase is fixed on the bunch's mean arrival time, so
  the result does not depend on the fiducial; LINEARIZE: the energy gain is
  linear in the arrival time and departs from the full sine at second order;
* CHANGE_T: the offset is the whole number of RF periods in row 0's arrival
  time (from 0.9 periods on), the cavity's kick is that of the unshifted beam
  (the sine is periodic), and the time coordinate drops by offset x c x beta
  (once per kick in the kick method)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from elegant_amd import configs, structs as S
from elegant_amd.lattice import Element, Lattice
from oracle import oracle

Po = configs.p_central_from_mev(1000.0)


def beam(n, seed, mean_dp=3e-3, sigma_dp=1e-3, s0=0.0):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, S.NPROP))
    rows[:, :4] = rng.normal(0, [1e-3, 1e-4, 8e-4, 8e-5], (n, 4))
    rows[:, 4] = s0 + rng.normal(0, 2e-3, n)
    rows[:, 5] = rng.normal(mean_dp, sigma_dp, n)
    rows[:, 6] = np.arange(1, n + 1)
    return rows


def momenta(rows, pc):
    return pc * (1 + rows[:, 5])


def exact_mean(values):
    return float(sum((Fraction(float(v)) for v in values), Fraction(0)) / len(values))


def ulp(x):
    return math.ulp(x)


def test_match_beamline_mean_is_kahan_accurate():
    b = beam(20001, 1)
    rows, pc = oracle.match_energy(b, Po, 0)
    P = momenta(b, Po)  # the rounded products the sum adds
    exact = exact_mean(P)
    assert abs(pc - exact) <= 2 * ulp(exact)
    # the check has teeth: a plain running sum misses by far more here
    naive = 0.0
    for v in P:
        naive += float(v)
    assert abs(naive / len(P) - exact) > 4 * ulp(exact)


def test_match_beamline_keeps_momenta_and_centres_delta():
    b = beam(5000, 2)
    rows, pc = oracle.match_energy(b, Po, 0)
    before, after = momenta(b, Po), momenta(rows, pc)
    assert np.max(np.abs(after - before) / before) < 4e-16
    assert abs(rows[:, 5].mean()) < 1e-15
    assert np.array_equal(rows[:, :5], b[:, :5])


def test_match_beamline_threshold_and_empty_beam():
    # a symmetric beam: the mean is the reference momentum to within rounding
    half = beam(1000, 3, mean_dp=0.0)
    b = np.concatenate([half, half])
    b[1000:, 5] = -half[:, 5]
    rows, pc = oracle.match_energy(b, Po, 0)
    assert pc == Po and np.array_equal(rows, b)
    # 5e-15 relative is below the threshold, 5e-14 above it
    for off, moved in ((5e-15, False), (5e-14, True)):
        c = b.copy()
        c[:, 5] += off
        rows, pc = oracle.match_energy(c, Po, 0)
        assert (pc != Po) == moved
        assert np.array_equal(rows, c) != moved
    rows, pc = oracle.match_energy(b[:0], Po, 0)
    assert pc == Po


def test_match_particles_shifts_every_momentum_and_keeps_time():
    b = beam(4000, 4)
    b[:, 4] = 12.5 + b[:, 4]
    rows, pc = oracle.match_energy(b, Po, 1)
    assert pc == Po
    before, after = momenta(b, Po), momenta(rows, Po)
    shift = Po - exact_mean(before)
    assert np.max(np.abs((after - before) - shift)) < 8 * ulp(Po)
    assert abs(exact_mean(after) - Po) < 4 * ulp(Po)

    def tof(r, P):
        return r[:, 4] / (P / np.sqrt(P * P + 1))
    assert np.max(np.abs(tof(rows, after) / tof(b, before) - 1)) < 1e-15
    assert np.array_equal(rows[:, :4], b[:, :4])


def rf_lattice(**kw):
    d = dict(L=0.3, VOLT=3e6, PHASE=160.0, FREQ=5e8, N_KICKS=0)
    d.update(kw)
    return Lattice([Element("RF", "RFCA", **d)])


def _set_fiducial(lat, phase_fiducial):
    st = lat.structs()[0]
    st.fiducial_seen, st.phase_fiducial = 1, phase_fiducial


def test_change_p0_is_the_cavity_then_the_matching():
    b = beam(3000, 5, mean_dp=0.0)
    b[::97, 5] = -0.9995  # a few slow particles, decelerated below rest at this phase
    outs = []
    for cp0 in (0, 1):
        lat = rf_lattice(CHANGE_P0=cp0, PHASE=141.0)
        _set_fiducial(lat, 0.0)
        outs.append(oracle.element(b, lat, 0, Po, 0.0, 0.3))
    (r0, n0), (r1, n1) = outs
    assert n0 == n1 < len(b)
    m, pc = oracle.match_energy(r0[:n0], Po, 0)
    st1 = rf_lattice(CHANGE_P0=1, PHASE=141.0)
    _set_fiducial(st1, 0.0)
    _, _, pc1 = oracle.element(b, st1, 0, Po, 0.0, 0.3, with_pc=True)
    assert pc1 == pc and pc != Po
    assert abs(pc - exact_mean(momenta(r0[:n0], Po))) <= 2 * ulp(pc)
    assert np.array_equal(r1[:n1], m)
    assert abs(r1[:n1, 5].mean()) < 1e-15
    # the lost rows are the cavity's own, untouched by the matching
    assert np.array_equal(r1[n1:], r0[n0:])


@pytest.mark.parametrize("nk", [0, 4])
def test_change_t_removes_whole_periods(nk):
    To = 1 / 5e8
    b = beam(500, 6, mean_dp=0.0, s0=0.0)
    beta0 = Po / math.sqrt(Po * Po + 1)
    k_row0 = 37
    b[:, 4] += k_row0 * To * 2.99792458e8 * beta0
    b[0, 4] += 0.3 * To * 2.99792458e8 * beta0  # row 0 decides: 37.3 periods rounds to 37
    b[1:, 4] += 5 * To * 2.99792458e8 * beta0   # the others sit 5 periods later
    res = {}
    for ct in (0, 1):
        lat = rf_lattice(CHANGE_T=ct, N_KICKS=nk)
        _set_fiducial(lat, 0.25)
        res[ct] = oracle.element(b, lat, 0, Po, 0.0, 0.3)
    (r0, n0), (r1, n1) = res[0], res[1]
    assert n0 == n1 == len(b)
    P = Po * (1 + r1[:, 5])
    beta = P / np.sqrt(P * P + 1)
    toff = k_row0 * To
    # the kick is the unshifted beam's: sin is periodic in the phase
    assert np.max(np.abs(r1[:, 5] - r0[:, 5])) < 1e-11
    assert np.max(np.abs(r1[:, :4] - r0[:, :4])) < 1e-12
    # s drops by the offset at the exit speed; the kick method takes it off at
    # every kick (each kick's shifted arrival time sets the time coordinate in
    # add_to_particle_energy, src/simple_rfca.c:586-660, :988-1006)
    times = max(nk, 1)
    assert np.max(np.abs((r0[:, 4] - r1[:, 4]) - times * toff * 2.99792458e8 * beta)) < 1e-9
    # the check has teeth: the particles' own periods (42) would shift s by 5 more
    assert np.min(np.abs((r0[1:, 4] - r1[1:, 4]) - times * 42 * To * 2.99792458e8 * beta[1:])) > 1.0


def test_change_t_below_09_periods_does_nothing():
    To = 1 / 5e8
    b = beam(300, 7, mean_dp=0.0)
    beta0 = Po / math.sqrt(Po * Po + 1)
    b[:, 4] = 0.5 * To * 2.99792458e8 * beta0 + b[:, 4] * 1e-3  # row 0 arrives after half a period
    out = []
    for ct in (0, 1):
        lat = rf_lattice(CHANGE_T=ct)
        _set_fiducial(lat, 0.25)
        out.append(oracle.element(b, lat, 0, Po, 0.0, 0.3))
    assert np.array_equal(out[0][0], out[1][0])


def test_energy_matching_modes_in_do_tracking():
    """ENERGY's MATCH_BEAMLINE and MATCH_PARTICLES go through the matching;
    both at once is refused"""
    b = beam(400, 8)
    els = [Element("D1", "DRIF", L=0.5), Element("E", "ENERGY", MATCH_BEAMLINE=1), Element("D2", "DRIF", L=0.5)]
    rows, n, _, pc = oracle.do_tracking(b, Lattice(els), 1, Po)
    assert abs(rows[:n, 5].mean()) < 1e-15 and pc != Po
    els[1] = Element("E", "ENERGY", MATCH_PARTICLES=1)
    rows, n, _, pc = oracle.do_tracking(b, Lattice(els), 1, Po)
    assert pc == Po and abs(exact_mean(momenta(rows[:n], Po)) - Po) < 4 * ulp(Po)
    els[1] = Element("E", "ENERGY", MATCH_PARTICLES=1, MATCH_BEAMLINE=1)
    with pytest.raises(RuntimeError, match="match_beamline AND match_particles"):
        oracle.do_tracking(b, Lattice(els), 1, Po)


def test_cavity_fill_time():
    """Q: the energy gain of a particle arriving dt after the fiducial time is
    the unloaded one times sqrt(1 - exp(-dt/tau)), tau = Q/omega
    (src/simple_rfca.c:399-401, :872): a particle arriving before the fiducial
    gains nothing, one arriving many fill times later the full gain"""
    freq = 5e8
    omega = 2 * math.pi * freq
    beta0 = Po / math.sqrt(Po * Po + 1)
    b = beam(6, 9, mean_dp=0.0, sigma_dp=0.0)
    b[:, :4] = 0.0
    c = 2.99792458e8
    # arrival times at the cavity centre: before the fiducial, just after, long after
    b[:, 4] = np.array([-3e-9, 0.0, 2e-10, 1e-9, 4e-6, 8e-6]) * c * beta0
    out = {}
    for q in (0.0, 300.0):
        lat = rf_lattice(Q=q, PHASE=90.0, L=0.0, N_KICKS=0)
        _set_fiducial(lat, 0.0)  # t0 = 0
        out[q] = oracle.element(b, lat, 0, Po, 0.0, 0.0)[0]
    g0 = out[0.0][:, 5] - b[:, 5]
    gq = out[300.0][:, 5] - b[:, 5]
    tau = 300.0 / omega
    # arriving at or before the fiducial time: the cavity is empty (the momentum
    # only goes through gamma and back, a rounding of 1e-16)
    assert abs(gq[0]) < 1e-15 and abs(gq[1]) < 1e-15 and abs(g0[0]) > 1e-4
    for k in (2, 3):
        dt = b[k, 4] / (c * beta0)
        assert abs(gq[k] / g0[k] - math.sqrt(1 - math.exp(-dt / tau))) < 1e-6
    assert abs(gq[5] / g0[5] - 1) < 1e-12
    with pytest.raises(RuntimeError, match="change_t!=0 && Q!=0"):
        oracle.element(b, rf_lattice(Q=300.0, CHANGE_T=1), 0, Po, 0.0, 0.3)


def _rf_out(b, fid, **kw):
    lat = rf_lattice(**kw)
    _set_fiducial(lat, fid)
    return oracle.element(b, lat, 0, Po, 0.0, 0.3)


def test_lock_phase_ignores_the_fiducial():
    b = beam(400, 10, mean_dp=0.0)
    b[:, 4] += 3.7
    for nk in (0, 1):
        r1, n1 = _rf_out(b, 0.0, LOCK_PHASE=1, N_KICKS=nk, PHASE=120.0)
        r2, n2 = _rf_out(b, 1.234, LOCK_PHASE=1, N_KICKS=nk, PHASE=120.0)
        assert n1 == n2 and np.array_equal(r1, r2)
        r3, _ = _rf_out(b, 1.234, N_KICKS=nk, PHASE=120.0)  # unlocked: the fiducial matters
        assert not np.array_equal(r1, r3)
    with pytest.raises(RuntimeError, match="n_kicks=1"):
        _rf_out(b, 0.0, LOCK_PHASE=1, N_KICKS=4)


def test_linearize_is_linear_in_arrival_time():
    """particles on a line in arrival time: the linearized gain has zero second
    differences; it matches the full sine at the mean arrival time to first
    order (the difference grows as the square of the spread)"""
    b = np.zeros((41, S.NPROP))
    b[:, 6] = np.arange(1, 42)
    d = []
    for spread in (1e-3, 2e-3):
        b[:, 4] = 5.0 + spread * np.linspace(-1, 1, 41)
        lin, _ = _rf_out(b, 0.0, LINEARIZE=1, PHASE=100.0, L=0.0)
        full, _ = _rf_out(b, 0.0, PHASE=100.0, L=0.0)
        g = lin[:, 5]
        assert np.max(np.abs(np.diff(g, 2))) < 1e-15
        d.append(np.max(np.abs(lin[:, 5] - full[:, 5])))
    assert 3.5 < d[1] / d[0] < 4.5, d


@pytest.mark.parametrize("etype,n,field", [("KQUAD", 1, "K1"), ("KSEXT", 2, "K2"), ("KOCT", 3, "K3")])
def test_bore_field_sets_the_strength(etype, n, field):
    """BORE and B (src/multipole.c:739-804): the strength is the field gradient
    n! B / BORE^n over the magnetic rigidity P_central m c / e, so the element
    tracks exactly as one given that K<n>"""
    e_mks, me_mks, c = 1.602176634e-19, 9.1093837015e-31, 2.99792458e8
    bore, B = 0.015, 0.3
    k = math.factorial(n) * B / bore ** n if n > 1 else B / bore
    k = k * (e_mks / (me_mks * c * Po))
    b = beam(300, 12, mean_dp=0.0)
    kw = dict(L=0.25, N_KICKS=8, FSE=1e-3)
    r1, n1 = oracle.element(b, Lattice([Element("M", etype, BORE=bore, B=B, **kw)]), 0, Po, 0.0, 0.25)
    r2, n2 = oracle.element(b, Lattice([Element("M", etype, **{field: k}, **kw)]), 0, Po, 0.0, 0.25)
    assert n1 == n2
    if n == 1:
        assert np.array_equal(r1, r2)
    else:  # n! B / BORE^n is formed in another order than here
        np.testing.assert_allclose(r1[:, :4], r2[:, :4], rtol=1e-12, atol=1e-18)
    r3, _ = oracle.element(b, Lattice([Element("M", etype, BORE=bore, B=2 * B, **kw)]), 0, Po, 0.0, 0.25)
    assert not np.array_equal(r1, r3)
