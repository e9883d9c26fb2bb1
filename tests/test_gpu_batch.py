"""Batched lattice variants: every variant's result equals that variant tracked
alone by the oracle, bit for bit, whatever the launch shape."""
import sys

import numpy as np
import pytest

from elegant_amd import configs, tracking
from oracle import oracle
from tests.helpers import assert_rows_identical

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tune", [dict(block=256, span=None, graph=True), dict(block=64, span=700, graph=False),
                                  dict(block=128, span=2083 * 2, graph=True)])
def test_batch_variants_vs_oracle(tune):
    print("batch", tune, flush=True)
    sys.stdout.flush()
    Po = configs.p_central_from_mev(3000.0)
    nvar = 5
    lats = configs.mba_variants(nvar, seed=7)
    counts = [300, 0, 1, 129, 64]
    beams = []
    for v, c in enumerate(counts):
        b = configs.mba_beam(c, seed=v, scale=400.0) if c else np.zeros((0, 10))
        beams.append(b)
    batch = tracking.Batch(beams)
    n, st = batch.track(lats, 3, Po, block=tune["block"], span_elements=tune["span"], use_graph=tune["graph"])
    print("survivors per variant:", [batch.n_alive(v) for v in range(nvar)], "total", n, flush=True)
    total_ref = 0
    for v in range(nvar):
        ref, nref, eff, _ = oracle.do_tracking(beams[v], configs.mba_variants(nvar, seed=7)[v], 3, Po)
        total_ref += nref
        got = batch.download(v)
        assert_rows_identical(got, ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total_ref


def test_batch_repeated_calls_reuse_tables():
    """a second track call on unchanged beamlines reuses the device tables and
    continues exactly like a second do_tracking call of the reference"""
    _say = print
    _say("batch repeated calls", flush=True)
    Po = configs.p_central_from_mev(3000.0)
    nvar = 3
    lats = configs.mba_variants(nvar, seed=11)
    refl = configs.mba_variants(nvar, seed=11)
    counts = [200, 70, 130]
    beams = [configs.mba_beam(c, seed=10 + v, scale=400.0) for v, c in enumerate(counts)]
    batch = tracking.Batch(beams)
    batch.track(lats, 2, Po)
    batch.track(lats, 1, Po)
    n, _ = batch.track(lats, 1, Po, use_graph=False)
    total = 0
    for v in range(nvar):
        r1, n1, _, _ = oracle.do_tracking(beams[v], refl[v], 2, Po)
        r2, n2, _, _ = oracle.do_tracking(r1[:n1], refl[v], 1, Po)
        r3, n3, _, _ = oracle.do_tracking(r2[:n2], refl[v], 1, Po)
        expect = np.concatenate([r3, r2[n2:], r1[n1:]])
        assert_rows_identical(batch.download(v), expect, batch.n_alive(v), n3, "variant %d" % v)
        total += n3
    assert n == total


def test_batch_more_elements_on_pass():
    """variants of a ring with MULT, ROTATE and MALIGN: the batched kernel applies
    ON_PASS elements on the absolute pass, as each variant tracked alone does"""
    from elegant_amd.lattice import Element, Lattice
    print("batch more elements", flush=True)
    sys.stdout.flush()
    Po = configs.p_central_from_mev(1500.0)

    def ring(v):
        return Lattice([
            Element("AP", "MAXAMP", X_MAX=0.03, Y_MAX=0.02),
            Element("QF", "KQUAD", L=0.3, K1=1.1 + 0.01 * v, N_KICKS=8),
            Element("M3", "MULT", L=0.1, KNL=300.0 * (1 + v), ORDER=3, N_SLICES=2, TILT=0.01 * v),
            Element("D1", "DRIF", L=1.0),
            Element("R1", "ROTATE", TILT=0.2, ON_PASS=1),
            Element("QD", "KQUAD", L=0.3, K1=-1.1, N_KICKS=8),
            Element("R2", "ROTATE", TILT=-0.2, ON_PASS=1),
            Element("MA", "MALIGN", DXP=2e-5 * v, DX=1e-5, ON_PASS=2),
            Element("D2", "EDRIFT", L=1.0),
        ])

    nvar = 3
    rng = np.random.default_rng(4)
    beams = []
    for v in range(nvar):
        b = np.zeros((400 + 50 * v, 10))
        b[:, :4] = rng.normal(0, [6e-3, 6e-4, 4e-3, 4e-4], (len(b), 4))
        b[:, 5] = rng.normal(0, 1e-3, len(b))
        b[:, 6] = np.arange(1, len(b) + 1)
        beams.append(b)
    batch = tracking.Batch(beams)
    n, _ = batch.track([ring(v) for v in range(nvar)], 4, Po)
    total = 0
    for v in range(nvar):
        ref, nref, _, _ = oracle.do_tracking(beams[v], ring(v), 4, Po)
        total += nref
        assert_rows_identical(batch.download(v), ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total


def test_batch_kick_multipole_entry_losses():
    """inherited loss positions at kick-multipole entries, settled per variant"""
    from tests.test_gpu_parity import entry_loss_beam, entry_loss_ring
    print("batch entry losses", flush=True)
    sys.stdout.flush()
    Po = configs.p_central_from_mev(2000.0)
    lats = [entry_loss_ring(v % 2 == 1) for v in range(3)]
    beams = [entry_loss_beam(300 + 40 * v, seed=40 + v) for v in range(3)]
    batch = tracking.Batch(beams)
    n, _ = batch.track(lats, 3, Po)
    total = 0
    for v in range(3):
        ref, nref, _, _ = oracle.do_tracking(beams[v], entry_loss_ring(v % 2 == 1), 3, Po)
        total += nref
        assert_rows_identical(batch.download(v), ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total


def test_batch_radiation():
    """radiating variants (different SYNCH_RAD rings) through the batched path"""
    from tests.test_gpu_parity import _beam_rows, radiation_ring
    print("batch radiation", flush=True)
    sys.stdout.flush()
    Po = configs.p_central_from_mev(1000.0)
    beams = [_beam_rows(300 + 100 * v, seed=50 + v) for v in range(2)]
    batch = tracking.Batch(beams)
    n, _ = batch.track([radiation_ring(), radiation_ring()], 3, Po)
    total = 0
    for v in range(2):
        ref, nref, _, _ = oracle.do_tracking(beams[v], radiation_ring(), 3, Po)
        total += nref
        assert_rows_identical(batch.download(v), ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total


def _rf_variants(nvar, volt, shared=False):
    from tests.test_batch_host import _rf_line
    out = []
    for v in range(nvar):
        lat, Po = _rf_line(volt=volt, shared=shared, two=shared)
        lat.set_param(1, "K1", 0.8 + 0.02 * v)
        out.append(lat)
    return out, Po


def _small_beams(counts, seed):
    rng = np.random.default_rng(seed)
    beams = []
    for c in counts:
        b = np.zeros((c, 10))
        b[:, :4] = rng.normal(0, [5e-3, 5e-4, 4e-3, 4e-4], (c, 4))
        b[:, 5] = rng.normal(0, 2e-3, c)
        b[:, 6] = np.arange(1, c + 1)
        beams.append(b)
    return beams


@pytest.mark.parametrize("volt", [0.0, 2e6])
def test_batch_two_calls_and_zero_pass_call(volt):
    """a call with n_passes=0, then 2 passes, then 1 more, on unchanged
    beamlines (device tables kept between calls unless an RF fiducial is still
    pending), each variant equal to the oracle's consecutive do_tracking calls"""
    print("batch two calls volt", volt, flush=True)
    nvar = 3
    lats, Po = _rf_variants(nvar, volt)
    refl, _ = _rf_variants(nvar, volt)
    counts = [150, 64, 97]
    beams = _small_beams(counts, 21)
    batch = tracking.Batch(beams)
    n0, _ = batch.track(lats, 0, Po)
    assert n0 == sum(counts)
    batch.track(lats, 2, Po)
    n, _ = batch.track(lats, 1, Po)
    total = 0
    for v in range(nvar):
        r1, n1, _, _ = oracle.do_tracking(beams[v], refl[v], 2, Po)
        r2, n2, _, _ = oracle.do_tracking(r1[:n1], refl[v], 1, Po)
        expect = np.concatenate([r2, r1[n1:]])
        assert_rows_identical(batch.download(v), expect, batch.n_alive(v), n2, "variant %d" % v)
        total += n2
    assert n == total


def test_batch_shared_cavity_struct():
    """one RF cavity struct at two positions of each line (and a second cavity):
    every position gets the fiducial the first one took"""
    print("batch shared cavity", flush=True)
    nvar = 3
    lats, Po = _rf_variants(nvar, 2e6, shared=True)
    refl, _ = _rf_variants(nvar, 2e6, shared=True)
    counts = [130, 1, 70]
    beams = _small_beams(counts, 22)
    batch = tracking.Batch(beams)
    n, _ = batch.track(lats, 3, Po, span_elements=4)
    total = 0
    for v in range(nvar):
        ref, nref, _, _ = oracle.do_tracking(beams[v], refl[v], 3, Po)
        total += nref
        assert_rows_identical(batch.download(v), ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total
