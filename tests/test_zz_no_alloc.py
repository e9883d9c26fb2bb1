"""eh_do_tracking_noalloc (descriptors, field terms and counters in the static
scratch, no device allocation) against eh_do_tracking in fused mode: bit for bit,
on two calls in a row, the second with a smaller input; a line larger than the
scratch is refused before anything runs."""
import sys

import numpy as np
import pytest

from elegant_amd import _lib, configs, tracking
from elegant_amd.lattice import Element, Lattice
from tests.helpers import assert_rows_identical

pytestmark = pytest.mark.gpu


def _pair(rows, mk_lat, turns, Po, tpl):
    a = tracking.Beam(rows)
    na, _, pa = tracking.track(a, mk_lat(), turns, Po, mode="fused", turns_per_launch=tpl)
    b = tracking.Beam(rows)
    nb, _, pb = tracking.track_noalloc(b, mk_lat(), turns, Po, turns_per_launch=tpl)
    ra, rb = a.download(), b.download()
    a.free()
    b.free()
    return (ra, na, pa), (rb, nb, pb)


def test_noalloc_equals_fused_on_two_calls():
    print("noalloc: MBA grid, then a smaller FODO beam", flush=True)
    tracking.init(0)
    Po = configs.p_central_from_mev(3000.0)
    rows = configs.da_grid(60, 50, xmax=0.025, ymax=0.02)  # 3000 particles, losses at the aperture
    (ra, na, pa), (rb, nb, pb) = _pair(rows, lambda: configs.mba_ring(Po), 3, Po, 2)
    assert na == nb and pa == pb
    assert_rows_identical(rb, ra, nb, na, "MBA")
    assert na < len(rows)
    Po1 = configs.p_central_from_mev(1000.0)
    rows1 = configs.fodo_beam(500, seed=9)
    rows1[:20, 0] *= 30.0  # some lost early
    (ra, na, pa), (rb, nb, pb) = _pair(rows1, configs.fodo_ring, 5, Po1, 3)
    assert na == nb and pa == pb
    assert_rows_identical(rb, ra, nb, na, "FODO")


def test_noalloc_refuses_a_line_beyond_the_scratch():
    print("noalloc: refusal", flush=True)
    tracking.init(0)
    Po = configs.p_central_from_mev(1000.0)
    lat = Lattice([Element("D%d" % i, "DRIF", L=0.01) for i in range(5000)])
    rows = configs.fodo_beam(64, seed=2)
    beam = tracking.Beam(rows)
    with pytest.raises(_lib.EHError, match="static scratch"):
        tracking.track_noalloc(beam, lat, 1, Po)
    assert beam.n_alive() == 64  # nothing ran
    beam.free()
