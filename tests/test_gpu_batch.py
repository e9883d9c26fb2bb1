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
    total_ref = 0
    for v in range(nvar):
        ref, nref, eff, _ = oracle.do_tracking(beams[v], configs.mba_variants(nvar, seed=7)[v], 3, Po)
        total_ref += nref
        got = batch.download(v)
        assert_rows_identical(got, ref, batch.n_alive(v), nref, "variant %d" % v)
    assert n == total_ref
