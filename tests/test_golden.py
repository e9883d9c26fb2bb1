"""The oracle reproduces the frozen fixtures in tests/golden bit for bit."""
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.make_golden import cases
from tests.helpers import assert_rows_identical

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("case", [c[0] for c in cases()])
def test_oracle_matches_fixture(case):
    name, lat, rows, turns, Po = [c for c in cases() if c[0] == case][0]
    f = np.load(os.path.join(HERE, "%s.npz" % name))
    np.testing.assert_array_equal(f["rows"], rows)
    out, n, _, _ = oracle.do_tracking(rows, lat, turns, Po)
    assert_rows_identical(out, f["out"], n, int(f["n"]), name)
