"""SDDS phase-space files: layout and the reference's t/p conversions."""
import math

import numpy as np

from elegant_amd import configs, sdds
from elegant_amd import structs as S


def test_round_trip(tmp_path):
    Po = configs.p_central_from_mev(1000.0)
    rows = configs.fodo_beam(500, seed=2)
    rows[:, 4] += 123.456  # path length well away from zero
    f = tmp_path / "beam.sdds"
    sdds.write_phase_space(f, rows, Po)
    back, pc = sdds.read_phase_space(f)
    assert pc == Po
    np.testing.assert_array_equal(back[:, :4], rows[:, :4])
    np.testing.assert_array_equal(back[:, 6], rows[:, 6])
    assert np.max(np.abs(back[:, 4] - rows[:, 4]) / np.abs(rows[:, 4])) < 4e-16
    assert np.max(np.abs(back[:, 5] - rows[:, 5])) < 4e-16
    assert (back[:, 7] == -1).all() and (back[:, 9] == 1).all()


def test_known_row(tmp_path):
    """s = 1 m, delta = 0 at pCentral = 1000: t = 1 / (c * 1000/sqrt(1000^2 + 1)), p = 1000"""
    rows = np.zeros((1, 10))
    rows[0, 4] = 1.0
    rows[0, 6] = 7
    f = tmp_path / "one.sdds"
    sdds.write_phase_space(f, rows, 1000.0, step=3)
    text = f.read_text().splitlines()
    assert text[0] == "SDDS1"
    assert [ln.split(",")[0] for ln in text if ln.startswith("&column")] == [
        "&column name=x", "&column name=xp", "&column name=y", "&column name=yp", "&column name=t",
        "&column name=p", "&column name=particleID"]
    vals = text[-1].split()
    t = float(vals[4])
    assert t == 1.0 / (S.C_MKS * 1000.0 / math.sqrt(1000.0 ** 2 + 1))
    assert float(vals[5]) == 1000.0 and int(vals[6]) == 7
    data = [ln for ln in text if not ln.startswith(("&", "!", "SDDS"))]
    assert data[:5] == ["3", "1000.0", "0.0", "1", "1"]


def test_multi_page_and_column_order(tmp_path):
    f = tmp_path / "pages.sdds"
    f.write_text("\n".join([
        "SDDS1",
        "&parameter name=pCentral, type=double &end",
        "&column name=particleID, type=long &end",
        "&column name=p, type=double &end",
        "&column name=t, type=double &end",
        "&column name=x, type=double &end",
        "&column name=xp, type=double &end",
        "&column name=y, type=double &end",
        "&column name=yp, type=double &end",
        "&data mode=ascii &end",
        "100.0", "1", "5 101.0 0.0 1e-3 2e-4 -1e-3 0.0",
        "100.0", "2", "6 99.0 1e-9 0 0 0 0", "9 100.0 0 0 0 0 0",
    ]) + "\n")
    rows, pc = sdds.read_phase_space(f)
    assert pc == 100.0
    assert list(rows[:, 6]) == [5, 6, 9]
    assert rows[0, 0] == 1e-3 and rows[0, 5] == (101.0 - 100.0) / 100.0
    assert rows[1, 4] == 1e-9 * S.C_MKS * 99.0 / math.sqrt(99.0 * 99.0 + 1)
