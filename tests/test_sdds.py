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


def test_binary_matches_ascii(tmp_path):
    """elegant's default binary flavour carries the same values as the ASCII one"""
    Po = configs.p_central_from_mev(1000.0)
    rows = configs.fodo_beam(300, seed=4)
    rows[:, 4] += 55.5
    sdds.write_phase_space(tmp_path / "a.sdds", rows, Po)
    sdds.write_phase_space(tmp_path / "b.sdds", rows, Po, binary=True)
    a, pa = sdds.read_phase_space(tmp_path / "a.sdds")
    b, pb = sdds.read_phase_space(tmp_path / "b.sdds")
    assert pa == pb == Po
    np.testing.assert_array_equal(a, b)
    raw = (tmp_path / "b.sdds").read_bytes()
    head, _, body = raw.partition(b"&data mode=binary, &end\n")
    assert head.startswith(b"SDDS1\n!# little-endian\n")
    assert len(body) == 4 + 4 + 8 + 8 + 4 + 4 + 300 * (6 * 8 + 8)  # rows, 5 parameters, 300 rows


def test_binary_reader_column_major_big_endian_pages(tmp_path):
    """a file packed by hand: big-endian, column-major rows, a string parameter,
    an extra float column, two pages"""
    import struct
    hdr = ["SDDS1", "!# big-endian", "&parameter name=Tag, type=string &end",
           "&parameter name=pCentral, type=double &end"]
    cols = [("x", "double", "d"), ("xp", "double", "d"), ("w", "float", "f"), ("y", "double", "d"),
            ("yp", "double", "d"), ("t", "double", "d"), ("p", "double", "d"), ("particleID", "long", "i")]
    hdr += ["&column name=%s, type=%s &end" % (c, t) for c, t, _ in cols]
    hdr.append("&data mode=binary, column_major_order=1, &end")
    body = b""
    pages = []
    for page, n in enumerate((3, 2)):
        vals = {c: [0.001 * (page + 1) * (k + 1) * (j + 2) for k in range(n)] for j, (c, _, _) in enumerate(cols)}
        vals["p"] = [2000.0 + k for k in range(n)]
        vals["particleID"] = [10 * page + k + 1 for k in range(n)]
        body += struct.pack(">i", n) + struct.pack(">i", 3) + b"abc" + struct.pack(">d", 1957.0)
        for c, _, f in cols:
            body += struct.pack(">%d%s" % (n, f), *vals[c])
        pages.append(vals)
    (tmp_path / "h.sdds").write_bytes(("\n".join(hdr) + "\n").encode() + body)
    rows, pc = sdds.read_phase_space(tmp_path / "h.sdds")
    assert pc == 1957.0 and len(rows) == 5
    x = pages[0]["x"] + pages[1]["x"]
    np.testing.assert_array_equal(rows[:, 0], x)
    np.testing.assert_array_equal(rows[:, 6], [1, 2, 3, 11, 12])
    p = np.array(pages[0]["p"] + pages[1]["p"])
    np.testing.assert_allclose(rows[:, 5], (p - 1957.0) / 1957.0, rtol=0, atol=1e-16)
