"""Seeded errors: the DLARAN stream of random_1_elegant and error_element's
binding/fractional rules."""
import numpy as np

from elegant_amd import errors
from elegant_amd.lattice import Element, Lattice


def test_dlaran_limbs_equal_48bit_lcg():
    """the 12-bit-limb arithmetic of DLARAN (src/drand_oag.c:160-186) is
    x <- 33952834046453 x mod 2^48, value x / 2^48 (exact in double)"""
    r = errors.Random1()
    x = 987654321  # odd already: (s/2)*2+1 = s
    for _ in range(2000):
        x = (x * 33952834046453) % (1 << 48)
        assert r() == x / float(1 << 48)


def test_multiplier_limbs():
    assert 494 * 2 ** 36 + 322 * 2 ** 24 + 2508 * 2 ** 12 + 2549 == 33952834046453


def test_even_seed_made_odd():
    a, b = errors.Random1(1000), errors.Random1(1001)
    assert [a() for _ in range(5)] == [b() for _ in range(5)]


def _ring():
    return Lattice([Element("Q", "KQUAD", L=0.3, K1=1.0, N_KICKS=8), Element("D", "DRIF", L=1.0),
                    Element("Q", "KQUAD", L=0.3, K1=1.0, N_KICKS=8), Element("S", "KSEXT", L=0.2, K2=10.0),
                    Element("Q", "KQUAD", L=0.3, K1=1.0, N_KICKS=8), Element("S", "KSEXT", L=0.2, K2=10.0)])


def test_bind_rules():
    lat = _ring()
    log = errors.ErrorSet([errors.ErrorItem("Q", "DX", 1e-4, bind=1)]).apply(lat)
    assert len({d for (_, _, d, _) in log}) == 1
    lat = _ring()
    log = errors.ErrorSet([errors.ErrorItem("Q", "DX", 1e-4, bind=0)]).apply(lat)
    assert len({d for (_, _, d, _) in log}) == 3
    lat = _ring()
    log = errors.ErrorSet([errors.ErrorItem("Q", "DX", 1e-4, bind=1, bind_number=2)]).apply(lat)
    ds = [d for (_, _, d, _) in log]
    assert ds[0] == ds[1] != ds[2]
    lat = _ring()
    log = errors.ErrorSet([errors.ErrorItem("Q", "DX", 1e-4, bind=-1)]).apply(lat)
    ds = [d for (_, _, d, _) in log]
    assert ds[1] == -ds[0] and ds[2] == ds[0]


def test_types_fractional_and_wildcards():
    lat = _ring()
    es = errors.ErrorSet([errors.ErrorItem("*", "FSE", 1e-3, type="uniform", bind=0, element_type="KQUAD"),
                          errors.ErrorItem("S", "K2", 0.01, type="plus_or_minus", fractional=True, bind=0)])
    log = es.apply(lat)
    fse = [v for (i, it, d, v) in log if it == "FSE"]
    assert len(fse) == 3 and all(abs(v) <= 1e-3 for v in fse)
    k2 = [(d, v) for (i, it, d, v) in log if it == "K2"]
    assert all(abs(d) == 0.01 and v == 10.0 * (1 + d) for d, v in k2)
    assert lat.structs()[3].k2 == k2[0][1]


def test_gaussian_cutoff():
    es = errors.ErrorSet([])
    v = np.array([es.gauss.lim(0.0, 1.0, 2.0) for _ in range(20000)])
    assert np.abs(v).max() <= 2.0
    assert abs(v.mean()) < 0.03 and 0.85 < v.std() < 0.9  # truncated at 2 sigma: std 0.8796
