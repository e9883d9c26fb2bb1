"""eh_math.h (shared by the checker and the GPU kernels) against the C library.

Pins the transcendental routines: a few ulp from the correctly rounded values
that numpy (glibc) returns, exact special values, and identities.
"""
import math

import numpy as np
import pytest


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    spacing = np.spacing(np.abs(b))
    return np.where(a == b, 0.0, np.abs(a - b) / spacing)


@pytest.fixture(scope="module")
def lib(oracle_mod):
    return oracle_mod.load()


def _vec(f, xs):
    return np.array([f(float(v)) for v in xs])


def test_sin_cos_accuracy(lib):
    rng = np.random.default_rng(1)
    for lo, hi in ((-4, 4), (-1e-6, 1e-6), (-1e4, 1e4), (-1e8, 1e8)):
        x = rng.uniform(lo, hi, 20000)
        assert ulps(_vec(lib.or_sin, x), np.sin(x)).max() <= 3, (lo, hi)
        assert ulps(_vec(lib.or_cos, x), np.cos(x)).max() <= 3, (lo, hi)


def test_asin_log_accuracy(lib):
    rng = np.random.default_rng(2)
    for lo, hi in ((-1, 1), (-0.01, 0.01), (0.45, 0.55), (0.99, 1.0)):
        x = rng.uniform(lo, hi, 20000)
        assert ulps(_vec(lib.or_asin, x), np.arcsin(x)).max() <= 3, (lo, hi)
    for x in (np.exp(rng.uniform(-700, 700, 20000)), 1 + rng.uniform(-1e-3, 1e-3, 20000)):
        assert ulps(_vec(lib.or_log, x), np.log(x)).max() <= 3


def test_special_values(lib):
    assert lib.or_sin(0.0) == 0.0 and lib.or_cos(0.0) == 1.0
    assert lib.or_asin(1.0) == math.pi / 2 and lib.or_asin(-1.0) == -math.pi / 2
    assert lib.or_asin(0.0) == 0.0
    assert lib.or_log(1.0) == 0.0
    assert lib.or_log(0.0) == -math.inf
    assert math.isnan(lib.or_log(-1.0))
    assert math.isnan(lib.or_asin(1.0000000000000002))
    assert math.isnan(lib.or_sin(math.inf)) and math.isnan(lib.or_cos(math.nan))
    # quadrant boundaries: sin(k pi/2) for small k
    for k in range(-8, 9):
        assert abs(lib.or_sin(k * math.pi / 2) - math.sin(k * math.pi / 2)) < 1e-15


def test_identities(lib):
    rng = np.random.default_rng(3)
    x = rng.uniform(-50, 50, 5000)
    s = _vec(lib.or_sin, x)
    c = _vec(lib.or_cos, x)
    assert np.max(np.abs(s * s + c * c - 1)) < 4e-16
    y = rng.uniform(-1.5, 1.5, 5000)
    assert np.max(np.abs(_vec(lib.or_asin, _vec(lib.or_sin, y)) - y)) < 4e-15  # asin is ill-conditioned near 1


def test_generated_constants_match_python():
    """the generated split constants reproduce pi/2 and ln 2"""
    import importlib.util
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(here, "elegant_amd", "csrc", "eh_math_consts.h")
    vals = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "#define":
            vals[parts[1]] = float.fromhex(parts[2])
    assert vals["EH_PIO2_HI"] == math.pi / 2
    assert vals["EH_PIO2_1"] + vals["EH_PIO2_2"] == pytest.approx(math.pi / 2, rel=1e-15)
    assert vals["EH_LN2_HI"] + vals["EH_LN2_LO"] == math.log(2)
    # 24-bit pieces: k * piece exact for |k| < 2^29
    for i in range(1, 5):
        m, e = math.frexp(vals["EH_PIO2_%d" % i])
        assert (m * 2 ** 24) == int(m * 2 ** 24)
    assert vals["EH_SIN_1"] == -1.0 / 6 and vals["EH_COS_1"] == -0.5
    assert importlib.util.find_spec("numpy") is not None


def test_exp_accuracy_and_special_values(lib):
    """eh_exp against the correctly rounded value (60-digit decimal exp rounded
    once) on the fringe map's small arguments and across the range; exact
    special values; exp(log x) = x to a few ulp"""
    from decimal import Decimal, localcontext

    def exact(v):
        with localcontext() as ctx:
            ctx.prec = 60
            return float(Decimal(v).exp())

    rng = np.random.default_rng(7)
    worst = 0.0
    for lo, hi in ((-1e-4, 1e-4), (-0.35, 0.35), (-5, 5), (-700, 700)):
        x = rng.uniform(lo, hi, 4000)
        got = _vec(lib.or_exp, x)
        ref = np.array([exact(float(v)) for v in x])
        worst = max(worst, ulps(got, ref).max())
        assert ulps(got, np.exp(x)).max() <= 1, (lo, hi)
    assert worst <= 1, worst
    # a planted Taylor series cut one term short misses the bar on |r| near ln2/2
    r = np.linspace(0.3, 0.346, 200)
    short = 1 + r + sum(r ** n / math.factorial(n) for n in range(2, 10))
    assert ulps(short, [exact(float(v)) for v in r]).max() > 1
    assert lib.or_exp(0.0) == 1.0 and lib.or_exp(-0.0) == 1.0
    assert lib.or_exp(1000.0) == math.inf and lib.or_exp(-1000.0) == 0.0
    assert math.isnan(lib.or_exp(math.nan))
    assert lib.or_exp(709.78) == pytest.approx(math.exp(709.78), rel=4e-16)
    assert lib.or_exp(-740.0) == math.exp(-740.0)  # subnormal result
    x = np.exp(rng.uniform(-30, 30, 2000))
    assert ulps(_vec(lib.or_exp, _vec(lib.or_log, x)), x).max() <= 64  # log's error times |log x|
