"""CSBEND edge methods in the oracle beyond the first-order kick: EDGE_EFFECTS 3
(applySimpleDipoleEdgeKick, src/csbend.c:5664-5690) and the edge kick limits
(src/csbend.c:932-937, :1130-1131, :1311-1312), pinned by properties that do
not restate the formulas:

* EDGE_EFFECTS 3 kicks the canonical momenta, so a bend with it is a
  symplectic map of (x, q_x, y, q_y) at fixed delta, while the first-order
  method, which kicks the slopes, is not (a planted variant that kicks slopes
  fails the check);
* both methods agree to second order in the amplitude: their difference grows
  as its cube;
* a kick limit caps the x' kick of each edge at the limit (times rho0/rho_actual
  with KICK_LIMIT_SCALING), seen on a bend so short that its body is the
  identity to 1e-12.

EDGE_EFFECTS 2 (K. Hwang's fringe map, dipoleFringeKHwang, src/csbend.c:4336-4410):

* without a gap it is the canonical kick of EDGE_EFFECTS 3 without the FINT
  correction, to rounding;
* with a gap g it departs from the first-order method with the FINT correction
  (psi) by O(g^2): the vertical term k2 y (1 + sin^2 e) g / ((1 + delta) rho^2
  cos^3 e) is the first-order expansion of tan(e - psi/(1 + delta)); a planted
  error in it would leave an O(g) difference;
* its EDGE_ORDER 2 terms are of second order in the amplitude.  Their
  coefficients are restated from the reference (parity unpinned beyond that
  scaling), with SDDS's ipow2/ipow3 (not vendored) as x*x and (x*x)*x.

The host build of the device maps (tests/batch_host) runs lines with each method
against the oracle bit for bit."""
import numpy as np
import pytest

from elegant_amd import configs, structs as S
from elegant_amd.lattice import Element, Lattice
from oracle import oracle
from tests.test_oracle import rows_of

Po = configs.p_central_from_mev(1500.0)


def bend(**kw):
    p = dict(L=0.8, ANGLE=0.1, K1=0.3, N_KICKS=16, E1=0.07, E2=0.05, HGAP=0.012, FINT=0.5)
    p.update(kw)
    return Lattice([Element("B", "CSBEND", **p)])


def run(lat, coords, lib=None):
    r, n = oracle.element(rows_of(coords), lat, 0, Po, 0.0, 0.8, lib=lib)
    assert n == len(r)
    order = np.argsort(r[:, 6])
    return r[order]


def canon(z):
    x, xp, y, yp, dp = z[0], z[1], z[2], z[3], z[5]
    f = (1 + dp) / np.sqrt(1 + xp * xp + yp * yp)
    return np.array([x, xp * f, y, yp * f])


def slopes(q, dp):
    f = 1 / np.sqrt((1 + dp) ** 2 - q[1] ** 2 - q[3] ** 2)
    return q[1] * f, q[3] * f


def jacobian(lat, z0, h=1e-7):
    """d(x, qx, y, qy)_out / d(x, qx, y, qy)_in at fixed delta, central differences"""
    q0 = canon(z0)
    J = np.zeros((4, 4))
    for k in range(4):
        cols = []
        for sgn in (1, -1):
            q = q0.copy()
            q[k] += sgn * h
            xp, yp = slopes(q, z0[5])
            cols.append([q[0], xp, q[2], yp, 0.0, z0[5]])
        out = run(lat, cols)
        J[:, k] = (canon(out[0]) - canon(out[1])) / (2 * h)
    return J


def symplectic_error(J):
    S4 = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=float)
    return np.max(np.abs(J.T @ S4 @ J - S4))


def test_method3_is_symplectic_and_method1_is_not():
    # steep slopes, where kicking x' instead of q_x departs from a symplectic map
    # (differences of h = 1e-7 leave ~5e-8 of noise; the body alone shows 7e-8)
    z0 = np.array([4e-3, 0.05, -3e-3, 0.03, 0.0, 2e-3])
    e3 = symplectic_error(jacobian(bend(EDGE1_EFFECTS=3, EDGE2_EFFECTS=3), z0))
    e1 = symplectic_error(jacobian(bend(EDGE1_EFFECTS=1, EDGE2_EFFECTS=1), z0))
    assert e3 < 5e-7, e3
    assert e1 > 1e-5, e1  # the first-order method kicks the slopes


def test_methods_agree_to_second_order():
    d = []
    for a in (2e-3, 4e-3, 8e-3):
        z = [[a, 0.5 * a * 0.1, -0.7 * a, 0.3 * a * 0.1, 0.0, 1e-3]]
        r1 = run(bend(EDGE1_EFFECTS=1, EDGE2_EFFECTS=1), z)[0]
        r3 = run(bend(EDGE1_EFFECTS=3, EDGE2_EFFECTS=3), z)[0]
        d.append(np.max(np.abs(r1[:4] - r3[:4])))
    # the difference is of third order in the amplitude (ratio 8 per doubling)
    assert 6.0 < d[1] / d[0] < 10.0 and 6.0 < d[2] / d[1] < 10.0, d


@pytest.mark.parametrize("method", [1, 3])
def test_kick_limit_caps_each_edge(method):
    """a bend of 1e-9 m and 1e-10 rad (rho = 10 m): the body changes x' by less
    than 1e-12, each edge kicks x' by tan(e) x / rho, capped at the limit"""
    e, lim = 0.3, 2e-4
    kw = dict(L=1e-9, ANGLE=1e-10, K1=0.0, N_KICKS=1, E1=e, E2=e, HGAP=0.0, FINT=0.0, EDGE1_EFFECTS=method,
              EDGE2_EFFECTS=method)
    xs = [1e-4, 1e-2, -1e-2]  # kicks of 3.1e-6 (free) and 3.1e-4 (capped) per edge at rho = 10 m
    z = [[x, 0, 0, 0, 0, 0] for x in xs]
    free = run(bend(**kw), z)
    capped = run(bend(EDGE1_KICK_LIMIT=lim, EDGE2_KICK_LIMIT=lim, **kw), z)
    np.testing.assert_array_equal(free[0], capped[0])  # below the limit: untouched

    def kicked(r):  # method 1 kicks the slope, method 3 the canonical momentum
        return r[1] if method == 1 else canon(r)[1]

    kick = np.tan(e) / 10.0
    for k in (1, 2):
        assert abs(kicked(free[k]) - 2 * kick * xs[k]) < 1e-9 * abs(xs[k]) + 1e-12
        assert abs(kicked(capped[k]) - 2 * lim * np.sign(xs[k])) < 1e-12
    # KICK_LIMIT_SCALING multiplies the limit by rho0 / rho_actual = 1 + FSE (the
    # body's field error then kicks x' by L FSE / rho0 = 2.5e-11)
    sc = run(bend(EDGE1_KICK_LIMIT=lim, EDGE2_KICK_LIMIT=lim, KICK_LIMIT_SCALING=1, FSE=0.25, **kw), z)
    assert abs(kicked(sc[1]) - 2 * lim * 1.25) < 5e-11


def test_method2_without_gap_is_the_canonical_kick():
    """HGAP = 0: Hwang's map keeps only tan(e) x / rho on q_x and -tan(e) y / rho
    on q_y, which is EDGE_EFFECTS 3 with psi = 0"""
    kw = dict(HGAP=0.0, FINT=0.5)
    z = [[2e-3, 1e-4, -1.5e-3, 2e-4, 0.0, 1e-3], [-4e-3, -3e-4, 2e-3, -1e-4, 0.0, -2e-3]]
    r2 = run(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2, **kw), z)
    r3 = run(bend(EDGE1_EFFECTS=3, EDGE2_EFFECTS=3, **kw), z)
    np.testing.assert_allclose(r2[:, :4], r3[:, :4], rtol=0, atol=1e-15)
    # the check has teeth: method 1, which kicks the slopes, is further away
    r1 = run(bend(EDGE1_EFFECTS=1, EDGE2_EFFECTS=1, **kw), z)
    assert np.max(np.abs(r1[:, :4] - r3[:, :4])) > 1e-12


def _gap_ratios(lib=None):
    z = [[0.0, 0.0, 1e-4, 0.0, 0.0, 1e-3]]
    d = []
    for g in (1e-3, 2e-3, 4e-3):
        kw = dict(HGAP=g, FINT=0.5, E1=0.2, E2=0.15)
        r1 = run(bend(EDGE1_EFFECTS=1, EDGE2_EFFECTS=1, **kw), z)[0]
        r2 = run(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2, **kw), z, lib)[0]
        r0 = run(bend(EDGE1_EFFECTS=1, EDGE2_EFFECTS=1, **dict(kw, FINT=0.0)), z)[0]
        assert abs(r1[3] - r0[3]) > 50 * abs(r2[3] - r1[3])  # the FINT term is first order
        d.append(abs(r2[3] - r1[3]))
    return d[1] / d[0], d[2] / d[1]


def test_method2_departs_from_method1_at_second_order_in_the_gap(tmp_path):
    """the vertical fringe focusing of method 2 equals method 1's FINT
    correction to first order in the gap: the difference at y alone grows as g^2"""
    q1, q2 = _gap_ratios()
    assert 3.0 < q1 < 5.0 and 3.0 < q2 < 5.0, (q1, q2)
    # a planted cos^2 for cos^3 in the FINT term leaves a first-order difference
    planted = oracle.build_variant(str(tmp_path / "liboracle_plant5.so"), ["OR_PLANT=5"])
    q1, q2 = _gap_ratios(planted)
    assert q1 < 2.5 and q2 < 2.5, (q1, q2)


def test_method2_second_order_terms_are_quadratic():
    d = []
    for a in (2e-3, 4e-3, 8e-3):
        z = [[a, 0.05 * a, -0.7 * a, 0.03 * a, 0.0, 1e-3]]
        r1 = run(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2), z)[0]
        r2 = run(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2, EDGE_ORDER=2), z)[0]
        d.append(np.max(np.abs(r1[:4] - r2[:4])))
    assert 3.0 < d[1] / d[0] < 5.0 and 3.0 < d[2] / d[1] < 5.0, d


def test_method2_refuses_higher_order_without_gap():
    with pytest.raises(RuntimeError, match="HGAP=0"):
        run(bend(EDGE1_EFFECTS=2, EDGE_ORDER=2, HGAP=0.0), [[1e-3, 0, 0, 0, 0, 0]])


def edge_ring(method, order=1):
    """four FODO cells whose bends use the given edge method"""
    els = [Element("AP", "MAXAMP", X_MAX=0.02, Y_MAX=0.012)]
    for c in range(4):
        els += [Element("Q%d" % c, "KQUAD", L=0.3, K1=1.5 if c % 2 else -1.5, N_KICKS=8),
                Element("D%d" % c, "DRIF", L=0.5),
                Element("B%d" % c, "CSBEND", L=0.8, ANGLE=0.06 if c % 2 else -0.04, K1=0.1, N_KICKS=8,
                        E1=0.1, E2=0.06, HGAP=0.015, FINT=0.55, H1=0.3, H2=-0.2, EDGE1_EFFECTS=method,
                        EDGE2_EFFECTS=method, EDGE_ORDER=order),
                Element("E%d" % c, "EDRIFT", L=0.5)]
    return Lattice(els)


def edge_beam(n, seed, sigma=6e-3):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, S.NPROP))
    b[:, :4] = rng.normal(0, [sigma, 0.1 * sigma, 0.6 * sigma, 0.06 * sigma], (n, 4))
    b[:, 4] = rng.normal(0, 1e-3, n)
    b[:, 5] = rng.normal(0, 2e-3, n)
    b[:, 6] = np.arange(1, n + 1)
    return b


def test_method4_is_symplectic_and_agrees_with_method2_to_first_order():
    """EDGE_EFFECTS 4 (Hwang's map in Lindberg's symplectic form,
    src/csbend.c:4413-4520) is a symplectic map of (x, q_x, y, q_y) at fixed
    delta, which method 2 at EDGE_ORDER 2 is not (edges of a bend so short that
    its body is the identity; without pole-face curvature, whose terms carry
    sec^2 e in the q_x kick and sec^3 e in the q_y kick in the reference and so
    derive from no common potential); the two share the linear terms (their
    difference vanishes faster than the amplitude)"""
    kw = dict(L=1e-9, ANGLE=1e-10, K1=0.0, N_KICKS=1, E1=0.3, E2=0.25, HGAP=0.03, FINT=0.5)
    z0 = np.array([3e-2, 0.02, -2e-2, 0.01, 0.0, 2e-3])
    e4 = symplectic_error(jacobian(bend(EDGE1_EFFECTS=4, EDGE2_EFFECTS=4, **kw), z0))
    e2 = symplectic_error(jacobian(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2, EDGE_ORDER=2, **kw), z0))
    assert e4 < 1e-7 and e2 > 1e-6, (e4, e2)
    d = []
    for a in (1e-3, 2e-3, 4e-3):
        z = [[a, 0.05 * a, -0.7 * a, 0.03 * a, 0.0, 1e-3]]
        r4 = run(bend(EDGE1_EFFECTS=4, EDGE2_EFFECTS=4), z)[0]
        r2 = run(bend(EDGE1_EFFECTS=2, EDGE2_EFFECTS=2), z)[0]
        d.append(np.max(np.abs(r4[:4] - r2[:4])))
    assert d[1] / d[0] > 3.0 and d[2] / d[1] > 3.0, d


@pytest.mark.parametrize("method,order", [(2, 1), (2, 2), (3, 1), (4, 1)])
def test_host_build_of_the_maps_matches_the_oracle(method, order):
    from tests.test_batch_host import _compare_with_oracle, _host_track
    beams = [edge_beam(n, 11 + method + order + v) for v, n in enumerate((200, 130))]
    mk = lambda: [edge_ring(method, order), edge_ring(method, order)]
    r, out6, lpass, cnt = _host_track(beams, mk(), 3, Po, None, 1)
    total = _compare_with_oracle(beams, mk, 3, Po, out6, lpass, cnt)
    assert r == total and 0 < total < sum(len(b) for b in beams)
