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
  identity to 1e-12."""
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


def run(lat, coords):
    r, n = oracle.element(rows_of(coords), lat, 0, Po, 0.0, 0.8)
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
