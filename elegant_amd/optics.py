"""Linear optics of a Lattice: hard-edge transfer matrices and periodic Twiss.

Used to match synthetic beams to a ring or cell, the role compute_periodic_twiss
(src/twiss.cc:123) plays when elegant builds a matched bunch
(src/bunched_beam.c:212-216, :547-566).  Kick multipoles and RF cavities act as
drifts at linear order; CSBEND is a sector bend with gradient (no edge angles
unless E1/E2 are given, which this module ignores).
"""
import math

import numpy as np

from . import structs as S


def _quad_block(K, L):
    """2x2 matrix of x'' = -K x over length L."""
    if abs(K) < 1e-15:
        return np.array([[1.0, L], [0.0, 1.0]])
    if K > 0:
        k = math.sqrt(K)
        return np.array([[math.cos(k * L), math.sin(k * L) / k], [-k * math.sin(k * L), math.cos(k * L)]])
    k = math.sqrt(-K)
    return np.array([[math.cosh(k * L), math.sinh(k * L) / k], [k * math.sinh(k * L), math.cosh(k * L)]])


def element_matrix(el):
    """5x5 linear map on (x, x', y, y', delta)."""
    M = np.eye(5)
    L = el.length()
    p = {k.upper(): v for k, v in el.params.items()}
    t = el.type
    if t in (S.T_MARK, S.T_MAXAMP):
        return M
    if t == S.T_KQUAD:
        K = p.get("K1", 0.0) * (1 + p.get("FSE", 0.0))
        M[0:2, 0:2] = _quad_block(K, L)
        M[2:4, 2:4] = _quad_block(-K, L)
        return M
    if t == S.T_CSBEND:
        ang = p.get("ANGLE", 0.0)
        K1 = p.get("K1", 0.0)
        if L == 0:
            return M
        h = ang / L
        kx2 = h * h + K1
        M[0:2, 0:2] = _quad_block(kx2, L)
        M[2:4, 2:4] = _quad_block(-K1, L)
        if abs(kx2) > 1e-15:
            if kx2 > 0:
                k = math.sqrt(kx2)
                M[0, 4] = h * (1 - math.cos(k * L)) / kx2
                M[1, 4] = h * math.sin(k * L) / k
            else:
                k = math.sqrt(-kx2)
                M[0, 4] = h * (math.cosh(k * L) - 1) / (-kx2)
                M[1, 4] = h * math.sinh(k * L) / k
        else:
            M[0, 4] = h * L * L / 2
            M[1, 4] = h * L
        return M
    # drifts, sextupoles, octupoles, RF cavities: drift at linear order
    M[0, 1] = L
    M[2, 3] = L
    return M


def one_turn(lattice):
    M = np.eye(5)
    for el in lattice.elements:
        M = element_matrix(el) @ M
    return M


def periodic_twiss(lattice):
    """dict with betax, alphax, betay, alphay, etax, etaxp, nux, nuy (fractional
    tunes of the period) or raises if the period is unstable."""
    M = one_turn(lattice)
    out = {}
    for plane, (i, j) in (("x", (0, 1)), ("y", (2, 3))):
        m = M[i:j + 1, i:j + 1]
        c = 0.5 * (m[0, 0] + m[1, 1])
        if abs(c) >= 1:
            raise ValueError("period is unstable in %s (cos mu = %g)" % (plane, c))
        s = math.copysign(math.sqrt(1 - c * c), m[0, 1])
        out["beta" + plane] = m[0, 1] / s
        out["alpha" + plane] = (m[0, 0] - m[1, 1]) / (2 * s)
        mu = math.atan2(s, c)
        if mu < 0:
            mu += 2 * math.pi
        out["nu" + plane] = mu / (2 * math.pi)
    A = np.eye(2) - M[0:2, 0:2]
    eta = np.linalg.solve(A, M[0:2, 4])
    out["etax"], out["etaxp"] = float(eta[0]), float(eta[1])
    return out


def momentum_compaction(lattice, twiss):
    """alpha_c = (1/C) * integral of h * eta over the bends (linear, sampled)."""
    eta = np.array([twiss["etax"], twiss["etaxp"], 0.0, 0.0, 1.0])
    v = np.array([twiss["etax"], twiss["etaxp"], 0.0, 0.0, 1.0])
    total = 0.0
    C = lattice.total_length()
    for el in lattice.elements:
        if el.type == S.T_CSBEND and el.length() > 0:
            L = el.length()
            h = el.params.get("ANGLE", el.params.get("angle", 0.0)) / L
            nsub = 16
            for k in range(nsub):
                sub = type(el)(el.name, el.type, **dict(el.params, L=L / nsub,
                                                        ANGLE=el.params.get("ANGLE", 0.0) / nsub))
                vm = element_matrix(sub) @ v
                total += h * 0.5 * (v[0] + vm[0]) * (L / nsub)
                v = vm
        else:
            v = element_matrix(el) @ v
    del eta
    return total / C
