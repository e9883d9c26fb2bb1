"""Algorithmic work per particle-element unit, by element type.

Counting rule: every IEEE double +, -, *, / and sqrt the reference routine
executes for one surviving particle counts as one flop; negation, fabs and
comparisons count zero; each call of sin, cos, asin or log counts as one
transcendental (T).  The "algorithmic" figure used for the roofline adds T as
one flop each, so it is a lower bound on the work the hardware has to do.
Counts follow the reference code paths cited per function (radiation off, no
error multipoles, no fringe, misalignment method 0); the radiation blocks the
reference runs with radiation off are counted.
"""
import math

from . import structs as S


def _get(params, *keys, default=0.0):
    for k in keys:
        for kk in (k, k.upper(), k.lower()):
            if kk in params:
                return params[kk]
    return default


def _integrator(order):
    return {2: 2, 4: 4, 6: 8}[int(order)]


def multipole(etype, p):
    """integrate_kick_multipole_ordn (src/multipole.c:1099-1456) plus the
    misalignment passes of multipole_tracking2; returns (flops, T)"""
    integ = int(_get(p, "INTEGRATION_ORDER", default=4))
    nsub = _integrator(integ)
    nk = int(_get(p, "N_KICKS", default=4 if etype == S.T_KOCT else 0))
    ns = int(_get(p, "N_SLICES", default=4 if etype == S.T_KOCT else 1))
    if nk > 0:
        ns = max(1, math.ceil(nk / integ)) if integ > 2 else nk
    orders = {S.T_KQUAD: [1], S.T_KSEXT: [2], S.T_KOCT: [3], S.T_KQUSE: [1, 2]}[etype][:]
    if etype == S.T_KSEXT:
        if _get(p, "K1"):
            orders.append(1)
        if _get(p, "J1"):
            orders.append(1)
    maxo = max(orders)
    f = 3 + 6 + 11 + 3 + 9  # scalings, beta0, slopes->momenta, KnL/n, momenta->slopes
    drift = 13  # dsh, x, y, s with sqrt, dzLoss
    kick = 2 * maxo  # power arrays
    for o in orders:
        kick += 1 + 2 * (o + 1) + o + 8  # strength, terms, sums, qx/qy and delta updates
    if _get(p, "HKICK") or _get(p, "xkick"):
        kick += 1 + 2 + 4
    if _get(p, "VKICK") or _get(p, "ykick"):
        kick += 1 + 2 + 4
    kick += 9  # momenta->slopes
    if _get(p, "L"):
        # the radiation block, run for any element with length even without
        # radiation (src/multipole.c:1360-1390): the delta_q sums of the kicks,
        # the round trip of qx, qy through 1 + dp, F2, the path-length factors
        # and a second momenta->slopes conversion
        kick += 4 * len(orders) + 39
    f += ns * (4 + nsub * drift + (nsub - 1) * kick) + 4 + 9 + 1
    if _get(p, "DX") or _get(p, "DY") or _get(p, "DZ"):
        f += 2 * 13
    if _get(p, "TILT"):
        f += 2 * 12
    return f, 0


def csbend(p):
    """integrate_csbend_ordn (src/csbend.c:1459-1720) with computeCSBENDFields
    (:99-172) and the per-particle frame changes of track_through_csbend"""
    if not _get(p, "ANGLE"):
        return 11, 0
    integ = int(_get(p, "INTEGRATION_ORDER", default=4))
    nsub = _integrator(integ)
    ns = int(_get(p, "N_SLICES", "N_KICKS", default=4))
    ks = [_get(p, "K%d" % i) for i in range(1, 9)]
    top = max([i + 1 for i, k in enumerate(ks) if k] or [0])
    eo = int(_get(p, "EXPANSION_ORDER", default=0)) or max(4, top + 2)
    eo1 = eo + 1
    if any(ks) or _get(p, "XKICK"):
        n_terms = eo1 * (eo1 + 1) // 2
        field = (eo1 - 1) + 3 * n_terms + eo1
    else:
        field = 0
    drift = 31
    # the kick, then the radiation block every bend runs (src/csbend.c:1627-1660):
    # slopes and path-length factor, F2, the round trip of QX, QY through 1 + DPoP
    kick = 1 + field + 12 + 36
    per_slice = 2 + nsub * drift + (nsub - 1) * kick
    f = ns * per_slice + 80 + 15  # + the time-of-flight rescaling (:1285-1296)
    t = ns * nsub * 4
    return f, t


def rfca(p):
    """trackRfCavityWithWakes matrix method (src/simple_rfca.c:854-941) plus
    removeInvalidParticles"""
    if not _get(p, "VOLT"):
        return 11, 0
    nk = int(_get(p, "N_KICKS", default=0))
    if nk <= 0 or not _get(p, "L"):
        return 78, 2  # sin, log
    return nk * 60 + 22, nk


def mult(p):
    """multipole_tracking (src/multipole.c:439-658), no BORE, no radiation"""
    if not (_get(p, "KNL") and _get(p, "FACTOR", default=1.0)):
        return (11, 0) if _get(p, "L") else (0, 0)
    order = int(_get(p, "ORDER", default=1))
    ns = int(_get(p, "N_SLICES", "N_KICKS", default=4))
    drift = 14 if _get(p, "L") else 0  # x, y and the path length with its square root
    kick = 2 * order + 3 * (order + 1) + 4 + 9  # powers, terms, qx/qy, momenta->slopes
    f = 8 + ns * (drift + kick) + drift + 1  # slopes->momenta, slices, final drift, s
    if _get(p, "DX") or _get(p, "DY") or _get(p, "DZ"):
        f += 2 * 13
    if _get(p, "TILT"):
        f += 2 * 12
    return f, 0


def malign(p):
    """offset_beam (src/do_tracking.c:3311-3372)"""
    f = 8 + (6 if _get(p, "DZ") else 0)
    if _get(p, "DT") or _get(p, "DP") or _get(p, "DE"):
        f += 12 + (6 if _get(p, "DP") else 0) + (7 if _get(p, "DE") else 0)
    return f, 0


def per_unit(el):
    """(flops, transcendentals) for one particle through one element"""
    t = el.type
    p = el.params
    if t in (S.T_MARK, S.T_MAXAMP):
        return 0, 0
    if t == S.T_DRIF:
        return 11, 0  # x, y and the second-order path length of the drift map
    if t == S.T_EDRIFT:
        return 11, 0
    if t in (S.T_KQUAD, S.T_KSEXT, S.T_KOCT, S.T_KQUSE):
        return multipole(t, p)
    if t == S.T_CSBEND:
        return csbend(p)
    if t == S.T_RFCA:
        return rfca(p)
    if t == S.T_MULT:
        return mult(p)
    if t == S.T_ROTATE:
        return 72, 0  # the 6x6 order-1 matrix loop of track_particles
    if t == S.T_MALIGN:
        return malign(p)
    if t == S.T_ENERGY:
        return 3, 0  # set_central_momentum: ((1 + delta) P - P_new) / P_new
    raise ValueError(t)


def lattice_census(lattice):
    """per-turn totals for one particle: dict with flops, T, units and the
    per-type breakdown"""
    out = {"flops": 0, "T": 0, "units": 0, "by_type": {}}
    for el in lattice.elements:
        f, t = per_unit(el)
        if el.type in (S.T_MARK, S.T_MAXAMP):
            continue
        out["flops"] += f
        out["T"] += t
        out["units"] += 1
        name = S.TYPE_NAMES[el.type]
        d = out["by_type"].setdefault(name, {"count": 0, "flops": 0, "T": 0})
        d["count"] += 1
        d["flops"] += f
        d["T"] += t
    out["algorithmic_per_unit"] = (out["flops"] + out["T"]) / max(1, out["units"])
    return out
