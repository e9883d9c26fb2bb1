"""Algorithmic work per particle-element unit, by element type.

Counting rule: every IEEE double +, -, *, / and sqrt the reference routine
executes for one surviving particle counts as one flop, except in the generic
transport-matrix loop of track_particles (DRIF's drift matrix, ROTATE's
rotation matrix), which multiplies through every zero and unit matrix entry
and is counted by its non-trivial terms (324 executed, 11 counted for an
order-2 drift); negation, fabs and comparisons count zero; each call of sin,
cos, tan, asin or log counts as one transcendental (T).  The "algorithmic"
figure used for the roofline adds T as one flop each, so it is a lower bound on
the work the hardware has to do.  The radiation blocks the reference runs with
radiation off are counted.

The formulas below were derived by instrumenting the checker: its counting
build (a C++ compile of the same source with counted doubles) tallies these
operations per particle, and a CPU test checks every formula against those
tallies on the FODO and MBA rings and on variants of every element type
(integration orders, slice counts, offsets, tilts, edges, expansion orders).
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


def _multipole_orders(etype, p):
    orders = {S.T_KQUAD: [1], S.T_KSEXT: [2], S.T_KOCT: [3], S.T_KQUSE: [1, 2]}[etype][:]
    if etype == S.T_KSEXT and (_get(p, "K1") or _get(p, "J1")):
        orders.append(1)
    return orders


# per-kick cost of one multipole order term (power arrays, expansion terms,
# sums, the qx/qy update); orders 1 and 2 read static coefficient tables, order
# 3 computes its coefficients (src/multipole.h:76-145)
_KICK_TERM = {1: 15, 2: 20, 3: 39}


def multipole(etype, p):
    """integrate_kick_multipole_ordn (src/multipole.c:1099-1456) plus the
    misalignment passes of multipole_tracking2; returns (flops, T).
    Per call 46 (scalings, beta0, slopes to momenta and back, KnL/n, exit
    checks); per slice 4 plus nsub drift sub-steps of 13 and nsub - 1 kicks of
    48 (momenta to slopes, and the radiation block, which runs for any element
    with length even without radiation, src/multipole.c:1360-1390) plus the
    order terms and 8 per steering kick; +26 for offsets, +30 and +6 per slice
    for a tilt.  Zero length: no drift work and no radiation block."""
    integ = int(_get(p, "INTEGRATION_ORDER", default=4))
    nsub = _integrator(integ)
    nk = int(_get(p, "N_KICKS", default=4 if etype == S.T_KOCT else 0))
    ns = int(_get(p, "N_SLICES", default=4 if etype == S.T_KOCT else 1))
    if nk > 0:
        ns = max(1, math.ceil(nk / integ)) if integ > 2 else nk
    orders = _multipole_orders(etype, p)
    if _get(p, "L"):
        kick = 48 + sum(_KICK_TERM[o] for o in orders)
        if _get(p, "HKICK") or _get(p, "xkick"):
            kick += 8
        if _get(p, "VKICK") or _get(p, "ykick"):
            kick += 8
        f = 46 + ns * (4 + nsub * 13 + (nsub - 1) * kick)
    else:
        m = max(orders)
        f = 46 + ns * (-(5 + 2 * m) + (9 + 2 * m) * nsub)
    if _get(p, "DX") or _get(p, "DY") or _get(p, "DZ"):
        f += 26
    if _get(p, "TILT"):
        f += 30 + 6 * ns
    return f, 0


def csbend(p):
    """track_through_csbend (src/csbend.c:1088-1403) with integrate_csbend_ordn
    (:1459-1720) and computeCSBENDFields (:99-172); returns (flops, T).
    Per call 94 (the entry and exit offsets and rotations, computed even when
    zero, the canonical conversions, the time-of-flight rescaling); per slice 2
    plus nsub exact-drift sub-steps of 31 flops and 4 transcendentals (asin,
    sin, cos, cos) and nsub - 1 kicks of 48 (with the radiation block every bend
    runs, :1627-1660) plus the field expansion, (EO1 - 1) + 3 n_terms + EO1 for
    EO1 = expansion order + 1 and n_terms = EO1 (EO1 + 1) / 2; per edge with
    EDGE_EFFECTS 11 flops and 2 T (EDGE_ORDER 1) or 69 and 3 T (EDGE_ORDER 2);
    a tilt adds 6 and 6 per slice."""
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
    per_slice = 2 + nsub * 31 + (nsub - 1) * (48 + field)
    f = 94 + ns * per_slice
    t = ns * nsub * 4
    edge_order = int(_get(p, "EDGE_ORDER", default=1))
    for k in ("EDGE1_EFFECTS", "EDGE2_EFFECTS"):
        if int(_get(p, k, default=1)):
            f += 69 if edge_order >= 2 else 11
            t += 3 if edge_order >= 2 else 2
    if _get(p, "TILT"):
        f += 6 + 6 * ns
    return f, t


def rfca(p):
    """trackRfCavityWithWakes (src/simple_rfca.c:295-985): matrix method
    (:854-941, sin and log) or N_KICKS kicks (:567-660, one sin each), plus
    removeInvalidParticles.  Per pass; the TMEAN fiducial of the first pass
    (13 per particle, once per run) is not per-unit work."""
    if not _get(p, "VOLT"):
        return 11, 0
    nk = int(_get(p, "N_KICKS", default=0))
    if not _get(p, "L"):
        return 55, 1
    if nk <= 0:
        return 72, 2
    return 4 + 86 * nk, nk


def mult(p):
    """multipole_tracking (src/multipole.c:439-658), no BORE, no radiation:
    per slice a kick of 16 + 5 n (orders 0-2, static coefficients) or
    n^2 + 6 n + 18 (computed coefficients) plus a drift of 14 with length;
    29 per call with length (18 without); +26 for offsets, +24 for a tilt"""
    if not (_get(p, "KNL") and _get(p, "FACTOR", default=1.0)):
        return (11, 0) if _get(p, "L") else (0, 0)
    order = int(_get(p, "ORDER", default=1))
    ns = int(_get(p, "N_SLICES", "N_KICKS", default=4))
    kick = 16 + 5 * order if order <= 2 else order * order + 6 * order + 18
    if _get(p, "L"):
        f = 29 + ns * (kick + 14)
    else:
        f = 18 + ns * kick
    if _get(p, "DX") or _get(p, "DY") or _get(p, "DZ"):
        f += 26
    if _get(p, "TILT"):
        f += 24
    return f, 0


def malign(p):
    """offset_beam (src/do_tracking.c:3311-3372): 9 for the transverse offsets,
    +6 for DZ; a time or energy change adds 11, +7 for DP, +8 for DE"""
    f = 9 + (6 if _get(p, "DZ") else 0)
    if _get(p, "DT") or _get(p, "DP") or _get(p, "DE"):
        f += 11 + (7 if _get(p, "DP") else 0) + (8 if _get(p, "DE") else 0)
    return f, 0


def per_unit(el):
    """(flops, transcendentals) for one particle through one element"""
    t = el.type
    p = el.params
    if t in (S.T_MARK, S.T_MAXAMP):
        return 0, 0
    if t == S.T_DRIF:
        # the generic matrix loop's non-trivial terms: x + L x', y + L y' and the
        # path length with its second-order terms (L/2) x'^2 and (L/2) y'^2; a
        # matrix entry of exactly +-1 (L or L/2) makes its products trivial
        L = _get(p, "L")
        if not L:
            return 0, 0
        if int(_get(p, "ORDER", default=0)) == 1:
            return 5 - (2 if abs(L) == 1 else 0), 0
        return 11 - (2 if abs(L) == 1 else 0) - (2 if abs(L / 2) == 1 else 0), 0
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
        return 12, 0  # the order-1 matrix loop's non-trivial terms
    if t == S.T_MALIGN:
        return malign(p)
    if t == S.T_ENERGY:
        # set_central_momentum: 4 per particle on the pass that changes the
        # reference momentum, nothing once it equals the target
        return 0, 0
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
