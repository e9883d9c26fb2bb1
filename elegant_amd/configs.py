"""Synthetic workloads: lattices and beams of the benchmark configurations.

fodo_ring()  -- 10-cell FODO ring, 1 GeV: per 10 m cell QF/QD KQUAD (L=0.3 m,
                K1=+-1.2 m^-2, N_KICKS=8), two KSEXT (L=0.2 m, K2=+-15 m^-3, N_KICKS=4),
                one CSBEND (L=1 m, ANGLE=2pi/20, N_KICKS=8, no edge effects, no
                radiation) and DRIF filling the cell, MAXAMP 20 mm x 20 mm.  The ten
                bends total pi, not 2pi: tracking does not need a closed ring, and
                the beam is matched to the periodic cell.
mba_ring()   -- 40-cell multi-bend-achromat-style ring, 3 GeV, 640 m: per 16 m cell
                5 CSBEND with gradients K1 in [-0.5, 0.5] m^-2 (L=0.8 m, N_KICKS=16),
                10 KQUAD (N_KICKS=8), 6 KSEXT, 2 KOCT (|K3| <= 1000 m^-4) and DRIF;
                one RFCA (L=0.3 m, V=2 MV, h=1296, PHASE=170 deg, matrix method) in
                cell 0; MAXAMP 20 mm x 20 mm.  Quadrupole strengths were fitted for a
                stable cell with alpha = 0 and beta ~ 13 m / 11 m at the cell start, beta_max ~ 16 m (tunes per cell 0.581 / 0.265).
da_grid()    -- dynamic-aperture grid of initial conditions (x, y) with all other
                coordinates zero, IDs row-major as in src/aperture_search.c:1358-1366.
gaussian_beam() -- 6D Gaussian matched to periodic Twiss, cut off at 2 sigma per
                coordinate (elegant's default distribution_cutoff, bunched_beam.nl:47),
                drawn with numpy's PCG64 from a seed (elegant's own generator is not
                reproduced).
"""
import math

import numpy as np

from . import optics
from . import structs as S
from .lattice import Element, Lattice

ME_MEV = S.ME_MEV


def p_central_from_mev(p_mev):
    """central beta*gamma for a momentum given in MeV/c (elegant's p_central_mev)"""
    return p_mev / ME_MEV


def fodo_cell(i=0):
    d = 1.6
    return [
        Element("QF", "KQUAD", L=0.3, K1=1.2, N_KICKS=8),
        Element("D1", "DRIF", L=d),
        Element("SF", "KSEXT", L=0.2, K2=15.0, N_KICKS=4),
        Element("D2", "DRIF", L=d),
        Element("B", "CSBEND", L=1.0, ANGLE=2 * math.pi / 20, N_KICKS=8, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0),
        Element("D3", "DRIF", L=d),
        Element("QD", "KQUAD", L=0.3, K1=-1.2, N_KICKS=8),
        Element("D4", "DRIF", L=d),
        Element("SD", "KSEXT", L=0.2, K2=-15.0, N_KICKS=4),
        Element("D5", "DRIF", L=d),
    ]


def fodo_ring(n_cells=10, aperture=0.02):
    els = [Element("AP", "MAXAMP", X_MAX=aperture, Y_MAX=aperture)]
    for i in range(n_cells):
        els.extend(fodo_cell(i))
    return Lattice(els)


MBA_KQ = [0.8526, -0.5177, 2.9022, -1.3801, 0.6626, 1.4626, 0.4974, -0.0237, -1.2771, -0.4511]
MBA_KB = [-0.5, -0.25, 0.1, -0.25, 0.5]
MBA_KS = [20.0, -25.0, 20.0, -25.0, 20.0, -25.0]
MBA_KO = [800.0, -1000.0]
MBA_CELL_LENGTH = 16.0
MBA_CELLS = 40
MBA_HARMONIC = 1296


def mba_cell(c, with_rf=False, rf=None):
    ang = 2 * math.pi / (5 * MBA_CELLS)
    els = []
    if with_rf:
        els += [Element("DL0", "DRIF", L=0.35), Element("RF", "RFCA", **rf), Element("DL1", "DRIF", L=0.35)]
    else:
        els.append(Element("DL", "DRIF", L=1.0))
    si = oi = 0
    for u in range(5):
        els.append(Element("QA%d" % u, "KQUAD", L=0.25, K1=MBA_KQ[2 * u], N_KICKS=8))
        els.append(Element("DA", "DRIF", L=0.2))
        if u < 3:
            els.append(Element("S%d" % si, "KSEXT", L=0.2, K2=MBA_KS[si], N_KICKS=4))
            si += 1
        else:
            els.append(Element("O%d" % oi, "KOCT", L=0.2, K3=MBA_KO[oi]))
            oi += 1
        els.append(Element("DB", "DRIF", L=0.2))
        els.append(Element("B%d" % u, "CSBEND", L=0.8, ANGLE=ang, K1=MBA_KB[u], N_KICKS=16, EDGE1_EFFECTS=0,
                           EDGE2_EFFECTS=0))
        els.append(Element("DC", "DRIF", L=0.2))
        if u < 3:
            els.append(Element("S%d" % si, "KSEXT", L=0.2, K2=MBA_KS[si], N_KICKS=4))
            si += 1
        else:
            els.append(Element("DD", "DRIF", L=0.2))
        els.append(Element("DE", "DRIF", L=0.2))
        els.append(Element("QB%d" % u, "KQUAD", L=0.25, K1=MBA_KQ[2 * u + 1], N_KICKS=8))
        els.append(Element("DF", "DRIF", L=0.3))
    els.append(Element("DL", "DRIF", L=1.0))
    return els


def mba_rf(p_central, volt=2.0e6, phase=170.0):
    beta = p_central / math.sqrt(p_central ** 2 + 1)
    f_rev = beta * S.C_MKS / (MBA_CELLS * MBA_CELL_LENGTH)
    return dict(L=0.3, VOLT=volt, PHASE=phase, FREQ=MBA_HARMONIC * f_rev, N_KICKS=0)


def mba_ring(p_central=None, n_cells=MBA_CELLS, aperture=0.02, rf=True):
    if p_central is None:
        p_central = p_central_from_mev(3000.0)
    els = [Element("AP", "MAXAMP", X_MAX=aperture, Y_MAX=aperture)]
    for c in range(n_cells):
        els.extend(mba_cell(c, with_rf=(rf and c == 0), rf=mba_rf(p_central)))
    return Lattice(els)


def mba_period():
    """one cell without RF, for matching"""
    return Lattice(mba_cell(0))


def da_grid(nx, ny, xmax=0.015, ymax=0.005):
    """(nx*ny, 10) rows: x in [-xmax, xmax], y in [-ymax, ymax], IDs 1.. row-major"""
    xs = np.linspace(-xmax, xmax, nx)
    ys = np.linspace(-ymax, ymax, ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    out = np.zeros((nx * ny, S.NPROP))
    out[:, 0] = X.ravel()
    out[:, 2] = Y.ravel()
    out[:, S.ID_INDEX] = np.arange(1, nx * ny + 1)
    return out


def _trunc_normal(rng, n, cutoff):
    out = rng.standard_normal(n)
    bad = np.abs(out) > cutoff
    while bad.any():
        out[bad] = rng.standard_normal(int(bad.sum()))
        bad = np.abs(out) > cutoff
    return out


def gaussian_beam(n, twiss, emitx, emity, sigma_dp, sigma_s, seed=1, cutoff=2.0):
    """matched 6D Gaussian, (n, 10) rows with IDs 1..n"""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, S.NPROP))
    for plane, emit, (ix, ixp) in (("x", emitx, (0, 1)), ("y", emity, (2, 3))):
        b, a = twiss["beta" + plane], twiss["alpha" + plane]
        u = _trunc_normal(rng, n, cutoff)
        v = _trunc_normal(rng, n, cutoff)
        out[:, ix] = math.sqrt(emit * b) * u
        out[:, ixp] = math.sqrt(emit / b) * (v - a * u)
    dp = sigma_dp * _trunc_normal(rng, n, cutoff)
    out[:, 5] = dp
    out[:, 4] = sigma_s * _trunc_normal(rng, n, cutoff)
    out[:, 0] += twiss.get("etax", 0.0) * dp
    out[:, 1] += twiss.get("etaxp", 0.0) * dp
    out[:, S.ID_INDEX] = np.arange(1, n + 1)
    return out


def fodo_beam(n=10000, seed=1):
    tw = optics.periodic_twiss(Lattice(fodo_cell()))
    return gaussian_beam(n, tw, 5e-9, 50e-12, 1e-3, 3e-3, seed=seed)


def mba_beam(n, seed=1, scale=1.0):
    tw = optics.periodic_twiss(mba_period())
    return gaussian_beam(n, tw, 5e-9 * scale, 50e-12 * scale, 1e-3, 3e-3, seed=seed)


def _trunc(rng, sigma, cutoff, n=None):
    v = _trunc_normal(rng, 1 if n is None else n, cutoff) * sigma
    return float(v[0]) if n is None else v


def mba_variants(n_variants, seed=1, sigma_fse=1e-3, sigma_dxy=30e-6, cutoff=2.0, p_central=None):
    """copies of mba_ring() with seeded errors on every KQUAD and KSEXT occurrence:
    relative strength error (FSE) and transverse offsets (DX, DY), Gaussian with a
    2-sigma cutoff, one generator stream per variant"""
    base = mba_ring(p_central)
    out = []
    for v in range(n_variants):
        rng = np.random.default_rng([seed, v])
        lat = base.copy()
        for el in lat.elements:
            if el.type in (S.T_KQUAD, S.T_KSEXT):
                el.params["FSE"] = _trunc(rng, sigma_fse, cutoff)
                el.params["DX"] = _trunc(rng, sigma_dxy, cutoff)
                el.params["DY"] = _trunc(rng, sigma_dxy, cutoff)
        out.append(lat)
    return out


def variant_counts(n_variants, seed=1, lo=1000, hi=100000, median=16000.0, sigma=0.8):
    """per-variant particle counts, lognormal clipped to [lo, hi]"""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.normal(math.log(median), sigma, n_variants))
    return np.clip(np.round(n), lo, hi).astype(np.int64)


def variant_beams(counts, seed=1):
    """6D Gaussian beams (config-0 emittances and spreads) matched to the MBA cell, at 3 GeV"""
    return [mba_beam(int(c), seed=seed * 100003 + v) for v, c in enumerate(counts)]
