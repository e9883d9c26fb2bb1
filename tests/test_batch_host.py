"""The batched path's host logic on the CPU (tests/batch_host): the work list
that k_batch_prep/k_batch compute (eh_batch_wave_starts, eh_batch_item_variant),
the descriptor tables across the pass-0 stop at an RF cavity (batch_tables,
batch_rf_update) and the rule for reusing device tables between calls
(batch_tables_reusable), and whole runs through an emulation of k_batch's spans
that takes the work items in shuffled order.  Variants are checked against the
oracle tracking each variant alone (src/do_tracking.c:2506-2524 semantics per
variant)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from elegant_amd import configs
from elegant_amd import structs as S
from elegant_amd.lattice import Element, Lattice
from oracle import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_host")
_h = None
INHERIT = 0x40000000


def harness():
    global _h
    if _h is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        _h = C.CDLL(os.path.join(HERE, "libbatch_host.so"))
        vp, L = C.c_void_p, C.c_long
        _h.bh_worklist.argtypes = [vp, C.c_int, vp, vp]
        _h.bh_worklist.restype = L
        _h.bh_lean_stage.argtypes = [C.c_int, C.c_int, vp]
        _h.bh_lean_stage.restype = L
        _h.bh_loss_words.argtypes = [C.c_int] * 5 + [C.c_int32, vp, vp]
        _h.bh_tables.argtypes = [vp, L, L, C.c_double, vp, vp, vp, vp]
        _h.bh_tables.restype = L
        _h.bh_track.argtypes = [vp, vp, L, vp, L, L, C.c_double, L, C.c_uint, vp, vp]
        _h.bh_track.restype = L
        _h.eh_last_error.restype = C.c_char_p
    return _h


def _lines(lattices):
    arrays, keep = [], []
    for lat in lattices:
        a, k = lat.c_beamline()
        arrays.append(a)
        keep.append(k)
    ptrs = (C.c_void_p * len(arrays))(*[C.cast(a, C.c_void_p) for a in arrays])
    return ptrs, keep


@pytest.mark.parametrize("counts", [[300, 0, 1, 129, 64], [0, 0, 5], [64], [65, 63, 0, 128, 1, 0],
                                    list(np.random.default_rng(3).integers(0, 400, 37))])
def test_worklist_covers_every_particle_once(counts):
    counts = np.array(counts, dtype=np.int32)
    nvar = len(counts)
    ws = np.zeros(nvar + 1, dtype=np.int32)
    items = np.zeros(int(((counts + 63) // 64).sum()) + 1, dtype=np.int32)
    total = harness().bh_worklist(counts.ctypes.data, nvar, ws.ctypes.data, items.ctypes.data)
    assert total == ((counts + 63) // 64).sum()
    assert ws[-1] == total
    seen = {v: np.zeros(c, dtype=int) for v, c in enumerate(counts)}
    for w in range(total):
        v = items[w]
        assert 0 <= v < nvar and counts[v] > 0, "item %d assigned to empty or invalid variant %d" % (w, v)
        assert ws[v] <= w < ws[v + 1]
        for lane in range(64):
            local = (w - ws[v]) * 64 + lane
            if local < counts[v]:
                seen[v][local] += 1
    for v, c in enumerate(counts):
        assert (seen[v] == 1).all(), "variant %d: particles not covered exactly once" % v


def _lean(t, ell=0):
    cap = C.c_int(-1)
    st = harness().bh_lean_stage(t, ell, C.byref(cap))
    return st, cap.value


def test_lean_map_refuses_full_only_types():
    for t in (S.T_MULT, S.T_ROTATE, S.T_MALIGN, S.T_ENERGY):
        assert _lean(t) == (8, 0), S.TYPE_NAMES[t]
    for t in (S.T_MARK, S.T_MAXAMP):
        assert _lean(t) == (-1, 1)
    # a descriptor with an elliptical aperture: refused, and never sent there
    for t in (S.T_DRIF, S.T_KQUAD, S.T_CSBEND, S.T_MARK):
        assert _lean(t, 1) == (8, 0), S.TYPE_NAMES[t]


def test_loss_code_round_trip():
    """the single loss word the fused kernels carry decodes to the pass and
    event the reference order needs: event_base + (dpass * nelem + e) * 3 + stage"""
    lp, le = C.c_int32(), C.c_int32()
    for dpass, nelem, ee, st, p0, base in ((0, 2082, 0, 0, 0, 0), (3, 2082, 2081, 2, 17, 1000),
                                          (15, 7, 6, 1 | 4, 3, 5), (1, 1, 0, 8, 0, 0)):
        harness().bh_loss_words(dpass, nelem, ee, st, p0, base, C.byref(lp), C.byref(le))
        assert lp.value == (p0 + dpass) | (0x40000000 if st & 4 else 0)
        assert le.value == base + (dpass * nelem + ee) * 3 + (st & 3)


def test_emulated_batch_elliptical_apertures():
    """variants behind elliptical MAXAMPs (exponents 2, 4 and 2 with YEXPONENT
    4): the full map's elliptical checks and sweeps against the oracle"""
    from tests.test_gpu_elliptical import _beam, elliptical_ring
    exps = ((2, 0), (4, 0), (2, 4))
    Po = configs.p_central_from_mev(1500.0)
    beams = [_beam(300 + 50 * v, seed=60 + v, xs=0.016, ys=0.009) for v in range(3)]
    r, out6, lpass, cnt = _host_track(beams, [elliptical_ring(*e) for e in exps], 3, Po, None, 1)
    total = _compare_with_oracle(beams, lambda: [elliptical_ring(*e) for e in exps], 3, Po, out6, lpass, cnt)
    assert r == total and 0 < total < sum(len(b) for b in beams)


def _rf_line(volt=2.0e6, shared=False, two=False, mode=0):
    """a small line with RF cavities: optionally one cavity struct at two
    positions (shared), or two different cavities (two)"""
    Po = configs.p_central_from_mev(3000.0)
    rf = configs.mba_rf(Po, volt=volt)
    rf["fiducial_mode"] = mode
    els = [Element("AP", "MAXAMP", X_MAX=0.02, Y_MAX=0.015),
           Element("Q1", "KQUAD", L=0.25, K1=0.8, N_KICKS=8, DX=1e-5),
           Element("D1", "DRIF", L=0.5),
           Element("RF", "RFCA", **rf),
           Element("B1", "CSBEND", L=0.8, ANGLE=0.03, K1=-0.3, N_KICKS=16, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0),
           Element("S1", "KSEXT", L=0.2, K2=20.0, N_KICKS=4),
           Element("D2", "DRIF", L=0.5)]
    if two:
        rf2 = dict(rf)
        rf2["PHASE"] = 175.0
        els.append(Element("RF2", "RFCA", **rf2))
    if shared:
        els.append(Element("RF", "RFCA", **rf))
    els += [Element("Q2", "KQUAD", L=0.25, K1=-0.7, N_KICKS=8), Element("D3", "DRIF", L=0.5)]
    lat = Lattice(els)
    if shared:
        structs = lat.structs()
        first = next(i for i, e in enumerate(lat.elements) if e.name == "RF")
        last = max(i for i, e in enumerate(lat.elements) if e.name == "RF")
        structs[last] = structs[first]
        lat._cbl = None
    return lat, Po


@pytest.mark.parametrize("case", ["plain", "volt0", "shared", "two", "mba"])
def test_tables_after_rf_stop_match_fresh_descriptors(case):
    """after batch_tables and the RF updates every (variant, position) holds the
    descriptor built for it alone; tables with a pending fiducial are never reused"""
    if case == "mba":
        lats = configs.mba_variants(3, seed=7)
        Po = configs.p_central_from_mev(3000.0)
    else:
        pairs = [_rf_line(volt=0.0 if case == "volt0" else 2e6, shared=case == "shared", two=case == "two")
                 for _ in range(3)]
        lats = [p[0] for p in pairs]
        Po = pairs[0][1]
        for v, lat in enumerate(lats):  # variants differ
            lat.set_param(1, "K1", 0.8 + 0.01 * v)
    ptrs, keep = _lines(lats)
    ne = len(lats[0])
    t_set = np.array([1e-9 * (v + 1) for v in range(len(lats))])
    rb, ra, npool = C.c_int32(), C.c_int32(), C.c_int32()
    bad = harness().bh_tables(ptrs, len(lats), ne, Po, t_set.ctypes.data, C.byref(rb), C.byref(ra), C.byref(npool))
    assert bad >= 0, harness().eh_last_error()
    assert bad == 0, "%d (variant, position) entries differ from a fresh descriptor" % bad
    pending = case != "volt0"
    assert rb.value == (0 if pending else 1), "tables with a pending RF fiducial must not be reused"
    assert ra.value == 1
    for lat in lats:
        for s in lat.structs():
            if isinstance(s, S.EH_RFCA) and s.volt != 0:
                assert s.fiducial_seen == 1


def _beams(counts, seed0=0, scale=400.0):
    return [configs.mba_beam(c, seed=seed0 + v, scale=scale) if c else np.zeros((0, S.NPROP))
            for v, c in enumerate(counts)]


def _host_track(beams, lats, n_passes, Po, span, seed):
    counts = np.array([len(b) for b in beams], dtype=np.int64)
    flat = np.ascontiguousarray(np.concatenate([b for b in beams if len(b)] or [np.zeros((0, S.NPROP))]))
    n = int(counts.sum())
    out6 = np.zeros((max(n, 1), 6))
    lpass = np.zeros(max(n, 1), dtype=np.int32)
    ptrs, keep = _lines(lats)
    ne = len(lats[0])
    r = harness().bh_track(flat.ctypes.data, counts.ctypes.data, len(lats), ptrs, ne, n_passes, Po,
                           span or ne, seed, out6.ctypes.data, lpass.ctypes.data)
    assert r >= 0, harness().eh_last_error()
    return r, out6, lpass, counts


def _compare_with_oracle(beams, mk_lats, n_passes, Po, out6, lpass, counts):
    off = 0
    total = 0
    for v, b in enumerate(beams):
        n = len(b)
        ref, nref, _, _ = oracle.do_tracking(b, mk_lats()[v], n_passes, Po)
        total += nref
        ids = b[:, 6]
        pos = {int(i): k for k, i in enumerate(ids)}
        mine_alive = lpass[off:off + n] == -1
        assert mine_alive.sum() == nref, "variant %d survivors: emulation %d oracle %d" % (v, mine_alive.sum(), nref)
        for row in range(n):
            k = pos[int(ref[row, 6])]
            got = out6[off + k]
            if row < nref:
                assert lpass[off + k] == -1, "variant %d id %d lost in emulation, alive in oracle" % (v, ref[row, 6])
                assert np.array_equal(got, ref[row, :6]), "variant %d id %d differs" % (v, ref[row, 6])
            else:
                lp = int(lpass[off + k])
                assert lp != -1, "variant %d id %d alive in emulation, lost in oracle" % (v, ref[row, 6])
                assert (lp & ~INHERIT) == int(ref[row, 7])
                cols = [0, 1, 2, 3, 5] if lp & INHERIT else [0, 1, 2, 3, 4, 5]
                same = (got[cols] == ref[row, cols]) | (np.isnan(got[cols]) & np.isnan(ref[row, cols]))
                assert same.all(), "variant %d lost id %d differs" % (v, ref[row, 6])
        off += n
    return total


@pytest.mark.parametrize("span,seed", [(None, 1), (700, 2), (2 * 2083, 3), (1, 4)])
def test_emulated_batch_matches_oracle(span, seed):
    """counts (300, 0, 1, 129, 64) of error-seeded MBA rings, 3 passes, the pass-0
    stop before the RF cavity, spans of one turn / 700 elements / two turns /
    one element, work items in shuffled order"""
    Po = configs.p_central_from_mev(3000.0)
    counts = [300, 0, 1, 129, 64]

    def mk():
        lats = configs.mba_variants(len(counts), seed=7)
        for lat in lats:
            for i, e in enumerate(lat.elements):
                if e.type == S.T_RFCA:
                    lat.set_param(i, "fiducial_mode", 1)  # LIGHT: no beam average needed
        return lats

    beams = _beams(counts)
    r, out6, lpass, cnt = _host_track(beams, mk(), 3, Po, span, seed)
    total = _compare_with_oracle(beams, mk, 3, Po, out6, lpass, cnt)
    assert r == total


def test_emulated_batch_shared_and_zero_volt_cavities():
    """a line holding one cavity struct at two positions, a second cavity, and
    variants whose cavities have no voltage"""
    counts = [200, 70, 0, 33]

    def mk():
        out = []
        for v in range(len(counts)):
            lat, Po = _rf_line(volt=0.0 if v == 3 else 2e6, shared=True, two=True, mode=1)
            lat.set_param(1, "K1", 0.8 + 0.02 * v)
            out.append(lat)
        return out

    Po = _rf_line()[1]
    rng = np.random.default_rng(9)
    beams = []
    for v, c in enumerate(counts):
        b = np.zeros((c, S.NPROP))
        b[:, :4] = rng.normal(0, [5e-3, 5e-4, 4e-3, 4e-4], (c, 4))
        b[:, 5] = rng.normal(0, 2e-3, c)
        b[:, 6] = np.arange(1, c + 1)
        beams.append(b)
    r, out6, lpass, cnt = _host_track(beams, mk(), 4, Po, 5, 11)
    total = _compare_with_oracle(beams, mk, 4, Po, out6, lpass, cnt)
    assert r == total


def test_unshifted_dense_blocks_lose_the_batch():
    """The batched wrong result once seen on the GPU (counts 300, 0, 1, 129, 64,
    three passes: all of variant 0's 300 particles reported lost) is what the
    batched tables give when the dense field blocks of a de-duplicated CSBEND
    keep the offsets of the bend's own term list instead of being shifted to
    where its terms start in the shared list: every bend after the first reads
    the first bend's coefficients.  Planted here on the CPU, the emulated batch
    loses every particle of variant 0 (and of the other variants); without the
    plant it matches the oracle (test_emulated_batch_matches_oracle)."""
    Po = configs.p_central_from_mev(3000.0)
    counts = [300, 0, 1, 129, 64]

    def mk():
        lats = configs.mba_variants(len(counts), seed=7)
        for lat in lats:
            for i, e in enumerate(lat.elements):
                if e.type == S.T_RFCA:
                    lat.set_param(i, "fiducial_mode", 1)
        return lats

    beams = _beams(counts)
    h = harness()
    h.bh_plant_unshifted_dense.argtypes = [C.c_int]
    h.bh_plant_unshifted_dense(1)
    try:
        r, out6, lpass, cnt = _host_track(beams, mk(), 3, Po, None, 1)
    finally:
        h.bh_plant_unshifted_dense(0)
    assert (lpass[:300] != -1).all(), "variant 0: %d of 300 survive the planted tables" % (lpass[:300] == -1).sum()
    assert r == 0
    ref, nref, _, _ = oracle.do_tracking(beams[0], mk()[0], 3, Po)
    assert nref == 300
