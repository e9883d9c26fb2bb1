"""Parity report: the HIP path against the oracle over long runs.

    python tests/parity_report.py [--out FILE] [--workers N]

For the 10-cell FODO ring (1e4 particles, 1 GeV, 1000 turns) and a strided
sample of the 40-cell MBA ring's 1000x1000 dynamic-aperture grid (3 GeV, RFCA),
tracks the same seeded particles on the GPU (fused kernels) and with the oracle
twice: once with the library's own sin/cos/asin/log (bit-identical to the GPU
by construction) and once with the C library's, i.e. as elegant itself is
built.  Reports, after 1 turn and after the full run, per coordinate the max and
RMS of |gpu - cpu| / max(|cpu|, floor) over survivors matched by particle ID
(floor = RMS of that coordinate over the survivors), the absolute deviation of
s, and the number of differences in the loss set (ID, loss pass, loss s).

Needs a GPU.  The oracle runs are split over worker processes by particle
and matched back by ID: the FODO ring has no RF cavity, and the MBA ring's RF
fiducial (the mean arrival time of the whole sample on the first pass) is
computed once and pinned in every run, so shards see the same cavity.
"""
import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from tests.helpers import COORD_NAMES, by_id, deviation_stats, loss_mask  # noqa: E402


def _say(msg):
    print(msg, flush=True)


def _oracle_worker(args):
    name, idx, turns, libm, fid = args
    from oracle import oracle
    lat, rows, Po = _case(name, fid)
    out, n, _, _ = oracle.do_tracking(rows[idx], lat, turns, Po, libm=libm)
    return out, n


def _case(name, fid=None):
    """(lattice, rows, Po); fid = (phase_fiducial, t_fiducial) pins every RFCA's
    first-encounter fiducial so that particle shards agree with the whole sample"""
    from elegant_amd import configs, structs as S
    if name == "fodo":
        Po = configs.p_central_from_mev(1000.0)
        return configs.fodo_ring(), configs.fodo_beam(10000), Po
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    if fid is not None:
        for e in lat.elements:
            if e.type == S.T_RFCA:
                e.params.update(fiducial_seen=1, phase_fiducial=fid[0], t_fiducial=fid[1])
    return lat, configs.da_grid(1000, 1000), Po


def fiducial(name, idx):
    """the TMEAN fiducial the whole sample sets at the RF cavity on its first pass"""
    if name != "mba":
        return None
    from elegant_amd import structs as S
    from oracle import oracle
    lat, rows, Po = _case(name)
    oracle.do_tracking(rows[idx], lat, 1, Po)
    rf = [s for s in lat.structs() if isinstance(s, S.EH_RFCA)]
    return (rf[0].phase_fiducial, rf[0].t_fiducial) if rf else None


def _merge(parts):
    alive = np.concatenate([p[0][:p[1]] for p in parts])
    lost = np.concatenate([p[0][p[1]:] for p in parts])
    return np.concatenate([alive, lost]), len(alive)


def _oracle(name, idx, turns, libm, workers, fid):
    shards = [s for s in np.array_split(idx, workers) if len(s)]
    with mp.get_context("fork").Pool(len(shards)) as pool:
        parts = pool.map(_oracle_worker, [(name, s, turns, libm, fid) for s in shards])
    return _merge(parts)


def _gpu(name, idx, turns, fid):
    from elegant_amd import tracking
    lat, rows, Po = _case(name, fid)
    beam = tracking.Beam(rows[idx])
    n, _, _ = tracking.track(beam, lat, turns, Po, mode="fused", turns_per_launch=8)
    out = beam.download()
    beam.free()
    return out, n


def compare(gpu, ng, ref, nr):
    """deviation stats of survivors, max |ds| of survivors, the number of
    differences in the loss mask (ID, loss pass) and max |ds| at loss over
    particles lost on the same pass in both"""
    stats, common = deviation_stats(gpu, ng, ref, nr)
    A, B = by_id(gpu, ng), by_id(ref, nr)
    ids = sorted(set(A) & set(B))
    ds = max((abs(A[i][4] - B[i][4]) for i in ids), default=0.0)
    mg = {(i, p): s for (i, p, s) in loss_mask(gpu, ng)}
    mr = {(i, p): s for (i, p, s) in loss_mask(ref, nr)}
    both = set(mg) & set(mr)
    dsl = max((abs(mg[k] - mr[k]) for k in both), default=0.0)
    return stats, common, ds, len(set(mg) ^ set(mr)), ng, nr, dsl


def _table(title, rows):
    lines = ["### " + title, "",
             "| turns | oracle math | survivors gpu/cpu | " + " | ".join("%s max / rms" % c for c in COORD_NAMES) +
             " | max abs Δs [m] | loss-mask diffs (ID, pass) | max abs Δ(loss s) [m] |",
             "|---|---|---|" + "---|" * len(COORD_NAMES) + "---|---|---|"]
    for turns, kind, (stats, common, ds, ndiff, ng, nr, dsl) in rows:
        cells = ["%.2e / %.2e" % stats[c] if c in stats else "-" for c in COORD_NAMES]
        lines.append("| %d | %s | %d / %d | %s | %.2e | %d | %.2e |" % (turns, kind, ng, nr, " | ".join(cells), ds,
                                                                        ndiff, dsl))
    return "\n".join(lines) + "\n"


def oracle_phase(name, idx, turn_list, workers):
    """oracle runs (worker processes are forked before this process touches the GPU)"""
    fid = fiducial(name, idx)
    refs = {}
    for turns in turn_list:
        for libm in (False, True):
            t0 = time.time()
            refs[turns, libm] = _oracle(name, idx, turns, libm, workers, fid)
            _say("%s: oracle[%s] %d turns, %d survivors (%.1f s)" % (
                name, "libm" if libm else "eh", turns, refs[turns, libm][1], time.time() - t0))
    return fid, refs


def gpu_phase(name, idx, turn_list, fid, refs):
    from elegant_amd import tracking
    tracking.init(0)
    rows = []
    for turns in turn_list:
        t0 = time.time()
        g, ng = _gpu(name, idx, turns, fid)
        _say("%s: gpu %d turns, %d survivors (%.1f s)" % (name, turns, ng, time.time() - t0))
        for libm in (False, True):
            r, nr = refs[turns, libm]
            kind = "C library (as elegant)" if libm else "library's own (bit-exact bar)"
            res = compare(g, ng, r, nr)
            _say("%s vs oracle[%s] after %d turns: max rel %s; loss-set diffs %d" % (
                name, "libm" if libm else "eh", turns,
                ", ".join("%s %.1e" % (c, res[0][c][0]) for c in COORD_NAMES if c in res[0]), res[3]))
            rows.append((turns, kind, res))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "gpurun_out", "parity_report.md"))
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--fodo-turns", type=int, nargs="+", default=[1, 1000])
    ap.add_argument("--mba-turns", type=int, nargs="+", default=[1, 1000])
    ap.add_argument("--mba-sample", type=int, default=256)
    args = ap.parse_args()
    out = ["# Parity report", ""]
    # strided sample of the DA grid restricted to |x| <= 8 mm, |y| <= 3 mm so that
    # most of it survives 1000 turns
    from elegant_amd import configs
    grid = configs.da_grid(1000, 1000)
    inner = np.flatnonzero((np.abs(grid[:, 0]) <= 8e-3) & (np.abs(grid[:, 2]) <= 3e-3))
    sel = inner[np.linspace(0, len(inner) - 1, args.mba_sample).astype(np.int64)]
    idx = np.arange(10000)
    o_mba = oracle_phase("mba", sel, args.mba_turns, args.workers)
    o_fodo = oracle_phase("fodo", idx, args.fodo_turns, args.workers)
    mba = gpu_phase("mba", sel, args.mba_turns, *o_mba)
    fodo = gpu_phase("fodo", idx, args.fodo_turns, *o_fodo)
    out.append(_table("10-cell FODO ring, 1 GeV, 1e4 particles (6D Gaussian, 2-sigma cutoff)", fodo))
    out.append(_table("40-cell MBA ring, 3 GeV, %d DA-grid particles (|x|<=8 mm, |y|<=3 mm, strided)" % len(sel),
                      mba))
    text = "\n".join(out)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write(text)
    print(text)


if __name__ == "__main__":
    main()
