#!/usr/bin/env python3
"""Throughput of elegant's element-tracking path on one MI355X.

Headline workload: the 40-cell multi-bend-achromat-style ring at 3 GeV (per
cell 5 gradient CSBENDs with N_KICKS=16, 10 KQUAD, 6 KSEXT, 2 KOCT; one RFCA of
2 MV at h=1296 and 170 deg; MAXAMP 20 mm) with a 1000 x 1000 dynamic-aperture
grid (x in +-15 mm, y in +-5 mm, other coordinates 0), i.e. 1e6 particles, in
FP64 with loss compaction, through the fused kernels.  One step = one turn of
the whole ring for the whole beam; W untimed turns, then K timed turns.
Particles are resident in HBM before the timed region starts.

Further legs in the same JSON line:
  batched  256 error-seeded copies of that ring (FSE 1e-3 on KQUAD/KSEXT,
           DX/DY 30 um, 2-sigma cutoff), lognormal particle counts in [1e3, 1e5]
           (4.9e6 in total), 6D Gaussian beams, all variants in one persistent
           launch per span captured in a HIP graph, launch shape from the
           autotuner;
  plain    the per-element drop-in path (one kernel per reference call, host
           synchronisation and compaction after each) on the 10-cell FODO ring
           with 1e4 particles: the launch-bound small-lattice case the batched
           path is compared with;
  cpu_baseline  the checker compiled like the reference (gcc -O3 -m64, C-library
           transcendentals) on a bounded strided sample of the headline
           workload, one single-threaded worker per host core.

The roofline entry is for the fused kernel of the headline leg: algorithmic
FP64 operations per launch (census per element type x particle-element units
per launch) over the average launch time, both measured inside this process
with HIP events on the library's stream.
"""
import argparse
import json
import multiprocessing as mp
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle·element·turns/s FP64 on 1 MI355X; batched vs plain; adjoint/forward"
UNIT = "particle·element·turns/s"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak, FMA counted as 2


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name, grid):
    from elegant_amd import configs
    if name == "mba":
        Po = configs.p_central_from_mev(3000.0)
        lat = configs.mba_ring(Po)
        rows = configs.da_grid(grid, grid)
        desc = ("40-cell MBA ring 3 GeV, %dx%d DA grid (x +-15 mm, y +-5 mm), RFCA 2 MV h=1296 170 deg, "
                "MAXAMP 20 mm" % (grid, grid))
    else:
        Po = configs.p_central_from_mev(1000.0)
        lat = configs.fodo_ring()
        rows = configs.fodo_beam(grid * grid)
        desc = "10-cell FODO ring 1 GeV, %d-particle matched 6D Gaussian, MAXAMP 20 mm" % (grid * grid)
    return lat, rows, Po, desc


_CPU = {}


def _cpu_init(name, grid, barrier):
    """worker set-up outside the timed window: the checker, the lattice and the
    grid are built once per worker before the barrier"""
    from oracle import oracle
    oracle.load(libm=True)
    lat, rows, Po, _ = workload(name, grid)
    _CPU.update(lat=lat, rows=rows, Po=Po, barrier=barrier)


def _cpu_worker(args):
    idx, turns = args
    from oracle import oracle
    sub = _CPU["rows"][idx]
    lat = _CPU["lat"].copy()
    _CPU["barrier"].wait()
    t0 = time.time()
    _, n, eff, _ = oracle.do_tracking(sub, lat, turns, _CPU["Po"], libm=True)
    return eff, t0, time.time()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(name, grid, sample, turns, repeats=2, single_sample=None):
    """the checker with C-library transcendentals on a strided sample of the same
    workload: one single-threaded worker per host core (particle shards, the
    reference's MPI decomposition), timed from a barrier that every worker
    reaches after building its lattice and grid (pool start-up excluded) to the
    last worker's end, median of `repeats` runs; and a genuine one-core run of
    a 1/cores share of the sample.  The pool is sized to this process's CPU
    affinity, at most 16 workers (the host-core share one GPU is given)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    _, rows, _, _ = workload(name, grid)
    n_total = len(rows)
    idx = np.linspace(0, n_total - 1, sample).astype(np.int64)
    shards = np.array_split(idx, cores)
    rates = []
    ctx = mp.get_context("fork")
    for _ in range(repeats):
        barrier = ctx.Barrier(cores)
        with ctx.Pool(cores, initializer=_cpu_init, initargs=(name, grid, barrier)) as pool:
            res = pool.map(_cpu_worker, [(s, turns) for s in shards], chunksize=1)
        units = sum(r[0] for r in res)
        wall = max(r[2] for r in res) - min(r[1] for r in res)
        rates.append(units / wall)
    # one core: a single worker process on its own share of the sample
    n1 = single_sample or max(1, sample // cores)
    idx1 = np.linspace(0, n_total - 1, n1).astype(np.int64)
    barrier = ctx.Barrier(1)
    with ctx.Pool(1, initializer=_cpu_init, initargs=(name, grid, barrier)) as pool:
        eff1, t0, t1 = pool.map(_cpu_worker, [(idx1, turns)])[0]
    one = eff1 / (t1 - t0)
    return {"value": float(np.median(rates)), "unit": UNIT, "cores": cores, "kind": "port",
            "sample": "%d of the %d particles (strided), %d turns, %d cores, median of %d runs timed from a "
                      "barrier after worker set-up; one-core run: %d particles, %d turns; oracle with "
                      "C-library sin/cos/asin/log, gcc -O3 -m64, no FMA" % (sample, n_total, turns, cores,
                                                                            repeats, n1, turns),
            "one_core_value": float(one), "one_core_particles": n1,
            "cpu": cpu_model()}


COUNTERS = os.path.join(ROOT, "profiles", "k_fused_mba_1e6_counters.md")


def counter_keys():
    """per-unit hardware FP64 operations and per-particle HBM traffic of k_fused,
    from the committed rocprofv3 counter summary (measured on the headline shape)"""
    try:
        for line in open(COUNTERS):
            if line.strip().startswith("bench-keys:"):
                return {k: float(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}
    except OSError:
        pass
    return {}


def leg_fused(args, census):
    from elegant_amd import tracking
    lat, rows, Po, desc = workload(args.config, args.grid)
    cen = census.lattice_census(lat)
    log("workload: %s; %d particles, %d counted elements/turn, %.4g flops + %d T per particle-turn" %
        (desc, len(rows), cen["units"], cen["flops"], cen["T"]))
    beam = tracking.Beam(rows)
    t0 = time.perf_counter()
    if args.warmup:
        n, _, _ = tracking.track(beam, lat, args.warmup, Po, mode=args.mode, turns_per_launch=args.turns_per_launch)
        log("warmup: %d turns, %d alive, %.3f s" % (args.warmup, n, time.perf_counter() - t0))
    t0 = time.perf_counter()
    n, st, _ = tracking.track(beam, lat, args.steps, Po, mode=args.mode, turns_per_launch=args.turns_per_launch)
    wall = time.perf_counter() - t0
    beam.free()
    log("timed: %d turns, %d alive, %.3f s wall, %.3f s in kernels, %d launches" %
        (args.steps, n, wall, st["kernel_ms"] / 1e3, st["launches"]))
    return desc, len(rows), n, wall, st, cen


def leg_lossy(args):
    """the headline ring with a grid reaching past the 20 mm MAXAMP and the
    dynamic aperture (x in +-30 mm, y in +-25 mm): particles are lost from the
    first element on, so every launch compacts survivors and appends lost
    particles to the graveyard; timed from the first turn (pass-0 RF fiducial
    included)"""
    from elegant_amd import configs, tracking
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    rows = configs.da_grid(args.grid, args.grid, xmax=0.030, ymax=0.025)
    beam = tracking.Beam(rows)
    t0 = time.perf_counter()
    n, st, _ = tracking.track(beam, lat, args.lossy_steps, Po, mode="fused", turns_per_launch=args.turns_per_launch)
    wall = time.perf_counter() - t0
    beam.free()
    units = st["particle_element_turns"]
    log("lossy timed: %d turns, %d of %d alive at the end, %.3f s wall" % (args.lossy_steps, n, len(rows), wall))
    return {"value": units / wall, "unit": UNIT, "turns": args.lossy_steps, "particles": len(rows), "survivors": n,
            "lost": len(rows) - n, "wall_s": round(wall, 4), "launches": st["launches"],
            "workload": "40-cell MBA ring 3 GeV, %dx%d grid x +-30 mm, y +-25 mm (past the 20 mm MAXAMP), fused, "
                        "timed from the first turn" % (args.grid, args.grid)}


def leg_batched(args):
    from elegant_amd import autotune, configs, tracking
    Po = configs.p_central_from_mev(3000.0)
    counts = configs.variant_counts(args.variants)
    t0 = time.perf_counter()
    lats = configs.mba_variants(args.variants, p_central=Po)
    beams = configs.variant_beams(counts)
    log("batched: %d variants, %d particles, setup %.1f s" % (args.variants, counts.sum(), time.perf_counter() - t0))
    t0 = time.perf_counter()
    tune, trials = autotune.tune_batch(beams, lats, Po, passes=1, cache=None)
    log("autotune: %s (%.1f s, %d trials)" % (tune, time.perf_counter() - t0, len(trials)))
    batch = tracking.Batch(beams)
    batch.track(lats, 1, Po, use_graph=True, **tune)  # warmup: fixes RF fiducials
    t0 = time.perf_counter()
    n, st = batch.track(lats, args.batch_steps, Po, use_graph=True, **tune)
    wall = time.perf_counter() - t0
    batch.free()
    units = st["particle_element_turns"]
    log("batched timed: %d turns, %d alive, %.3f s wall, %.3f s on the stream, %d launches" %
        (args.batch_steps, n, wall, st["kernel_ms"] / 1e3, st["launches"]))
    return {"value": units / wall, "unit": UNIT, "turns": args.batch_steps, "variants": args.variants,
            "particles": int(counts.sum()), "survivors": n, "wall_s": round(wall, 4),
            "stream_s": round(st["kernel_ms"] / 1e3, 4), "launches": st["launches"], "tune": tune,
            "workload": "256 error-seeded MBA rings (FSE 1e-3 KQUAD/KSEXT, DX/DY 30 um, 2-sigma cutoff), "
                        "lognormal counts in [1e3, 1e5], 6D Gaussian beams, 3 GeV, one graph-captured "
                        "persistent launch per span"}


def leg_plain(args):
    from elegant_amd import tracking
    lat, rows, Po, desc = workload("fodo", 100)
    beam = tracking.Beam(rows)
    tracking.track(beam, lat, 1, Po, mode="plain")
    t0 = time.perf_counter()
    n, st, _ = tracking.track(beam, lat, args.plain_steps, Po, mode="plain")
    wall = time.perf_counter() - t0
    beam.free()
    units = st["particle_element_turns"]
    log("plain timed: %d turns, %d alive, %.3f s wall, %d launches" % (args.plain_steps, n, wall, st["launches"]))
    return {"value": units / wall, "unit": UNIT, "turns": args.plain_steps, "particles": len(rows),
            "survivors": n, "wall_s": round(wall, 4), "launches": st["launches"],
            "workload": desc + ", per-element drop-in path"}


def leg_adjoint(args):
    """adjoint/forward cost: L = mean over survivors of x^2/beta_x + y^2/beta_y after
    the turns, gradient with respect to every KQUAD K1, KSEXT K2 and KOCT K3 of the
    MBA ring (720 strengths), checkpoints every --adj-ckpt turns"""
    from elegant_amd import adjoint, configs, optics, tracking
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    tw = optics.periodic_twiss(configs.mba_period())
    bx, by = tw["betax"], tw["betay"]
    rows = configs.mba_beam(args.adj_particles, seed=21, scale=args.adj_scale)
    spec = adjoint.ParamSpec.strengths(lat)
    turns = args.adj_turns
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=args.adj_ckpt)
    adj.fix_rf(rows)
    # plain forward tracking of the same beam, same turns: the reference cost
    beam = tracking.Beam(rows)
    t0 = time.perf_counter()
    n_f, st_f, _ = tracking.track(beam, lat, turns, Po, mode="fused", turns_per_launch=args.turns_per_launch)
    t_fwd = time.perf_counter() - t0
    beam.free()
    t0 = time.perf_counter()
    final, alive = adj.forward(rows)
    t_rec = time.perf_counter() - t0
    n = alive.sum()
    G = np.zeros((len(rows), 6))
    G[alive, 0] = 2 * final[alive, 0] / bx / n
    G[alive, 2] = 2 * final[alive, 2] / by / n
    t0 = time.perf_counter()
    gp, _ = adj.backward(G)
    t_bwd = time.perf_counter() - t0
    st = adj.stats
    adj.free()
    # the same gradient with every anchor kept by the forward pass (no re-tracking):
    # 48 B per particle per anchor per turn, affordable in HBM at this size
    adj1 = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=1)
    t0 = time.perf_counter()
    adj1.forward(rows)
    gp1, _ = adj1.backward(G)
    t_all1 = time.perf_counter() - t0
    st1 = adj1.stats
    adj1.free()
    log("adjoint, anchors kept (checkpoint every turn): %.2f s forward+backward, %.1f MB of anchors; "
        "gradient agrees to %.2e" % (t_all1, st1["anchor_bytes"] / 1e6,
                                     np.linalg.norm(gp1 - gp) / max(np.linalg.norm(gp), 1e-300)))
    log("adjoint: %d turns, %d particles, %d survivors; forward %.2f s, recorded forward %.2f s, backward %.2f s "
        "(GPU: %.2f fwd, %.2f re-track, %.2f reverse); recon err %.2e" % (
            turns, len(rows), n, t_fwd, t_rec, t_bwd, st["forward_ms"] / 1e3, st["retrack_ms"] / 1e3,
            st["backward_ms"] / 1e3, st["max_recon_err"]))
    return {"ratio": round((t_rec + t_bwd) / t_fwd, 3), "backward_over_forward": round(t_bwd / t_fwd, 3),
            "forward_s": round(t_fwd, 3), "recorded_forward_s": round(t_rec, 3), "backward_s": round(t_bwd, 3),
            "gpu_forward_s": round(st["forward_ms"] / 1e3, 3), "gpu_retrack_s": round(st["retrack_ms"] / 1e3, 3),
            "gpu_reverse_s": round(st["backward_ms"] / 1e3, 3), "turns": turns, "particles": len(rows),
            "survivors": int(n), "parameters": len(spec), "checkpoint_every": args.adj_ckpt,
            "max_recon_err": st["max_recon_err"], "grad_norm": float(np.linalg.norm(gp)),
            "ratio_anchors_kept": round(t_all1 / t_fwd, 3), "anchors_kept_bytes": st1["anchor_bytes"],
            "grad_rel_diff_anchors_kept": float(np.linalg.norm(gp1 - gp) / max(np.linalg.norm(gp), 1e-300)),
            "workload": "MBA ring 3 GeV, 6D Gaussian beam (emittances x%g of 5 nm / 50 pm), L = mean(x^2/bx + "
                        "y^2/by) after the turns, d/d(720 K1/K2/K3)" % args.adj_scale}


def probe_elements(particles=1000000, turns=4):
    """time k_fused on rings of 100 copies of one element type (per-type cost)"""
    import math
    from elegant_amd import configs, tracking
    from elegant_amd.lattice import Element, Lattice
    Po = configs.p_central_from_mev(3000.0)
    ang = 2 * math.pi / 200
    cases = {
        "CSBEND K1 N_KICKS=16": Element("B", "CSBEND", L=0.8, ANGLE=ang, K1=-0.5, N_KICKS=16, EDGE1_EFFECTS=0,
                                        EDGE2_EFFECTS=0),
        "CSBEND N_KICKS=16": Element("B", "CSBEND", L=0.8, ANGLE=ang, N_KICKS=16, EDGE1_EFFECTS=0, EDGE2_EFFECTS=0),
        "KQUAD N_KICKS=8": Element("Q", "KQUAD", L=0.25, K1=0.8, N_KICKS=8),
        "KSEXT N_KICKS=4": Element("S", "KSEXT", L=0.2, K2=20.0, N_KICKS=4),
        "KOCT": Element("O", "KOCT", L=0.2, K3=800.0),
        "DRIF": Element("D", "DRIF", L=0.2),
        "EDRIFT": Element("E", "EDRIFT", L=0.2),
    }
    side = int(round(math.sqrt(particles)))
    rows = configs.da_grid(side, side, xmax=1e-3, ymax=1e-3)
    out = {}
    for name, el in cases.items():
        els = []
        for i in range(100):
            prm = dict(el.params)
            if "K1" in prm:
                prm["K1"] = prm["K1"] * (1 if i % 2 else -1)  # alternating gradients keep the beam
            els.append(Element(el.name + str(i), el.type, **prm))
        beam = tracking.Beam(rows)
        lat = Lattice(els)
        tracking.track(beam, lat, 1, Po, mode="fused", turns_per_launch=1)
        n, st, _ = tracking.track(beam, lat, turns, Po, mode="fused", turns_per_launch=turns)
        beam.free()
        ns = st["main_ms"] * 1e6 / max(1.0, st["main_units"])
        out[name] = round(ns, 4)
        log("probe %-22s %8d alive  %.4f ns per particle-element" % (name, n, ns))
    return out


def leg_study(args):
    """error-seed sextupole study (batched adjoint): --study-seeds seeds of the
    MBA ring (FSE/DX/DY errors), --study-particles each, aperture penalty
    softplus((r - r_ap)/1e-4) after --study-turns turns, twelve sextupole
    families (six names split over even and odd cells), checkpoints every
    --study-ckpt turns, --study-steps gradient steps; every evaluation is one
    batched forward + backward over all seeds"""
    from elegant_amd import adjoint, configs, studies
    Po = configs.p_central_from_mev(3000.0)
    lats = configs.mba_variants(args.study_seeds, seed=17, p_central=Po)
    names = ["S0", "S1", "S2", "S3", "S4", "S5"]
    spec = adjoint.ParamSpec.split_families(lats[0], names, groups=2)
    beams = [configs.mba_beam(args.study_particles, seed=500 + v, scale=400.0) for v in range(args.study_seeds)]
    obj = studies.aperture_penalty(r_ap=4e-3, width=1e-4)
    trk = adjoint.BatchAdjointTracker(lats, spec, Po, n_turns=args.study_turns, checkpoint_every=args.study_ckpt)
    theta0 = spec.values(lats[0])
    trk.forward(beams, theta0)  # RF fiducials

    def log_step(h):
        log("study step %d: objective %.6g, |grad| %.3g, step %.3g" % h)

    t0 = time.perf_counter()
    theta, hist = studies.gradient_tuning(trk, beams, obj, theta0, steps=args.study_steps, step_size=0.5,
                                          log=log_step)
    wall = time.perf_counter() - t0
    n_eval = args.study_steps + 1
    # forward particle-element units of one evaluation (every element but the MAXAMP marker)
    units = args.study_seeds * args.study_particles * args.study_turns * float(len(lats[0]) - 1)
    per_eval = wall / n_eval
    trk.free()
    log("study: %d seeds x %d particles x %d turns, %d evaluations in %.2f s (%.2f s each)" %
        (args.study_seeds, args.study_particles, args.study_turns, n_eval, wall, per_eval))
    return {"seconds_per_evaluation": round(per_eval, 3), "evaluations": n_eval, "wall_s": round(wall, 3),
            "forward_backward_units_per_s": units / per_eval, "objective_first": hist[0][1],
            "objective_last": hist[-1][1], "seeds": args.study_seeds, "particles_per_seed": args.study_particles,
            "turns": args.study_turns, "checkpoint_every": args.study_ckpt, "parameters": len(spec),
            "workload": "MBA ring 3 GeV, error seeds (FSE 1e-3, DX/DY 30 um), 6D Gaussian beams (emittances x400), "
                        "L = mean softplus((r - 4 mm)/0.1 mm), 12 sextupole families, batched adjoint"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed turns of the headline leg")
    ap.add_argument("--warmup", type=int, default=1, help="untimed turns before timing")
    ap.add_argument("--config", choices=["mba", "fodo"], default="mba")
    ap.add_argument("--grid", type=int, default=1000, help="grid side (mba) / sqrt(particles) (fodo)")
    ap.add_argument("--mode", choices=["fused", "plain"], default="fused")
    ap.add_argument("--turns-per-launch", type=int, default=4)
    ap.add_argument("--variants", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--lossy-steps", type=int, default=8)
    ap.add_argument("--no-lossy", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=10000)
    ap.add_argument("--cpu-turns", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-turns", type=int, default=100)
    ap.add_argument("--adj-particles", type=int, default=100000)
    ap.add_argument("--adj-ckpt", type=int, default=50)
    ap.add_argument("--adj-scale", type=float, default=400.0, help="beam emittance scale (amplitude ~ DA)")
    ap.add_argument("--study", action="store_true",
                    help="also run the error-seed sextupole study leg (batched adjoint)")
    ap.add_argument("--study-seeds", type=int, default=64)
    ap.add_argument("--study-particles", type=int, default=20000)
    ap.add_argument("--study-turns", type=int, default=20)
    ap.add_argument("--study-ckpt", type=int, default=10)
    ap.add_argument("--study-steps", type=int, default=2)
    ap.add_argument("--probe-elements", action="store_true",
                    help="also time rings of one element type each (per-type cost of the fused kernel)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="first run a plain copy kernel over N doubles (known HBM traffic, for counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (particles and lattice variants "
              "would shard across GPUs by range). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline forks its workers before this process initialises the GPU
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(args.config, args.grid, args.cpu_sample, args.cpu_turns)
        log("cpu baseline: %.4g %s on %d cores (%.1f s)" % (cpu["value"], UNIT, cpu["cores"],
                                                            time.perf_counter() - t0))

    from elegant_amd import _lib, census, tracking
    tracking.init(0)
    if args.copy_calibration:
        _lib.check(tracking.lib().eh_diag_copy(args.copy_calibration, 1))
        log("copy kernel: %d doubles (%d bytes read, %d written)" % (
            args.copy_calibration, 8 * args.copy_calibration, 8 * args.copy_calibration))
    desc, n0, n, wall, st, cen = leg_fused(args, census)
    units = st["particle_element_turns"]
    value = units / wall

    alg_per_unit = cen["algorithmic_per_unit"]
    launches = max(1, st["main_launches"])
    avg_launch_s = st["main_ms"] / 1e3 / launches
    flops_per_launch = st["main_units"] / launches * alg_per_unit
    achieved = flops_per_launch / avg_launch_s / 1e12 if avg_launch_s > 0 else 0.0
    keys = counter_keys() if (args.config == "mba" and args.mode == "fused") else {}
    traffic = None
    hw = None
    if keys:
        traffic = round(keys["traffic_bytes_per_particle_launch"] * n0)
        hw_t = keys["hw_fp64_ops_per_unit"] * st["main_units"] / launches / avg_launch_s / 1e12
        hw = {"achieved": round(hw_t, 3), "frac": round(hw_t / FP64_PEAK_TFLOPS, 4),
              "source": "hardware FP64 ops per unit from profiles/k_fused_mba_1e6_counters.md (FMA = 2) x "
                        "units per launch / live launch time"}
    roofline = {"bound": "fp64_valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                "traffic_source": "FETCH_SIZE + WRITE_SIZE scaled by a copy kernel, "
                                  "profiles/k_fused_mba_1e6_counters.md" if traffic else None,
                "algorithmic_traffic": 112 * n0,
                "hw_fp64": hw,
                "pipe": "FP64 VALU, no MFMA (compute-bound: the vector FP64 pipe; peak is MI355X's FP64 "
                        "vector peak)",
                "kernel": "k_fused" if args.mode == "fused" else "k_plain",
                "algorithmic_flops_per_unit": round(alg_per_unit, 2),
                "units_per_launch": round(st["main_units"] / launches),
                "launches": st["main_launches"],
                "avg_launch_ms": round(avg_launch_s * 1e3, 3)}

    out = {"metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (generated DA grid / seeded Gaussian beams; no reference data files)",
           "config": {"workload": desc, "particles": n0, "turns_timed": args.steps,
                      "elements_per_turn": cen["units"], "mode": args.mode,
                      "turns_per_launch": args.turns_per_launch, "survivors": n},
           "roofline": roofline}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if not args.no_lossy and args.config == "mba":
        out["lossy"] = leg_lossy(args)
    if not args.no_batch:
        out["batched"] = leg_batched(args)
    if not args.no_plain:
        out["plain"] = leg_plain(args)
        if "batched" in out:
            out["batched_over_plain"] = round(out["batched"]["value"] / out["plain"]["value"], 2)
    if args.probe_elements:
        out["element_probe_ns_per_unit"] = probe_elements()
    if not args.no_adjoint:
        out["adjoint"] = leg_adjoint(args)
        out["adjoint_over_forward"] = out["adjoint"]["ratio"]
    if args.study:
        out["study"] = leg_study(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
