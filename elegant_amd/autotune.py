"""Launch-shape autotuner for the batched and fused tracking kernels.

The batched path (one persistent launch per span of elements over every lattice
variant) has three knobs: the workgroup size, the number of resident waves per
CU the persistent grid is sized for, and the number of elements each launch
covers (span).  The single-lattice fused path has one: turns per launch.  The
best values depend on the lattice (register pressure of the element mix, how
often particles are lost) and on the particle count, so they are measured, not
guessed: each candidate tracks a copy of the real beams for a few turns and the
one with the least device time per particle-element-turn wins.

Trials run on fresh copies of the beams, so the caller's beams are untouched; RF
cavity fiducials are fixed first (one untimed turn) so every candidate times the
same steady-state work.  Winners are cached per (device, lattice signature,
particle-count bucket) in a JSON file.
"""
import hashlib
import json
import math
import os

import numpy as np

from . import structs as S
from . import tracking

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "elegant_amd", "autotune.json")

BLOCKS = (64, 128, 256)
WAVES_PER_CU = (4, 8, 16, 24)
SPAN_FACTORS = (0.25, 0.5, 1, 2, 4)   # elements per launch, in units of the lattice length


def lattice_signature(lat):
    h = hashlib.sha1()
    for e in lat.elements:
        h.update(S.TYPE_NAMES[e.type].encode())
        for k in sorted(e.params):
            if k.upper() in ("N_KICKS", "N_SLICES", "INTEGRATION_ORDER", "EXPANSION_ORDER", "ORDER"):
                h.update(("%s=%s" % (k, e.params[k])).encode())
    return h.hexdigest()[:16]


def _bucket(n):
    return int(2 ** round(math.log2(max(1, n))))


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _time_batch(beams, lattices, p_central, passes, tune):
    b = tracking.Batch(beams)
    try:
        n, st = b.track(lattices, passes, p_central, block=tune["block"], span_elements=tune["span_elements"],
                        waves_per_cu=tune["waves_per_cu"], use_graph=True)
    finally:
        b.free()
    units = max(1.0, st["particle_element_turns"])
    return st["kernel_ms"] / units, st


def tune_batch(beams, lattices, p_central, passes=2, cache=CACHE, device_name="", verbose=False):
    """Pick block, waves_per_cu and span_elements for Batch.track on these inputs.

    Two stages, so the search stays short: workgroup size x resident waves at a
    one-lattice span, then the span at the winning shape.  Returns the tune dict
    (keys block, waves_per_cu, span_elements) and the trial table."""
    ne = len(lattices[0])
    total = int(sum(len(b) for b in beams))
    key = "batch|%s|%s|%d|%d" % (device_name, lattice_signature(lattices[0]), len(beams), _bucket(total))
    data = _load_cache(cache) if cache else {}
    if key in data:
        return data[key]["tune"], data[key]["trials"]
    # fix RF fiducials once (first-encounter semantics), untimed
    warm = tracking.Batch(beams)
    try:
        warm.track(lattices, 1, p_central)
    finally:
        warm.free()
    trials = []
    best = None
    for block in BLOCKS:
        for w in WAVES_PER_CU:
            tune = {"block": block, "waves_per_cu": w, "span_elements": ne}
            t, _ = _time_batch(beams, lattices, p_central, passes, tune)
            trials.append(dict(tune, ns_per_unit=t * 1e6))
            if verbose:
                print("autotune", trials[-1], flush=True)
            if best is None or t < best[0]:
                best = (t, tune)
    shape = dict(best[1])
    for f in SPAN_FACTORS:
        span = max(1, int(round(ne * f)))
        if span == ne:
            continue
        tune = dict(shape, span_elements=span)
        t, _ = _time_batch(beams, lattices, p_central, passes, tune)
        trials.append(dict(tune, ns_per_unit=t * 1e6))
        if verbose:
            print("autotune", trials[-1], flush=True)
        if t < best[0]:
            best = (t, tune)
    if cache:
        data = _load_cache(cache)
        data[key] = {"tune": best[1], "trials": trials}
        _store_cache(cache, data)
    return best[1], trials


def tune_fused(rows, lattice, p_central, passes=4, candidates=(1, 2, 4, 8, 16), cache=CACHE, device_name="",
               verbose=False):
    """Pick turns_per_launch for tracking.track(mode="fused")."""
    key = "fused|%s|%s|%d" % (device_name, lattice_signature(lattice), _bucket(len(rows)))
    data = _load_cache(cache) if cache else {}
    if key in data:
        return data[key]["tune"], data[key]["trials"]
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    warm = tracking.Beam(rows)
    tracking.track(warm, lattice, 1, p_central, mode="fused", turns_per_launch=1)
    warm.free()
    trials = []
    best = None
    for tpl in candidates:
        beam = tracking.Beam(rows)
        try:
            _, st, _ = tracking.track(beam, lattice, passes, p_central, mode="fused", turns_per_launch=tpl)
        finally:
            beam.free()
        t = st["kernel_ms"] / max(1.0, st["particle_element_turns"])
        trials.append({"turns_per_launch": tpl, "ns_per_unit": t * 1e6})
        if verbose:
            print("autotune", trials[-1], flush=True)
        if best is None or t < best[0]:
            best = (t, {"turns_per_launch": tpl})
    if cache:
        data = _load_cache(cache)
        data[key] = {"tune": best[1], "trials": trials}
        _store_cache(cache, data)
    return best[1], trials
