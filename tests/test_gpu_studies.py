"""Study drivers on the GPU: grid dynamic aperture and gradient tuning."""
import numpy as np
import pytest

from elegant_amd import adjoint, configs, optics, studies, tracking
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_da_grid_matches_oracle_survivors():
    print("da grid driver", flush=True)
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    da = studies.dynamic_aperture_grid(lat, Po, 21, 11, 0.03, 0.012, 3)
    rows = configs.da_grid(21, 11, 0.03, 0.012)
    ref, n, _, _ = oracle.do_tracking(rows, configs.mba_ring(Po), 3, Po)
    alive = np.zeros(len(rows), dtype=bool)
    alive[ref[:n, 6].astype(int) - 1] = True
    assert (da["mask"].ravel() == alive).all()
    assert da["survivors"] == n and 0 < n < len(rows)
    assert da["area"] > 0


def test_gradient_tuning_descends():
    print("gradient tuning", flush=True)
    Po = configs.p_central_from_mev(1000.0)
    tw = optics.periodic_twiss(configs.Lattice(configs.fodo_cell()))
    bx, by = tw["betax"], tw["betay"]
    lats = [configs.fodo_ring() for _ in range(2)]
    spec = [adjoint.ParamSpec.families(l, ["SF", "SD"]) for l in lats]
    beams = [configs.fodo_beam(512, seed=s)[:, :] * 1.0 for s in (1, 2)]
    for b in beams:
        b[:, :4] *= 8.0
    trackers = [adjoint.AdjointTracker(l, s, Po, n_turns=20, checkpoint_every=10) for l, s in zip(lats, spec)]

    def objective(final, alive):
        f = final[alive]
        n = max(1, alive.sum())
        L = float(np.mean(f[:, 0] ** 2 / bx + f[:, 2] ** 2 / by)) if len(f) else 0.0
        G = np.zeros_like(final)
        G[alive, 0] = 2 * final[alive, 0] / bx / n
        G[alive, 2] = 2 * final[alive, 2] / by / n
        return L, G

    theta0 = spec[0].values(lats[0])
    theta, hist = studies.gradient_tuning(trackers, beams, objective, theta0, steps=4, step_size=1.0)
    vals = [h[1] for h in hist]
    assert all(b <= a for a, b in zip(vals, vals[1:]))
    assert vals[-1] < vals[0]


def test_batched_seed_tuning_matches_per_seed_trackers():
    """the sextupole-family study on error seeds (aperture penalty, twelve
    split families) gives the same descent with one BatchAdjointTracker as with
    one AdjointTracker per seed"""
    print("batched seed tuning", flush=True)
    Po = configs.p_central_from_mev(3000.0)
    nseed, npart, turns = 4, 256, 6
    names = ["S0", "S1", "S2", "S3", "S4", "S5"]
    beams = [configs.mba_beam(npart, seed=60 + v, scale=400.0) for v in range(nseed)]
    obj = studies.aperture_penalty(r_ap=4e-3, width=1e-4)

    def run(batched):
        lats = configs.mba_variants(nseed, seed=9, p_central=Po)
        spec = adjoint.ParamSpec.split_families(lats[0], names, groups=2)
        theta0 = spec.values(lats[0])
        if batched:
            trk = adjoint.BatchAdjointTracker(lats, spec, Po, n_turns=turns, checkpoint_every=3)
        else:
            trk = [adjoint.AdjointTracker(l, spec, Po, n_turns=turns, checkpoint_every=3) for l in lats]
        theta, hist = studies.gradient_tuning(trk, beams, obj, theta0, steps=3, step_size=0.5)
        return theta, hist

    tb, hb = run(True)
    ts, hs = run(False)
    vb = np.array([h[1] for h in hb])
    vs = np.array([h[1] for h in hs])
    print("batched", vb, "per seed", vs, flush=True)
    assert np.allclose(vb, vs, rtol=1e-10, atol=0)
    assert np.allclose(tb, ts, rtol=1e-10, atol=1e-14)
    assert vb[-1] < vb[0]


def test_batched_seed_gradient_vs_autograd_fixture():
    """the batched HIP adjoint's gradient of the error-seed sextupole-tuning
    objective (error-seeded MBA rings, softplus aperture penalty averaged over
    seeds, twelve split sextupole families, a checkpoint inside the run) equals
    PyTorch autograd through the independent restatement of the maps, frozen in
    tests/golden/cfg4_gradient.npz by tests/golden/make_cfg4_gradient.py"""
    import os
    import sys
    print("batched seed gradient vs autograd fixture", flush=True)
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_cfg4_gradient as mk
    ref = np.load(os.path.join(here, "cfg4_gradient.npz"))
    Po, lats, spec, beams = mk.workload()
    np.testing.assert_array_equal(spec.values(lats[0]), ref["theta"])
    bt = adjoint.BatchAdjointTracker(lats, spec, Po, n_turns=mk.TURNS, checkpoint_every=2)
    res = bt.forward(beams, ref["theta"])
    for v in range(mk.NSEED):
        np.testing.assert_array_equal(mk.rf_fiducials(lats[v]), ref["fiducials"][v])  # TMEAN, bit for bit
        assert res[v][1].all()
        fin, tf = res[v][0], ref["finals"][v]
        assert np.linalg.norm(fin - tf) <= 1e-8 * np.linalg.norm(tf)  # the restatement is pinned to ~1e-10
    obj = studies.aperture_penalty(r_ap=float(ref["r_ap"]), width=float(ref["width"]))
    Ls, Gs = zip(*[obj(f, a) for f, a in res])
    gp, _ = bt.backward(list(Gs))
    gp = gp / mk.NSEED
    L = float(np.mean(Ls))
    dl = abs(L - float(ref["L"])) / abs(float(ref["L"]))
    dg = np.linalg.norm(gp - ref["grad"]) / np.linalg.norm(ref["grad"])
    print("L", L, ref["L"], "rel", dl, "grad rel (norm)", dg, flush=True)
    # the softplus argument is (r - r_ap) / 1e-4: the ~1e-10 relative agreement of
    # the final coordinates with the restatement becomes ~1e-10 in L
    assert dl <= 1e-8
    assert np.linalg.norm(gp - ref["grad"]) <= 1e-7 * np.linalg.norm(ref["grad"])
    assert np.all(np.abs(gp - ref["grad"]) <= 1e-6 * np.abs(ref["grad"]) + 1e-7 * np.linalg.norm(ref["grad"]))
    bt.free()
