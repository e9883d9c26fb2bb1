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
