"""Host-side pieces of the study drivers: the aperture penalty objective and
the split sextupole families (no GPU)."""
import math

import numpy as np

from elegant_amd import adjoint, configs, studies


def test_aperture_penalty_value_and_gradient():
    obj = studies.aperture_penalty(2e-3, 1e-4)
    rng = np.random.default_rng(0)
    f = rng.normal(0, 2e-3, size=(60, 6))
    alive = rng.random(60) > 0.15
    L, G = obj(f, alive)
    r = np.hypot(f[:, 0], f[:, 2])
    u = (r - 2e-3) / 1e-4
    expect = (sum(math.log1p(math.exp(v)) for v in u[alive]) + math.log1p(math.exp(10.0)) * (~alive).sum()) / 60
    assert abs(L - expect) <= 1e-12 * abs(expect)
    assert not G[~alive].any() and not G[:, [1, 3, 4, 5]].any()
    for i in np.flatnonzero(alive)[:8]:
        for c in (0, 2):
            h = 1e-9
            fp, fm = f.copy(), f.copy()
            fp[i, c] += h
            fm[i, c] -= h
            fd = (obj(fp, alive)[0] - obj(fm, alive)[0]) / (2 * h)
            # the difference quotient carries ~eps * L / h of rounding noise
            assert abs(fd - G[i, c]) <= 1e-5 * abs(G[i, c]) + 1e-16 * abs(L) / h * 50, (i, c, fd, G[i, c])


def test_aperture_penalty_far_inside_is_flat():
    obj = studies.aperture_penalty(1e-2, 1e-4)
    f = np.zeros((10, 6))
    f[:, 0] = 1e-4
    L, G = obj(f, np.ones(10, dtype=bool))
    assert L < 1e-40 and np.abs(G).max() < 1e-30


def test_split_families_twelve_knobs():
    lat = configs.mba_ring(configs.p_central_from_mev(3000.0))
    names = ["S0", "S1", "S2", "S3", "S4", "S5"]
    spec = adjoint.ParamSpec.split_families(lat, names, groups=2)
    assert len(spec) == 12
    seen = set()
    for occ in spec.slots:
        assert len(occ) == 20  # 40 cells, every other one
        for i, kind in occ:
            assert kind == "K" and lat.elements[i].name in names
            assert i not in seen
            seen.add(i)
    assert len(seen) == 240
    arr = spec.slot_array(lat)
    assert (arr[:, 0] >= 0).sum() == 240
