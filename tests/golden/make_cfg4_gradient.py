"""Writes cfg4_gradient.npz: the gradient of the error-seed sextupole-tuning
objective, computed by PyTorch autograd through the independent restatement of
the maps (tests/torch_model.py), for tests/test_gpu_studies.py to check the
batched HIP adjoint against.

    python tests/golden/make_cfg4_gradient.py      (about 3 minutes of CPU)

The workload has the shape of the batched sextupole-tuning study: error-seeded
copies of the 40-cell MBA ring (relative strength errors 1e-3 and offsets of
30 um on every KQUAD and KSEXT), 6D Gaussian beams, the soft survival penalty
mean(softplus((r - r_ap) / 1e-4)) averaged over seeds, and twelve knobs (the six
sextupole families split over even and odd cells); it is cut to 4 seeds x 32
particles x 3 turns so that autograd finishes in minutes.  RF fiducials are
taken by the oracle's first pass (the reference's TMEAN rule), as the GPU path
takes them; the test checks that both agree bit for bit."""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from elegant_amd import adjoint, configs  # noqa: E402
from elegant_amd import structs as S  # noqa: E402
from oracle import oracle  # noqa: E402
from tests.torch_model import TorchRing  # noqa: E402

NSEED, NPART, TURNS = 4, 32, 3
ERR_SEED, BEAM_SEED0, SCALE = 9, 60, 25.0
R_AP, WIDTH = 1.5e-3, 1e-4
FAMILIES = ["S0", "S1", "S2", "S3", "S4", "S5"]


def workload():
    Po = configs.p_central_from_mev(3000.0)
    lats = configs.mba_variants(NSEED, seed=ERR_SEED, p_central=Po)
    spec = adjoint.ParamSpec.split_families(lats[0], FAMILIES, groups=2)
    beams = [configs.mba_beam(NPART, seed=BEAM_SEED0 + v, scale=SCALE) for v in range(NSEED)]
    return Po, lats, spec, beams


def rf_fiducials(lat):
    return np.array([s.phase_fiducial for s in lat.structs() if isinstance(s, S.EH_RFCA)])


def main():
    Po, lats, spec, beams = workload()
    theta0 = spec.values(lats[0])
    fids, L_seeds, grads, finals = [], [], [], []
    for v in range(NSEED):
        oracle.do_tracking(beams[v], lats[v], 1, Po)  # first pass fixes the RF fiducial in the structs
        fids.append(rf_fiducials(lats[v]))
        ring = TorchRing(lats[v], Po, spec.slot_array(lats[v]), len(spec))
        theta = torch.tensor(theta0, requires_grad=True)
        out = ring.track(torch.tensor(beams[v][:, :6]), theta, TURNS,
                         progress=lambda t: print("seed %d turn %d" % (v, t + 1), flush=True))
        r = torch.sqrt(out[:, 0] ** 2 + out[:, 2] ** 2)
        L = torch.nn.functional.softplus((r - R_AP) / WIDTH).mean()
        L.backward()
        L_seeds.append(float(L.detach()))
        grads.append(theta.grad.numpy().copy())
        finals.append(out.detach().numpy())
        print("seed %d L %.6g |grad| %.3g" % (v, L_seeds[-1], np.linalg.norm(grads[-1])), flush=True)
    np.savez(os.path.join(HERE, "cfg4_gradient.npz"), theta=theta0, fiducials=np.array(fids),
             L=np.mean(L_seeds), grad=np.mean(grads, axis=0), finals=np.array(finals),
             r_ap=R_AP, width=WIDTH)


if __name__ == "__main__":
    main()
