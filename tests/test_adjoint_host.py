"""The adjoint arithmetic (elegant_amd/csrc/eh_adjoint.h) compiled for the host
CPU by a test harness (tests/adjoint_host), checked against PyTorch autograd
through the independent restatement in tests/torch_model.py.  The GPU kernels
run the same functions; tests/test_gpu_adjoint.py repeats these checks there."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import torch

from elegant_amd import adjoint, configs
from elegant_amd.lattice import Element, Lattice
from oracle import oracle
from tests.torch_model import TorchRing, forward_mode_projection

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_host")
_h = None


def harness():
    global _h
    if _h is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        _h = C.CDLL(os.path.join(HERE, "libadj_host.so"))
        vp, L = C.c_void_p, C.c_long
        _h.adjh_gradient.argtypes = [vp, L, vp, L, vp, L, L, C.c_double, vp, vp, vp, vp, vp, vp]
        _h.adjh_gradient.restype = L
    return _h


def host_gradient(lat, spec, Po, rows, turns, W):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    n = len(rows)
    slots = np.ascontiguousarray(spec.slot_array(lat), dtype=np.int32)
    W = np.ascontiguousarray(W, dtype=np.float64)
    final = np.zeros((n, 6))
    alive = np.zeros(n, dtype=np.int32)
    gp = np.zeros(max(1, len(spec)))
    gz = np.zeros((n, 6))
    err = C.c_double()
    arr, keep = lat.c_beamline()
    r = harness().adjh_gradient(rows.ctypes.data, n, C.cast(arr, C.c_void_p), len(lat), slots.ctypes.data,
                                len(spec), turns, Po, W.ctypes.data, final.ctypes.data, alive.ctypes.data,
                                gp.ctypes.data, gz.ctypes.data, C.byref(err))
    assert r == n
    return final, alive.astype(bool), gp[:len(spec)], gz, err.value


def torch_gradient(lat, spec, Po, rows, turns, W):
    ring = TorchRing(lat, Po, spec.slot_array(lat), len(spec))
    theta = torch.tensor(spec.values(lat), requires_grad=True)
    z0 = torch.tensor(rows[:, :6], requires_grad=True)
    out = ring.track(z0, theta, turns)
    (out * torch.tensor(W)).sum().backward()
    return theta.grad.numpy(), z0.grad.numpy(), out.detach().numpy()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def mix_ring():
    return Lattice([
        Element("Q", "KQUAD", L=0.3, K1=1.1, N_KICKS=8, DX=1e-4, DY=-2e-4, DZ=3e-4, TILT=0.02, FSE=1e-3,
                HKICK=2e-5, VKICK=-1e-5),
        Element("D", "DRIF", L=0.5),
        Element("S", "KSEXT", L=0.2, K2=-30.0, K1=0.05, J1=0.03, N_KICKS=4, DX=-5e-5),
        Element("B", "CSBEND", L=0.9, ANGLE=0.06, K1=-0.3, K2=3.0, N_KICKS=8, E1=0.02, E2=0.03, HGAP=0.012,
                FINT=0.6),
        Element("E", "EDRIFT", L=0.4),
        Element("QS", "KQUSE", L=0.2, K1=0.4, K2=-12.0, FSE1=1e-3, FSE2=-2e-3, N_KICKS=8, DX=3e-5, DY=1e-5,
                TILT=0.01),
        Element("O", "KOCT", L=0.2, K3=2000.0, TILT=-0.1),
        Element("QD", "KQUAD", L=0.3, K1=-1.0, N_KICKS=8),
        Element("D2", "DRIF", L=0.3, ORDER=1),
    ])


WEIGHTS = np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3])


@pytest.mark.parametrize("ring", ["mixed", "mba"])
def test_host_forward_equals_oracle(ring):
    """the maps compiled for the host reproduce the oracle bit for bit"""
    if ring == "mixed":
        Po = configs.p_central_from_mev(2000.0)
        lat = mix_ring()
        rows = configs.fodo_beam(50, seed=8)
        turns = 4
    else:
        Po = configs.p_central_from_mev(3000.0)
        lat = configs.mba_ring(Po)
        rows = configs.mba_beam(40, seed=12, scale=25.0)
        oracle.do_tracking(rows, lat, 1, Po)  # RF fiducial
        turns = 2
    spec = adjoint.ParamSpec.strengths(lat)
    final, alive, _, _, _ = host_gradient(lat, spec, Po, rows, turns, np.zeros((len(rows), 6)))
    ref, n, _, _ = oracle.do_tracking(rows, lat, turns, Po)
    got = {int(r[6]): r[:6] for r in ref[:n]}
    assert alive.sum() == n
    for i in np.flatnonzero(alive):
        np.testing.assert_array_equal(final[i], got[int(rows[i, 6])])


def test_mixed_ring_gradients():
    Po = configs.p_central_from_mev(2000.0)
    lat = mix_ring()
    spec = adjoint.ParamSpec.strengths(lat) + adjoint.ParamSpec.offsets(lat)
    rows = configs.fodo_beam(24, seed=8)
    rows[:, :4] *= 2.0
    W = np.random.default_rng(1).normal(size=(24, 6)) * WEIGHTS
    final, alive, gp, gz, err = host_gradient(lat, spec, Po, rows, 5, W)
    assert alive.all()
    tp, tz, tout = torch_gradient(lat, spec, Po, rows, 5, W)
    assert rel(final, tout) < 1e-9
    assert rel(gp, tp) < 1e-8, np.c_[gp, tp]
    assert rel(gz, tz) < 1e-8
    assert err < 1e-13


def _elementwise(gp, tp, tol):
    scale = np.maximum(np.abs(tp), 1e-3 * np.max(np.abs(tp)))
    return np.max(np.abs(gp - tp) / scale) < tol


def test_bend_k1_and_fse_gradients():
    """a bend's K1 and FSE as parameters (field per unit K1 and FSE, 1/rho_actual
    and the edge angles growing with FSE) against autograd through the torch
    restatement, parameter by parameter"""
    Po = configs.p_central_from_mev(2000.0)
    lat = mix_ring()
    spec = adjoint.ParamSpec.bends(lat) + adjoint.ParamSpec.strengths(lat)
    assert spec.names[:2] == ["B#3.K1", "B#3.FSE"]
    rows = configs.fodo_beam(24, seed=8)
    rows[:, :4] *= 2.0
    W = np.random.default_rng(3).normal(size=(24, 6)) * WEIGHTS
    final, alive, gp, gz, err = host_gradient(lat, spec, Po, rows, 5, W)
    assert alive.all()
    tp, tz, tout = torch_gradient(lat, spec, Po, rows, 5, W)
    assert rel(final, tout) < 1e-9
    assert _elementwise(gp, tp, 1e-7), np.c_[gp, tp]
    assert rel(gz, tz) < 1e-8


def test_mba_gradients_with_rf():
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.strengths(lat) + adjoint.ParamSpec.offsets(lat)
    rows = configs.mba_beam(6, seed=9, scale=25.0)
    oracle.do_tracking(rows, lat, 1, Po)  # RF fiducial
    W = np.random.default_rng(2).normal(size=(6, 6)) * WEIGHTS
    final, alive, gp, gz, err = host_gradient(lat, spec, Po, rows, 2, W)
    assert alive.all()
    tp, tz, tout = torch_gradient(lat, spec, Po, rows, 2, W)
    assert rel(final, tout) < 1e-8
    assert rel(gp, tp) < 1e-7
    assert rel(gz, tz) < 1e-7
    assert err < 1e-11


def test_forward_mode_ten_turns_host():
    """10 turns of the MBA ring: the adjoint's gradient projected on a random
    direction equals the complex-step (forward-mode) derivative along it"""
    from elegant_amd import optics
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.strengths(lat)
    rows = configs.mba_beam(4, seed=10, scale=25.0)
    oracle.do_tracking(rows, lat, 1, Po)  # RF fiducial
    tw = optics.periodic_twiss(configs.mba_period())
    bx, by = tw["betax"], tw["betay"]
    turns = 10
    W0 = np.zeros((4, 6))
    final, alive, _, _, _ = host_gradient(lat, spec, Po, rows, turns, W0)
    assert alive.all()
    G = np.zeros((4, 6))
    G[:, 0] = 2 * final[:, 0] / bx / 4
    G[:, 2] = 2 * final[:, 2] / by / 4
    _, _, gp, _, err = host_gradient(lat, spec, Po, rows, turns, G)
    v = np.random.default_rng(3).normal(size=len(spec)) * np.abs(spec.values(lat))
    dL = forward_mode_projection(lat, spec, Po, rows, turns, bx, by, v)
    assert abs(float(np.dot(gp, v)) - dL) <= 1e-7 * abs(dL), (float(np.dot(gp, v)), dL)
    assert err < 1e-11


def test_batch_tables_match_single_lattice():
    """the batched path's de-duplicated descriptor tables (host part of
    eh_batch_track) give every variant's elements exactly the single-lattice
    descriptors and field values"""
    Po = configs.p_central_from_mev(3000.0)
    lats = configs.mba_variants(3, seed=5)
    arrs = [l.c_beamline() for l in lats]
    ptrs = (C.c_void_p * 3)(*[C.cast(a, C.c_void_p) for a, _ in arrs])
    h = harness()
    h.adjh_batch_tables_check.argtypes = [C.c_void_p, C.c_long, C.c_long, C.c_double]
    h.adjh_batch_tables_check.restype = C.c_long
    assert h.adjh_batch_tables_check(C.cast(ptrs, C.c_void_p), 3, len(lats[0]), Po) == 0


def test_division_through_reciprocal_pair():
    """eh_div_recip (the CSBEND kick's divisions by rho0 and rho_actual) equals
    IEEE division bit for bit: 4e6 random pairs, half of them with quotients
    next to a rounding midpoint, plus signed zeros"""
    h = harness()
    h.adjh_div_recip_check.argtypes = [C.c_long, C.c_uint64]
    h.adjh_div_recip_check.restype = C.c_long
    assert h.adjh_div_recip_check(4000000, 12345) == 0
