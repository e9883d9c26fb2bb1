"""Reverse-mode adjoint of multi-turn tracking on the GPU.

Checks, in order of strength:
  * the adjoint's forward pass equals ordinary tracking bit for bit (survivors,
    loss set);
  * gradients with respect to strengths, offsets and initial coordinates equal
    PyTorch autograd through an independent restatement of the maps
    (tests/torch_model.py) to 1e-7 relative (norm-wise), with checkpoints every
    turn and with re-tracking between checkpoints;
  * a forward-mode directional derivative (complex step through the same
    restatement) over 10 turns of the MBA ring equals the adjoint's gradient
    projected on the same direction;
  * central finite differences of the GPU forward pass (step 1e-7 relative)
    agree with the adjoint on randomly chosen parameters.
"""
import sys

import numpy as np
import pytest
import torch

from elegant_amd import adjoint, configs, tracking
from elegant_amd.lattice import Element, Lattice
from tests.torch_model import TorchRing, forward_mode_projection

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mix_ring():
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


def _torch_grads(lat, spec, Po, rows, turns, W):
    ring = TorchRing(lat, Po, spec.slot_array(lat), len(spec))
    theta = torch.tensor(spec.values(lat), requires_grad=True)
    z0 = torch.tensor(rows[:, :6], requires_grad=True)
    out = ring.track(z0, theta, turns)
    (out * torch.tensor(W)).sum().backward()
    return theta.grad.numpy(), z0.grad.numpy(), out.detach().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_forward_matches_tracking():
    _say("adjoint forward vs fused tracking")
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    rows = configs.da_grid(40, 25, xmax=0.03, ymax=0.012)
    spec = adjoint.ParamSpec.strengths(lat)
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=5, checkpoint_every=2)
    final, alive = adj.forward(rows)
    ref_lat = configs.mba_ring(Po)
    beam = tracking.Beam(rows)
    tracking.track(beam, ref_lat, 1, Po, mode="fused")  # RF fiducial from pass 0, as fix_rf does
    beam.free()
    beam = tracking.Beam(rows)
    n, _, _ = tracking.track(beam, ref_lat, 5, Po, mode="fused")
    out = beam.download()
    assert alive.sum() == n
    assert 0 < n < len(rows)
    ids = out[:n, 6].astype(int) - 1
    assert set(ids) == set(np.flatnonzero(alive))
    np.testing.assert_array_equal(final[ids], out[:n, :6])


@pytest.mark.parametrize("ckpt", [1, 2])
def test_gradients_vs_autograd_mixed_ring(ckpt):
    _say("adjoint vs autograd, element mix, checkpoint every %d" % ckpt)
    Po = configs.p_central_from_mev(2000.0)
    lat = _mix_ring()
    spec = adjoint.ParamSpec.strengths(lat) + adjoint.ParamSpec.offsets(lat)
    rows = configs.fodo_beam(64, seed=8)
    rows[:, :4] *= 2.0
    turns = 5
    rng = np.random.default_rng(1)
    W = rng.normal(size=(len(rows), 6)) * np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3])
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=ckpt)
    final, alive = adj.forward(rows)
    assert alive.all()
    gp, gz = adj.backward(W)
    tp, tz, tout = _torch_grads(lat, spec, Po, rows, turns, W)
    assert _rel(final, tout) < 1e-9
    assert _rel(gp, tp) < 1e-7, (gp, tp)
    assert _rel(gz, tz) < 1e-7
    assert adj.stats["max_recon_err"] < 1e-12


@pytest.mark.parametrize("ckpt", [1, 2])
def test_bend_gradients_vs_autograd(ckpt):
    """K1 and FSE of the MBA ring's bends (and the quadrupoles' K), parameter by
    parameter against autograd through the torch restatement"""
    _say("adjoint vs autograd, MBA bends K1 and FSE, checkpoint every %d" % ckpt)
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.bends(lat) + adjoint.ParamSpec.strengths(lat)
    rows = configs.mba_beam(32, seed=10, scale=25.0)
    turns = 2
    W = np.random.default_rng(5).normal(size=(len(rows), 6)) * np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3])
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=ckpt)
    final, alive = adj.forward(rows)
    assert alive.all()
    gp, gz = adj.backward(W)
    tp, tz, tout = _torch_grads(lat, spec, Po, rows, turns, W)
    assert _rel(final, tout) < 1e-8
    nb = len(adjoint.ParamSpec.bends(lat))
    scale = np.maximum(np.abs(tp[:nb]), 1e-3 * np.max(np.abs(tp[:nb])))
    assert np.max(np.abs(gp[:nb] - tp[:nb]) / scale) < 1e-6, np.c_[gp[:nb], tp[:nb]][:8]
    assert _rel(gp, tp) < 1e-7
    assert _rel(gz, tz) < 1e-7


@pytest.mark.parametrize("ckpt", [1, 2])
def test_gradients_vs_autograd_mba(ckpt):
    _say("adjoint vs autograd, MBA ring with RF, checkpoint every %d" % ckpt)
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.strengths(lat) + adjoint.ParamSpec.offsets(lat)
    rows = configs.mba_beam(32, seed=9, scale=25.0)
    turns = 3
    rng = np.random.default_rng(2)
    W = rng.normal(size=(len(rows), 6)) * np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3])
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=ckpt)
    final, alive = adj.forward(rows)
    assert alive.all()
    gp, gz = adj.backward(W)
    tp, tz, tout = _torch_grads(lat, spec, Po, rows, turns, W)
    assert _rel(final, tout) < 1e-8
    assert _rel(gp, tp) < 1e-7
    assert _rel(gz, tz) < 1e-7
    assert adj.stats["max_recon_err"] < 1e-11


def test_forward_mode_ten_turns():
    _say("adjoint vs forward mode (complex step), MBA ring, 10 turns")
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.strengths(lat)
    rows = configs.mba_beam(4, seed=10, scale=25.0)
    turns = 10
    tw = configs.optics.periodic_twiss(configs.mba_period())
    bx, by = tw["betax"], tw["betay"]
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=5)
    final, alive = adj.forward(rows)
    assert alive.all()
    n = len(rows)
    G = np.zeros((n, 6))
    G[:, 0] = 2 * final[:, 0] / bx / n
    G[:, 2] = 2 * final[:, 2] / by / n
    gp, _ = adj.backward(G)
    v = np.random.default_rng(3).normal(size=len(spec)) * np.abs(spec.values(lat))
    dL = forward_mode_projection(lat, spec, Po, rows, turns, bx, by, v,
                                 progress=lambda t: _say("  complex-step turn %d" % (t + 1)))
    proj = float(np.dot(gp, v))
    assert abs(proj - dL) <= 1e-7 * abs(dL), (proj, dL)


def test_finite_differences():
    _say("adjoint vs central finite differences, MBA ring, 20 turns")
    Po = configs.p_central_from_mev(3000.0)
    lat = configs.mba_ring(Po)
    spec = adjoint.ParamSpec.strengths(lat)
    rows = configs.mba_beam(256, seed=11, scale=25.0)
    turns = 20
    tw = configs.optics.periodic_twiss(configs.mba_period())
    bx, by = tw["betax"], tw["betay"]
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=10)

    def loss(final, alive):
        f = final[alive]
        return float(np.mean(f[:, 0] ** 2 / bx + f[:, 2] ** 2 / by))

    theta = spec.values(lat)
    final, alive = adj.forward(rows, theta)
    assert alive.all()
    n = len(rows)
    G = np.zeros((n, 6))
    G[:, 0] = 2 * final[:, 0] / bx / n
    G[:, 2] = 2 * final[:, 2] / by / n
    gp, _ = adj.backward(G)
    pick = np.random.default_rng(4).choice(len(spec), 20, replace=False)

    def central(q, rel):
        h = rel * abs(theta[q])
        tp = theta.copy()
        tp[q] += h
        lp = loss(*adj.forward(rows, tp))
        tm = theta.copy()
        tm[q] -= h
        lm = loss(*adj.forward(rows, tm))
        return (lp - lm) / (2 * h)

    fd7, fd5 = [], []
    for q in pick:
        _say("  finite differences of %s" % spec.names[q])
        fd7.append(central(q, 1e-7))
        fd5.append(central(q, 1e-5))
    adj.forward(rows, theta)
    fd7, fd5 = np.array(fd7), np.array(fd5)
    scale = np.maximum(np.abs(gp[pick]), 1e-3 * np.max(np.abs(gp)))
    err7 = np.abs(fd7 - gp[pick]) / scale
    err5 = np.abs(fd5 - gp[pick]) / scale
    _say("  max deviation: h=1e-7 %.2e, h=1e-5 %.2e" % (err7.max(), err5.max()))
    # at h = 1e-7 the difference quotient is dominated by the rounding noise of 20
    # turns of tracking (the change of the loss is ~1e-14 of its size); at 1e-5 it
    # has converged to the adjoint's value
    assert err7.max() < 0.1, list(zip(pick, fd7, gp[pick]))
    assert err5.max() < 2e-3, list(zip(pick, fd5, gp[pick]))


@pytest.mark.parametrize("ckpt", [1, 2])
def test_batched_variants_equal_individual(ckpt):
    """eh_adjoint_forward_batch over error-seeded variants (uneven counts, one
    variant empty) equals an AdjointTracker per variant: final coordinates and
    d/d(initial) bit for bit, the shared-parameter gradient equal to the sum of
    the per-variant gradients to rounding"""
    _say("batched adjoint vs per-variant adjoint, checkpoint every %d" % ckpt)
    Po = configs.p_central_from_mev(3000.0)
    nvar = 4
    counts = [70, 0, 129, 33]
    lats = configs.mba_variants(nvar, seed=21, p_central=Po)
    spec = adjoint.ParamSpec.families(lats[0], ["S0", "S1", "S2", "S3", "S4", "S5"]) + \
        adjoint.ParamSpec.families(lats[0], ["QA0", "QB1"], kind="DX")
    beams = [configs.mba_beam(c, seed=40 + v, scale=25.0) if c else np.zeros((0, 10)) for v, c in enumerate(counts)]
    turns = 3
    rng = np.random.default_rng(5)
    Ws = [rng.normal(size=(c, 6)) * np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3]) for c in counts]
    bt = adjoint.BatchAdjointTracker(lats, spec, Po, n_turns=turns, checkpoint_every=ckpt)
    theta = spec.values(lats[0])
    res = bt.forward(beams, theta)
    gp, gz = bt.backward(Ws)
    total = np.zeros(len(spec))
    for v in range(nvar):
        if counts[v] == 0:
            assert len(res[v][0]) == 0
            continue
        ref_lat = configs.mba_variants(nvar, seed=21, p_central=Po)[v]
        one = adjoint.AdjointTracker(ref_lat, spec, Po, n_turns=turns, checkpoint_every=ckpt)
        final, alive = one.forward(beams[v], theta)
        g1, z1 = one.backward(Ws[v])
        one.free()
        np.testing.assert_array_equal(res[v][1], alive)
        np.testing.assert_array_equal(res[v][0][alive], final[alive])
        np.testing.assert_array_equal(gz[v], z1)
        total += g1
    assert _rel(gp, total) < 1e-12, (gp, total)
    assert bt.stats["n_survivors"] == sum(int(r[1].sum()) for r in res)
    bt.free()


def test_track_turns_autograd_function():
    """loss.backward() through track_turns gives the gradients of
    AdjointTracker.backward and of PyTorch autograd through the independent
    restatement; a second forward on the same tracker before backward is refused"""
    _say("track_turns torch.autograd.Function")
    Po = configs.p_central_from_mev(2000.0)
    lat = _mix_ring()
    spec = adjoint.ParamSpec.strengths(lat) + adjoint.ParamSpec.offsets(lat)
    rows = configs.fodo_beam(64, seed=12)
    rows[:, :4] *= 2.0
    turns = 4
    W = np.random.default_rng(6).normal(size=(len(rows), 6)) * np.array([1e3, 1e4, 1e3, 1e4, 1.0, 1e3])
    adj = adjoint.AdjointTracker(lat, spec, Po, n_turns=turns, checkpoint_every=2)
    theta = torch.tensor(spec.values(lat), dtype=torch.float64, requires_grad=True)
    z0 = torch.tensor(rows[:, :6], dtype=torch.float64, requires_grad=True)
    final, alive = adjoint.track_turns(theta, z0, adj)
    assert bool(alive.all())
    (final * torch.tensor(W)).sum().backward()
    gp, gz = adj.backward(W)  # the same forward pass, directly
    np.testing.assert_array_equal(theta.grad.numpy(), gp)
    np.testing.assert_array_equal(z0.grad.numpy(), gz)
    tp, tz, tout = _torch_grads(_mix_ring(), spec, Po, rows, turns, W)
    assert _rel(final.detach().numpy(), tout) < 1e-9
    assert _rel(theta.grad.numpy(), tp) < 1e-7
    assert _rel(z0.grad.numpy(), tz) < 1e-7
    # two graph nodes on one tracker: the first node's checkpoints are gone
    th2 = theta.detach().clone().requires_grad_(True)
    f1, _ = adjoint.track_turns(th2, z0.detach(), adj)
    f2, _ = adjoint.track_turns(th2 * 1.0, z0.detach(), adj)
    with pytest.raises(RuntimeError, match="another forward pass"):
        (f1 * torch.tensor(W)).sum().backward()
    adj.free()
