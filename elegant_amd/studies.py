"""Study drivers on top of the tracking kernels.

dynamic_aperture_grid: the grid mode of elegant's aperture search
(src/aperture_search.c:1236-1520): particles start on an (x, y) grid at
delta = 0 (IDs row-major, x outer, as :1358-1366), all are tracked together,
and the survivors after n turns form the aperture.  The boundary reported here
is, for each x column, the largest |y| reached from y = 0 without a lost grid
point in between, and the area is the sum of the surviving cells' areas.

gradient_tuning: an optimize-style loop (src/optimize.c:912, :2002) that steps
element strengths along the adjoint gradient of a tracking objective instead of
simplex moves, judging each step with the plain forward tracker.  It takes
either a list of AdjointTrackers (one per error seed) or one
BatchAdjointTracker that runs all seeds in one launch per segment.

aperture_penalty: a smooth loss-of-aperture objective, the mean over particles
of softplus((r - r_ap) / width) with r = sqrt(x^2 + y^2) after the turns; a
lost particle counts as a fixed penalty and carries no gradient.
"""
import math

import numpy as np

from . import configs, tracking


def dynamic_aperture_grid(lattice, p_central, nx, ny, xmax, ymax, turns, turns_per_launch=8):
    rows = configs.da_grid(nx, ny, xmax, ymax)
    beam = tracking.Beam(rows)
    n, st, _ = tracking.track(beam, lattice, turns, p_central, mode="fused", turns_per_launch=turns_per_launch)
    out = beam.download()
    beam.free()
    alive = np.zeros(nx * ny, dtype=bool)
    alive[out[:n, 6].astype(np.int64) - 1] = True
    mask = alive.reshape(nx, ny)
    xs = np.linspace(-xmax, xmax, nx)
    ys = np.linspace(-ymax, ymax, ny)
    j0 = int(np.argmin(np.abs(ys)))
    ytop = np.zeros(nx)
    ybot = np.zeros(nx)
    for i in range(nx):
        if not mask[i, j0]:
            continue
        j = j0
        while j + 1 < ny and mask[i, j + 1]:
            j += 1
        ytop[i] = ys[j]
        j = j0
        while j - 1 >= 0 and mask[i, j - 1]:
            j -= 1
        ybot[i] = ys[j]
    dx = xs[1] - xs[0] if nx > 1 else 2 * xmax
    dy = ys[1] - ys[0] if ny > 1 else 2 * ymax
    return {"mask": mask, "x": xs, "y": ys, "y_top": ytop, "y_bottom": ybot, "area": float(mask.sum() * dx * dy),
            "survivors": int(n), "stats": st}


def aperture_penalty(r_ap, width=1e-4, lost=None):
    """objective(final, alive) -> (L, dL/dfinal) with L the mean over all
    particles of softplus((r - r_ap) / width) for survivors and `lost` (default
    softplus(10)) for lost particles"""
    lost_value = math.log1p(math.exp(10.0)) if lost is None else float(lost)

    def objective(final, alive):
        n = len(final)
        G = np.zeros_like(final)
        if n == 0:
            return 0.0, G
        x, y = final[:, 0], final[:, 2]
        r = np.sqrt(x * x + y * y)
        u = (r - r_ap) / width
        sp = np.maximum(u, 0) + np.log1p(np.exp(-np.abs(u)))
        sig = 0.5 * (1 + np.tanh(0.5 * u))
        L = (np.sum(sp[alive]) + lost_value * np.count_nonzero(~alive)) / n
        w = np.where(alive & (r > 0), sig / (width * np.where(r > 0, r, 1.0)) / n, 0.0)
        G[:, 0] = w * x
        G[:, 2] = w * y
        return float(L), G

    return objective


def gradient_tuning(trackers, beams, objective, theta0, steps, step_size, judge=None, log=None):
    """Gradient descent on the mean of objective(final, alive) over several
    error seeds sharing one ParamSpec: `trackers` is a list of AdjointTrackers
    (one per seed, beams a list of particle arrays) or one BatchAdjointTracker
    (all seeds in one launch per segment).

    objective(final, alive) -> (L, dL/dfinal); judge(theta) -> float, the value
    used to accept a step (default: the mean objective of the forward passes).
    Steps that do not improve the main value halve the step size."""
    theta = np.array(theta0, dtype=np.float64)
    batched = hasattr(trackers, "lats")

    def evaluate(th):
        tot, grad = 0.0, np.zeros_like(th)
        if batched:
            res = trackers.forward(beams, th)
            Gs = []
            for final, alive in res:
                L, G = objective(final, alive)
                tot += L
                Gs.append(G)
            gp, _ = trackers.backward(Gs)
            return tot / len(res), gp / len(res)
        for adj, rows in zip(trackers, beams):
            final, alive = adj.forward(rows, th)
            L, G = objective(final, alive)
            gp, _ = adj.backward(G)
            tot += L
            grad += gp
        return tot / len(trackers), grad / len(trackers)

    L, g = evaluate(theta)
    best = judge(theta) if judge else L
    history = [(0, best, float(np.linalg.norm(g)), step_size)]
    for k in range(1, steps + 1):
        trial = theta - step_size * g / max(np.linalg.norm(g), 1e-300)
        Lt, gt = evaluate(trial)
        val = judge(trial) if judge else Lt
        if val < best:
            theta, g, best = trial, gt, val
        else:
            step_size *= 0.5
        history.append((k, best, float(np.linalg.norm(g)), step_size))
        if log:
            log(history[-1])
    return theta, history
