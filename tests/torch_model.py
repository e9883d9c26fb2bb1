"""Differentiable PyTorch (float64, CPU) restatement of the tracking maps for
surviving particles, used to check the HIP adjoint.

TEST INFRASTRUCTURE ONLY.  It follows the same reference routines as the
oracle (src/multipole.c:660-1456, src/csbend.c:654-1720, src/simple_rfca.c:824-967,
src/matrix.c order-2 DRIF, src/csbend.c:3792-3816 EDRIFT), written with torch
operations so that autograd (reverse mode) and complex-step evaluation (forward
mode: every map is analytic, so Im f(z + i h v) / h is the exact directional
derivative, without subtractive cancellation) give exact derivatives.  Losses and apertures are not modelled: callers use
particles that survive.  Element constants are taken from the lattice's element
structs (the same parameter values the library reads); CSBEND field
coefficients come from the oracle's field table.
"""
import math

import numpy as np
import torch

from elegant_amd import structs as S
from oracle import oracle

C_MKS = 2.99792458e8
BETA = 2.0 ** (1.0 / 3.0)
DRIFT4 = [0.5 / (2 - BETA), (1 - BETA) / (2 - BETA) / 2, (1 - BETA) / (2 - BETA) / 2, 0.5 / (2 - BETA)]
KICK4 = [1.0 / (2 - BETA), -BETA / (2 - BETA), 1.0 / (2 - BETA), 0.0]
DRIFT2 = [0.5, 0.5]
KICK2 = [1.0, 0.0]


def _tables(order):
    if order == 4:
        return DRIFT4, KICK4
    if order == 2:
        return DRIFT2, KICK2
    raise ValueError("integration order %d not modelled" % order)


def _cpow(n, x, y):
    """(x+iy)^n / n! as (re, im)"""
    re, im = torch.ones_like(x), torch.zeros_like(x)
    for k in range(1, n + 1):
        re, im = (re * x - im * y) / k, (re * y + im * x) / k
    return re, im


def _s2m(xp, yp, dp):
    w = torch.sqrt(1 + xp * xp + yp * yp)
    return (1 + dp) * xp / w, (1 + dp) * yp / w


def _m2s(qx, qy, dp):
    den = torch.sqrt((1 + dp) * (1 + dp) - qx * qx - qy * qy)
    return qx / den, qy / den


def _powers(v, n):
    """[v^0, ..., v^n] by repeated multiplication (as the reference's power
    arrays; also keeps complex-step evaluation free of branch cuts)"""
    out = [torch.ones_like(v)]
    for _ in range(n):
        out.append(out[-1] * v)
    return out


def _rot(x, y, xp, yp, c, s):
    return x * c + y * s, -x * s + y * c, xp * c + yp * s, -xp * s + yp * c


class TorchRing:
    """Maps of a Lattice for one do_tracking pass (pass index > 0 semantics do not
    matter without losses).  param_slots: (n_elements, 3) int array of gradient
    slots for strength, DX, DY (-1: fixed)."""

    def __init__(self, lattice, p_central, param_slots=None, n_params=0):
        self.lat = lattice
        self.Po = p_central
        self.structs = lattice.structs()
        ne = len(lattice)
        self.slots = np.full((ne, 4), -1, dtype=np.int64) if param_slots is None else np.asarray(param_slots)
        self.n_params = n_params
        self.csb = {}
        for i, (el, st) in enumerate(zip(lattice.elements, self.structs)):
            if el.type == S.T_CSBEND and st.angle != 0:
                self.csb[i] = self._csbend_setup(st)

    def theta0(self):
        """current parameter values in slot order"""
        th = np.zeros(self.n_params)
        for i, st in enumerate(self.structs):
            for k in range(4):
                q = self.slots[i, k]
                if q >= 0:
                    th[q] = getattr(st, self._field(self.lat.elements[i].type, k))
        return th

    @staticmethod
    def _field(etype, k):
        if k == 1:
            return "dx"
        if k == 2:
            return "dy"
        if k == 3:
            return "fse"
        return {S.T_KQUAD: "k1", S.T_KSEXT: "k2", S.T_KOCT: "k3", S.T_KQUSE: "k1", S.T_CSBEND: "k1"}[etype]

    def _csbend_setup(self, cb):
        if cb.tilt or cb.etilt or cb.dx or cb.dy or cb.dz or cb.use_bn or cb.xKick or cb.yKick:
            raise ValueError("CSBEND option not modelled")
        angle = cb.angle
        rho0 = cb.length / angle
        b = np.zeros(9)
        c = np.zeros(9)
        for j in range(1, 9):
            b[j] = getattr(cb, "k%d" % j) * rho0
        b[1] *= (1 + cb.fse + cb.fseQuadrupole) / (1 + cb.fse + cb.fseDipole)
        b[2] *= (1 + cb.fse) / (1 + cb.fse + cb.fseDipole)
        e1, e2 = cb.e1, cb.e2
        if angle < 0:
            raise ValueError("negative-angle CSBEND not modelled")
        h = 1 / rho0
        fse = cb.fse + cb.fseDipole
        rho_actual = 1 / ((1 + fse) * h)
        Kg = 2 * cb.hgap * cb.fint
        psi1 = Kg / abs(rho_actual) / math.cos(e1) * (1 + math.sin(e1) ** 2)
        psi2 = Kg / abs(rho_actual) / math.cos(e2) * (1 + math.sin(e2) ** 2)
        Fx, Fy, eo1 = oracle.field_table(b, c, h, cb.nonlinear, cb.expansionOrder)
        terms_x = [(i, j, Fx[i, j]) for i in range(eo1) for j in range(eo1 - i) if Fx[i, j] != 0]
        terms_y = [(i, j, Fy[i, j]) for i in range(eo1) for j in range(eo1 - i) if Fy[i, j] != 0]
        field = any(b[1:] != 0)
        edge1 = bool(cb.edgeFlags & 1) and cb.edge1_effects == 1
        edge2 = bool(cb.edgeFlags & 2) and cb.edge2_effects == 1
        # for K1 and FSE as parameters: the table is linear in b[], so it is the
        # table without b1 and b2 plus b1 and b2 times their unit tables (the
        # dipole's constant 1 - b0 only in the first)
        rest = b.copy()
        rest[1] = rest[2] = 0
        F0x, F0y, _ = oracle.field_table(rest, c, h, cb.nonlinear, eo1 - 1)
        unit = []
        for n in (1, 2):
            e = np.zeros(9)
            e[n] = 1.0
            ux, uy, _ = oracle.field_table(e, np.zeros(9), h, cb.nonlinear, eo1 - 1)
            uy[0, 0] -= 1
            unit.append((ux, uy))
        return dict(rho0=rho0, rho_a=rho_actual, tx=terms_x, ty=terms_y, field=field, e1=e1, e2=e2, psi1=psi1,
                    psi2=psi2, edge1=edge1, edge2=edge2, eo1=eo1, F0=(F0x, F0y), unit=unit, h=h,
                    k1=cb.k1, k2=cb.k2, fse=cb.fse, fd=cb.fseDipole, fq=cb.fseQuadrupole, Kg=Kg)

    def _bend_params(self, i, theta):
        """the bend's rho_actual, edge angles and field terms as functions of its
        K1 and FSE (src/csbend.c:800-857, :940-1000) when either is a parameter"""
        cs = self.csb[i]
        if self.slots[i, 0] < 0 and self.slots[i, 3] < 0:
            return cs["rho_a"], cs["psi1"], cs["psi2"], cs["tx"], cs["ty"], cs["field"]
        k1 = self._param(theta, i, 0, cs["k1"])
        F = self._param(theta, i, 3, cs["fse"])
        rho0, fd, fq = cs["rho0"], cs["fd"], cs["fq"]
        b1 = k1 * rho0 * (1 + F + fq) / (1 + F + fd)
        b2 = cs["k2"] * rho0 * (1 + F) / (1 + F + fd)
        rho_a = 1 / ((1 + F + fd) * cs["h"])
        psi1 = cs["Kg"] / rho_a / math.cos(cs["e1"]) * (1 + math.sin(cs["e1"]) ** 2)
        psi2 = cs["Kg"] / rho_a / math.cos(cs["e2"]) * (1 + math.sin(cs["e2"]) ** 2)
        (F0x, F0y), (u1x, u1y), (u2x, u2y) = cs["F0"], cs["unit"][0], cs["unit"][1]
        eo1 = cs["eo1"]
        tx, ty = [], []
        for a in range(eo1):
            for j in range(eo1 - a):
                if F0x[a, j] or u1x[a, j] or u2x[a, j]:
                    tx.append((a, j, F0x[a, j] + b1 * u1x[a, j] + b2 * u2x[a, j]))
                if F0y[a, j] or u1y[a, j] or u2y[a, j]:
                    ty.append((a, j, F0y[a, j] + b1 * u1y[a, j] + b2 * u2y[a, j]))
        return rho_a, psi1, psi2, tx, ty, True

    # ---- element maps -----------------------------------------------------
    def _param(self, theta, i, k, default):
        q = self.slots[i, k]
        return theta[q] if q >= 0 else default

    def _multipole(self, z, i, st, etype, theta):
        x, xp, y, yp, s, dp = z
        K = self._param(theta, i, 0, getattr(st, self._field(etype, 0)))
        dx = self._param(theta, i, 1, st.dx)
        dy = self._param(theta, i, 2, st.dy)
        order = {S.T_KQUAD: 1, S.T_KSEXT: 2, S.T_KOCT: 3, S.T_KQUSE: 1}[etype]
        fse = st.fse1 if etype == S.T_KQUSE else st.fse
        terms = [(order, 0, K * st.length * (1 + fse))]
        if etype == S.T_KQUSE:
            terms.append((2, 0, st.k2 * st.length * (1 + st.fse2)))
        if etype == S.T_KSEXT:
            if st.k1:
                terms.append((1, 0, st.k1 * st.length))
            if st.j1:
                terms.append((1, 1, st.j1 * st.length))
        xk = yk = 0.0
        if etype in (S.T_KQUAD, S.T_KSEXT):
            xk = st.xkick * st.xKickCalibration
            yk = st.ykick * st.yKickCalibration
        if st.expandHamiltonian:
            raise ValueError("expanded Hamiltonian not modelled")
        integ = st.integration_order
        dfr, kfr = _tables(integ)
        nsub = len(dfr)
        ns = st.nSlices
        if st.n_kicks > 0:
            ns = max(1, math.ceil(st.n_kicks / integ)) if integ > 2 else st.n_kicks
        L = st.length
        dz = st.dz
        tilt = st.tilt
        has_off = bool(dx if not torch.is_tensor(dx) else True) or bool(dy if not torch.is_tensor(dy) else True) \
            or dz != 0
        if has_off:
            w = torch.sqrt(1 + xp * xp + yp * yp)
            s = s + dz * w
            x = x - dx + dz * xp
            y = y - dy + dz * yp
        if tilt:
            x, y, xp, yp = _rot(x, y, xp, yp, math.cos(tilt), math.sin(tilt))
        qx, qy = _s2m(xp, yp, dp)
        xp, yp = _m2s(qx, qy, dp)
        h0 = L / ns
        for _ in range(ns):
            for j in range(nsub):
                hh = h0 * dfr[j]
                x = x + xp * hh
                y = y + yp * hh
                s = s + hh * torch.sqrt(1 + xp * xp + yp * yp)
                if j == nsub - 1:
                    break
                for (n, skew, KL) in terms:
                    k = KL / ns * kfr[j]
                    re, im = _cpow(n, x, y)
                    Fy, Fx = (im, -re) if skew else (re, im)
                    qx = qx - k * Fy
                    qy = qy + k * Fx
                if xk:
                    qx = qx + xk / ns * kfr[j]
                if yk:
                    qy = qy + yk / ns * kfr[j]
                xp, yp = _m2s(qx, qy, dp)
        if tilt:
            x, y, xp, yp = _rot(x, y, xp, yp, math.cos(-tilt), math.sin(-tilt))
        if has_off:
            w = torch.sqrt(1 + xp * xp + yp * yp)
            s = s - dz * w
            x = x + dx - dz * xp
            y = y + dy - dz * yp
        return [x, xp, y, yp, s, dp]

    @staticmethod
    def _fields(field, tx, ty, X, Y):
        if not field:
            return torch.zeros_like(X), torch.ones_like(X)
        px = _powers(X, 10)
        py = _powers(Y, 10)
        Fx = sum(c * px[i] * py[j] for (i, j, c) in tx) if tx else torch.zeros_like(X)
        Fy = sum(c * px[i] * py[j] for (i, j, c) in ty) if ty else torch.zeros_like(X)
        return Fx, Fy

    def _csbend(self, z, i, st, theta):
        x, xp, y, yp, s, dp = z
        cs = self.csb[i]
        rho0 = cs["rho0"]
        rho_a, psi1, psi2, tx, ty, field = self._bend_params(i, theta)
        if cs["edge1"]:
            rho = (1 + dp) * rho_a
            xp = xp + math.tan(cs["e1"]) / rho * x
            yp = yp - torch.tan(cs["e1"] - psi1 / (1 + dp)) / rho * y
        QX, QY = _s2m(xp, yp, dp)
        X, Y = x, y
        dfr, kfr = _tables(st.integration_order)
        nsub = len(dfr)
        sl = st.length / st.nSlices
        dist = torch.zeros_like(X)
        for _ in range(st.nSlices):
            for j in range(nsub):
                th = sl * dfr[j] / rho0
                f = torch.sqrt((1 + dp) * (1 + dp) - QY * QY)
                phi = torch.asin(QX / f)
                psi = th + phi
                QXn = f * torch.sin(psi)
                R = rho0 + X
                F = R * math.sin(th) / (f * torch.cos(psi))
                Y = Y + QY * F
                dist = dist + F * (1 + dp)
                X = R * torch.cos(phi) / torch.cos(psi) - rho0
                QX = QXn
                if j == nsub - 1:
                    break
                ds = sl * kfr[j]
                Fx, Fy = self._fields(field, tx, ty, X, Y)
                QX = QX - ds * (1 + X / rho0) * Fy / rho_a
                QY = QY + ds * (1 + X / rho0) * Fx / rho_a
        s = s + dist
        xp, yp = _m2s(QX, QY, dp)
        x, y = X, Y
        if cs["edge2"]:
            rho = (1 + dp) * rho_a
            xp = xp + math.tan(cs["e2"]) / rho * x
            yp = yp - torch.tan(cs["e2"] - psi2 / (1 + dp)) / rho * y
        return [x, xp, y, yp, s, dp]

    def _rfca(self, z, st):
        x, xp, y, yp, s, dp = z
        Pc = self.Po
        if st.volt == 0:
            L = st.length
            return [x + xp * L, xp, y + yp * L, yp, s + L * torch.sqrt(1 + xp * xp + yp * yp), dp]
        if st.nKicks > 0 and st.length:
            raise ValueError("RFCA kick method not modelled")
        length = st.length
        omega = 2 * math.pi * st.freq
        volt = st.volt / (1e6 * S.ME_MEV)
        phase = st.phase_fiducial + st.phase * math.pi / 180.0
        x = x - st.dx
        y = y - st.dy
        P = Pc * (1 + dp)
        gamma = torch.sqrt(P * P + 1)
        beta_i = P / gamma
        ds1 = length / 2 * torch.sqrt(1 + xp * xp + yp * yp)
        t = (s + ds1) / (C_MKS * beta_i)
        dgamma = volt * torch.sin(omega * t + phase) if omega else volt * math.sin(phase) * torch.ones_like(t)
        if st.end1Focus and length:
            iF = dgamma / (2 * gamma * length)
            xp = xp - x * iF
            yp = yp - y * iF
        g1 = gamma + dgamma
        dP = torch.sqrt(g1 * g1 - 1) - P
        R22 = 1 / (1 + dP / P)
        R12 = length * (P / dP * torch.log(1 + dP / P))
        s = s + ds1
        x, xp = x + xp * R12, xp * R22
        y, yp = y + yp * R12, yp * R22
        s = s + length / 2 * torch.sqrt(1 + xp * xp + yp * yp)
        dp = (P + dP - Pc) / Pc
        if st.end2Focus and length:
            iF = -dgamma / (2 * g1 * length)
            xp = xp - x * iF
            yp = yp - y * iF
        s = (P + dP) / g1 * s / beta_i
        return [x + st.dx, xp, y + st.dy, yp, s, dp]

    def turn(self, z, theta):
        for i, (el, st) in enumerate(zip(self.lat.elements, self.structs)):
            t = el.type
            if t == S.T_DRIF:
                L = st.length
                T = 0.0 if st.order == 1 else L / 2
                x, xp, y, yp, s, dp = z
                z = [L * xp + x, xp, L * yp + y, yp, ((L + s) + T * yp * yp) + T * xp * xp, dp]
            elif t == S.T_EDRIFT:
                x, xp, y, yp, s, dp = z
                L = st.length
                z = [x + xp * L, xp, y + yp * L, yp, s + L * torch.sqrt(1 + xp * xp + yp * yp), dp]
            elif t in (S.T_KQUAD, S.T_KSEXT, S.T_KOCT, S.T_KQUSE):
                z = self._multipole(z, i, st, t, theta)
            elif t == S.T_CSBEND:
                if st.angle == 0:
                    x, xp, y, yp, s, dp = z
                    L = st.length
                    z = [x + xp * L, xp, y + yp * L, yp, s + L * torch.sqrt(1 + xp * xp + yp * yp), dp]
                else:
                    z = self._csbend(z, i, st, theta)
            elif t == S.T_RFCA:
                z = self._rfca(z, st)
        return z

    def track(self, z0, theta, n_turns, progress=None):
        """z0: (n, 6) tensor; theta: (n_params,) tensor; returns (n, 6).  Real
        tensors give autograd derivatives; complex ones a complex-step
        (forward-mode) derivative: every map here is analytic in its inputs."""
        z = [z0[:, k] for k in range(6)]
        for t in range(n_turns):
            z = self.turn(z, theta)
            if progress:
                progress(t)
        return torch.stack(z, dim=1)


def forward_mode_projection(lat, spec, Po, rows, turns, bx, by, v, progress=None):
    """d/dh of L(theta + h v), L = mean(x^2/bx + y^2/by) after `turns`, by complex
    step (h = 1e-30) through this restatement; spec is an adjoint.ParamSpec"""
    ring = TorchRing(lat, Po, spec.slot_array(lat), len(spec))
    h = 1e-30
    theta = torch.tensor(spec.values(lat), dtype=torch.complex128) + 1j * h * torch.tensor(v, dtype=torch.complex128)
    z0 = torch.tensor(rows[:, :6], dtype=torch.complex128)
    out = ring.track(z0, theta, turns, progress=progress)
    L = (out[:, 0] * out[:, 0] / bx + out[:, 2] * out[:, 2] / by).mean()
    return float(L.imag) / h
