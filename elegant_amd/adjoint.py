"""Gradients of multi-turn tracking with respect to element strengths and offsets.

    spec = ParamSpec.strengths(lattice)             # every KQUAD K1, KSEXT K2, KOCT K3, KQUSE K1
    adj = AdjointTracker(lattice, spec, p_central, n_turns=1000, checkpoint_every=50)
    theta = torch.tensor(spec.values(lattice), dtype=torch.float64, requires_grad=True)
    final, alive = track_turns(theta, z0, adj)      # z0: (n, 6) float64 tensor
    loss = objective(final[alive]); loss.backward()  # theta.grad = dL/dtheta

The forward pass runs the library's tracking maps (bit-identical to
eh_do_tracking for survivors) and keeps checkpoints; the backward pass is the
hand-written reverse-mode kernel (include/elegant_hip.h, eh_adjoint_*).  Lost
particles carry no gradient: the loss set is piecewise constant in the
parameters.  torch is used for the autograd plumbing only; the arithmetic runs
in the HIP library.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import structs as S
from . import tracking

KIND_FIELD = {(S.T_KQUAD, "K"): "k1", (S.T_KSEXT, "K"): "k2", (S.T_KOCT, "K"): "k3", (S.T_KQUSE, "K"): "k1",
              (S.T_CSBEND, "K"): "k1"}
# gradient slots per element, in the order of the library's param_slots rows
KINDS = ("K", "DX", "DY", "FSE")
MULTIPOLES = (S.T_KQUAD, S.T_KSEXT, S.T_KOCT, S.T_KQUSE)
# the kinds each element type can be differentiated with respect to (a bend's
# K is its K1)
ALLOWED = {**{t: ("K", "DX", "DY") for t in MULTIPOLES}, S.T_CSBEND: ("K", "FSE")}


class ParamSpec:
    """Which element parameters are differentiated: a list of slots, each a list of
    (element index, kind) pairs that share one value (kind: "K", "DX" or "DY")."""

    def __init__(self, slots, names=None):
        self.slots = [list(s) for s in slots]
        self.names = names or ["p%d" % i for i in range(len(self.slots))]

    def __len__(self):
        return len(self.slots)

    @classmethod
    def strengths(cls, lat, types=MULTIPOLES):
        slots, names = [], []
        for i, e in enumerate(lat.elements):
            if e.type in types:
                slots.append([(i, "K")])
                names.append("%s#%d.K" % (e.name, i))
        return cls(slots, names)

    @classmethod
    def offsets(cls, lat, types=MULTIPOLES):
        slots, names = [], []
        for i, e in enumerate(lat.elements):
            if e.type in types:
                for k in ("DX", "DY"):
                    slots.append([(i, k)])
                    names.append("%s#%d.%s" % (e.name, i, k))
        return cls(slots, names)

    @classmethod
    def families(cls, lat, names, kind="K"):
        """one slot per element name: all occurrences share it (e.g. sextupole families)"""
        slots = []
        for nm in names:
            occ = [(i, kind) for i, e in enumerate(lat.elements) if e.name == nm and e.type in MULTIPOLES]
            if not occ:
                raise KeyError("no multipole named %s" % nm)
            slots.append(occ)
        return cls(slots, ["%s.%s" % (n, kind) for n in names])

    @classmethod
    def split_families(cls, lat, names, groups=2, kind="K"):
        """each named family split into `groups` families by occurrence
        (occurrence k goes to group k mod groups), e.g. the six sextupole names
        of a ring split over even and odd cells: twelve knobs"""
        slots, labels = [], []
        for nm in names:
            occ = [(i, kind) for i, e in enumerate(lat.elements) if e.name == nm and e.type in MULTIPOLES]
            if not occ:
                raise KeyError("no multipole named %s" % nm)
            for g in range(groups):
                part = occ[g::groups]
                if part:
                    slots.append(part)
                    labels.append("%s/%d.%s" % (nm, g, kind))
        return cls(slots, labels)

    @classmethod
    def bends(cls, lat, kinds=("K", "FSE")):
        """K1 and FSE of every CSBEND with an angle"""
        slots, names = [], []
        for i, e in enumerate(lat.elements):
            if e.type == S.T_CSBEND and e.params.get("ANGLE", 0):
                for k in kinds:
                    slots.append([(i, k)])
                    names.append("%s#%d.%s" % (e.name, i, "K1" if k == "K" else k))
        return cls(slots, names)

    def __add__(self, other):
        return ParamSpec(self.slots + other.slots, self.names + other.names)

    def slot_array(self, lat):
        arr = np.full((len(lat), len(KINDS)), -1, dtype=np.int32)
        for q, occ in enumerate(self.slots):
            for (i, kind) in occ:
                if kind not in ALLOWED.get(lat.elements[i].type, ()):
                    raise ValueError("element %d (%s): no %s gradient for %s" % (
                        i, lat.elements[i].name, kind, S.TYPE_NAMES[lat.elements[i].type]))
                k = KINDS.index(kind)
                if arr[i, k] >= 0:
                    raise ValueError("element %d %s appears in two slots" % (i, kind))
                arr[i, k] = q
        return arr

    @staticmethod
    def _field(etype, kind):
        return KIND_FIELD[(etype, kind)] if kind == "K" else kind.lower()

    def values(self, lat):
        """current values (first occurrence of each slot)"""
        st = lat.structs()
        out = np.zeros(len(self.slots))
        for q, occ in enumerate(self.slots):
            i, kind = occ[0]
            out[q] = getattr(st[i], self._field(lat.elements[i].type, kind))
        return out

    def apply(self, lat, theta):
        """write values into the lattice (element params and structs)"""
        theta = np.asarray(theta, dtype=np.float64)
        param_name = {"K": None, "DX": "DX", "DY": "DY", "FSE": "FSE"}
        strength = {S.T_KQUAD: "K1", S.T_KSEXT: "K2", S.T_KOCT: "K3", S.T_KQUSE: "K1", S.T_CSBEND: "K1"}
        for q, occ in enumerate(self.slots):
            for (i, kind) in occ:
                e = lat.elements[i]
                name = param_name[kind] or strength[e.type]
                lat.set_param(i, name, float(theta[q]))


class AdjointStats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double), ("retrack_ms", C.c_double),
                ("forward_units", C.c_double), ("max_recon_err", C.c_double), ("max_recon_err_s", C.c_double),
                ("anchor_bytes", C.c_double), ("checkpoint_bytes", C.c_double), ("n_survivors", C.c_long)]


def _lib_adj():
    L = _lib.load()
    if not getattr(L, "_adj_types", False):
        vp, Lg, dp = C.c_void_p, C.c_long, C.c_double
        L.eh_adjoint_create.argtypes = []
        L.eh_adjoint_create.restype = vp
        L.eh_adjoint_free.argtypes = [vp]
        L.eh_adjoint_free.restype = None
        L.eh_adjoint_forward.argtypes = [vp, vp, Lg, vp, Lg, vp, Lg, Lg, dp, Lg, vp, vp]
        L.eh_adjoint_forward.restype = Lg
        L.eh_adjoint_backward.argtypes = [vp, vp, vp, vp, C.POINTER(AdjointStats)]
        L.eh_adjoint_backward.restype = Lg
        L.eh_adjoint_forward_batch.argtypes = [vp, vp, vp, Lg, vp, Lg, vp, Lg, Lg, dp, Lg, vp, vp]
        L.eh_adjoint_forward_batch.restype = Lg
        L._adj_types = True
    return L


def fix_rf(lat, coords, p_central):
    """first-encounter RF fiducials, as tracking sets them: one pass on a copy"""
    needs = any(isinstance(s, S.EH_RFCA) and not s.fiducial_seen and s.volt != 0 for s in lat.structs())
    if needs and len(coords):
        b = tracking.Beam(np.asarray(coords, dtype=np.float64))
        tracking.track(b, lat, 1, p_central, mode="fused", turns_per_launch=1)
        b.free()


class AdjointTracker:
    """Forward (with checkpoints) and reverse-mode backward of n_turns of tracking."""

    def __init__(self, lattice, spec, p_central, n_turns, checkpoint_every=50):
        self.lat = lattice
        self.spec = spec
        self.Po = p_central
        self.n_turns = int(n_turns)
        self.C = int(checkpoint_every)
        self.slots = spec.slot_array(lattice)
        L = _lib_adj()
        self._h = L.eh_adjoint_create()
        if not self._h:
            raise _lib.EHError("eh_adjoint_create failed: %s" % L.eh_last_error().decode())
        self.stats = None
        self.n = 0
        self.generation = 0  # forward passes so far; the device holds the last one's checkpoints

    def fix_rf(self, coords):
        fix_rf(self.lat, coords, self.Po)

    def forward(self, coords, theta=None):
        """coords (n, 10) or (n, 6); returns final (n, 6) and alive (n,) bool"""
        coords = np.asarray(coords, dtype=np.float64)
        if coords.shape[1] == 6:
            full = np.zeros((coords.shape[0], S.NPROP))
            full[:, :6] = coords
            full[:, 6] = np.arange(1, coords.shape[0] + 1)
            coords = full
        coords = np.ascontiguousarray(coords)
        if theta is not None:
            self.spec.apply(self.lat, theta)
        self.fix_rf(coords)
        n = coords.shape[0]
        self.n = n
        final = np.empty((n, 6))
        alive = np.empty(n, dtype=np.int32)
        arr, keep = self.lat.c_beamline()
        L = _lib_adj()
        self.generation += 1
        r = L.eh_adjoint_forward(self._h, coords.ctypes.data, n, C.cast(arr, C.c_void_p), len(self.lat),
                                 self.slots.ctypes.data, len(self.spec), self.n_turns, self.Po, self.C,
                                 final.ctypes.data, alive.ctypes.data)
        _lib.check(r)
        del keep
        return final, alive.astype(bool)

    def backward(self, grad_final):
        g = np.ascontiguousarray(grad_final, dtype=np.float64).reshape(self.n, 6)
        gp = np.zeros(max(1, len(self.spec)))
        gz = np.zeros((self.n, 6))
        st = AdjointStats()
        r = _lib_adj().eh_adjoint_backward(self._h, g.ctypes.data, gp.ctypes.data, gz.ctypes.data, C.byref(st))
        _lib.check(r)
        self.stats = {k: getattr(st, k) for k, _ in AdjointStats._fields_}
        return gp[:len(self.spec)], gz

    def free(self):
        if self._h:
            _lib_adj().eh_adjoint_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _rows10(coords):
    coords = np.asarray(coords, dtype=np.float64)
    if coords.shape[1] == 6:
        full = np.zeros((coords.shape[0], S.NPROP))
        full[:, :6] = coords
        full[:, 6] = np.arange(1, coords.shape[0] + 1)
        coords = full
    return np.ascontiguousarray(coords)


class BatchAdjointTracker:
    """AdjointTracker for several variants of one lattice structure (error seeds,
    say) that share one ParamSpec: all variants run in one launch per
    checkpoint segment (eh_adjoint_forward_batch), and the parameter gradient is
    summed over every variant's particles, i.e. it is the gradient of the sum of
    the per-variant objectives."""

    def __init__(self, lattices, spec, p_central, n_turns, checkpoint_every=50):
        self.lats = list(lattices)
        if not self.lats:
            raise ValueError("no lattices")
        types = [e.type for e in self.lats[0].elements]
        for lat in self.lats[1:]:
            if [e.type for e in lat.elements] != types:
                raise ValueError("variants must share the element sequence")
        self.spec = spec
        self.Po = p_central
        self.n_turns = int(n_turns)
        self.C = int(checkpoint_every)
        self.slots = spec.slot_array(self.lats[0])
        L = _lib_adj()
        self._h = L.eh_adjoint_create()
        if not self._h:
            raise _lib.EHError("eh_adjoint_create failed: %s" % L.eh_last_error().decode())
        self.stats = None
        self.counts = []
        self.generation = 0  # forward passes so far; the device holds the last one's checkpoints

    def forward(self, coords_list, theta=None):
        """coords_list: one (n_v, 10) or (n_v, 6) array per variant; returns a
        list of (final (n_v, 6), alive (n_v,) bool)"""
        rows = [_rows10(c) for c in coords_list]
        if len(rows) != len(self.lats):
            raise ValueError("one particle array per variant")
        if theta is not None:
            for lat in self.lats:
                self.spec.apply(lat, theta)
        for lat, r in zip(self.lats, rows):
            fix_rf(lat, r, self.Po)
        self.counts = [len(r) for r in rows]
        allrows = np.ascontiguousarray(np.concatenate(rows))
        n = len(allrows)
        final = np.empty((n, 6))
        alive = np.empty(n, dtype=np.int32)
        arrays, keep = [], []
        for lat in self.lats:
            a, k = lat.c_beamline()
            arrays.append(a)
            keep.append(k)
        ptrs = (C.c_void_p * len(arrays))(*[C.cast(a, C.c_void_p) for a in arrays])
        counts = np.array(self.counts, dtype=np.int64)
        self.generation += 1
        r = _lib_adj().eh_adjoint_forward_batch(self._h, allrows.ctypes.data, counts.ctypes.data, len(self.lats),
                                               ptrs, len(self.lats[0]), self.slots.ctypes.data, len(self.spec),
                                               self.n_turns, self.Po, self.C, final.ctypes.data, alive.ctypes.data)
        _lib.check(r)
        del keep
        out, k = [], 0
        for c in self.counts:
            out.append((final[k:k + c], alive[k:k + c].astype(bool)))
            k += c
        return out

    def backward(self, grad_list):
        """grad_list: dL/d(final) per variant, (n_v, 6) each; returns (dL/dtheta,
        [dL/d(initial) per variant])"""
        g = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 6) for x in grad_list]))
        n = sum(self.counts)
        if g.shape[0] != n:
            raise ValueError("gradient rows do not match the forward pass")
        gp = np.zeros(max(1, len(self.spec)))
        gz = np.zeros((n, 6))
        st = AdjointStats()
        r = _lib_adj().eh_adjoint_backward(self._h, g.ctypes.data, gp.ctypes.data, gz.ctypes.data, C.byref(st))
        _lib.check(r)
        self.stats = {k: getattr(st, k) for k, _ in AdjointStats._fields_}
        out, k = [], 0
        for c in self.counts:
            out.append(gz[k:k + c])
            k += c
        return gp[:len(self.spec)], out

    def free(self):
        if self._h:
            _lib_adj().eh_adjoint_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _torch():
    import torch
    return torch


def _make_function():
    torch = _torch()

    class TrackTurns(torch.autograd.Function):
        """(theta, z0) -> (final coordinates, survival mask) through the HIP tracker;
        backward through the HIP adjoint.  The tracker keeps the checkpoints of
        its last forward pass only (on the device), so a graph node can be
        differentiated only while its forward pass is the tracker's last one;
        forward also writes theta into the tracker's lattice."""

        @staticmethod
        def forward(ctx, theta, z0, tracker):
            th = theta.detach().cpu().numpy().astype(np.float64)
            z = z0.detach().cpu().numpy().astype(np.float64)
            final, alive = tracker.forward(z, th)
            ctx.tracker = tracker
            ctx.generation = tracker.generation
            ctx.theta_dev = theta.device
            ctx.z_dev = z0.device
            out = torch.from_numpy(final).to(z0.device)
            mask = torch.from_numpy(alive).to(z0.device)
            ctx.mark_non_differentiable(mask)
            return out, mask

        @staticmethod
        def backward(ctx, grad_final, grad_mask):
            if ctx.tracker.generation != ctx.generation:
                raise RuntimeError("track_turns: the tracker has run another forward pass since this one; "
                                   "call backward before reusing the tracker (or use one tracker per graph node)")
            g = grad_final.detach().cpu().numpy().astype(np.float64)
            gp, gz = ctx.tracker.backward(g)
            return torch.from_numpy(gp).to(ctx.theta_dev), torch.from_numpy(gz).to(ctx.z_dev), None

    return TrackTurns


_TrackTurns = None


def track_turns(theta, z0, tracker):
    """differentiable tracking: returns (final (n, 6), alive (n,) bool)"""
    global _TrackTurns
    if _TrackTurns is None:
        _TrackTurns = _make_function()
    return _TrackTurns.apply(theta, z0, tracker)
