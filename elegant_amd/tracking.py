"""Python face of the gfx950 tracking library.

    beam = Beam(coords)              # coords: (n, 10) float64 rows, elegant layout
    n = track(beam, lattice, n_passes, p_central, mode="fused")
    out = beam.download()            # rows in the reference's order, lost rows above n

Everything here is a thin ctypes layer over include/elegant_hip.h.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import structs as S

MODES = {"plain": 0, "fused": 1}


def lib():
    return _lib.load()


def init(device=0):
    _lib.check(lib().eh_init(device))


class Beam:
    """Particles resident on the GPU (SoA), created from an (n, 10) AoS array."""

    def __init__(self, coords, capacity=None):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        if coords.ndim != 2 or coords.shape[1] != S.NPROP:
            raise ValueError("coords must have shape (n, %d)" % S.NPROP)
        self.n = coords.shape[0]
        L = lib()
        self._h = L.eh_beam_create(max(1, capacity or self.n))
        if not self._h:
            raise _lib.EHError("eh_beam_create failed: %s" % L.eh_last_error().decode())
        _lib.check(L.eh_beam_upload_flat(self._h, coords.ctypes.data, self.n))

    @property
    def handle(self):
        return self._h

    def n_alive(self):
        return lib().eh_beam_n_alive(self._h)

    def download(self):
        out = np.empty((self.n, S.NPROP), dtype=np.float64)
        _lib.check(lib().eh_beam_download_flat(self._h, out.ctypes.data, self.n))
        return out

    def row_order(self):
        """upload row of the particle a download puts in each row (how an
        `accepted` array moves with the particles)"""
        out = np.empty(self.n, dtype=np.int32)
        _lib.check(lib().eh_beam_row_order(self._h, out.ctypes.data, self.n))
        return out

    def track_particles(self, vmatrix):
        """apply a transport map (S.EH_VMATRIX) to the surviving particles"""
        return _lib.check(lib().eh_track_particles(self._h, self.n_alive(), C.byref(vmatrix)))

    def match_energy(self, p_central, change_beam):
        """do_match_energy on the surviving particles; returns the reference
        momentum afterwards"""
        pc = C.c_double(p_central)
        _lib.check(lib().eh_do_match_energy(self._h, self.n_alive(), C.byref(pc), int(change_beam)))
        return pc.value

    def free(self):
        if self._h:
            lib().eh_beam_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def track(beam, lattice, n_passes, p_central, mode="fused", turns_per_launch=16):
    """do_tracking over `lattice` for n_passes; returns (n_alive, stats dict, p_central)."""
    arr, keep = lattice.c_beamline()
    pc = C.c_double(p_central)
    st = S.eh_track_stats()
    n = lib().eh_do_tracking(beam.handle, arr, len(lattice), n_passes, C.byref(pc), MODES[mode],
                             int(turns_per_launch), C.byref(st))
    _lib.check(n)
    del keep
    return n, stats_dict(st), pc.value


def track_noalloc(beam, lattice, n_passes, p_central, turns_per_launch=16):
    """track() in fused mode through eh_do_tracking_noalloc (no device
    allocation: descriptors, field terms and counters in the static scratch)"""
    arr, keep = lattice.c_beamline()
    pc = C.c_double(p_central)
    st = S.eh_track_stats()
    n = lib().eh_do_tracking_noalloc(beam.handle, arr, len(lattice), n_passes, C.byref(pc), int(turns_per_launch),
                                     C.byref(st))
    _lib.check(n)
    del keep
    return n, stats_dict(st), pc.value


def delete_phase_references():
    """forget every phase-reference group's phase in the library's table
    (delete_phase_references, src/phase_reference.c:115-127)"""
    _lib.check(lib().eh_delete_phase_references())


def get_phase_reference(ref):
    """(status, phase) of a phase-reference group: status 1 set, 2 not set, 3 unknown"""
    ph = C.c_double(0)
    st = lib().eh_get_phase_reference(ref, C.byref(ph))
    return st, ph.value


def stats_dict(st):
    return {k: getattr(st, k) for k, _ in S.eh_track_stats._fields_}


def csbend_field_table(b, c, h, nonlinear=1, expansion_order=0):
    """CSBEND field-expansion coefficients (host routine of the library)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    Fx = np.zeros(121)
    Fy = np.zeros(121)
    hn, hs = C.c_int(), C.c_int()
    dp = C.POINTER(C.c_double)
    r = lib().eh_csbend_field_table(b.ctypes.data_as(dp), c.ctypes.data_as(dp), h, nonlinear, expansion_order,
                                    Fx.ctypes.data_as(dp), Fy.ctypes.data_as(dp), C.byref(hn), C.byref(hs))
    _lib.check(r)
    return Fx.reshape(11, 11), Fy.reshape(11, 11), r, hn.value, hs.value


class Batch:
    """Variants of one beamline, each with its own particles, tracked together."""

    def __init__(self, beams):
        beams = [np.ascontiguousarray(b, dtype=np.float64).reshape(-1, S.NPROP) for b in beams]
        self.n = [len(b) for b in beams]
        L = lib()
        counts = np.array(self.n, dtype=np.int64)
        self._h = L.eh_batch_create(len(beams), counts.ctypes.data)
        if not self._h:
            raise _lib.EHError("eh_batch_create failed: %s" % L.eh_last_error().decode())
        for v, b in enumerate(beams):
            _lib.check(L.eh_batch_upload(self._h, v, b.ctypes.data, len(b)))

    def track(self, lattices, n_passes, p_central, block=256, span_elements=None, waves_per_cu=8, use_graph=True):
        """track every variant n_passes turns; lattices[v] is variant v's Lattice"""
        if len(lattices) != len(self.n):
            raise ValueError("one lattice per variant")
        ne = len(lattices[0])
        arrays, keep = [], []
        for lat in lattices:
            if len(lat) != ne:
                raise ValueError("variants must have the same number of elements")
            a, k = lat.c_beamline()
            arrays.append(a)
            keep.append(k)
        ptrs = (C.c_void_p * len(arrays))(*[C.cast(a, C.c_void_p) for a in arrays])
        tune = S.eh_batch_tune(block, int(span_elements or ne), waves_per_cu, int(bool(use_graph)))
        pc = C.c_double(p_central)
        st = S.eh_track_stats()
        n = lib().eh_batch_track(self._h, ptrs, ne, n_passes, C.byref(pc), C.byref(tune), C.byref(st))
        _lib.check(n)
        del keep
        return n, stats_dict(st)

    def n_alive(self, v):
        return lib().eh_batch_n_alive(self._h, v)

    def download(self, v):
        out = np.empty((self.n[v], S.NPROP), dtype=np.float64)
        _lib.check(lib().eh_batch_download(self._h, v, out.ctypes.data, self.n[v]))
        return out

    def free(self):
        if self._h:
            lib().eh_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
