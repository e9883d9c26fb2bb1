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
    return n, {"particle_element_turns": st.particle_element_turns, "kernel_ms": st.kernel_ms,
               "launches": st.launches, "n_alive": st.n_alive}, pc.value


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
