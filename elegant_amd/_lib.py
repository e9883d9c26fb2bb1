"""Loader for the in-tree gfx950 library (libelegant_hip.so).

The library is the product: there is no CPU fallback.  If it is missing,
importing the tracking API raises, so a GPU run can never silently compute
somewhere else.
"""
import ctypes as C
import os

from . import structs as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libelegant_hip.so")

_lib = None

# every symbol declared in include/elegant_hip.h
EXPORTS = [
    "eh_last_error", "eh_init", "eh_beam_create", "eh_beam_free", "eh_beam_upload", "eh_beam_download",
    "eh_beam_upload_flat", "eh_beam_download_flat", "eh_beam_n_alive", "eh_beam_n_total", "eh_beam_set_pass",
    "eh_multipole_tracking2", "eh_track_through_csbend", "eh_simple_rf_cavity", "eh_exactDrift", "eh_drift_beam",
    "eh_limit_amplitudes", "eh_removeInvalidParticles", "eh_do_tracking", "eh_csbend_field_table",
    "eh_replay_order", "eh_selftest_math", "eh_batch_create", "eh_batch_free", "eh_batch_upload",
    "eh_batch_download", "eh_batch_n_alive", "eh_batch_track", "eh_diag_copy", "eh_adjoint_create",
    "eh_adjoint_free", "eh_adjoint_forward", "eh_adjoint_backward", "eh_adjoint_forward_batch",
    "eh_multipole_tracking", "eh_offset_beam", "eh_rotate_beam", "eh_replay_sequence", "eh_set_central_momentum",
    "eh_track_particles", "eh_beam_row_order", "eh_do_tracking_noalloc", "eh_elimit_amplitudes",
    "eh_get_phase_reference", "eh_set_phase_reference", "eh_delete_phase_references", "eh_do_match_energy",
]


class EHError(RuntimeError):
    pass


def load():
    """Load libelegant_hip.so and declare its signatures; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EHError("%s is not built: run `python -m elegant_amd.build` (or __graft_entry__.build())" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    vp, L, D, I = C.c_void_p, C.c_long, C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    lib.eh_last_error.restype = C.c_char_p
    lib.eh_init.argtypes = [I]
    lib.eh_init.restype = I
    lib.eh_beam_create.argtypes = [L]
    lib.eh_beam_create.restype = vp
    lib.eh_beam_free.argtypes = [vp]
    lib.eh_beam_free.restype = None
    lib.eh_beam_upload.argtypes = [vp, C.POINTER(dp), L]
    lib.eh_beam_upload.restype = L
    lib.eh_beam_download.argtypes = [vp, C.POINTER(dp), L]
    lib.eh_beam_download.restype = L
    lib.eh_beam_upload_flat.argtypes = [vp, vp, L]
    lib.eh_beam_upload_flat.restype = L
    lib.eh_beam_download_flat.argtypes = [vp, vp, L]
    lib.eh_beam_download_flat.restype = L
    lib.eh_beam_n_alive.argtypes = [vp]
    lib.eh_beam_n_alive.restype = L
    lib.eh_beam_n_total.argtypes = [vp]
    lib.eh_beam_n_total.restype = L
    lib.eh_beam_set_pass.argtypes = [vp, L]
    lib.eh_beam_set_pass.restype = None
    elp = C.POINTER(S.EH_ELEMENT_LIST)
    lib.eh_multipole_tracking2.argtypes = [vp, L, elp, D, D, D, C.POINTER(S.EH_MAXAMP), L]
    lib.eh_multipole_tracking2.restype = L
    lib.eh_track_through_csbend.argtypes = [vp, L, C.POINTER(S.EH_CSBEND), D, D, D, C.POINTER(S.EH_MAXAMP), L]
    lib.eh_track_through_csbend.restype = L
    lib.eh_simple_rf_cavity.argtypes = [vp, L, C.POINTER(S.EH_RFCA), dp, D]
    lib.eh_simple_rf_cavity.restype = L
    lib.eh_multipole_tracking.argtypes = [vp, L, C.POINTER(S.EH_MULT), D, D, D]
    lib.eh_multipole_tracking.restype = L
    lib.eh_offset_beam.argtypes = [vp, L, C.POINTER(S.EH_MALIGN), D]
    lib.eh_offset_beam.restype = L
    lib.eh_rotate_beam.argtypes = [vp, L, C.POINTER(S.EH_ROTATE)]
    lib.eh_rotate_beam.restype = L
    lib.eh_set_central_momentum.argtypes = [vp, L, D, dp]
    lib.eh_set_central_momentum.restype = L
    lib.eh_do_match_energy.argtypes = [vp, L, dp, L]
    lib.eh_do_match_energy.restype = L
    lib.eh_track_particles.argtypes = [vp, L, C.POINTER(S.EH_VMATRIX)]
    lib.eh_track_particles.restype = L
    lib.eh_beam_row_order.argtypes = [vp, vp, L]
    lib.eh_beam_row_order.restype = L
    lib.eh_exactDrift.argtypes = [vp, L, D]
    lib.eh_exactDrift.restype = L
    lib.eh_drift_beam.argtypes = [vp, L, D, L]
    lib.eh_drift_beam.restype = L
    lib.eh_limit_amplitudes.argtypes = [vp, D, D, L, D, D, L, L]
    lib.eh_limit_amplitudes.restype = L
    lib.eh_elimit_amplitudes.argtypes = [vp, D, D, L, D, D, L, L, L, L]
    lib.eh_elimit_amplitudes.restype = L
    lib.eh_get_phase_reference.argtypes = [C.c_int64, dp]
    lib.eh_get_phase_reference.restype = L
    lib.eh_set_phase_reference.argtypes = [C.c_int64, D]
    lib.eh_set_phase_reference.restype = L
    lib.eh_delete_phase_references.argtypes = []
    lib.eh_delete_phase_references.restype = L
    lib.eh_removeInvalidParticles.argtypes = [vp, L, D, D]
    lib.eh_removeInvalidParticles.restype = L
    lib.eh_do_tracking.argtypes = [vp, elp, L, L, dp, I, I, C.POINTER(S.eh_track_stats)]
    lib.eh_do_tracking.restype = L
    lib.eh_do_tracking_noalloc.argtypes = [vp, elp, L, L, dp, I, C.POINTER(S.eh_track_stats)]
    lib.eh_do_tracking_noalloc.restype = L
    lib.eh_csbend_field_table.argtypes = [dp, dp, D, I, I, dp, dp, C.POINTER(I), C.POINTER(I)]
    lib.eh_csbend_field_table.restype = L
    lib.eh_replay_order.argtypes = [L, vp, vp, L, vp]
    lib.eh_replay_order.restype = L
    lib.eh_replay_sequence.argtypes = [L, vp, vp, L, vp, vp]
    lib.eh_replay_sequence.restype = L
    lib.eh_selftest_math.argtypes = [vp, vp, L, vp]
    lib.eh_selftest_math.restype = L
    lib.eh_diag_copy.argtypes = [L, C.c_int]
    lib.eh_diag_copy.restype = L
    lib.eh_batch_create.argtypes = [L, vp]
    lib.eh_batch_create.restype = vp
    lib.eh_batch_free.argtypes = [vp]
    lib.eh_batch_free.restype = None
    lib.eh_batch_upload.argtypes = [vp, L, vp, L]
    lib.eh_batch_upload.restype = L
    lib.eh_batch_download.argtypes = [vp, L, vp, L]
    lib.eh_batch_download.restype = L
    lib.eh_batch_n_alive.argtypes = [vp, L]
    lib.eh_batch_n_alive.restype = L
    lib.eh_batch_track.argtypes = [vp, vp, L, L, dp, C.POINTER(S.eh_batch_tune), C.POINTER(S.eh_track_stats)]
    lib.eh_batch_track.restype = L
    _lib = lib
    return lib


def check(ret):
    """Raise EHError for a negative return code."""
    if ret is None or (isinstance(ret, int) and ret < 0):
        raise EHError("elegant_hip error %s: %s" % (ret, _lib.eh_last_error().decode()))
    return ret
