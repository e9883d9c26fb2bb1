"""ctypes wrapper of the CPU checker (liboracle.so / liboracle_libm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(libm=False, count=False):
    name = "liboracle_count.so" if count else ("liboracle_libm.so" if libm else "liboracle.so")
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    _libs[name] = load_path(path)
    return _libs[name]


def build_variant(out_path, defines):
    """compile the checker with extra -D definitions (e.g. a planted mistake,
    OR_PLANT=n) into out_path; returns the loaded library"""
    cmd = ["gcc", "-m64", "-O3", "-mfma", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-I" + os.path.join(HERE, "..", "include"),
           "-I" + os.path.join(HERE, "..", "elegant_amd", "csrc")] + ["-D" + d for d in defines] + \
        ["-shared", "-o", out_path, os.path.join(HERE, "elegant_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    return load_path(out_path)


def load_path(path):
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.or_do_tracking_flat.argtypes = [vp, C.c_long, vp, C.c_long, C.c_long, dp, dp]
    lib.or_do_tracking_flat.restype = C.c_long
    lib.or_element_flat.argtypes = [vp, C.c_long, vp, dp, C.c_double, C.c_double, vp]
    lib.or_element_flat.restype = C.c_long
    lib.or_limit_amplitudes_flat.argtypes = [vp, C.c_double, C.c_double, C.c_long, C.c_double, C.c_double, C.c_long]
    lib.or_limit_amplitudes_flat.restype = C.c_long
    lib.or_elimit_amplitudes_flat.argtypes = [vp, C.c_double, C.c_double, C.c_long, C.c_double, C.c_double, C.c_long,
                                              C.c_long, C.c_long]
    lib.or_elimit_amplitudes_flat.restype = C.c_long
    lib.or_do_match_energy_flat.argtypes = [vp, C.c_long, dp, C.c_long]
    lib.or_do_match_energy_flat.restype = C.c_long
    lib.or_ipow.argtypes = [C.c_double, C.c_long]
    lib.or_ipow.restype = C.c_double
    lib.or_get_phase_reference.argtypes = [dp, C.c_long]
    lib.or_get_phase_reference.restype = C.c_long
    lib.or_set_phase_reference.argtypes = [C.c_long, C.c_double]
    lib.or_set_phase_reference.restype = C.c_long
    lib.or_delete_phase_references.argtypes = []
    lib.or_csbend_field_table.argtypes = [dp, dp, C.c_double, C.c_int, C.c_int, dp, dp]
    lib.or_csbend_field_table.restype = C.c_long
    lib.or_last_error.restype = C.c_char_p
    lib.or_track_particles_flat.argtypes = [vp, C.c_long, vp]
    lib.or_track_particles_flat.restype = C.c_long
    if hasattr(lib, "or_count_read"):
        lib.or_count_read.argtypes = [vp]
        lib.or_count_reset.argtypes = []
        lib.or_targ_start.argtypes = [C.c_long]
        lib.or_targ_read.argtypes = [vp, vp, C.c_long]
        lib.or_targ_read.restype = C.c_long
    for f in ("or_sin", "or_cos", "or_asin", "or_log", "or_exp"):
        getattr(lib, f).argtypes = [C.c_double]
        getattr(lib, f).restype = C.c_double
    return lib


OP_NAMES = ("add", "sub", "mul", "div", "sqrt", "T")


def count_ops(coords, lattice, n_passes, p_central):
    """Run do_tracking in the counting build; returns (counts, rows, n_alive,
    effort).  counts[op] for op in OP_NAMES is the census tally of the + - * /
    and sqrt on doubles and the sin, cos, tan, asin and log calls (every
    operation executed, the generic matrix loops counted by their non-trivial
    terms, see or_count.h); counts["all_" + op] every operation executed."""
    lib = load(count=True)
    lib.or_count_reset()
    rows = np.array(coords, dtype=np.float64, order="C", copy=True)
    arr, keep = lattice.c_beamline()
    pc = C.c_double(p_central)
    eff = C.c_double(0)
    n = _check(lib, lib.or_do_tracking_flat(rows.ctypes.data, rows.shape[0], C.cast(arr, C.c_void_p), len(lattice),
                                            n_passes, C.byref(pc), C.byref(eff)))
    out = np.zeros(12, dtype=np.uint64)
    lib.or_count_read(out.ctypes.data)
    del keep
    counts = dict(zip(OP_NAMES, (int(v) for v in out[:6])))
    counts.update(("all_" + k, int(v)) for k, v in zip(OP_NAMES, out[6:]))
    return counts, rows, n, eff.value


T_KINDS = ("sin", "cos", "tan", "asin", "log", "exp")


def transcendental_args(coords, lattice, n_passes, p_central, cap=10_000_000):
    """the arguments of every sin, cos, tan, asin and log call do_tracking makes
    (counting build): returns (args, kinds) with kinds indexing T_KINDS"""
    lib = load(count=True)
    lib.or_targ_start(cap)
    count_ops(coords, lattice, n_passes, p_central)
    n = lib.or_targ_read(None, None, 0)
    n = min(n, cap)
    args = np.zeros(n)
    kinds = np.zeros(n, dtype=np.int32)
    lib.or_targ_read(args.ctypes.data, kinds.ctypes.data, n)
    lib.or_targ_start(0)
    return args, kinds


def _check(lib, r):
    if r < 0:
        raise RuntimeError("oracle error %d: %s" % (r, lib.or_last_error().decode()))
    return r


def do_tracking(coords, lattice, n_passes, p_central, libm=False):
    """Track a copy of coords (n, 10); returns (rows, n_alive, effort, p_central)."""
    lib = load(libm)
    rows = np.array(coords, dtype=np.float64, order="C", copy=True)
    arr, keep = lattice.c_beamline()
    pc = C.c_double(p_central)
    eff = C.c_double(0)
    n = _check(lib, lib.or_do_tracking_flat(rows.ctypes.data, rows.shape[0], C.cast(arr, C.c_void_p), len(lattice),
                                            n_passes, C.byref(pc), C.byref(eff)))
    del keep
    return rows, n, eff.value, pc.value


def delete_phase_references(libm=False):
    """forget every phase-reference group's phase (delete_phase_references,
    src/phase_reference.c:115-127), as elegant does between steps"""
    load(libm).or_delete_phase_references()


def get_phase_reference(ref, libm=False):
    """(status, phase) of a phase-reference group: status 1 set, 2 not set, 3 unknown"""
    ph = C.c_double(0)
    st = load(libm).or_get_phase_reference(C.byref(ph), ref)
    return st, ph.value


def element(coords, lattice, index, p_central, z_start, z_end, maxamp=None, libm=False, lib=None, with_pc=False):
    """Apply one element routine (no sweeps) to a copy of coords (lib: a checker
    build other than the default, e.g. from build_variant); with_pc: also return
    the reference momentum afterwards."""
    lib = lib or load(libm)
    rows = np.array(coords, dtype=np.float64, order="C", copy=True)
    arr, keep = lattice.c_beamline()
    pc = C.c_double(p_central)
    amp = C.cast(C.pointer(maxamp), C.c_void_p) if maxamp is not None else None
    el = C.cast(C.byref(arr, index * C.sizeof(arr._type_)), C.c_void_p)
    n = _check(lib, lib.or_element_flat(rows.ctypes.data, rows.shape[0], el, C.byref(pc), z_start, z_end, amp))
    del keep
    return (rows, n, pc.value) if with_pc else (rows, n)


def match_energy(coords, p_central, change_beam, libm=False):
    """do_match_energy on a copy of coords (every row taken as a survivor);
    returns (rows, p_central)"""
    lib = load(libm)
    rows = np.array(coords, dtype=np.float64, order="C", copy=True)
    pc = C.c_double(p_central)
    _check(lib, lib.or_do_match_energy_flat(rows.ctypes.data, rows.shape[0], C.byref(pc), change_beam))
    return rows, pc.value


def field_table(b, c, h, nonlinear=1, expansion_order=0):
    lib = load()
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    Fx = np.zeros(121)
    Fy = np.zeros(121)
    dp = C.POINTER(C.c_double)
    r = _check(lib, lib.or_csbend_field_table(b.ctypes.data_as(dp), c.ctypes.data_as(dp), h, nonlinear,
                                              expansion_order, Fx.ctypes.data_as(dp), Fy.ctypes.data_as(dp)))
    return Fx.reshape(11, 11), Fy.reshape(11, 11), r


def track_particles(coords, vmatrix, libm=False):
    """track_particles with an EH_VMATRIX (ctypes struct) on a copy of coords"""
    lib = load(libm)
    rows = np.array(coords, dtype=np.float64, order="C", copy=True)
    _check(lib, lib.or_track_particles_flat(rows.ctypes.data, rows.shape[0], C.cast(C.pointer(vmatrix), C.c_void_p)))
    return rows
