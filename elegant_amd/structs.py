"""ctypes mirrors of the element structs in include/elegant_hip.h.

Field names follow elegant's PARAMETER tables (src/track_data.c), so a lattice
written with elegant's parameter names maps one-to-one.  Defaults are the
reference defaults (src/track_data.c:1125, :1169, :1227, :1898, :579, :994, :3270, :486-502,
:929-944, :1005-1010).
"""
import ctypes as C

NPROP = 10
ID_INDEX = 6
LOSS_PASS_INDEX = 7

(T_MARK, T_DRIF, T_EDRIFT, T_KQUAD, T_KSEXT, T_KOCT, T_CSBEND, T_RFCA, T_MAXAMP, T_KQUSE, T_MULT, T_ROTATE,
 T_MALIGN, T_ENERGY) = range(14)
TYPE_NAMES = {T_MARK: "MARK", T_DRIF: "DRIF", T_EDRIFT: "EDRIFT", T_KQUAD: "KQUAD", T_KSEXT: "KSEXT",
              T_KOCT: "KOCT", T_CSBEND: "CSBEND", T_RFCA: "RFCA", T_MAXAMP: "MAXAMP", T_KQUSE: "KQUSE",
              T_MULT: "MULT", T_ROTATE: "ROTATE", T_MALIGN: "MALIGN", T_ENERGY: "ENERGY"}
TYPE_CODES = {v: k for k, v in TYPE_NAMES.items()}

BEND_EDGE1_EFFECTS = 1
BEND_EDGE2_EFFECTS = 2
FID_TMEAN, FID_LIGHT, FID_FIRST, FID_PMAX = 0, 1, 2, 3

C_MKS = 2.99792458e8
ME_MEV = 0.51099895000

i32 = C.c_int32
dbl = C.c_double


class EH_DRIFT(C.Structure):
    _fields_ = [("length", dbl), ("order", i32), ("pad_", i32)]


class EH_EDRIFT(C.Structure):
    _fields_ = [("length", dbl)]


class EH_KQUAD(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "k1", "tilt", "dx", "dy", "dz", "fse", "xkick", "ykick",
                                   "xKickCalibration", "yKickCalibration")] + \
               [(n, i32) for n in ("n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad")] + \
               [(n, dbl) for n in ("bore", "B")]


class EH_KSEXT(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "k2", "k1", "j1", "tilt", "dx", "dy", "dz", "fse", "xkick", "ykick",
                                   "xKickCalibration", "yKickCalibration")] + \
               [(n, i32) for n in ("n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad")] + \
               [(n, dbl) for n in ("bore", "B")]


class EH_KOCT(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "k3", "tilt", "dx", "dy", "dz", "fse")] + \
               [(n, i32) for n in ("n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad")] + \
               [(n, dbl) for n in ("bore", "B")]


class EH_KQUSE(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "k1", "k2", "tilt", "dx", "dy", "dz", "fse1", "fse2")] + \
               [(n, i32) for n in ("n_kicks", "nSlices", "integration_order", "expandHamiltonian", "synch_rad")]


class EH_CSBEND(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "angle", "k1", "k2", "k3", "k4", "k5", "k6", "k7", "k8",
                                   "e1", "e2", "tilt", "h1", "h2", "hgap", "fint",
                                   "dx", "dy", "dz", "xKick", "yKick",
                                   "fse", "fseDipole", "fseQuadrupole", "etilt",
                                   "b1", "b2", "b3", "b4", "b5", "b6", "b7", "b8")] + \
               [(n, i32) for n in ("nSlices", "nonlinear", "edge1_effects", "edge2_effects", "edge_order",
                                   "integration_order", "expandHamiltonian", "use_bn", "expansionOrder",
                                   "etiltSign", "malignMethod", "edgeFlags", "synch_rad", "kick_limit_scaling")] + \
               [("edge1_kick_limit", dbl), ("edge2_kick_limit", dbl)]


class EH_RFCA(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "volt", "phase", "freq", "Q", "dx", "dy", "tReference")] + \
               [(n, i32) for n in ("nKicks", "change_p0", "change_t", "end1Focus", "end2Focus", "standingWave",
                                   "linearize", "lockPhase", "fiducial_mode")] + \
               [("phase_reference", C.c_int64), ("fiducial_seen", i32), ("pad_", i32),
                ("phase_fiducial", dbl), ("t_fiducial", dbl)]


class EH_MULT(C.Structure):
    _fields_ = [(n, dbl) for n in ("length", "KnL", "tilt", "bore", "BTipL", "dx", "dy", "dz", "factor")] + \
               [(n, i32) for n in ("order", "synch_rad", "expandHamiltonian", "nSlices")]


class EH_VMATRIX(C.Structure):
    """first- or second-order transport map; T packed by rows, T[i][j*(j+1)/2 + k], k <= j"""
    _fields_ = [("order", i32), ("pad_", i32), ("C", dbl * 6), ("R", (dbl * 6) * 6), ("T", (dbl * 21) * 6)]


def vmatrix(C0, R, T=None):
    """EH_VMATRIX from C (6), R (6, 6) and optionally T (6, 6, 6), of which the
    entries T[i, j, k] with k <= j are used (elegant's storage)"""
    m = EH_VMATRIX()
    m.order = 2 if T is not None else 1
    for i in range(6):
        m.C[i] = float(C0[i])
        for j in range(6):
            m.R[i][j] = float(R[i][j])
            if T is not None:
                for k in range(j + 1):
                    m.T[i][j * (j + 1) // 2 + k] = float(T[i][j][k])
    return m


class EH_ROTATE(C.Structure):
    _fields_ = [("tilt", dbl)] + [(n, i32) for n in ("on_pass", "exclude_floor", "exclude_optics")]


class EH_MALIGN(C.Structure):
    _fields_ = [(n, dbl) for n in ("dxp", "dyp", "dx", "dy", "dz", "dt", "dp", "de")] + \
               [(n, i32) for n in ("on_pass", "startPID", "endPID")]


class EH_ENERGY(C.Structure):
    _fields_ = [("central_energy", dbl), ("central_momentum", dbl), ("match_beamline", i32),
                ("match_particles", i32)]


class EH_MAXAMP(C.Structure):
    _fields_ = [("x_max", dbl), ("y_max", dbl), ("elliptical", i32), ("exponent", i32), ("y_exponent", i32),
                ("pad_", i32)]


class EH_ELEMENT_LIST(C.Structure):
    _fields_ = [("type", i32), ("occurence", i32), ("name", C.c_char_p), ("p_elem", C.c_void_p)]


class eh_track_stats(C.Structure):
    _fields_ = [("particle_element_turns", dbl), ("kernel_ms", dbl), ("launches", C.c_long), ("n_alive", C.c_long),
                ("main_ms", dbl), ("main_launches", C.c_long), ("main_units", dbl)]


class eh_batch_tune(C.Structure):
    _fields_ = [("block", i32), ("span_elements", i32), ("waves_per_cu", i32), ("use_graph", i32)]


STRUCT_OF = {T_DRIF: EH_DRIFT, T_EDRIFT: EH_EDRIFT, T_KQUAD: EH_KQUAD, T_KSEXT: EH_KSEXT, T_KOCT: EH_KOCT,
             T_CSBEND: EH_CSBEND, T_RFCA: EH_RFCA, T_MAXAMP: EH_MAXAMP, T_KQUSE: EH_KQUSE, T_MULT: EH_MULT,
             T_ROTATE: EH_ROTATE, T_MALIGN: EH_MALIGN, T_ENERGY: EH_ENERGY}

# reference defaults (src/track_data.c)
DEFAULTS = {
    T_DRIF: dict(order=0),
    T_KQUAD: dict(xKickCalibration=1.0, yKickCalibration=1.0, n_kicks=0, nSlices=1, integration_order=4),
    T_KSEXT: dict(xKickCalibration=1.0, yKickCalibration=1.0, n_kicks=0, nSlices=1, integration_order=4),
    T_KOCT: dict(n_kicks=4, nSlices=4, integration_order=4),
    T_KQUSE: dict(n_kicks=0, nSlices=1, integration_order=4),
    T_CSBEND: dict(fint=0.5, nSlices=4, nonlinear=1, edge1_effects=1, edge2_effects=1, edge_order=1,
                   integration_order=4, etiltSign=1),
    T_RFCA: dict(freq=500.0e6, tReference=-1.0),
    T_MAXAMP: dict(exponent=2),
    T_MULT: dict(factor=1.0, order=1, nSlices=4),
    T_ROTATE: dict(on_pass=-1),
    T_MALIGN: dict(on_pass=-1, startPID=-1, endPID=-1),
}

# elegant parameter names -> struct fields where they differ
ALIASES = {
    T_KQUAD: {"B": "B", "N_KICKS": "n_kicks", "N_SLICES": "nSlices", "HKICK": "xkick", "VKICK": "ykick", "SYNCH_RAD": "synch_rad",
              "HCALIBRATION": "xKickCalibration", "VCALIBRATION": "yKickCalibration",
              "INTEGRATION_ORDER": "integration_order", "EXPAND_HAMILTONIAN": "expandHamiltonian"},
    T_KSEXT: {"B": "B", "N_KICKS": "n_kicks", "N_SLICES": "nSlices", "HKICK": "xkick", "VKICK": "ykick", "SYNCH_RAD": "synch_rad",
              "HCALIBRATION": "xKickCalibration", "VCALIBRATION": "yKickCalibration",
              "INTEGRATION_ORDER": "integration_order", "EXPAND_HAMILTONIAN": "expandHamiltonian"},
    T_KOCT: {"B": "B", "N_KICKS": "n_kicks", "N_SLICES": "nSlices", "INTEGRATION_ORDER": "integration_order", "SYNCH_RAD": "synch_rad",
             "EXPAND_HAMILTONIAN": "expandHamiltonian"},
    T_KQUSE: {"N_KICKS": "n_kicks", "N_SLICES": "nSlices", "INTEGRATION_ORDER": "integration_order", "SYNCH_RAD": "synch_rad",
              "EXPAND_HAMILTONIAN": "expandHamiltonian"},
    T_CSBEND: {"N_KICKS": "nSlices", "N_SLICES": "nSlices", "XKICK": "xKick", "YKICK": "yKick",
               "FSE_DIPOLE": "fseDipole", "FSE_QUADRUPOLE": "fseQuadrupole",
               "EDGE1_EFFECTS": "edge1_effects", "EDGE2_EFFECTS": "edge2_effects", "EDGE_ORDER": "edge_order",
               "INTEGRATION_ORDER": "integration_order", "EXPAND_HAMILTONIAN": "expandHamiltonian",
               "USE_BN": "use_bn", "EXPANSION_ORDER": "expansionOrder", "ETILT_SIGN": "etiltSign",
               "MALIGN_METHOD": "malignMethod", "NONLINEAR": "nonlinear", "SYNCH_RAD": "synch_rad",
               "EDGE1_KICK_LIMIT": "edge1_kick_limit", "EDGE2_KICK_LIMIT": "edge2_kick_limit",
               "KICK_LIMIT_SCALING": "kick_limit_scaling"},
    T_RFCA: {"N_KICKS": "nKicks", "CHANGE_P0": "change_p0", "CHANGE_T": "change_t", "END1_FOCUS": "end1Focus",
             "PHASE_REFERENCE": "phase_reference",
             "END2_FOCUS": "end2Focus", "STANDING_WAVE": "standingWave", "LINEARIZE": "linearize",
             "LOCK_PHASE": "lockPhase", "T_REFERENCE": "tReference", "VOLT": "volt"},
    T_MULT: {"N_SLICES": "nSlices", "N_KICKS": "nSlices", "KNL": "KnL", "BTIPL": "BTipL",
             "SYNCH_RAD": "synch_rad", "EXPAND_HAMILTONIAN": "expandHamiltonian"},
    T_ROTATE: {"ON_PASS": "on_pass", "EXCLUDE_FLOOR": "exclude_floor", "EXCLUDE_OPTICS": "exclude_optics"},
    T_MALIGN: {"ON_PASS": "on_pass", "START_PID": "startPID", "END_PID": "endPID"},
    T_ENERGY: {"CENTRAL_ENERGY": "central_energy", "CENTRAL_MOMENTUM": "central_momentum",
               "MATCH_BEAMLINE": "match_beamline", "MATCH_PARTICLES": "match_particles"},
    T_MAXAMP: {"YEXPONENT": "y_exponent"},
    T_DRIF: {},
    T_EDRIFT: {},
}


def make_struct(etype, params):
    """Build the ctypes struct for an element from elegant-style parameters."""
    cls = STRUCT_OF[etype]
    s = cls()
    vals = dict(DEFAULTS.get(etype, {}))
    alias = ALIASES.get(etype, {})
    fields = {f[0] for f in cls._fields_}
    for k, v in params.items():
        key = alias.get(k.upper(), None)
        if key is None:
            key = k if k in fields else k.lower()
            if key == "l":
                key = "length"
        if key not in fields:
            raise KeyError("parameter %s not supported for %s" % (k, TYPE_NAMES[etype]))
        vals[key] = v
    for k, v in vals.items():
        setattr(s, k, v)
    return s
