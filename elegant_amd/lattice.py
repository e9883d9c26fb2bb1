"""Beamline description for the tracking path.

A Lattice is an ordered list of elements with elegant's type names and
parameter names (src/track_data.c PARAMETER tables).  It marshals to the
EH_ELEMENT_LIST array the C ABI takes, the analogue of elegant's linked
ELEMENT_LIST (src/track.h:310-349).
"""
import copy
import ctypes as C

from . import structs as S


class Element:
    def __init__(self, name, etype, **params):
        if isinstance(etype, str):
            etype = S.TYPE_CODES[etype.upper()]
        self.name = name
        self.type = etype
        self.params = dict(params)

    def length(self):
        if self.type in (S.T_MARK, S.T_MAXAMP, S.T_ROTATE, S.T_MALIGN, S.T_ENERGY):
            return 0.0
        for k in ("L", "length"):
            if k in self.params:
                return float(self.params[k])
        return 0.0

    def __repr__(self):
        return "%s: %s %r" % (self.name, S.TYPE_NAMES[self.type], self.params)


class Lattice:
    def __init__(self, elements=None):
        self.elements = list(elements or [])
        self._structs = None
        self._cbl = None

    def append(self, el):
        self.elements.append(el)
        self._structs = None
        self._cbl = None
        return self

    def extend(self, els):
        for e in els:
            self.append(e)
        return self

    def copy(self):
        return Lattice([copy.deepcopy(e) for e in self.elements])

    def __len__(self):
        return len(self.elements)

    def total_length(self):
        return sum(e.length() for e in self.elements)

    def n_counted(self):
        """elements that count as particle-element units (not MARK/MAXAMP)"""
        return sum(1 for e in self.elements if e.type not in (S.T_MARK, S.T_MAXAMP))

    def _edge_flags(self, i):
        """BEND_EDGE1/2_EFFECTS as determine_bend_flags sets them (src/bend_matrix.c:1330-1393):
        an edge is active unless the same bend (type and name) is adjacent on that side,
        skipping MARK elements."""
        el = self.elements[i]
        flags = 0

        def neighbour(step):
            j = i + step
            while 0 <= j < len(self.elements):
                o = self.elements[j]
                if o.type == el.type and o.name == el.name:
                    return True
                if o.type == S.T_MARK:
                    j += step
                    continue
                return False
            return False

        p = el.params
        e1 = p.get("EDGE1_EFFECTS", p.get("edge1_effects", 1))
        e2 = p.get("EDGE2_EFFECTS", p.get("edge2_effects", 1))
        if e1 and not neighbour(-1):
            flags |= S.BEND_EDGE1_EFFECTS
        if e2 and not neighbour(+1):
            flags |= S.BEND_EDGE2_EFFECTS
        return flags

    def structs(self):
        """ctypes element structs (cached; RFCA fiducial state lives here between calls)."""
        if self._structs is None:
            out = []
            for i, e in enumerate(self.elements):
                if e.type == S.T_MARK:
                    out.append(None)
                    continue
                s = S.make_struct(e.type, e.params)
                if e.type == S.T_CSBEND:
                    s.edgeFlags = self._edge_flags(i)
                out.append(s)
            self._structs = out
        return self._structs

    def set_param(self, i, name, value):
        """change one parameter of element i, in its struct too (tracking sees it)"""
        el = self.elements[i]
        el.params[name] = value
        if self._structs is not None and self._structs[i] is not None:
            new = S.make_struct(el.type, el.params)
            if el.type == S.T_CSBEND:
                new.edgeFlags = self._edge_flags(i)
            if el.type == S.T_RFCA:  # keep the fiducial state the run has set
                for f in ("fiducial_seen", "phase_fiducial", "t_fiducial"):
                    setattr(new, f, getattr(self._structs[i], f))
            C.memmove(C.addressof(self._structs[i]), C.addressof(new), C.sizeof(new))

    def reset_state(self):
        """forget RFCA fiducials (a new run)"""
        for s in self.structs():
            if isinstance(s, S.EH_RFCA):
                s.fiducial_seen = 0
                s.phase_fiducial = 0.0
                s.t_fiducial = 0.0

    def c_beamline(self):
        """(EH_ELEMENT_LIST array, keepalive) for the C ABI; the same structs are shared,
        so state written by tracking (RFCA fiducials) persists in this Lattice.  Built
        once and cached (the structs are updated in place by set_param)."""
        if self._cbl is None or self._structs is None:
            self._cbl = self._build_c_beamline()
        return self._cbl

    def _build_c_beamline(self):
        structs = self.structs()
        arr = (S.EH_ELEMENT_LIST * max(1, len(self.elements)))()
        keep = [arr]
        occ = {}
        for i, (e, s) in enumerate(zip(self.elements, structs)):
            name = e.name.encode()
            keep.append(name)
            occ[e.name] = occ.get(e.name, 0) + 1
            arr[i].type = e.type
            arr[i].occurence = occ[e.name]
            arr[i].name = name
            arr[i].p_elem = C.cast(C.pointer(s), C.c_void_p) if s is not None else None
        return arr, keep
