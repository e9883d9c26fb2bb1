"""SDDS (ASCII) particle files in elegant's phase-space layout.

The same seeded particles can be handed to an unmodified elegant (its
sdds_beam input) and its output (watch points, final coordinates) read back for
comparison.  Layout and conversions follow the reference:

* columns x, xp, y, yp, t, p, particleID and parameters Step, pCentral, Charge,
  Particles, IDSlotsPerBunch (src/sdds_support.c:153-171);
* written rows: p = pCentral (1 + delta), t = s / (c p / sqrt(p^2 + 1))
  (dump_phase_space, src/sdds_support.c:1589-1594);
* read rows: s = t c p / sqrt(p^2 + 1), delta = (p - pCentral) / pCentral, loss
  pass, bunch = -1, weight = 1 (new_sdds_beam, src/sdds_beam.c:418-437).

Files are written as ASCII or, as elegant writes them by default, binary (SDDS1
header with "!# little-endian", then per page the row count as a 32-bit
integer, the parameter values and the rows, each value in its declared binary
type, strings as a 32-bit length and the characters).  The reader takes either
flavour (binary rows in row or column major order, either byte order), these
columns in any order, extra columns and parameters ignored, several pages
concatenated.
"""
import math
import struct

import numpy as np

from . import structs as S

COLUMNS = (("x", "units=m, type=double"), ("xp", "symbol=\"x'\", type=double"),
           ("y", "units=m, type=double"), ("yp", "symbol=\"y'\", type=double"),
           ("t", "units=s, type=double"), ("p", "units=\"m$be$nc\", type=double"),
           ("particleID", "type=ulong64"))


def write_phase_space(path, rows, p_central, step=1, charge=0.0, n=None, binary=False):
    """rows: (n, 10) elegant-layout coordinates; the first n (default all) are
    written, as ASCII or (binary=True) binary SDDS"""
    rows = np.asarray(rows, dtype=np.float64)
    n = len(rows) if n is None else int(n)
    lines = ["SDDS1"]
    if binary:
        lines.append("!# little-endian")
    lines.append('&parameter name=Step, type=long, description="Simulation step" &end')
    lines.append('&parameter name=pCentral, symbol="p$bcen$n", units="m$be$nc", type=double, '
                 'description="Reference beta*gamma" &end')
    lines.append('&parameter name=Charge, type=double, units=C, description="Bunch charge before sampling" &end')
    lines.append('&parameter name=Particles, type=long, description="Number of particles before sampling" &end')
    lines.append('&parameter name=IDSlotsPerBunch, type=long, '
                 'description="Number of particle ID slots reserved to a bunch" &end')
    for name, attrs in COLUMNS:
        lines.append("&column name=%s, %s &end" % (name, attrs))
    if binary:
        lines.append("&data mode=binary, &end")
        p = p_central * (1 + rows[:n, 5])
        t = rows[:n, 4] / (S.C_MKS * p / np.sqrt(p * p + 1))
        rec = np.zeros(n, dtype=_ROW_DTYPE)
        for k, c in enumerate(("x", "xp", "y", "yp")):
            rec[c] = rows[:n, k]
        rec["t"], rec["p"] = t, p
        rec["particleID"] = rows[:n, 6].astype(np.uint64)
        with open(path, "wb") as f:
            f.write(("\n".join(lines) + "\n").encode())
            f.write(struct.pack("<iiddii", n, int(step), float(p_central), float(charge), n, n))
            f.write(rec.tobytes())
        return
    lines.append("&data mode=ascii, &end")
    lines.append("! page number 1")
    lines += [str(int(step)), repr(float(p_central)), repr(float(charge)), str(n), str(n)]
    lines.append("          %d" % n)
    for r in rows[:n]:
        p = p_central * (1 + r[5])
        t = r[4] / (S.C_MKS * p / math.sqrt(p * p + 1))
        lines.append("%s %s %s %s %s %s %d" % (repr(float(r[0])), repr(float(r[1])), repr(float(r[2])),
                                               repr(float(r[3])), repr(float(t)), repr(float(p)), int(r[6])))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


_ROW_DTYPE = np.dtype([("x", "<f8"), ("xp", "<f8"), ("y", "<f8"), ("yp", "<f8"), ("t", "<f8"), ("p", "<f8"),
                       ("particleID", "<u8")])
_TYPES = {"double": "f8", "float": "f4", "long": "i4", "ulong": "u4", "long64": "i8", "ulong64": "u8",
          "short": "i2", "ushort": "u2", "character": "S1"}


def _binary_header(raw):
    """(params [(name, type)], cols [(name, type)], byte order, column major,
    offset of the binary data) of a binary SDDS file"""
    params, cols, order, colmajor, pos = [], [], "<", False, 0
    first = True
    while True:
        end = raw.index(b"\n", pos)
        ln = raw[pos:end].decode("latin-1").strip()
        pos = end + 1
        if first:
            if not ln.startswith("SDDS"):
                raise ValueError("not an SDDS file")
            first = False
            continue
        if ln.startswith("!#"):
            order = ">" if "big-endian" in ln else "<"
        elif ln.startswith("&parameter"):
            params.append((_field(ln, "name"), _field(ln, "type")))
        elif ln.startswith("&column"):
            cols.append((_field(ln, "name"), _field(ln, "type")))
        elif ln.startswith("&data"):
            compact = ln.replace(" ", "")
            colmajor = "column_major_order=1" in compact
            if "mode=binary" not in compact:
                return None
            return params, cols, order, colmajor, pos
        elif ln.startswith("&array"):
            raise ValueError("SDDS arrays are not supported")


def _read_binary(raw, hdr):
    """rows (x, xp, y, yp, t, p, id) and the first pCentral of a binary SDDS file"""
    params, cols, order, colmajor, pos = hdr
    names = [c for c, _ in cols]
    need = ("x", "xp", "y", "yp", "t", "p")
    out, pc_file = [], None

    def scalar(typ, pos):
        if typ == "string":
            (ln,) = struct.unpack_from(order + "i", raw, pos)
            return raw[pos + 4:pos + 4 + ln].decode("latin-1"), pos + 4 + ln
        dt = np.dtype(order + _TYPES[typ])
        return np.frombuffer(raw, dt, 1, pos)[0], pos + dt.itemsize

    while pos < len(raw):
        (nrows,) = struct.unpack_from(order + "i", raw, pos)
        pos += 4
        for name, typ in params:
            v, pos = scalar(typ, pos)
            if name == "pCentral" and pc_file is None:
                pc_file = float(v)
        if any(t == "string" for _, t in cols):
            raise ValueError("string columns are not supported in binary SDDS")
        dts = [np.dtype(order + _TYPES[t]) for _, t in cols]
        if colmajor:
            data = {}
            for (name, _), dt in zip(cols, dts):
                data[name] = np.frombuffer(raw, dt, nrows, pos)
                pos += dt.itemsize * nrows
        else:
            rec = np.dtype([(name, dt) for (name, _), dt in zip(cols, dts)])
            arr = np.frombuffer(raw, rec, nrows, pos)
            pos += rec.itemsize * nrows
            data = {name: arr[name] for name in names}
        ids = data["particleID"].astype(np.float64) if "particleID" in data else np.arange(1, nrows + 1)
        out.append(np.column_stack([np.asarray(data[c], dtype=np.float64) for c in need] + [ids]))
    return (np.concatenate(out) if out else np.zeros((0, 7))), pc_file


def _header(lines):
    params, cols, i = [], [], 0
    if not lines or not lines[0].startswith("SDDS"):
        raise ValueError("not an SDDS file")
    i = 1
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if ln.startswith("&parameter"):
            params.append(_field(ln, "name"))
        elif ln.startswith("&column"):
            cols.append(_field(ln, "name"))
        elif ln.startswith("&data"):
            if "mode=ascii" not in ln.replace(" ", ""):
                raise ValueError("only ASCII SDDS data is supported")
            if "no_row_counts=1" in ln.replace(" ", ""):
                raise ValueError("files without row counts are not supported")
            return params, cols, i
    raise ValueError("no &data line")


def _field(line, key):
    for part in line.replace("&end", "").split(","):
        part = part.strip()
        if part.startswith("&parameter") or part.startswith("&column"):
            part = part.split(None, 1)[1] if " " in part else ""
        if part.startswith(key + "="):
            return part.split("=", 1)[1].strip().strip('"')
    raise ValueError("no %s in %r" % (key, line))


def read_phase_space(path, p_central=None):
    """returns (rows (n, 10), p_central) for all pages of the file; p_central
    defaults to the file's pCentral parameter (first page)"""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = _binary_header(raw)
    if hdr is not None:
        arr, pc_file = _read_binary(raw, hdr)
        return _to_rows(arr, p_central if p_central is not None else pc_file)
    lines = raw.decode("latin-1").splitlines()
    params, cols, i = _header(lines)
    need = ("x", "xp", "y", "yp", "t", "p")
    idx = [cols.index(c) for c in need]
    id_col = cols.index("particleID") if "particleID" in cols else None
    data = [ln for ln in lines[i:] if ln.strip() and not ln.lstrip().startswith("!")]
    out, pc_file, k = [], None, 0
    while k < len(data):
        pvals = data[k:k + len(params)]
        k += len(params)
        if "pCentral" in params and pc_file is None:
            pc_file = float(pvals[params.index("pCentral")])
        nrows = int(data[k])
        k += 1
        for j in range(nrows):
            tok = data[k + j].split()
            out.append([float(tok[c]) for c in idx] + [float(tok[id_col]) if id_col is not None else j + 1])
        k += nrows
    return _to_rows(np.array(out, dtype=np.float64).reshape(-1, 7), p_central if p_central is not None else pc_file)


def _to_rows(arr, pc):
    if pc is None:
        raise ValueError("no pCentral in the file and none given")
    rows = np.zeros((len(arr), S.NPROP))
    rows[:, :4] = arr[:, :4]
    p = arr[:, 5]
    gamma = np.sqrt(p * p + 1)
    rows[:, 4] = arr[:, 4] * S.C_MKS * p / gamma
    rows[:, 5] = (p - pc) / pc
    rows[:, 6] = arr[:, 6]
    rows[:, 7] = -1
    rows[:, 8] = -1
    rows[:, 9] = 1
    return rows, pc
