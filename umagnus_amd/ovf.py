"""OVF 1.0 / 2.0 field files (OOMMF), the exchange format of the reference.

Restates the reference's reader and writers (oommf/oommf.go:17-214,
oommf/ovf1.go, oommf/ovf2.go):

- OVF 2.0 "Binary 4"/"Binary 8" data is little endian and starts with the control
  number 1234567.0 / 123456789012345.0; OVF 1.0 binary data is big endian.
- Text data lists one cell per line, components in order, x fastest.
- Header keys are matched case-insensitively; unknown keys are an error, as in
  readHeader (oommf.go:84-170).
- `valuemultiplier` is parsed but not applied, again as in the reference.

Arrays are numpy arrays shaped (ncomp, Nz, Ny, Nx), the SoA layout of the
simulation handle.
"""
import io

import numpy as np

CONTROL4 = 1234567.0
CONTROL8 = 123456789012345.0

_IGNORED = {"oommf", "segment count", "begin", "meshtype", "xbase", "ybase", "zbase", "xmin", "ymin", "zmin",
            "xmax", "ymax", "zmax", "valuerangeminmag", "valuerangemaxmag", "end", "", "valuelabels",
            "valueunit", "meshunit"}


class OvfError(ValueError):
    pass


def _split(line):
    """'# key: value' -> ('key', 'value'), both stripped of '#' and blanks (parseHeaderLine)."""
    key, sep, value = line.partition(":")
    key = key.strip("# \t\r")
    return key, (value.strip("# \t\r") if sep else "")


def _read_line(f):
    b = f.readline()
    if not b:
        return None
    return b.decode("latin-1").rstrip("\n")


def read(path_or_file, keep_float64=False):
    """Read an OVF1/OVF2 file.  Returns (data, meta): data (ncomp, Nz, Ny, Nx) in
    float32, like the reference's float32 data.Slice (Binary 8 values are rounded,
    ovf2.go readOVF2DataBinary8); keep_float64 keeps Binary 8 data in float64.
    meta: title, ncomp, size, cellsize, unit, time, format, ovf, desc,
    valuemultiplier."""
    f = open(path_or_file, "rb") if isinstance(path_or_file, (str, bytes)) or hasattr(path_or_file, "__fspath__") \
        else path_or_file
    try:
        first = _read_line(f)
        if first is None:
            raise OvfError("empty file")
        tag = first.strip().lower()
        if tag == "# oommf ovf 2.0":
            ovf, ncomp = 2, None
        elif tag == "# oommf: rectangular mesh v1.0":
            ovf, ncomp = 1, 3   # OVF1 only supports vectors
        else:
            raise OvfError("unknown header: " + first)
        meta = {"ovf": ovf, "title": "", "unit": "", "size": [0, 0, 0], "cellsize": [0.0, 0.0, 0.0],
                "desc": {}, "time": 0.0, "valuemultiplier": 1.0}
        while True:
            line = _read_line(f)
            if line is None:
                raise OvfError("missing 'Begin: Data'")
            norm = line.strip("# \t\r").lower().replace(" ", "")
            if norm.startswith("begin:data"):
                break
            key, value = _split(line)
            k = key.lower()
            if k in _IGNORED:
                continue
            if k == "title":
                meta["title"] = value
            elif k == "valueunits":
                meta["unit"] = value.split(" ")[0]
            elif k == "valuedim":
                ncomp = int(value)
            elif k in ("xnodes", "ynodes", "znodes"):
                meta["size"]["xyz".index(k[0])] = int(value)
            elif k in ("xstepsize", "ystepsize", "zstepsize"):
                meta["cellsize"]["xyz".index(k[0])] = float(value)
            elif k == "valuemultiplier":
                meta["valuemultiplier"] = float(value)
            elif k == "desc":
                dk, _, dv = value.partition(":")
                meta["desc"][dk.strip("# ")] = dv.strip("# ")
            else:
                raise OvfError("Unknown key: " + key)
        _, value = _split(line)
        words = value.strip().split(" ", 2)
        if len(words) < 2 or words[0].lower() != "data":
            raise OvfError("Expected: Begin: Data")
        fmt = ("binary " + words[2].strip()).lower() if len(words) >= 3 else "text"
        if "Time (s)" in meta["desc"]:
            meta["time"] = float(meta["desc"]["Time (s)"] or 0)
        if "Total simulation time" in meta["desc"]:
            meta["time"] = float(meta["desc"]["Total simulation time"].split()[0])
        if meta["cellsize"] == [0.0, 0.0, 0.0]:
            meta["cellsize"] = [1.0, 1.0, 1.0]   # default (unitless) cell size
        meta["ncomp"] = ncomp
        meta["format"] = fmt
        Nx, Ny, Nz = meta["size"]
        count = ncomp * Nx * Ny * Nz
        if fmt == "text":
            vals = []
            while len(vals) < count:
                line = _read_line(f)
                if line is None:
                    raise OvfError("unexpected end of text data")
                if line.startswith("#"):
                    continue
                vals.extend(float(v) for v in line.split())
            flat = np.asarray(vals[:count], np.float32)
        elif fmt in ("binary 4", "binary 8"):
            w = 4 if fmt == "binary 4" else 8
            order = ">" if ovf == 1 else "<"
            dt = np.dtype(f"{order}f{w}")
            ctrl = np.frombuffer(f.read(w), dt)[0]
            if ctrl != (CONTROL4 if w == 4 else CONTROL8):
                raise OvfError(f"invalid OVF{ovf} control number: {ctrl}")
            raw = f.read(w * count)
            if len(raw) != w * count:
                raise OvfError("unexpected end of binary data")
            flat = np.frombuffer(raw, dt).astype(np.float64 if (w == 8 and keep_float64) else np.float32)
        else:
            raise OvfError(f"unknown format: OVF{ovf} {fmt}")
        data = flat.reshape(Nz, Ny, Nx, ncomp).transpose(3, 0, 1, 2).copy()
        return data, meta
    finally:
        if f is not path_or_file:
            f.close()


def _hdr(out, key, *values):
    # hdr(): "# key: " then the values separated by blanks (oommf.go:203-208)
    out.write(f"# {key}: " + " ".join(_fmt(v) for v in values) + "\n")


def go_float(v):
    """Go's %v of a float64 (strconv 'g', precision -1): the shortest decimal
    that reads back to v, in exponent form (two exponent digits at least) when
    the decimal exponent is below -4 or at least 6; integral values carry no
    '.0'."""
    import decimal
    import math
    if v != v:
        return "NaN"
    if math.isinf(v):
        return "+Inf" if v > 0 else "-Inf"
    if v == 0:
        return "-0" if math.copysign(1.0, v) < 0 else "0"
    sign, digits, e = decimal.Decimal(repr(v)).as_tuple()   # repr: shortest round-trip digits
    digits = list(digits)
    while len(digits) > 1 and digits[-1] == 0:
        digits.pop()
        e += 1
    while len(digits) > 1 and digits[0] == 0:
        digits.pop(0)
    d = "".join(map(str, digits))
    exp = len(d) - 1 + e
    out = "-" if sign else ""
    if exp < -4 or exp >= 6:
        out += d[0] + ("." + d[1:] if len(d) > 1 else "") + ("e-%02d" % -exp if exp < 0 else "e+%02d" % exp)
    elif exp >= 0:
        ip = d[:exp + 1].ljust(exp + 1, "0")
        out += ip + ("." + d[exp + 1:] if len(d) > exp + 1 else "")
    else:
        out += "0." + "0" * (-exp - 1) + d
    return out


def _fmt(v):
    # fmt.Fprintln's spelling of an operand
    return go_float(v) if isinstance(v, float) else str(v)


def write(path, data, cellsize, name="m", unit="1", time=0.0, fmt="binary 4", ovf=2):
    """Write (ncomp, Nz, Ny, Nx) data as OVF2 (ovf=2: 'binary 4' or 'text') or OVF1
    (ovf=1, vectors only: 'binary 4' or 'text'), following writeOVF2Header /
    writeOVF1Header field by field."""
    data = np.asarray(data)
    if data.ndim == 3:
        data = data[None]
    ncomp, Nz, Ny, Nx = data.shape
    fmt = fmt.lower()
    if fmt == "binary":
        fmt = "binary 4"
    if fmt not in ("binary 4", "text"):
        raise OvfError(f"Illegal OVF data format: {fmt}. Options are: Text, Binary 4")
    canon = "Binary 4" if fmt == "binary 4" else "Text"
    cx, cy, cz = (float(c) for c in cellsize)
    out = io.StringIO()
    if ovf == 2:
        out.write("# OOMMF OVF 2.0\n")
        _hdr(out, "Segment count", "1")
        _hdr(out, "Begin", "Segment")
        _hdr(out, "Begin", "Header")
        _hdr(out, "Title", name)
        _hdr(out, "meshtype", "rectangular")
        _hdr(out, "meshunit", "m")
        for k in ("xmin", "ymin", "zmin"):
            _hdr(out, k, 0)
        _hdr(out, "xmax", cx * Nx)
        _hdr(out, "ymax", cy * Ny)
        _hdr(out, "zmax", cz * Nz)
        labels = [name] if ncomp == 1 else [f"{name}_{chr(ord('x') + i)}" for i in range(ncomp)]
        _hdr(out, "valuedim", ncomp)
        _hdr(out, "valuelabels", *labels)
        _hdr(out, "valueunits", *([unit or "1"] * (1 if ncomp == 1 else 3)))
        _hdr(out, "Desc", "Total simulation time: ", time, " s")
        _hdr(out, "xbase", cx / 2)
        _hdr(out, "ybase", cy / 2)
        _hdr(out, "zbase", cz / 2)
        _hdr(out, "xnodes", Nx)
        _hdr(out, "ynodes", Ny)
        _hdr(out, "znodes", Nz)
        _hdr(out, "xstepsize", cx)
        _hdr(out, "ystepsize", cy)
        _hdr(out, "zstepsize", cz)
        _hdr(out, "End", "Header")
    elif ovf == 1:
        if ncomp != 3:
            raise OvfError("Cannot save the quantity: the OVF1 format only supports 3D-vector fields.")
        _hdr(out, "OOMMF", "rectangular mesh v1.0")
        _hdr(out, "Segment count", "1")
        _hdr(out, "Begin", "Segment")
        _hdr(out, "Begin", "Header")
        _hdr(out, "Desc", "Time (s)", ": ", time)
        _hdr(out, "Title", name)
        _hdr(out, "meshtype", "rectangular")
        _hdr(out, "meshunit", "m")
        _hdr(out, "xbase", cx / 2)
        _hdr(out, "ybase", cy / 2)
        _hdr(out, "zbase", cz / 2)
        _hdr(out, "xstepsize", cx)
        _hdr(out, "ystepsize", cy)
        _hdr(out, "zstepsize", cz)
        for k in ("xmin", "ymin", "zmin"):
            _hdr(out, k, 0)
        _hdr(out, "xmax", cx * Nx)
        _hdr(out, "ymax", cy * Ny)
        _hdr(out, "zmax", cz * Nz)
        _hdr(out, "xnodes", Nx)
        _hdr(out, "ynodes", Ny)
        _hdr(out, "znodes", Nz)
        _hdr(out, "ValueRangeMinMag", 1e-08)
        _hdr(out, "ValueRangeMaxMag", 1)
        _hdr(out, "valueunit", unit)
        _hdr(out, "valuemultiplier", 1)
        _hdr(out, "End", "Header")
    else:
        raise OvfError("ovf must be 1 or 2")
    _hdr(out, "Begin", "Data " + canon)
    cells = data.astype(np.float32).transpose(1, 2, 3, 0)   # (Nz, Ny, Nx, ncomp), x fastest
    with open(path, "wb") as fh:
        fh.write(out.getvalue().encode("latin-1"))
        if fmt == "text":
            rows = cells.reshape(-1, ncomp)
            fh.write("".join(" ".join(repr(float(v)) for v in r) + " \n" for r in rows).encode("latin-1"))
        else:
            dt = np.dtype(">f4") if ovf == 1 else np.dtype("<f4")
            fh.write(np.asarray([CONTROL4], dt).tobytes())
            fh.write(cells.astype(dt).tobytes())
        tail = io.StringIO()
        _hdr(tail, "End", "Data " + canon)
        _hdr(tail, "End", "Segment")
        fh.write(tail.getvalue().encode("latin-1"))
