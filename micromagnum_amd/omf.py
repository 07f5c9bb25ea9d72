"""OMF / OVF 1.0 rectangular-mesh vector-field reader and writer.

Follows the file format handled by the reference's OMF importer/exporter
(src/magneto/mmm/io/OMFImport.cpp:240-262, OMFHeader.h:28-52,
OMFExport.cpp:72): a '#'-prefixed header with xnodes/ynodes/znodes,
x/y/zstepsize and valuemultiplier, then "Data Text", "Data Binary 4"
(check value 1234567.0) or "Data Binary 8" (check value 123456789012345.0),
values in x-fastest order, one (x, y, z) triple per cell.

Returns numpy arrays in the SoA layout used everywhere in this package:
shape (3, nz, ny, nx), float64.
"""
import numpy as np

__all__ = ["read_omf", "write_omf", "OMFField"]


class OMFField:
    def __init__(self, data, delta, header):
        self.data = data          # (3, nz, ny, nx) float64, A/m (multiplier applied)
        self.delta = delta        # (dx, dy, dz) in metres
        self.header = header      # dict of raw header entries (lower-case keys)

    @property
    def num_nodes(self):
        _, nz, ny, nx = self.data.shape
        return nx, ny, nz


def read_omf(path):
    with open(path, "rb") as f:
        raw = f.read()
    header = {}
    pos = 0
    mode = None
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("latin-1").strip()
        pos = end + 1
        if not line.startswith("#"):
            continue
        body = line[1:].strip()
        low = body.lower()
        if low.startswith("begin: data"):
            mode = low[len("begin: data"):].strip()
            break
        if ":" in body:
            k, v = body.split(":", 1)
            header[k.strip().lower()] = v.strip()
    nx, ny, nz = (int(header[k]) for k in ("xnodes", "ynodes", "znodes"))
    delta = tuple(float(header[k]) for k in ("xstepsize", "ystepsize", "zstepsize"))
    mult = float(header.get("valuemultiplier", "1.0"))
    n = nx * ny * nz
    if mode == "text":
        stop = raw.index(b"# End: Data", pos) if b"# End: Data" in raw[pos:] else len(raw)
        vals = np.array(raw[pos:stop].split(), dtype=np.float64)
    elif mode in ("binary 4", "binary 8"):
        w = 4 if mode == "binary 4" else 8
        # OVF 1.0 binary is big-endian (MSB first)
        dt = np.dtype(">f4") if w == 4 else np.dtype(">f8")
        check = np.frombuffer(raw, dtype=dt, count=1, offset=pos)[0]
        expect = 1234567.0 if w == 4 else 123456789012345.0
        if check != expect:
            raise ValueError("OMF binary check value mismatch: %r" % check)
        vals = np.frombuffer(raw, dtype=dt, count=3 * n, offset=pos + w).astype(np.float64)
    else:
        raise ValueError("unsupported OMF data mode: %r" % mode)
    if vals.size < 3 * n:
        raise ValueError("OMF file truncated: %d values for %d cells" % (vals.size, n))
    vals = vals[: 3 * n].reshape(nz, ny, nx, 3) * mult
    data = np.ascontiguousarray(np.moveaxis(vals, 3, 0))
    return OMFField(data, delta, header)


def write_omf(path, data, delta, fmt="binary 8", title="micromagnum_amd field"):
    """Write a (3, nz, ny, nx) field as OVF 1.0 ("text", "binary 4", "binary 8")."""
    data = np.asarray(data, dtype=np.float64)
    _, nz, ny, nx = data.shape
    dx, dy, dz = delta
    mag = np.sqrt((data ** 2).sum(axis=0))
    lines = [
        "# OOMMF: rectangular mesh v1.0",
        "# Segment count: 1",
        "# Begin: Segment",
        "# Begin: Header",
        "# Title: %s" % title,
        "# meshunit: m",
        "# valueunit: A/m",
        "# valuemultiplier: 1.0",
        "# xmin: 0", "# ymin: 0", "# zmin: 0",
        "# xmax: %r" % (nx * dx), "# ymax: %r" % (ny * dy), "# zmax: %r" % (nz * dz),
        "# ValueRangeMaxMag: %r" % float(mag.max()),
        "# ValueRangeMinMag: %r" % float(mag.min()),
        "# meshtype: rectangular",
        "# xbase: %r" % (dx / 2), "# ybase: %r" % (dy / 2), "# zbase: %r" % (dz / 2),
        "# xstepsize: %r" % dx, "# ystepsize: %r" % dy, "# zstepsize: %r" % dz,
        "# xnodes: %d" % nx, "# ynodes: %d" % ny, "# znodes: %d" % nz,
        "# End: Header",
    ]
    flat = np.moveaxis(data, 0, 3).reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode())
        if fmt == "text":
            f.write(b"# Begin: Data Text\n")
            for v in flat:
                f.write(("%r %r %r\n" % (float(v[0]), float(v[1]), float(v[2]))).encode())
            f.write(b"# End: Data Text\n")
        else:
            w = 4 if fmt == "binary 4" else 8
            dt = ">f4" if w == 4 else ">f8"
            f.write(("# Begin: Data Binary %d\n" % w).encode())
            chk = 1234567.0 if w == 4 else 123456789012345.0
            f.write(np.array([chk], dtype=dt).tobytes())
            f.write(flat.astype(dt).tobytes())
            f.write(("\n# End: Data Binary %d\n" % w).encode())
        f.write(b"# End: Segment\n")
