"""Snapshot files in the reference's stream layout (input_output.f90:231-384
reader, :393-544 writer): for x, y, z, xm, ym, zm an int32 count then the
float64 values; for U, V, W (and P) three int32 extents then the
Fortran-ordered float64 data.  All z planes are written (the reference's
one-rank writer drops the last one, so its files do not round-trip).  An
optional trailer holds t, dPdx, Qflow_x_0 for restarts.  Byte order is
native unless big_endian (the reference ran with F_UFMTENDIAN=big)."""
import numpy as np


def _centres(f):
    return 0.5 * (f[:-1] + f[1:])


def write_snapshot(path, x, y, z, U, V, W, P=None, trailer=None, big_endian=False):
    e = ">" if big_endian else "<"
    i4, f8 = np.dtype(e + "i4"), np.dtype(e + "f8")
    with open(path, "wb") as f:
        for a in (x, y, z, _centres(x), _centres(y), _centres(z)):
            np.array([len(a)], dtype=i4).tofile(f)
            np.asarray(a, dtype=f8).tofile(f)
        for a in (U, V, W) + ((P,) if P is not None else ()):
            np.array(a.shape, dtype=i4).tofile(f)
            np.asarray(a, dtype=f8).ravel(order="F").tofile(f)
        if trailer is not None:
            np.asarray(trailer, dtype=f8).tofile(f)


def read_snapshot(path, big_endian=False, with_pressure=None):
    e = ">" if big_endian else "<"
    i4, f8 = np.dtype(e + "i4"), np.dtype(e + "f8")
    raw = open(path, "rb").read()
    pos = 0

    def take(dt, n):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dt, count=n, offset=pos)
        pos += a.nbytes
        return a

    grids = []
    for _ in range(6):
        n = int(take(i4, 1)[0])
        grids.append(take(f8, n).copy())
    fields = []
    while pos + 12 <= len(raw) and len(fields) < 4:
        shp = tuple(int(v) for v in take(i4, 3))
        cnt = shp[0] * shp[1] * shp[2]
        if cnt <= 0 or pos + cnt * 8 > len(raw):
            pos -= 12
            break
        fields.append(take(f8, cnt).reshape(shp, order="F").astype(np.float64))
    trailer = take(f8, (len(raw) - pos) // 8).copy() if pos < len(raw) else None
    out = dict(x=grids[0], y=grids[1], z=grids[2], U=fields[0], V=fields[1], W=fields[2])
    if len(fields) > 3:
        out["P"] = fields[3]
    out["trailer"] = trailer
    return out
