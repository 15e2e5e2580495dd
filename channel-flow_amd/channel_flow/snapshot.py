"""Snapshot files in the reference's stream layout (input_output.f90:231-384
reader, :393-544 writer): for x, y, z, xm, ym, zm an int32 count then the
float64 values; then U, V, W, P and nu_t, each as three int32 extents then
the Fortran-ordered float64 data.  All z planes are written (the reference's
one-rank writer drops the last one, so its files do not round-trip).

Restart state the format has no room for (the 7 device scalars
t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0 and tau_eqwm_initial) goes to
a side file <snapshot>.restart, which a reference reader never opens.
Byte order is native unless big_endian (the reference ran with
F_UFMTENDIAN=big)."""
import os

import numpy as np

SCALARS = ["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"]


def _centres(f):
    return 0.5 * (f[:-1] + f[1:])


def write_snapshot(path, x, y, z, U, V, W, P=None, nu_t=None, big_endian=False):
    """The reference's record sequence; P and nu_t are written when given
    (a reference file always has both)."""
    e = ">" if big_endian else "<"
    i4, f8 = np.dtype(e + "i4"), np.dtype(e + "f8")
    with open(path, "wb") as f:
        for a in (x, y, z, _centres(x), _centres(y), _centres(z)):
            np.array([len(a)], dtype=i4).tofile(f)
            np.asarray(a, dtype=f8).tofile(f)
        for a in (U, V, W, P, nu_t):
            if a is None:
                continue
            np.array(a.shape, dtype=i4).tofile(f)
            np.asarray(a, dtype=f8).ravel(order="F").tofile(f)


def read_snapshot(path, big_endian=False):
    """Grids and fields of a snapshot.  After U, V and W, a P record and a
    nu_t record are taken when present (each recognised by its extents
    being those of a cell-centred field); anything after them is ignored,
    as the reference's reader ignores everything after W."""
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
    nx, ny, nz = len(grids[0]), len(grids[1]), len(grids[2])
    shapes = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz),
              (nx + 1, ny + 1, nz + 1), (nx + 1, ny + 1, nz + 1)]
    names = ["U", "V", "W", "P", "nu_t"]
    out = dict(x=grids[0], y=grids[1], z=grids[2])
    for name, shp in zip(names, shapes):
        if pos + 12 > len(raw):
            break
        got = tuple(int(v) for v in np.frombuffer(raw, dtype=i4, count=3, offset=pos))
        if got != shp:
            if name in ("U", "V", "W"):
                raise ValueError("wrong size in input file (%s): %s" % (name, got))
            break
        cnt = shp[0] * shp[1] * shp[2]
        if pos + 12 + cnt * 8 > len(raw):
            raise ValueError("truncated %s record" % name)
        pos += 12
        out[name] = take(f8, cnt).reshape(shp, order="F").astype(np.float64)
    return out


def restart_path(path):
    return str(path) + ".restart"


def write_restart(path, scalars, tau_init=None, big_endian=False):
    """Side file <path>.restart: the 7 scalars (SCALARS order), then the
    extents and values of tau_eqwm_initial (0 0 0 when absent)."""
    e = ">" if big_endian else "<"
    with open(restart_path(path), "wb") as f:
        np.asarray([scalars[k] for k in SCALARS], dtype=e + "f8").tofile(f)
        if tau_init is None:
            np.zeros(3, dtype=e + "i4").tofile(f)
        else:
            np.array(tau_init.shape, dtype=e + "i4").tofile(f)
            np.asarray(tau_init, dtype=e + "f8").ravel(order="F").tofile(f)


def read_restart(path, big_endian=False):
    """(scalars dict, tau_eqwm_initial or None), or None without a side file."""
    p = restart_path(path)
    if not os.path.exists(p):
        return None
    e = ">" if big_endian else "<"
    raw = open(p, "rb").read()
    s = np.frombuffer(raw, dtype=e + "f8", count=7)
    shp = tuple(int(v) for v in np.frombuffer(raw, dtype=e + "i4", count=3, offset=56))
    tau = None
    if shp[0] * shp[1] * shp[2] > 0:
        tau = np.frombuffer(raw, dtype=e + "f8", count=shp[0] * shp[1] * shp[2], offset=68).reshape(shp, order="F")
    return dict(zip(SCALARS, (float(v) for v in s))), tau
