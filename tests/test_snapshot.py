"""Snapshot format (CPU): round trip, the reference's exact record sequence,
and the restart side file."""
import numpy as np
import pytest

from channel_flow import cases
from channel_flow.snapshot import read_restart, read_snapshot, write_restart, write_snapshot


def _fields(cells=(8, 6, 4)):
    cfg = cases.case_dns64()
    cfg.cells = cells
    x, y, z = cfg.grid()
    U, V, W = cases.initial_field(cfg)
    rng = np.random.default_rng(0)
    shp = (U.shape[0] + 1, U.shape[1], U.shape[2])
    P = np.asfortranarray(rng.standard_normal(shp))
    nt = np.asfortranarray(np.abs(rng.standard_normal(shp)) * 1e-3)
    return x, y, z, U, V, W, P, nt


def test_snapshot_roundtrip(tmp_path):
    x, y, z, U, V, W, P, nt = _fields()
    for be in (False, True):
        p = tmp_path / ("s%d" % be)
        write_snapshot(p, x, y, z, U, V, W, P, nt, big_endian=be)
        d = read_snapshot(p, big_endian=be)
        for k, a in dict(x=x, y=y, z=z, U=U, V=V, W=W, P=P, nu_t=nt).items():
            np.testing.assert_array_equal(d[k], a)
    raw = open(tmp_path / "s0", "rb").read()
    assert np.frombuffer(raw[:4], "<i4")[0] == len(x)


def test_reads_the_reference_record_sequence(tmp_path):
    """A file laid out record by record as the reference writes it
    (input_output.f90:411-544: Shape(x), x ... then nx,nyg,nzg, U ...,
    P, then nu_t in the P slot's extents), composed here byte by byte: the
    nu_t record is read as nu_t, and nothing is taken for restart state."""
    x, y, z, U, V, W, P, nt = _fields((10, 8, 6))
    i4 = lambda *v: np.array(v, dtype="<i4").tobytes()
    f8 = lambda a: np.asarray(a, dtype="<f8").ravel(order="F").tobytes()
    xm, ym, zm = (0.5 * (a[1:] + a[:-1]) for a in (x, y, z))
    nx, ny, nz = len(x), len(y), len(z)
    raw = b"".join([i4(nx), f8(x), i4(ny), f8(y), i4(nz), f8(z), i4(nx - 1), f8(xm), i4(ny - 1), f8(ym),
                    i4(nz - 1), f8(zm),
                    i4(nx, ny + 1, nz + 1), f8(U), i4(nx + 1, ny, nz + 1), f8(V), i4(nx + 1, ny + 1, nz), f8(W),
                    i4(nx + 1, ny + 1, nz + 1), f8(P), i4(nx + 1, ny + 1, nz + 1), f8(nt)])
    p = tmp_path / "ref_layout.100"
    p.write_bytes(raw)
    d = read_snapshot(p)
    np.testing.assert_array_equal(d["P"], P)
    np.testing.assert_array_equal(d["nu_t"], nt)
    assert read_restart(p) is None
    # the writer produces the same bytes
    write_snapshot(tmp_path / "ours", x, y, z, U, V, W, P, nt)
    assert (tmp_path / "ours").read_bytes() == raw


def test_wrong_extents_are_refused(tmp_path):
    x, y, z, U, V, W, P, nt = _fields()
    write_snapshot(tmp_path / "s", x, y, z, U, V[:, :-1, :], W)
    with pytest.raises(ValueError, match=r"\(V\)"):
        read_snapshot(tmp_path / "s")


def test_restart_side_file(tmp_path):
    s = dict(t=0.5, dt=1e-3, dPdx=2e-3, dPdz=0.0, Qflow_x_0=1.01, dPdy=-1e-5, Qflow_y_0=0.0)
    tau = np.asfortranarray(np.random.default_rng(1).uniform(1e-4, 1e-2, (10, 2, 7)))
    write_restart(tmp_path / "r", s, tau)
    s2, tau2 = read_restart(tmp_path / "r")
    assert s2 == s
    np.testing.assert_array_equal(tau2, tau)
    write_restart(tmp_path / "q", s)
    assert read_restart(tmp_path / "q")[1] is None
