"""Snapshot format round trip (CPU)."""
import numpy as np

from channel_flow import cases
from channel_flow.snapshot import read_snapshot, write_snapshot


def test_snapshot_roundtrip(tmp_path):
    cfg = cases.case_dns64()
    cfg.cells = (8, 6, 4)
    x, y, z = cfg.grid()
    U, V, W = cases.initial_field(cfg)
    P = np.asfortranarray(np.random.default_rng(0).standard_normal((U.shape[0] + 1, U.shape[1], U.shape[2])))
    for be in (False, True):
        p = tmp_path / ("s%d" % be)
        write_snapshot(p, x, y, z, U, V, W, P, trailer=[0.5, 1e-3, 1.0], big_endian=be)
        d = read_snapshot(p, big_endian=be)
        for k, a in dict(x=x, y=y, z=z, U=U, V=V, W=W, P=P).items():
            np.testing.assert_array_equal(d[k], a)
        np.testing.assert_array_equal(d["trailer"], [0.5, 1e-3, 1.0])
    # header layout: int32 count then the values
    raw = open(tmp_path / "s0", "rb").read()
    assert np.frombuffer(raw[:4], "<i4")[0] == len(x)
