"""Grids whose transform lengths carry the prime factors 5 and 7 (the LDS
transforms' radix-5 and radix-7 passes; FFTW, initialization.f90:235-278,
takes any length): ten RK3 steps against the oracle, whose DFT is a plain
mixed-radix restatement for any length.  Half x lengths 20 = 4*5, 35 = 5*7
(odd, so the real-to-complex split pairs k with M-k around an odd centre)
and 15 = 3*5; z lengths 28 = 4*7, 50 = 2*5*5 and 42 = 2*3*7."""
import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("cells", [(40, 24, 28), (70, 16, 50), (30, 20, 42),
                                   (2048, 8, 12), (12, 8, 2048), (1400, 8, 10)])
def test_ten_steps_with_radix_5_and_7(cells):
    """Also the largest transform lengths cf_create accepts (half x length
    1024, z length 2048, run-time plans) and 1400 = 2^3 5^2 7 in x."""
    from channel_flow.solver import ChannelFlow
    cfg = cases.case_dns64()
    cfg.cells = cells
    print("cells %s" % (cells,), flush=True)
    g = ChannelFlow(cfg)
    o = Oracle(cfg, jref=g.jref_eqwm)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    _, dmax = g.check_divergence()
    assert dmax < 1e-10
    g.close()


def test_mixed_precision_radix_5_and_7_tracks_fp64():
    """The FP32 transforms of the mixed path on the same lengths: the state
    after ten steps stays within FP32 round-off of the FP64 run."""
    from channel_flow.solver import ChannelFlow
    out = []
    for prec in ("fp64", "mixed"):
        cfg = cases.case_dns64()
        cfg.cells = (70, 16, 50)
        cfg.precision = prec
        g = ChannelFlow(cfg)
        g.initialise(*cases.initial_field(cfg))
        g.run(10, graph=True)
        out.append(g.get_state())
        _, dmax = g.check_divergence()
        assert dmax < (1e-10 if prec == "fp64" else 1e-4)
        g.close()
    for a, b in zip(out[1], out[0]):
        assert relerr(a, b) < 1e-5
