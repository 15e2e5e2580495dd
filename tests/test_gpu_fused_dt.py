"""compute_dt (time_integration.f90:179-227) taken from the last projection
of each step (the fused form the graph-replayed steps use) against the
standalone compute_dt before every step (the reference driver's order,
main.f90:70-73): the same dt bit for bit, so the same fields bit for bit.
The flow is made fast (|U| near 50, |W| near 30) on a uniform y grid so the
convective limit binds and dt changes from step to step."""
import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def fast_cfg(precision="fp64", x_mass_cte=1):
    c = cases.case_dns64()
    c.cells = (32, 24, 16)
    c.stretch = 0.0
    c.nu = 1e-4
    c.cfl = 0.5
    c.dPdx = 0.0
    c.x_mass_cte = x_mass_cte
    c.precision = precision
    return c


def fast_field(cfg):
    U, V, W = cases.initial_field(cfg)
    U = U + 48.5
    W = W + 30.0
    return np.asfortranarray(U), V, np.asfortranarray(W)


@pytest.mark.parametrize("precision,xm", [("fp64", 1), ("fp64", 0), ("mixed", 1)])
def test_fused_dt_equals_standalone(precision, xm):
    from channel_flow.solver import ChannelFlow
    cfg = fast_cfg(precision, xm)
    a, b = ChannelFlow(cfg), ChannelFlow(cfg)
    U, V, W = fast_field(cfg)
    a.initialise(U, V, W)
    b.initialise(U, V, W)
    dts = []
    for _ in range(6):
        a.run(1, graph=True)                  # dt from the previous step's projection
        dts.append(b.compute_dt())            # standalone reduction
        b.compute_time_step_RK3()
        assert a.scalars["dt"] == dts[-1]
        for x, y in zip(a.get_state(), b.get_state()):
            np.testing.assert_array_equal(x, y)
    assert max(dts) < 5e-3 and len(set(dts)) > 1, dts  # the convective limit binds and moves


def test_fused_dt_steps_match_oracle():
    from channel_flow.solver import ChannelFlow
    cfg = fast_cfg()
    g = ChannelFlow(cfg)
    o = Oracle(cfg)
    U, V, W = fast_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for x, y in zip(g.get_state(), o.get_state()):
        assert float(np.max(np.abs(x - y)) / np.max(np.abs(y))) <= 1e-10
    assert abs(g.scalars["t"] - o.scalars["t"]) <= 1e-14 * o.scalars["t"]
    # the two trajectories differ at round-off, so their last dt by an ulp
    assert abs(g.scalars["dt"] - o.scalars["dt"]) <= 1e-13 * o.scalars["dt"]
