"""The mass-flow options of time_integration.f90:152-171 on the GPU against
the oracle: constant centreline velocity (x_mass_cte = 2,
mass_flow.f90:192-211) and constant bulk flow in y (y_mass_cte = 1,
mass_flow.f90:147-176), alone and together, FP64 at the 1e-10 bar; the
mixed path tracks FP64."""
import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def cfg_for(xm, ym, precision="fp64", cells=(32, 24, 16)):
    c = cases.case_dns64()
    c.cells = cells
    c.x_mass_cte, c.y_mass_cte = xm, ym
    c.u_centerline = 1.3
    c.dPdx = 2e-3
    c.precision = precision
    return c


@pytest.mark.parametrize("xm,ym", [(2, 0), (0, 1), (2, 1), (1, 1)])
def test_mass_flow_options_match_oracle(xm, ym):
    from channel_flow.solver import ChannelFlow
    cfg = cfg_for(xm, ym)
    g = ChannelFlow(cfg)
    o = Oracle(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) <= 1e-10, relerr(a, b)
    sg, so = g.scalars, o.scalars
    dt = so["dt"]
    assert abs(sg["dPdx"] - so["dPdx"]) * dt <= 1e-13
    assert abs(sg["dPdy"] - so["dPdy"]) * dt <= 1e-13
    if xm == 2:
        # the corrected centreline mean of the GPU state is the target
        # (mass_flow.f90:200-209: U(1:nx-2, ny/2+1, 1:nzg-2))
        Ug = g.get_state()[0]
        jc = g.ny // 2 + 1
        assert abs(Ug[0:g.nx - 2, jc - 1, 0:g.nzg - 2].mean() - 1.3) <= 1e-12
    if ym == 1:
        # a correction was made (the projected field carries a bulk V)
        assert abs(so["dPdy"]) > 0


@pytest.mark.parametrize("xm,ym", [(2, 1)])
def test_mass_flow_options_mixed_tracks_fp64(xm, ym):
    from channel_flow.solver import ChannelFlow
    a = ChannelFlow(cfg_for(xm, ym, "fp64"))
    b = ChannelFlow(cfg_for(xm, ym, "mixed"))
    U, V, W = cases.initial_field(a.cfg)
    a.initialise(U, V, W)
    b.initialise(U, V, W)
    a.run(10, graph=True)
    b.run(10, graph=True)
    for x, y in zip(b.get_state(), a.get_state()):
        assert relerr(x, y) <= 1e-4, relerr(x, y)
    assert abs(b.scalars["dPdx"] - a.scalars["dPdx"]) <= 1e-3 * abs(a.scalars["dPdx"]) + 1e-6


def test_bulk_v_and_centerline_means_match_oracle():
    """cf_mean_mass_flow_v (mass_flow.f90:61-88) and cf_mean_centerline_u
    (mass_flow.f90:200-207) of a raw random field."""
    from channel_flow.solver import ChannelFlow
    cfg = cfg_for(0, 0)
    g = ChannelFlow(cfg)
    o = Oracle(cfg)
    rng = np.random.default_rng(4)
    U = np.asfortranarray(1.0 + rng.standard_normal(g.shape_u))
    V = np.asfortranarray(0.3 + rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    g.set_state(U, V, W)
    o.set_state(U, V, W)
    assert abs(g.compute_mean_mass_flow_V() - o.compute_mean_mass_flow_V()) <= 1e-13
    assert abs(g.compute_mean_centerline_U() - o.compute_mean_centerline_U()) <= 1e-13
