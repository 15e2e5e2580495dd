"""With Dirichlet walls the FP64 projection kernels write each projected
value to the periodic and wall ghost cells it is the source of, instead of a
ghost pass after the projection (the mixed path keeps the pass).  The state after ten steps, ghosts
included, is bit-identical to the ghost pass (CF_GHOST_COPY=0, read at
graph capture) for the step's variants that read ghosts in different ways:
plain DNS, bulk and centreline mass flow, y mass flow, rotation, CFL time
steps, mixed precision, Smagorinsky and dynamic Smagorinsky (their kernels
read the ghosts), and the lane y solve with the compiled length 384."""
import numpy as np
import pytest

from channel_flow import cases

pytestmark = pytest.mark.gpu


def _run(cfg, copies, monkeypatch, nsteps=10):
    from channel_flow.solver import ChannelFlow
    monkeypatch.setenv("CF_GHOST_COPY", "1" if copies else "0")
    g = ChannelFlow(cfg)
    g.initialise(*cases.initial_field(cfg))
    g.run(nsteps, graph=True)
    out = g.get_state(), g.scalars, (g.nu_t if cfg.les_model else None)
    g.close()
    return out


@pytest.mark.parametrize("variant", ["dns", "bulk_flow", "centreline", "y_flow", "rotation", "cfl", "mixed",
                                     "smagorinsky", "dynamic", "lane_384"])
def test_ghost_copies_are_bitwise_the_ghost_pass(variant, monkeypatch):
    cfg = cases.case_dns64()
    cfg.cells = (32, 24, 16)
    if variant == "bulk_flow":
        cfg.x_mass_cte = 1
    elif variant == "centreline":
        cfg.x_mass_cte = 2
        cfg.u_centerline = 1.2
    elif variant == "y_flow":
        cfg.x_mass_cte = 1
        cfg.y_mass_cte = 1
    elif variant == "rotation":
        cfg.omega_z = 0.05
    elif variant == "cfl":
        cfg.cfl = 0.5
    elif variant == "mixed":
        cfg.precision = "mixed"
        cfg.x_mass_cte = 1
    elif variant == "smagorinsky":
        cfg.les_model = 1
        cfg.nu = 1e-4
    elif variant == "dynamic":
        cfg.les_model = 2
        cfg.nu = 1e-4
    elif variant == "lane_384":
        monkeypatch.setenv("CF_TH_LANE", "1")
        cfg.cells = (64, 20, 384)
    print("variant %s cells %s" % (variant, cfg.cells), flush=True)
    (a, sa, na), (b, sb, nb) = _run(cfg, False, monkeypatch), _run(cfg, True, monkeypatch)
    for x, y in zip(a, b):
        assert np.array_equal(x, y), np.max(np.abs(x - y))
    for k in ("t", "dt", "dPdx", "Qflow_x_0", "dPdy"):
        assert sa[k] == sb[k], k
    if na is not None:
        assert np.array_equal(na, nb)
