"""GPU parity at the workloads' own sizes and against the reference's own
EQWM build.

- 64^3 DNS exactly as configured (tanh 2.0, nu = 1/2800, dt = 2e-3, constant
  mass flow): ten RK3 steps against the oracle at the FP64 bar.
- 128x64x128 wall-modelled LES (Smagorinsky Cs = 0.1, EQWM at h = 0.1 with
  30 cells, tol 1e-8) on the FP64 path: RK3 steps against the oracle.
- cf_eqwm_batch on every row of tests/golden/eqwm_ref.json, which the
  reference's eqwm.f90 (solve_u_eqwm, eqwm.f90:90-223) produced when built
  with flang: same iteration count, tau within 1% of the Picard tolerance
  (tol * tau_lam, the quantity eqwm.f90's stopping test uses).
- check_divergence (projection.f90:234-256): the reference's signed maximum,
  floored at 0, on fields whose divergence is not round-off.
"""
import json
import os

import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(cfg):
    from channel_flow.solver import ChannelFlow
    g = ChannelFlow(cfg)
    o = Oracle(cfg, jref=g.jref_eqwm)
    return g, o


def test_dns64_ten_steps_as_configured():
    cfg = cases.case_dns64()
    assert cfg.cells == (64, 64, 64) and cfg.stretch == 2.0 and cfg.cfl == -2e-3
    assert abs(cfg.nu - 1 / 2800.0) < 1e-18 and cfg.x_mass_cte == 1
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    errs = [relerr(a, b) for a, b in zip(g.get_state(), o.get_state())]
    print("dns64 10 steps: rel err U V W", errs)
    assert max(errs) <= 1e-10, errs
    _, dmax = g.check_divergence()
    assert dmax <= 1e-10, dmax
    sg, so = g.scalars, o.scalars
    assert abs(sg["t"] - so["t"]) <= 1e-14
    assert abs(sg["dPdx"] - so["dPdx"]) * so["dt"] <= 1e-13


@pytest.mark.parametrize("nsteps", [10])
def test_wmles128_fp64_steps(nsteps):
    cfg = cases.case_wmles128()
    cfg.precision = "fp64"
    assert cfg.cells == (128, 64, 128) and cfg.les_model == 1 and cfg.iwall_model == 3
    assert cfg.eqwm_ncv == 30 and cfg.eqwm_tol == 1e-8 and cfg.cs == 0.1
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(nsteps, graph=True)
    o.run(nsteps)
    errs = [relerr(a, b) for a, b in zip(g.get_state(), o.get_state())]
    print("wmles128 fp64 %d steps: rel err U V W" % nsteps, errs)
    assert max(errs) <= 1e-10, errs
    _, dmax = g.check_divergence()
    assert dmax <= 1e-10, dmax
    wg, wo = g.wall_state(), o.wall_state()
    werr = [relerr(a, b) for a, b in zip(wg, wo)]
    print("wmles128 wall alpha_x alpha_z tau:", werr)
    assert max(werr) <= 1e-10, werr
    assert relerr(g.nu_t, o.nu_t) <= 1e-9


def _golden_rows():
    with open(os.path.join(HERE, "golden", "eqwm_ref.json")) as f:
        return json.load(f)["rows"]


def test_eqwm_batch_matches_reference_golden_rows():
    from channel_flow.solver import eqwm_batch
    rows = _golden_rows()
    assert len(rows) == 122
    groups = {}
    for r in rows:
        groups.setdefault((r["ncv"], r["stretch"]), []).append(r)
    checked = 0
    worst = 0.0
    for (ncv, stretch), rs in sorted(groups.items()):
        u = np.array([r["u"] for r in rs])
        nu = np.array([r["nu"] for r in rs])
        h = np.array([r["h"] for r in rs])
        tin = np.array([r["tau_in"] for r in rs])
        tau, it = eqwm_batch(u, nu, h, tin, ncv=ncv, stretch=stretch, tol=1e-7, max_iter=100)
        for q, r in enumerate(rs):
            assert int(it[q]) == r["iter"], (ncv, stretch, q, int(it[q]), r["iter"])
            tau_lam = r["nu"] * r["u"] / r["h"]
            d = abs(tau[q] - r["tau"]) / (1e-7 * tau_lam)
            worst = max(worst, d)
            assert d <= 1e-2, (ncv, stretch, q, tau[q], r["tau"])
            checked += 1
    print("eqwm golden: %d rows, worst |dtau| = %.3e tol*tau_lam" % (checked, worst))
    assert checked == 122


def test_eqwm_batch_matches_oracle_iterations():
    """The product-form GPU solve against the oracle's restatement of the
    reference solve (stored arrays): same iteration count, tau to round-off."""
    from channel_flow.solver import eqwm_batch
    from oracle_lib import eqwm as orc_eqwm
    rng = np.random.default_rng(5)
    n = 257
    u = rng.uniform(0.2, 2.0, n)
    nu = 10 ** rng.uniform(-5, -3, n)
    h = rng.uniform(0.05, 0.2, n)
    tin = rng.uniform(1e-4, 1e-2, n)
    for ncv, tol in ((30, 1e-8), (100, 1e-7)):
        tau, it = eqwm_batch(u, nu, h, tin, ncv=ncv, stretch=1.1, tol=tol, max_iter=100)
        for q in range(n):
            t_o, i_o = orc_eqwm(u[q], nu[q], h[q], tin[q], ncv=ncv, stretch=1.1, tol=tol, max_iter=100)
            assert int(it[q]) == i_o, (ncv, q)
            assert abs(tau[q] - t_o) <= 1e-11 * abs(t_o), (ncv, q, tau[q], t_o)


def test_check_divergence_signed_maximum():
    """The reference's monitor value is max(0, max_ijk div) (projection.f90:
    241-248), not max |div|.  A raw random field (divergence of order one,
    both signs) and a field with div = -1 in every cell (V = -y, U = W = 0:
    signed maximum exactly 0, max |div| = 1)."""
    cfg = cases.case_dns64()
    cfg.cells = (32, 24, 16)
    g, o = pair(cfg)
    rng = np.random.default_rng(7)
    U = np.asfortranarray(rng.standard_normal(g.shape_u))
    V = np.asfortranarray(rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    g.set_state(U, V, W)
    o.set_state(U, V, W)
    sg, ag = g.check_divergence()
    so, ao = o.check_divergence()
    assert so > 1.0 and ao >= so
    assert abs(sg - so) <= 1e-13 * so, (sg, so)
    assert abs(ag - ao) <= 1e-13 * ao, (ag, ao)
    Z = np.zeros
    Vn = np.asfortranarray(np.broadcast_to(-g.y[None, :, None], g.shape_v))
    g.set_state(Z(g.shape_u, order="F"), Vn, Z(g.shape_w, order="F"))
    o.set_state(Z(g.shape_u, order="F"), Vn, Z(g.shape_w, order="F"))
    sg, ag = g.check_divergence()
    so, ao = o.check_divergence()
    assert sg == 0.0 and so == 0.0
    assert abs(ag - 1.0) <= 1e-13 and abs(ao - 1.0) <= 1e-13
