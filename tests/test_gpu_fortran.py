"""The Fortran host (modules with the reference's interfaces, driver with
main.f90's loop) drives the HIP library through ISO_C_BINDING; its snapshot
equals the Python host's result for the same run, and a restart from that
snapshot continues it."""
import os
import subprocess

import numpy as np
import pytest

from channel_flow import cases
from channel_flow.snapshot import read_snapshot

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "channel-flow_amd", "lib", "channel_fd")


def input_text(cfg, nsteps, nsave, nstats, nmon, random_init, filein, fileout):
    nx, ny, nz = cfg.faces
    Lx, Ly, Lz = cfg.lengths
    return "\n".join([
        "! domain size nx ny nz", "%d %d %d" % (nx, ny, nz),
        "! CFL", repr(cfg.cfl),
        "! viscosity", "0 %r" % cfg.nu,
        "! Pressure gradient dPdx, dPdz, x_mass_cte, y_mass_cte", "%r %r %d %d" % (cfg.dPdx, cfg.dPdz, cfg.x_mass_cte, cfg.y_mass_cte),
        "! LES model", str(cfg.les_model),
        "! wall model", str(cfg.iwall_model),
        "! slip lengths", "0.1 0.2 0.3 1 1",
        "! nsteps, nsave, nstats, nmonitor", "%d %d %d %d" % (nsteps, nsave, nstats, nmon),
        "! input flow field file", "'%s'" % filein,
        "! output flow field file", "'%s'" % fileout,
        "! first number for output flow field", "0",
        "! random initial condition", str(random_init),
        "! Rossby number", "0",
        "! BC for eddy viscosity", str(cfg.dirichlet_nu_t),
        "! matching y for EQWM, Tij estimation", "0 0",
        "! filter ratio", "1.6",
        "! for generalized wall model", "!   cu c_test diag", "1 1 0", "!   ai coefficients:", "-1 0 0 1 1 -1",
        "! domain Lx Ly Lz, y stretching, seed, perturbation amplitude",
        "%r %r %r %r %d %r" % (Lx, Ly, Lz, cfg.stretch, cfg.seed, cfg.amp),
        "! precision Cs eqwm_ncv eqwm_stretch eqwm_tol eqwm_max_iter use_graph",
        "0 %r %d %r %r %d 1" % (cfg.cs, cfg.eqwm_ncv, cfg.eqwm_stretch, cfg.eqwm_tol, cfg.eqwm_max_iter),
        ""])


def run_driver(tmp_path, text):
    inp = tmp_path / "input"
    inp.write_text(text)
    r = subprocess.run([EXE], stdin=open(inp), cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_fortran_driver_matches_python_host(tmp_path):
    from channel_flow.solver import ChannelFlow
    cfg = cases.case_dns64()
    cfg.cells = (32, 24, 16)
    out = run_driver(tmp_path, input_text(cfg, 6, 6, 3, 3, 1, "dummy", "run"))
    assert "step" in out
    snap = read_snapshot(tmp_path / "run.6")
    g = ChannelFlow(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    g.run(6, graph=True)
    # the Fortran main program builds its own tanh grid (flang's libm vs numpy: 1 ulp)
    np.testing.assert_allclose(snap["y"], g.y, rtol=0, atol=4e-16)
    for name, a in zip("UVW", g.get_state()):
        assert np.max(np.abs(snap[name] - a)) <= 1e-12 * np.max(np.abs(a))
    t, dPdx, q0 = snap["trailer"]
    assert t == pytest.approx(6 * 2e-3, rel=1e-14)
    # restart from the snapshot for 4 more steps == 10 steps straight
    (tmp_path / "r2").mkdir()
    run_driver(tmp_path / "r2", input_text(cfg, 4, 4, 4, 4, 0, str(tmp_path / "run.6"), "cont"))
    snap2 = read_snapshot(tmp_path / "r2" / "cont.4")
    g.run(4, graph=True)
    for name, a in zip("UVW", g.get_state()):
        assert np.max(np.abs(snap2[name] - a)) <= 1e-12 * np.max(np.abs(a))
    _, dmax = g.check_divergence()
    assert dmax < 1e-10


def test_fortran_driver_runs_reference_example(tmp_path):
    """The reference's example input (dynamic Smagorinsky, Neumann nu_t,
    CFL 0.5, constant dPdx) through the Fortran driver: its snapshot equals
    the Python host's run of the same case."""
    from channel_flow.solver import ChannelFlow
    cfg = cases.case_example()
    out = run_driver(tmp_path, input_text(cfg, 5, 5, 5, 5, 1, "dummy", "ex"))
    assert "step" in out
    snap = read_snapshot(tmp_path / "ex.5")
    g = ChannelFlow(cfg)
    g.initialise(*cases.initial_field(cfg))
    g.run(5, graph=True)
    for name, a in zip("UVW", g.get_state()):
        assert np.max(np.abs(snap[name] - a)) <= 1e-12 * np.max(np.abs(a))
