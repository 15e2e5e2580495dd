"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle on
the same seeded inputs.  FP64 path tolerance: max relative field difference
1e-10 (steps) and round-off (single operators)."""
import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _solver():
    from channel_flow.solver import ChannelFlow
    return ChannelFlow


def small_cfg(**kw):
    cfg = cases.case_dns64()
    cfg.cells = kw.pop("cells", (32, 24, 16))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(cfg):
    CF = _solver()
    g = CF(cfg)
    o = Oracle(cfg, jref=g.jref_eqwm)
    return g, o


def interior_u(g, A):
    return A[1:g.nx - 1, 1:g.nyg - 1, 1:g.nzg - 2]


def test_rhs_random_fields():
    cfg = small_cfg()
    g, o = pair(cfg)
    rng = np.random.default_rng(1)
    U = np.asfortranarray(rng.standard_normal(g.shape_u))
    V = np.asfortranarray(rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    ru, rv, rw = g.compute_rhs(U, V, W)
    su, sv, sw = o.compute_rhs(U, V, W)
    for a, b in ((ru, su), (rv, sv), (rw, sw)):
        assert relerr(a, b) < 1e-12


def test_rhs_with_eddy_viscosity_and_rotation():
    cfg = small_cfg(omega_z=0.05)
    g, o = pair(cfg)
    rng = np.random.default_rng(2)
    U = np.asfortranarray(rng.standard_normal(g.shape_u))
    V = np.asfortranarray(rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    nt = np.asfortranarray(np.abs(rng.standard_normal(g.shape_c)) * 1e-3)
    g.set_nu_t(nt)
    o.set_nu_t(nt)
    ru, rv, rw = g.compute_rhs(U, V, W)
    su, sv, sw = o.compute_rhs(U, V, W)
    for a, b in ((ru, su), (rv, sv), (rw, sw)):
        assert relerr(a, b) < 1e-12


def test_boundary_conditions_bitexact():
    cfg = small_cfg()
    g, o = pair(cfg)
    rng = np.random.default_rng(3)
    U = np.asfortranarray(rng.standard_normal(g.shape_u))
    V = np.asfortranarray(rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    g.set_state(U, V, W)
    o.set_state(U, V, W)
    g.apply_boundary_conditions()
    o.apply_boundary_conditions()
    for a, b in zip(g.get_state(), o.get_state()):
        np.testing.assert_array_equal(a, b)


def test_projection_divergence_free():
    cfg = small_cfg()
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    _, gmax = g.check_divergence()
    assert gmax < 1e-12
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-12
    assert abs(g.scalars["Qflow_x_0"] - o.scalars["Qflow_x_0"]) < 1e-14


@pytest.mark.parametrize("cells", [(32, 24, 16), (48, 32, 24)])
def test_ten_rk3_steps(cells):
    cfg = small_cfg(cells=cells)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=False)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10
    _, dmax = g.check_divergence()
    assert dmax < 1e-10
    # dPdx = correction/dt (time_integration.f90:156): compare the correction,
    # a difference of two O(1) bulk flows summed in different orders
    dt = o.scalars["dt"]
    assert abs(g.scalars["dPdx"] - o.scalars["dPdx"]) * dt < 1e-13


@pytest.mark.parametrize("cells", [(128, 24, 96), (64, 20, 384), (384, 12, 64)])
def test_ten_rk3_steps_large_grid_kernels(cells, monkeypatch):
    """The kernels the 384x256x384 workload runs, on grids the oracle steps
    in seconds: x and z transform lengths compiled with fixed plans (half
    length 64, 32 or 192; z length 96, 384 or 64) and the lane form of the
    y solve, forced here (it is chosen by default from 65,536 modes up), with
    the z transforms and the y solve as separate launches (the fused form is
    the default from 128 x wavenumbers up)."""
    monkeypatch.setenv("CF_TH_LANE", "1")
    monkeypatch.setenv("CF_ZSOLVE", "0")
    cfg = small_cfg(cells=cells)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    _, dmax = g.check_divergence()
    assert dmax < 1e-10


def test_graph_replay_matches_eager():
    cfg = small_cfg()
    CF = _solver()
    a, b = CF(cfg), CF(cfg)
    U, V, W = cases.initial_field(cfg)
    a.initialise(U, V, W)
    b.initialise(U, V, W)
    for _ in range(3):  # graph replays with eager work and syncs in between
        a.run(2, graph=True)
        b.run(2, graph=False)
        a.check_divergence()
        for x, y in zip(a.get_state(), b.get_state()):
            np.testing.assert_array_equal(x, y)


def test_cfl_dt_matches():
    cfg = small_cfg(cfl=0.5)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    assert g.compute_dt() == o.compute_dt()


def test_statistics_match():
    cfg = small_cfg()
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(2, graph=True)
    o.run(2)
    gs, os_ = g.compute_statistics(), o.compute_statistics()
    assert set(gs) == set(os_)
    for k in os_:
        np.testing.assert_allclose(gs[k], os_[k], rtol=1e-12, atol=1e-14, err_msg=k)
    assert os_["utau"] > 0 and os_["Retau"] > 0


@pytest.mark.parametrize("les", [0, 1])
def test_statistics_of_random_fields_match(les):
    """Plane statistics, u_tau, Re_tau and bulk flows of a field that varies
    in x and z (the reference's sum ranges with their duplicated periodic
    images matter) with an eddy viscosity at the walls."""
    cfg = small_cfg(les_model=les, cs=0.1)
    g, o = pair(cfg)
    rng = np.random.default_rng(11)
    U = np.asfortranarray(1.0 + 0.3 * rng.standard_normal(g.shape_u))
    V = np.asfortranarray(0.2 * rng.standard_normal(g.shape_v))
    W = np.asfortranarray(0.2 * rng.standard_normal(g.shape_w))
    # a wall-normal gradient that makes utau^2 positive
    U[:, 0, :] -= 3.0
    U[:, -1, :] -= 3.0
    g.set_state(U, V, W)
    o.set_state(U, V, W)
    if les:
        nt = np.asfortranarray(np.abs(rng.standard_normal(g.shape_c)) * 1e-3)
        g.set_nu_t(nt)
        o.set_nu_t(nt)
    gs, os_ = g.compute_statistics(), o.compute_statistics()
    for k in os_:
        np.testing.assert_allclose(gs[k], os_[k], rtol=1e-12, atol=1e-14, err_msg=k)
    assert np.isfinite(os_["utau"]) and os_["utau"] > 0


def test_statistics_nan_is_an_error():
    """statistics.f90:165-167: a NaN anywhere in U, V or W (ghosts included)
    makes compute_statistics fail with 'Error: NaNs!'; the state itself is
    only read."""
    cfg = small_cfg()
    CF = _solver()
    g = CF(cfg)
    U, V, W = cases.initial_field(cfg)
    for comp, idx in ((0, (3, 4, 5)), (1, (0, 0, 0)), (2, (-1, -1, -1))):
        F = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F")]
        F[comp][idx] = np.nan
        g.set_state(*F)
        with pytest.raises(RuntimeError, match="NaNs"):
            g.compute_statistics()
    g.set_state(U, V, W)
    g.compute_statistics()  # a finite state passes


def test_smagorinsky_matches():
    cfg = small_cfg(les_model=1, cs=0.1, stretch=0.0)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.compute_eddy_viscosity()
    o.compute_eddy_viscosity()
    assert relerr(g.nu_t, o.nu_t) < 1e-12


@pytest.mark.parametrize("cells,dirichlet", [((32, 24, 16), 1), ((36, 22, 18), 0), ((144, 20, 72), 1),
                                             ((64, 96, 384), 0)])
def test_dynamic_smagorinsky_matches(cells, dirichlet):
    """sgs_dynamic_Smagorinsky (subgrid.f90:376-538) on a stretched grid.
    (36, 22, 18): x and z tiles and the y row chunks of the kernels are
    ragged.  (144, 20, 72): three 64-column x tiles, the last 16 columns
    wide, so halos and ghost copies cross tile edges.  (64, 96, 384): long
    marches, 22 chunks of 18 z planes in the strain kernel and 6 chunks of
    18 rows in the LM kernel, the last of each ragged.  The plane sums are
    tree-ordered on the GPU: tolerance 1e-11."""
    cfg = small_cfg(les_model=2, cells=cells, dirichlet_nu_t=dirichlet)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.compute_eddy_viscosity()
    o.compute_eddy_viscosity()
    ng, no = g.nu_t, o.nu_t
    assert np.count_nonzero(no[1:-1, 1:-1, 1:-1]) > 0.1 * no[1:-1, 1:-1, 1:-1].size  # model active
    assert relerr(ng, no) < 1e-11, relerr(ng, no)


@pytest.mark.parametrize("les,dirichlet", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_avg_nu_t_matches_oracle(les, dirichlet):
    """avg_nu_t after an eddy-viscosity evaluation (cf_get_avg_nu_t): the
    clipped plane mean (Smagorinsky) or the dynamic coefficient, wall rows
    by the nu_t rule."""
    cfg = small_cfg(les_model=les, cs=0.1, dirichlet_nu_t=dirichlet)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.compute_eddy_viscosity()
    o.compute_eddy_viscosity()
    ag, ao = g.avg_nu_t, o.avg_nu_t
    assert np.any(ao != 0)
    assert relerr(ag, ao) < 1e-11, relerr(ag, ao)


@pytest.mark.parametrize("wall", [0, 3])
def test_dynamic_smagorinsky_steps_match(wall):
    """Ten RK3 steps with the dynamic model (graph-replayed on the GPU),
    without and with the equilibrium wall model, at the FP64 bar."""
    kw = dict(iwall_model=3, stretch=0.0, eqwm_ncv=30, eqwm_tol=1e-8, nu=1e-4, cfl=-1e-3,
              dirichlet_nu_t=0) if wall else dict(dPdx=2e-3)
    cfg = small_cfg(les_model=2, **kw)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    assert relerr(g.nu_t, o.nu_t) < 1e-9


def test_reference_example_ten_steps():
    """The reference's own example input (dynamic Smagorinsky, CFL 0.5,
    constant dPdx): ten RK3 steps at the FP64 bar, divergence-free."""
    cfg = cases.case_example()
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    assert relerr(g.nu_t, o.nu_t) < 1e-8
    _, dmax = g.check_divergence()
    assert dmax < 1e-10
    assert abs(g.scalars["t"] - o.scalars["t"]) < 1e-12 * o.scalars["t"]


def test_wall_model_matches():
    cfg = small_cfg(les_model=1, iwall_model=3, stretch=0.0, eqwm_ncv=30, eqwm_tol=1e-8, nu=1e-4)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.compute_eddy_viscosity()
    o.compute_eddy_viscosity()
    g.compute_wall_model()
    o.compute_wall_model()
    for a, b in zip(g.wall_state(), o.wall_state()):
        assert relerr(a, b) < 1e-12, relerr(a, b)


def test_wmles_steps_match():
    """Ten WMLES steps (Smagorinsky + EQWM, eager launches): the Picard
    iteration takes the same number of steps at every wall point on both
    sides, so the FP64 bar holds for the fields and the wall state."""
    cfg = small_cfg(les_model=1, iwall_model=3, stretch=0.0, eqwm_ncv=30, eqwm_tol=1e-8,
                    nu=1e-4, cfl=-1e-3)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=False)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    for a, b in zip(g.wall_state(), o.wall_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)


def test_saved_pressure_matches_oracle():
    """P saved at the first substep of the last step (projection.f90:129-137):
    phi/(dt a11) with neighbour-copy ghosts, after three steps."""
    cfg = small_cfg(dPdx=2e-3)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(3, graph=True)
    o.run(3)
    Pg, Po = g.P, o.P
    assert np.all(np.isfinite(Pg))
    assert relerr(Pg, Po) < 1e-10, relerr(Pg, Po)


@pytest.mark.parametrize("les", [0, 1, 2])
def test_actual_pressure_matches_oracle(les):
    """compute_pressure (pressure.f90:21-341) on the same state: the wall
    conditions from V and the plane-averaged eddy viscosity, the divergence
    of the momentum right-hand sides, and the Poisson solve with wall
    sources.  The GPU adds the wall sources before the transforms (the
    reference after them), which changes rounding only: FP64 tolerance 1e-10."""
    cfg = small_cfg(les_model=les, cs=0.1, dPdx=2e-3)
    g, o = pair(cfg)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(2, graph=False)
    o.run(2)
    if les:
        g.compute_eddy_viscosity()
        o.compute_eddy_viscosity()
    Pg = g.compute_pressure()
    Po = o.compute_pressure()
    assert np.all(np.isfinite(Pg))
    assert relerr(Pg, Po) < 1e-10, relerr(Pg, Po)
