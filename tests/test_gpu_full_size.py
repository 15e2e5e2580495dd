"""The 384x256x384 DNS workload at its full size, checked through properties
that do not depend on an oracle run at this size: the projection's
divergence bound, the mass-flow correction, the closed-form time step,
translation equivariance across the kernels' tile edges, Galilean
invariance of the dynamic Smagorinsky model, and the mixed path's
divergence and window-statistics bounds against the FP64 path.  The same
kernels are compared with the oracle on smaller grids in the parity tests."""
import numpy as np
import pytest

from channel_flow import cases

pytestmark = pytest.mark.gpu

SHIFT = 64  # x columns: one tile of the stencil, filter and model kernels


def solver():
    from channel_flow.solver import ChannelFlow
    return ChannelFlow


def cfg(precision="fp64", **kw):
    c = cases.case_dns384()
    c.precision = precision
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.fixture(scope="module")
def field():
    """The seeded raw initial field of the workload (about 5 s on the host)."""
    return cases.initial_field(cfg())


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def roll_x(a, s):
    """Shift the unique x columns (1-based 2..385 of faces and centres) by s;
    the ghosts are rebuilt by the boundary conditions."""
    b = a.copy()
    b[1:385] = np.roll(a[1:385], s, axis=0)
    return b


def test_fp64_invariants_after_three_steps(field):
    """max |div u| <= 1e-10 after the projection; after the mass-flow
    correction the bulk velocity equals Qflow_x_0 to round-off
    (mass_flow.f90:130); dt is the closed form of compute_dt
    (time_integration.f90:179-227): with the convective limit built on a
    maximum it never binds here, so dt = min(CFL d_min^2 / nu, 5e-3)."""
    s = solver()(cfg())
    try:
        s.initialise(*field)
        s.run(3, graph=True)
        _, dmax = s.check_divergence()
        assert dmax < 1e-10, dmax
        sc = s.scalars
        q = s.compute_mean_mass_flow_U()
        assert abs(q - sc["Qflow_x_0"]) < 1e-13 * abs(sc["Qflow_x_0"]), (q, sc["Qflow_x_0"])
        c = s.cfg
        dmin = [np.diff(cases.centres(a)).min() for a in (s.x, s.y, s.z)]
        dt = min(min(c.cfl * d * d / c.nu for d in dmin), 5e-3)
        assert abs(sc["dt"] - dt) <= 1e-14 * dt, (sc["dt"], dt)
    finally:
        s.close()


def test_fp64_x_shift_equivariance(field):
    """Two RK3 steps of the field and of the field shifted by one 64-column
    tile give shifted results: every stencil, FFT and y-solve tile edge sits
    at a different place in the flow.  Only the FFT rounding differs."""
    CF = solver()
    a, b = CF(cfg()), CF(cfg())
    try:
        a.initialise(*field)
        b.initialise(*(roll_x(f, SHIFT) for f in field))
        a.run(2, graph=True)
        b.run(2, graph=True)
        for fa, fb in zip(a.get_state(), b.get_state()):
            e = relerr(roll_x(fa, SHIFT)[1:385], fb[1:385])
            assert e < 1e-10, e
    finally:
        a.close()
        b.close()


def test_dynamic_smagorinsky_galilean_invariance(field):
    """The dynamic model at full size (six x tiles, 32-plane and 32-row
    marches): a uniform shift of U, ghosts included, leaves Lij and Mij and
    hence nu_t unchanged, since every filter weight set sums to one
    (subgrid.f90:134-192).  The model is active and nu_t is clipped at 0."""
    s = solver()(cfg(les_model=2, dirichlet_nu_t=1))
    try:
        s.initialise(*field)
        U, V, W = s.get_state()
        s.compute_eddy_viscosity()
        n0, c0 = s.nu_t, s.avg_nu_t
        s.set_state(U + 0.5, V, W)
        s.compute_eddy_viscosity()
        n1, c1 = s.nu_t, s.avg_nu_t
        inner = n0[1:-1, 1:-1, 1:-1]
        assert np.all(np.isfinite(n0)) and np.all(inner >= 0.0)
        assert np.count_nonzero(inner) > 0.1 * inner.size
        assert relerr(n1, n0) < 1e-9, relerr(n1, n0)
        assert relerr(c1[1:-1], c0[1:-1]) < 1e-9, relerr(c1[1:-1], c0[1:-1])
    finally:
        s.close()


def _window(g, steps, start, every):
    g.run(start, graph=True)
    g.stats_reset()
    for _ in range((steps - start) // every):
        g.run(every, graph=True)
        g.stats_accumulate()
    return g.stats_window()


def test_mixed_bounds_at_full_size(field):
    """The mixed path on the full workload: max |div u| <= 1e-5, and the mean
    profile and Reynolds stresses of a window of steps 20-60 within 1% of
    the FP64 path (relative to each quantity's maximum over y)."""
    CF = solver()
    g64, g32 = CF(cfg("fp64")), CF(cfg("mixed"))
    try:
        g64.initialise(*field)
        g32.initialise(*field)
        w64 = _window(g64, 60, 20, 5)
        w32 = _window(g32, 60, 20, 5)
        for k in ("Umean", "uu", "vv", "ww", "uv"):
            e = float(np.max(np.abs(w32[k] - w64[k])) / np.max(np.abs(w64[k])))
            assert e < 1e-2, (k, e)
        _, div = g32.check_divergence()
        assert div < 1e-5, div
    finally:
        g64.close()
        g32.close()
