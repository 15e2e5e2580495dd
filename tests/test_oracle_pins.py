"""Pins for the CPU oracle: each check is chosen so that a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails it.

The reference ships no tests or golden vectors for this path; its only
arithmetic dependencies (FFTW 3.3.4, LAPACK 3.2.1 zgtsv) are pinned by the
mathematics: numpy's FFT and dense solves, closed forms, symmetries and
invariants."""
import math

import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle, lib, P, eqwm


def cfg_small(**kw):
    c = cases.case_dns64()
    c.cells = kw.pop("cells", (16, 12, 8))
    c.x_mass_cte = kw.pop("x_mass_cte", 0)
    c.dPdx = kw.pop("dPdx", 0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# ---------------------------------------------------------------- FFT / zgtsv
@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 96, 384])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(n, sign):
    rng = np.random.default_rng(n)
    re, im = rng.standard_normal(n), rng.standard_normal(n)
    x = re + 1j * im
    lib().orc_fft(n, P(re), P(im), sign)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    np.testing.assert_allclose(re + 1j * im, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_fft2d_matches_numpy():
    rng = np.random.default_rng(0)
    n1, n2 = 12, 8
    a = rng.standard_normal((n2, n1)) + 0j
    re = np.ascontiguousarray(a.real.ravel())
    im = np.zeros_like(re)
    lib().orc_fft2d(n1, n2, P(re), P(im), -1)
    np.testing.assert_allclose((re + 1j * im).reshape(n2, n1), np.fft.fft2(a), atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_zgtsv_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    n = 9
    d = rng.standard_normal(n)
    dl = rng.standard_normal(n - 1) * (3.0 if seed % 2 else 0.3)  # forces row swaps when large
    du = rng.standard_normal(n - 1)
    A = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    br, bi = np.ascontiguousarray(b.real), np.ascontiguousarray(b.imag)
    info = lib().orc_zgtsv(n, P(dl.copy()), P(d.copy()), P(du.copy()), P(br), P(bi))
    assert info == 0
    np.testing.assert_allclose(br + 1j * bi, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_zgtsv_singular_reports_info():
    d = np.zeros(3)
    dl, du = np.zeros(2), np.ones(2)
    br, bi = np.ones(3), np.zeros(3)
    assert lib().orc_zgtsv(3, P(dl), P(d), P(du), P(br), P(bi)) == 1


# ------------------------------------------------------------ grid operators
def test_modified_wavenumbers_are_dft_eigenvalues():
    o = Oracle(cfg_small())
    g = o.grid()
    n = o.nx - 2
    dx = g["dmin"][0]
    # eigenvalues of the periodic 3-point Laplacian = DFT of its stencil
    stencil = np.zeros(n)
    stencil[0], stencil[1], stencil[-1] = -2 / dx ** 2, 1 / dx ** 2, 1 / dx ** 2
    np.testing.assert_allclose(g["kxx"], np.fft.fft(stencil).real, atol=1e-9 / dx ** 2)


def test_y_weights_interpolate_linear_exactly():
    o = Oracle(cfg_small())
    g = o.grid()
    yg = g["yg"]
    yf = o.y
    # w0*f(yg_j) + w1*f(yg_{j+1}) == f(y_j) for linear f
    f = lambda t: 3.0 * t - 0.7
    interp = g["w0"] * f(yg[:-1]) + g["w1"] * f(yg[1:])
    np.testing.assert_allclose(interp, f(yf), atol=1e-13)


def test_dyy_is_neumann_laplacian():
    o = Oracle(cfg_small())
    d, l, u = o.dyy()
    A = np.diag(d) + np.diag(l, -1) + np.diag(u, 1)
    # Neumann closure: constants are in the null space, quadratics give 2
    np.testing.assert_allclose(A @ np.ones(len(d)), 0.0, atol=1e-9)
    yg = o.grid()["yg"][1:-1]
    q = A @ (yg ** 2)
    # interior rows of a face-centred 3-point operator on the cell centres
    yf = o.y
    expect = (yg[2:] - yg[:-2]) / (yf[2:-1] - yf[1:-2])
    np.testing.assert_allclose(q[1:-1], expect, rtol=1e-10)


# ------------------------------------------------------------------- RHS
def fields(o, fu=None, fv=None, fw=None):
    U, V, W = (np.zeros(s, order="F") for s in (o.shape_u, o.shape_v, o.shape_w))
    g = o.grid()
    X, XG, Y, YG, Z, ZG = o.x, g["xg"], o.y, g["yg"], o.z, g["zg"]
    if fu:
        U[:] = fu(X[:, None, None], YG[None, :, None], ZG[None, None, :])
    if fv:
        V[:] = fv(XG[:, None, None], Y[None, :, None], ZG[None, None, :])
    if fw:
        W[:] = fw(XG[:, None, None], YG[None, :, None], Z[None, None, :])
    return U, V, W


def test_rhs_uniform_flow_is_forcing_only():
    c = cfg_small(dPdx=0.3, dPdz=-0.2, stretch=0.0)
    o = Oracle(c)
    U, V, W = fields(o, lambda x, y, z: 1.7 + 0 * x, None, lambda x, y, z: -0.4 + 0 * x)
    ru, rv, rw = o.compute_rhs(U, V, W)
    np.testing.assert_allclose(ru, 0.3, atol=1e-12)
    np.testing.assert_allclose(rv, 0.0, atol=1e-12)
    np.testing.assert_allclose(rw, -0.2, atol=1e-12)


def test_rhs_x_mode_convection_and_diffusion_closed_form():
    c = cfg_small(nu=0.01, dPdx=0.0)
    o = Oracle(c)
    g = o.grid()
    dx = o.x[2] - o.x[1]
    k = 2 * math.pi / (o.x[-2] - o.x[0]) * 2  # second harmonic of the period
    U, V, W = fields(o, lambda x, y, z: np.sin(k * x) + 0 * y * z)
    ru, _, _ = o.compute_rhs(U, V, W)
    xi = o.x[1:-1]
    lam = -4.0 / dx ** 2 * math.sin(k * dx / 2) ** 2  # modified wavenumber
    conv = -math.cos(k * dx / 2) ** 2 * math.sin(k * dx) / dx * np.sin(2 * k * xi)
    visc = 2.0 * c.nu * lam * np.sin(k * xi)  # d(2 nu du/dx)/dx
    expect = (conv + visc)[:, None, None]
    np.testing.assert_allclose(ru, np.broadcast_to(expect, ru.shape), atol=1e-10)


def test_rhs_x_z_exchange_symmetry():
    """u and w equations are mirror images under x <-> z on a cube-like grid."""
    c = cfg_small(cells=(12, 10, 12), lengths=(2.0, 2.0, 2.0), nu=0.02)
    o = Oracle(c)
    rng = np.random.default_rng(7)
    A = rng.standard_normal(o.shape_u)      # U(i_face, j, k_ctr)
    B = rng.standard_normal(o.shape_v)      # V(i_ctr, j, k_ctr)
    # W(i_ctr, j, k_face) := U^T, U := W^T, V := V^T with x <-> z
    U = np.asfortranarray(A)
    W = np.asfortranarray(np.transpose(rng.standard_normal(o.shape_u), (2, 1, 0)))
    V = np.asfortranarray(B)
    ru, rv, rw = o.compute_rhs(U, V, W)
    U2 = np.asfortranarray(np.transpose(W, (2, 1, 0)))
    W2 = np.asfortranarray(np.transpose(U, (2, 1, 0)))
    V2 = np.asfortranarray(np.transpose(V, (2, 1, 0)))
    ru2, rv2, rw2 = o.compute_rhs(U2, V2, W2)
    # rhs_w(i,j,k) on 2..nxg-1 x 2..nz-1 vs rhs_u of the swapped field on 2..nx-1 x 2..nzg-1
    np.testing.assert_allclose(np.transpose(rw, (2, 1, 0)), ru2, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.transpose(rv, (2, 1, 0)), rv2, rtol=1e-12, atol=1e-12)


def test_rhs_linear_shear_in_y_has_no_viscous_term():
    c = cfg_small(nu=0.05, dPdx=0.0)
    o = Oracle(c)
    U, V, W = fields(o, lambda x, y, z: 2.0 * y + 0 * x * z)
    ru, rv, rw = o.compute_rhs(U, V, W)
    np.testing.assert_allclose(ru, 0.0, atol=1e-11)
    np.testing.assert_allclose(rw, 0.0, atol=1e-11)


# --------------------------------------------------------- RK3 amplification
def test_rk3_amplification_factor():
    """W = sin(k x): divergence free, convection free; one RK3 step multiplies
    it mid-channel by R(z) = 1 + z + z^2/2 + z^3/6, z = nu*lambda*dt."""
    c = cfg_small(cells=(16, 24, 8), nu=0.05, cfl=-0.01, dPdx=0.0, stretch=0.0)
    o = Oracle(c)
    k = 2 * math.pi / (o.x[-2] - o.x[0])
    U, V, W = fields(o, None, None, lambda x, y, z: np.sin(k * x) + 0 * y * z)
    o.set_state(U, V, W)
    o.apply_boundary_conditions()
    o.set_scalars(dt=0.01)
    o.compute_time_step_RK3()
    _, _, W1 = o.get_state()
    g = o.grid()
    dx = o.x[2] - o.x[1]
    lam = -4.0 / dx ** 2 * math.sin(k * dx / 2) ** 2
    z = c.nu * lam * 0.01
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6
    jm = o.nyg // 2
    xg = g["xg"]
    np.testing.assert_allclose(W1[1:-2, jm, 3], R * np.sin(k * xg[1:-2]), atol=1e-12)


# ----------------------------------------------------------------- projection
def test_projection_makes_divergence_zero_and_preserves_solenoidal_part():
    c = cfg_small(cells=(16, 16, 12))
    o = Oracle(c)
    U, V, W = cases.initial_field(c)
    o.initialise(U, V, W)
    _, dmax = o.check_divergence()
    assert dmax < 1e-12
    a = o.get_state()
    o.apply_boundary_conditions()
    o.compute_projection_step(0)
    o.apply_boundary_conditions()
    for x, y in zip(a, o.get_state()):  # projecting twice changes nothing
        np.testing.assert_allclose(x, y, atol=1e-12)


def test_mass_flow_correction_restores_bulk_flow():
    c = cfg_small(cells=(16, 16, 12), x_mass_cte=1)
    o = Oracle(c)
    U, V, W = cases.initial_field(c)
    o.initialise(U, V, W)
    q0 = o.scalars["Qflow_x_0"]
    o.set_scalars(dt=2e-3)
    o.compute_time_step_RK3()
    assert abs(o.compute_mean_mass_flow_U() - q0) < 1e-13


def test_ghost_identities():
    c = cfg_small()
    o = Oracle(c)
    rng = np.random.default_rng(3)
    o.set_state(rng.standard_normal(o.shape_u), rng.standard_normal(o.shape_v), rng.standard_normal(o.shape_w))
    o.apply_boundary_conditions()
    U, V, W = o.get_state()
    nx, nxg, ny, nyg, nz, nzg = o.nx, o.nxg, o.ny, o.nyg, o.nz, o.nzg
    np.testing.assert_array_equal(U[0, 1:-1, 1:-2], U[nx - 2, 1:-1, 1:-2])   # U(1)=U(nx-1)
    np.testing.assert_array_equal(U[:, 0, :], -U[:, 1, :])                   # Dirichlet
    np.testing.assert_array_equal(V[:, 0, :], 0.0)
    np.testing.assert_array_equal(V[:, ny - 1, :], 0.0)
    np.testing.assert_array_equal(W[:, :, 0], W[:, :, nz - 2])               # W(1)=W(nz-1)
    np.testing.assert_array_equal(V[:, :, nzg - 2], V[:, :, 1])              # V(nzg-1)=V(2)


# --------------------------------------------------------------------- dt
def test_dt_closed_forms():
    c = cfg_small(cfl=-3e-3)
    o = Oracle(c)
    assert o.compute_dt() == 3e-3
    c = cfg_small(cfl=0.5, nu=0.01)
    o = Oracle(c)
    U, V, W = cases.initial_field(c)
    o.initialise(U, V, W)
    g = o.grid()
    dm = g["dmin"]
    # the convective bound uses Max (never binds) -> viscous limit or cap
    expect = min(min(0.5 * dm[0] ** 2 / 0.01, 0.5 * dm[1] ** 2 / 0.01, 0.5 * dm[2] ** 2 / 0.01), 5e-3)
    assert o.compute_dt() == pytest.approx(expect, rel=1e-15)


# -------------------------------------------------------------- Smagorinsky
def test_smagorinsky_simple_shear():
    c = cfg_small(les_model=1, cs=0.1, stretch=0.0, cells=(12, 12, 8))
    o = Oracle(c)
    gam = 0.8
    U, V, W = fields(o, lambda x, y, z: gam * y + 0 * x * z)
    o.set_state(U, V, W)
    o.apply_boundary_conditions()
    o.compute_eddy_viscosity()
    nt = o.nu_t
    g = o.grid()
    dx, dz = g["dmin"][0], g["dmin"][2]
    dy = o.y[1] - o.y[0]
    expect = (0.1 * (dx * dz * dy) ** (1 / 3)) ** 2 * gam
    # interior away from the walls (the wall ghost breaks the linear profile)
    np.testing.assert_allclose(nt[2:-2, 3:-3, 2:-2], expect, rtol=1e-12)


@pytest.mark.parametrize("dirichlet", [0, 1])
def test_smagorinsky_ghost_sequence(dirichlet):
    """The completed Smagorinsky (subgrid.f90:564-619) with the dynamic
    model's ghost sequence (subgrid.f90:521-536): periodic copies in x and
    z, clipping, then odd (Dirichlet) or even (Neumann) wall rows; the
    plane average is clipped and gets the same wall rule."""
    c = cfg_small(les_model=1, cs=0.1, cells=(12, 10, 8), dirichlet_nu_t=dirichlet)
    o = Oracle(c)
    rng = np.random.default_rng(11)
    o.set_state(*(rng.standard_normal(s) for s in (o.shape_u, o.shape_v, o.shape_w)))
    o.apply_boundary_conditions()
    o.compute_eddy_viscosity()
    nt, av = o.nu_t, o.avg_nu_t
    # centres: 1 <- nxg-2, nxg-1 <- 2, nxg <- 3 (0-based -3, 1, 2), the same in z
    for ax in (0, 2):
        a = np.moveaxis(nt, ax, 0)
        np.testing.assert_array_equal(a[0], a[-3])
        np.testing.assert_array_equal(a[-2], a[1])
        np.testing.assert_array_equal(a[-1], a[2])
    assert np.all(nt[:, 1:-1, :] >= 0)
    sgn = -1.0 if dirichlet else 1.0
    np.testing.assert_array_equal(nt[:, 0, :], sgn * nt[:, 1, :])
    np.testing.assert_array_equal(nt[:, -1, :], sgn * nt[:, -2, :])
    # plane mean over the reference's range 2..nxg-1 x 2..nzg-1, clipped
    np.testing.assert_allclose(av[1:-1], np.maximum(nt[1:-1, 1:-1, 1:-1].mean(axis=(0, 2)), 0.0), rtol=1e-14)
    np.testing.assert_array_equal(av[[0, -1]], sgn * av[[1, -2]])


def test_pressure_output_scaling():
    """P saved at the first substep (projection.f90:129-137) is phi/(dt a11):
    the velocity correction of the same projection equals dt a11 grad P, so a
    wrong scale or a shifted P fails.  Ghosts are copies of the neighbours."""
    c = cfg_small(cells=(12, 10, 8))
    o = Oracle(c)
    rng = np.random.default_rng(12)
    o.set_state(*(rng.standard_normal(s) for s in (o.shape_u, o.shape_v, o.shape_w)))
    o.apply_boundary_conditions()
    dt = 3.7e-3
    o.set_scalars(dt=dt)
    U0, V0, _ = o.get_state()
    o.compute_projection_step(rk_step=1)
    U1, V1, _ = o.get_state()
    P = o.P
    g = o.grid()
    a11 = 8.0 / 15.0
    xg, yg = g["xg"], g["yg"]
    # U faces i = 2..nx-1, V faces j = 2..ny-1 (1-based); interior j, k
    du = (U0 - U1)[1:-1, 1:-1, 1:-1] * (xg[2:-1] - xg[1:-2])[:, None, None]
    np.testing.assert_allclose(du, dt * a11 * (P[2:-1, 1:-1, 1:-1] - P[1:-2, 1:-1, 1:-1]), rtol=1e-9,
                               atol=1e-12 * np.abs(du).max())
    dv = (V0 - V1)[1:-1, 1:-1, 1:-1] * (yg[2:-1] - yg[1:-2])[None, :, None]
    np.testing.assert_allclose(dv, dt * a11 * (P[1:-1, 2:-1, 1:-1] - P[1:-1, 1:-2, 1:-1]), rtol=1e-9,
                               atol=1e-12 * np.abs(dv).max())
    for ax in (0, 1, 2):
        a = np.moveaxis(P, ax, 0)
        np.testing.assert_array_equal(a[0], a[1])
        np.testing.assert_array_equal(a[-1], a[-2])


# ------------------------------------------------------ dynamic Smagorinsky
def _hat_z(a):
    """Simpson filter along one periodic direction: the x-z weights
    (1/36 1/9 1/36; 1/9 4/9 1/9) on a field constant in x, and the y weights
    on a field constant in y, both reduce to (1/6 2/3 1/6)."""
    return np.roll(a, 1) / 6 + 2 * a / 3 + np.roll(a, -1) / 6


def _dsm_z_only(uc, wf, dz):
    """Lilly's dynamic coefficient for U = U(z) at the centres and W = W(z) on
    the z faces (centre m between faces m-1 and m), V = 0, derived by hand:
    the only nonzero strains are S33 (centres) and S13 (edge values
    1/2 dU/dz, averaged to the centres); L and M reduce to their 13 and 33
    components.  The plane sums run over centres 2..nzg-1 of the reference,
    which counts the first centre twice (its periodic image at nzg-1)."""
    wc = 0.5 * (np.roll(wf, 1) + wf)

    def strain(u, w):
        s33 = (w - np.roll(w, 1)) / dz
        e = 0.5 * (u - np.roll(u, 1)) / dz
        s13 = 0.5 * (e + np.roll(e, -1))
        return s33, s13, np.sqrt(2 * s33 ** 2 + 4 * s13 ** 2)

    L33 = _hat_z(wc * wc) - _hat_z(wc) ** 2
    L13 = _hat_z(uc * wc) - _hat_z(uc) * _hat_z(wc)
    s33, s13, S = strain(uc, wf)
    f33, f13, Sf = strain(_hat_z(uc), _hat_z(wf))
    M33 = 4 * Sf * f33 - _hat_z(S * s33)   # filter width ratio 2, squared
    M13 = 4 * Sf * f13 - _hat_z(S * s13)
    lm, mm = L33 * M33 + 2 * L13 * M13, M33 ** 2 + 2 * M13 ** 2
    avg = (lm.sum() + lm[0]) / (mm.sum() + mm[0])
    return avg, np.maximum(-0.5 * avg * S, 0.0)


@pytest.mark.parametrize("dirichlet", [0, 1])
def test_dynamic_smagorinsky_z_only_field(dirichlet):
    """sgs_dynamic_Smagorinsky (subgrid.f90:376-538) on U(z), W(z): the 3-D
    oracle agrees with the one-dimensional reduction above, row by row, and
    the wall ghosts follow the Dirichlet (odd) or Neumann (even) rule."""
    c = cfg_small(les_model=2, cells=(8, 10, 16), dirichlet_nu_t=dirichlet)
    o = Oracle(c)
    g = o.grid()
    Lz = c.lengths[2]
    U, V, W = fields(o, lambda x, y, z: 0.9 * np.sin(2 * np.pi * z / Lz + 0.3) + 0 * x * y, None,
                     lambda x, y, z: 0.5 * np.cos(4 * np.pi * z / Lz + 1.1) + 0 * x * y)
    o.set_state(U, V, W)
    o.compute_eddy_viscosity()
    nt, av = o.nu_t, o.avg_nu_t
    N = o.nz - 2
    avg, n1 = _dsm_z_only(U[0, 0, 1:1 + N], W[0, 0, 1:1 + N], o.z[1] - o.z[0])
    assert avg < 0 and n1.max() > 0          # the model is active
    np.testing.assert_allclose(av[1:-1], avg, rtol=1e-12)
    np.testing.assert_allclose(nt[1:-1, 1:-1, 1:1 + N], np.broadcast_to(n1, nt[1:-1, 1:-1, 1:1 + N].shape),
                               rtol=1e-11, atol=1e-18)
    sgn = -1.0 if dirichlet else 1.0
    np.testing.assert_array_equal(av[[0, -1]], sgn * av[[1, -2]])
    np.testing.assert_array_equal(nt[:, 0, :], sgn * nt[:, 1, :])
    np.testing.assert_array_equal(nt[:, -1, :], sgn * nt[:, -2, :])


def _random_bc_state(o, seed):
    rng = np.random.default_rng(seed)
    U, V, W = (rng.standard_normal(s) for s in (o.shape_u, o.shape_v, o.shape_w))
    o.set_state(U, V, W)
    o.apply_boundary_conditions()
    return o.get_state()


def _dsm(o, U, V, W):
    o.set_state(U, V, W)
    o.compute_eddy_viscosity()
    return o.nu_t, o.avg_nu_t


def test_dynamic_smagorinsky_x_z_exchange():
    """Exchanging the roles of x and z (U <-> W transposed, equal periodic
    grids) transposes nu_t and keeps the plane averages: pins the pairing of
    the 12/13/23 components, the x and z face offsets and the filter axes."""
    c = cfg_small(les_model=2, cells=(12, 10, 12), lengths=(math.pi, 2.0, math.pi))
    o = Oracle(c)
    U, V, W = _random_bc_state(o, 3)
    nt, av = _dsm(o, U, V, W)
    t = (2, 1, 0)
    nt2, av2 = _dsm(o, W.transpose(t), V.transpose(t), U.transpose(t))
    assert (nt[1:-2, 1:-1, 1:-2] > 0).mean() > 0.2
    np.testing.assert_allclose(av2, av, rtol=1e-12)
    np.testing.assert_allclose(nt2, nt.transpose(t), rtol=1e-11, atol=1e-14 * np.abs(nt).max())


def test_dynamic_smagorinsky_wall_mirror():
    """Mirroring y about the channel centre (U, W reversed, V reversed and
    negated, on a stretched grid symmetric by construction) reverses nu_t and
    the averages: pins the one-sided filter rows, the V-to-centre
    interpolation and the y offsets of the edge strains."""
    c = cfg_small(les_model=2, cells=(12, 14, 8))
    x, y, z = c.grid()
    half = y[: (len(y) + 1) // 2]
    y = np.concatenate([half, 2.0 - half[::-1][1:]]) if len(y) % 2 else np.concatenate([half, 2.0 - half[::-1]])
    o = Oracle(c, x, y, z)
    U, V, W = _random_bc_state(o, 5)
    nt, av = _dsm(o, U, V, W)
    nt2, av2 = _dsm(o, U[:, ::-1, :], -V[:, ::-1, :], W[:, ::-1, :])
    np.testing.assert_allclose(av2, av[::-1], rtol=1e-9)
    np.testing.assert_allclose(nt2, nt[:, ::-1, :], rtol=1e-9, atol=1e-12 * np.abs(nt).max())


def test_dynamic_smagorinsky_galilean_invariance():
    """Lij = hat(ui uj) - hat(ui) hat(uj) and Mij are unchanged by a uniform
    velocity added to U, V and W: the filters preserve constants (one-sided
    wall rows included) and the products pair the same components."""
    c = cfg_small(les_model=2, cells=(12, 10, 8))
    o = Oracle(c)
    U, V, W = _random_bc_state(o, 7)
    nt, av = _dsm(o, U, V, W)
    nt2, av2 = _dsm(o, U + 0.7, V - 0.3, W + 0.45)
    np.testing.assert_allclose(av2, av, rtol=1e-10)
    np.testing.assert_allclose(nt2, nt, rtol=1e-10, atol=1e-12 * np.abs(nt).max())


# -------------------------------------------------------------------- EQWM
def test_eqwm_laminar_limit():
    """Tiny u: mixing length negligible, linear profile, tau = nu u / h."""
    tau, it = eqwm(1e-6, 1e-3, 0.1, 1e-8, ncv=30, tol=1e-12, max_iter=200)
    assert tau == pytest.approx(1e-3 * 1e-6 / 0.1, rel=1e-6)


def test_eqwm_similarity():
    """The discrete system is scale free: tau h^2/nu^2 depends on u h/nu only."""
    t1, _ = eqwm(1.0, 1e-4, 0.1, 1e-3, ncv=30, tol=1e-12, max_iter=300)
    t2, _ = eqwm(2.0, 2e-4, 0.1, 4e-3, ncv=30, tol=1e-12, max_iter=300)   # nu x2, u x2
    t3, _ = eqwm(0.5, 1e-4, 0.2, 2.5e-4, ncv=30, tol=1e-12, max_iter=300)  # h x2, u /2
    assert t2 * 0.1 ** 2 / 2e-4 ** 2 == pytest.approx(t1 * 0.1 ** 2 / 1e-4 ** 2, rel=1e-9)
    assert t3 * 0.2 ** 2 / 1e-4 ** 2 == pytest.approx(t1 * 0.1 ** 2 / 1e-4 ** 2, rel=1e-9)


def test_eqwm_log_law_range():
    """At high Re the converged stress follows the log law within a few %."""
    u, nu, h = 20.0, 1e-5, 0.1
    tau, it = eqwm(u, nu, h, 1.0, ncv=60, tol=1e-10, max_iter=500)
    ut = math.sqrt(tau)
    uplus = u / ut
    yplus = h * ut / nu
    assert uplus == pytest.approx(math.log(yplus) / 0.41 + 5.2, rel=0.06)


# -------------------------------------------------------------- statistics
def test_statistics_of_analytic_field():
    c = cfg_small(cells=(12, 16, 8))
    o = Oracle(c)
    U, V, W = fields(o, lambda x, y, z: 1.5 * (1 - (y - 1) ** 2) + 0 * x * z,
                     lambda x, y, z: 0.25 + 0 * x * y * z)
    o.set_state(U, V, W)
    s = o.compute_statistics()
    # U interpolated to the faces y_j (linear weights of a quadratic)
    np.testing.assert_allclose(s["Vmean"], 0.25, rtol=1e-14)
    np.testing.assert_allclose(s["V2mean"], 0.0625, rtol=1e-14)
    g = o.grid()
    yg = g["yg"]
    uf = g["w0"] * 1.5 * (1 - (yg[:-1] - 1) ** 2) + g["w1"] * 1.5 * (1 - (yg[1:] - 1) ** 2)
    np.testing.assert_allclose(s["Umean"], uf, rtol=1e-13)
    np.testing.assert_allclose(s["UVmean"], 0.25 * uf, rtol=1e-13)


def test_wall_model_closed_form():
    """compute_equilibrium_wall_model (wallmodel.f90:591-685) on fields whose
    wall-parallel velocity is known in closed form: U = a + b x at the
    matching height and W linear in z (the 3-point Lagrange interpolations
    reproduce both exactly), nu_t = 0.  Then u_les = sqrt(U^2 + W(zg)^2) at
    each U point, tau = the ODE stress of u_les, alpha_x = tau U/(u_les nu),
    and alpha_z (4-point average of tau W/(u_les nu) on the U points) is
    checked against the same average evaluated directly."""
    c = cfg_small(cells=(12, 16, 10), stretch=0.0, iwall_model=3, les_model=0, eqwm_ncv=30,
                  eqwm_tol=1e-10, eqwm_max_iter=300, nu=1e-4, dPdx=1e-3)
    o = Oracle(c)
    U, V, W = fields(o, lambda x, y, z: 0.8 + 0.05 * x + 0 * y * z, None,
                     lambda x, y, z: 0.1 + 0.03 * z + 0 * x * y)
    o.set_state(U, V, W)
    o.set_tau_init(np.full((o.nx, 2, o.nzg), 1e-3, order="F"))
    o.compute_wall_model()
    ax, az, ti = o.wall_state()
    g = o.grid()
    zg = g["zg"]
    ul = np.sqrt((0.8 + 0.05 * o.x)[:, None] ** 2 + (0.1 + 0.03 * zg)[None, :] ** 2)  # (nx, nzg)
    tau = np.vectorize(lambda u: eqwm(u, 1e-4, g["yg"][o.p.jref_eqwm - 1], 1e-3, ncv=30, stretch=1.1,
                                      tol=1e-10, max_iter=300)[0])(ul)
    for wall in (0, 1):
        np.testing.assert_allclose(ti[:, wall, :], tau, rtol=1e-12)
        np.testing.assert_allclose(ax[:, wall, :], tau * (0.8 + 0.05 * o.x)[:, None] / ul / 1e-4, rtol=1e-12)
    azt = tau * (0.1 + 0.03 * zg)[None, :] / ul / 1e-4  # at U points (i, k)
    # interior W points: average of the 4 surrounding U-point values
    exp = 0.25 * (azt[:-1, :-1] + azt[1:, :-1] + azt[:-1, 1:] + azt[1:, 1:])  # (nx-1, nzg-1) -> W (i, k)
    np.testing.assert_allclose(az[1:o.nx - 2, 0, 1:o.nz - 2], exp[1:o.nx - 2, 1:o.nz - 2], rtol=1e-12)


# ------------------------------------------------------------- actual pressure
def test_pressure_wall_condition_closed_form():
    """pressure.f90 compute_pressure_BC on a uniform y grid: for
    V = c (y - y1)(y - yN) the one-sided formulas are exact, so
    d/dy(nu dV/dy) = 2 c nu at both walls (no eddy viscosity)."""
    c = cfg_small(cells=(8, 12, 8), stretch=0.0)
    o = Oracle(c)
    U, V, W = (np.zeros(s, order="F") for s in (o.shape_u, o.shape_v, o.shape_w))
    y = o.y
    cc = 0.7
    V[:] = (cc * (y - y[0]) * (y - y[-1]))[None, :, None]
    o.set_state(U, V, W)
    _, _, b1, b2 = o.compute_pressure(details=True)
    np.testing.assert_allclose(b1, 2 * cc * c.nu, rtol=1e-9)
    np.testing.assert_allclose(b2, 2 * cc * c.nu, rtol=1e-9)


@pytest.mark.parametrize("les", [False, True])
def test_pressure_solves_poisson_with_wall_conditions(les):
    """The returned P satisfies the finite-volume Laplacian = rhs_p at every
    interior cell (ghost rows from the wall conditions included), up to the
    gauge row: the plane-mean residual is zero except at j = 2, where the
    (0,0) mode's diagonal is scaled by 3/2 (residual -D(2)/2 <P(:,2,:)>)."""
    c = cfg_small(cells=(16, 16, 12), les_model=1 if les else 0, dPdx=1e-3)
    o = Oracle(c)
    U, V, W = cases.initial_field(c)
    o.initialise(U, V, W)
    o.apply_boundary_conditions()
    if les:
        o.compute_eddy_viscosity()
    Pa, rp, b1, b2 = o.compute_pressure(details=True)
    g = o.grid()
    x, y, z, yg = o.x, o.y, o.z, g["yg"]
    dx, dz = x[1] - x[0], z[1] - z[0]
    nxg, nyg, nzg = o.nxg, o.nyg, o.nzg
    Pi = Pa[1:nxg - 2, :, 1:nzg - 2]            # unique cells i = 2..nxg-2, k = 2..nzg-2 (all j)
    Lx = (np.roll(Pi, -1, 0) - 2 * Pi + np.roll(Pi, 1, 0)) / dx ** 2
    Lz = (np.roll(Pi, -1, 2) - 2 * Pi + np.roll(Pi, 1, 2)) / dz ** 2
    j = np.arange(1, nyg - 1)                   # 0-based rows of j = 2..nyg-1
    dyf = (y[j] - y[j - 1])[None, :, None]
    up = (Pi[:, j + 1, :] - Pi[:, j, :]) / (yg[j + 1] - yg[j])[None, :, None]
    dn = (Pi[:, j, :] - Pi[:, j - 1, :]) / (yg[j] - yg[j - 1])[None, :, None]
    L = Lx[:, j, :] + (up - dn) / dyf + Lz[:, j, :]
    res = L - rp[:nxg - 3, :, :nzg - 3]
    scale = np.abs(rp).max()
    mean = res.mean(axis=(0, 2))
    np.testing.assert_allclose(res - mean[None, :, None], 0.0, atol=1e-9 * scale)
    np.testing.assert_allclose(mean[1:], 0.0, atol=1e-9 * scale)
    b = -(1.0 / (y[1] - y[0])) * (1.0 / (yg[2] - yg[1]) + 1.0 / (yg[1] - yg[0]))
    cdn = 1.0 / (y[1] - y[0]) / (yg[1] - yg[0])
    assert abs(mean[0] - (-0.5 * (b + cdn) * Pi[:, 1, :].mean())) < 1e-9 * scale
    # ghost rows carry the wall conditions, periodic copies the duplicates
    np.testing.assert_allclose((Pa[1:-1, 1, 1:-1] - Pa[1:-1, 0, 1:-1]) / (yg[1] - yg[0]), b1, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(Pa[nxg - 2], Pa[1])
    np.testing.assert_allclose(Pa[:, :, nzg - 2], Pa[:, :, 1])
