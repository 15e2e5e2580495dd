"""Closed-form checks of the oracle's right-hand sides (equations.f90:24-498)
and plane statistics (statistics.f90:22-171) on a tanh-stretched y grid.

Each check raises AssertionError when the oracle at `path` (default: the
normal build) disagrees with an answer derived by hand from the reference's
discrete operators.  tests/test_oracle_pins_closed_form.py runs them on the
normal build; tests/test_oracle_planted.py runs them on builds with one
planted mistake each and expects them to fail.

Notation: y(j) are the wall-normal faces, yg(j) the centres (yg(1) and
yg(nyg) the mirrored ghosts), so yg(j) = (y(j-1) + y(j))/2 inside.  The
reference interpolates centre values to the faces with weight_y_0/1
(initialization.f90:386-388), which is exact for fields linear in y; face
values go to the centres as plain averages (interpolation.f90:47-61).
"""
import math

import numpy as np

from channel_flow import cases
from oracle_lib import Oracle


def cfg(**kw):
    c = cases.case_dns64()
    c.cells = kw.pop("cells", (12, 14, 8))
    c.stretch = kw.pop("stretch", 2.0)
    c.x_mass_cte = 0
    c.dPdx = kw.pop("dPdx", 0.0)
    c.dPdz = kw.pop("dPdz", 0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def close(a, b, tol, what):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = max(float(np.max(np.abs(b))), 1.0)
    err = float(np.max(np.abs(a - b)))
    assert err <= tol * scale, "%s: max error %.3e (scale %.3e)" % (what, err, scale)


def _zeros(o):
    return (np.zeros(o.shape_u, order="F"), np.zeros(o.shape_v, order="F"), np.zeros(o.shape_w, order="F"))


def _profile_fields(o, fu=None, fv=None, fw=None, fnt=None):
    """Fields that depend on y only: U, W and nu_t on the centres yg (ghost
    rows included), V on the faces y."""
    yg = o.grid()["yg"]
    U, V, W = _zeros(o)
    if fu:
        U[:] = fu(yg)[None, :, None]
    if fv:
        V[:] = fv(o.y)[None, :, None]
    if fw:
        W[:] = fw(yg)[None, :, None]
    nt = np.zeros(o.shape_c, order="F")
    if fnt:
        nt[:] = fnt(yg)[None, :, None]
    return U, V, W, nt


# ------------------------------------------------------------------ RHS
def check_rhs_linear_profiles(path=None):
    """U = cu*yg, W = cw*yg, V = b*y, nu_t = a + bt*yg, rotation Om:
      rhs_u(j) = -2 b cu yg(j) + cu bt + 2 Om b yg(j)
      rhs_v(j) = -b^2 (yg(j) + yg(j+1)) + 2 b bt - 2 Om cu y(j)
      rhs_w(j) = -2 b cw yg(j) + cw bt
    -d(uv)/dy: uv at face j is cu b y(j)^2 (weights exact on linear U), so
    its difference over y(j)-y(j-1) is cu b (y(j)+y(j-1)) = 2 cu b yg(j);
    d((nu+nu_t) du/dy)/dy: du/dy = cu, nu_t at face j is a + bt y(j), so the
    term is cu bt; -dv^2/dy on the centres: b^2 (yg(j+1)^2 - yg(j)^2) over
    yg(j+1)-yg(j); 2 d((nu+nu_t) dv/dy)/dy with nu_t at the centres:
    2 b bt; Coriolis: V averaged to the U location is b yg(j), U interpolated
    to the V location is cu y(j).  Every x and z derivative vanishes."""
    cu, cw, b, a, bt, om = 1.3, -0.6, 0.7, 0.01, 0.02, 0.3
    o = Oracle(cfg(nu=0.004, omega_z=om), path=path)
    yg, y = o.grid()["yg"], o.y
    U, V, W, nt = _profile_fields(o, lambda t: cu * t, lambda t: b * t, lambda t: cw * t, lambda t: a + bt * t)
    o.set_nu_t(nt)
    ru, rv, rw = o.compute_rhs(U, V, W)
    ygi = yg[1:-1]
    close(ru, np.broadcast_to((-2 * b * cu * ygi + cu * bt + 2 * om * b * ygi)[None, :, None], ru.shape), 1e-12,
          "rhs_u of linear profiles")
    yi = y[1:-1]
    close(rv, np.broadcast_to((-b * b * (yg[1:-2] + yg[2:-1]) + 2 * b * bt - 2 * om * cu * yi)[None, :, None],
                              rv.shape), 1e-12, "rhs_v of linear profiles")
    close(rw, np.broadcast_to((-2 * b * cw * ygi + cw * bt)[None, :, None], rw.shape), 1e-12,
          "rhs_w of linear profiles")


def check_rhs_v_stretched_viscous(path=None):
    """V = c*y^2 (U = W = 0, nu_t = 0): the viscous part of rhs_v,
    2 d(nu dv/dy)/dy with dv/dy on the centres over the face spacings
    y(j)-y(j-1) and the outer difference over yg(j+1)-yg(j), is
    2 nu c (y(j+1)-y(j-1)) / ((y(j+1)-y(j-1))/2) = 4 nu c exactly on any
    grid.  The convective part does not depend on nu, so two viscosities
    isolate it."""
    c = 0.9
    outs = []
    for nu in (0.05, 0.01):
        o = Oracle(cfg(nu=nu), path=path)
        U, V, W, _ = _profile_fields(o, fv=lambda t: c * t * t)
        outs.append(o.compute_rhs(U, V, W))
    dv = outs[0][1] - outs[1][1]
    close(dv, np.full(dv.shape, 4 * c * (0.05 - 0.01)), 1e-12, "viscous part of rhs_v for V = c y^2")
    for q in (0, 2):
        close(outs[0][q] - outs[1][q], 0.0, 1e-12, "rhs_u/rhs_w for V = c y^2")


def check_rhs_v_nu_t_on_edges(path=None):
    """V = sin(k xg) or sin(k zg) (U = W = 0), nu_t = a + bt*yg: the x and z
    viscous terms of rhs_v use nu_t interpolated in y to the face y(j)
    (a + bt y(j), exact) and the 3-point second difference, whose
    eigenvalue is -4 sin^2(k d/2)/d^2.  All other terms vanish:
      rhs_v = (nu + a + bt y(j)) * lam * V."""
    a, bt, nu = 0.003, 0.01, 0.002
    o = Oracle(cfg(nu=nu), path=path)
    g = o.grid()
    xg, zg, y = g["xg"], g["zg"], o.y
    _, _, _, nt = _profile_fields(o, fnt=lambda t: a + bt * t)
    o.set_nu_t(nt)
    nuf = (nu + a + bt * y[1:-1])[None, :, None]
    dx = o.x[2] - o.x[1]
    kx = 2 * math.pi * 2 / (o.x[-2] - o.x[0])
    U, V, W = _zeros(o)
    V[:] = np.sin(kx * xg)[:, None, None]
    _, rv, _ = o.compute_rhs(U, V, W)
    lam = -4.0 / dx ** 2 * math.sin(kx * dx / 2) ** 2
    close(rv, nuf * lam * np.sin(kx * xg[1:-1])[:, None, None], 1e-12, "rhs_v x viscous with nu_t(y)")
    dz = o.z[2] - o.z[1]
    kz = 2 * math.pi * 1 / (o.z[-2] - o.z[0])
    U, V, W = _zeros(o)
    V[:] = np.sin(kz * zg)[None, None, :]
    _, rv, _ = o.compute_rhs(U, V, W)
    lam = -4.0 / dz ** 2 * math.sin(kz * dz / 2) ** 2
    close(rv, nuf * lam * np.sin(kz * zg[1:-1])[None, None, :], 1e-12, "rhs_v z viscous with nu_t(y)")


def check_rhs_wall_mirror(path=None):
    """On the symmetric tanh grid, mirroring the flow in y (U, W, nu_t
    reversed; V reversed and negated; rotation reversed) mirrors the right-
    hand sides: rhs_u, rhs_w reversed, rhs_v reversed and negated."""
    om = 0.2
    rng = np.random.default_rng(3)
    oa = Oracle(cfg(nu=0.01, omega_z=om), path=path)
    ob = Oracle(cfg(nu=0.01, omega_z=-om), path=path)
    U = np.asfortranarray(rng.standard_normal(oa.shape_u))
    V = np.asfortranarray(rng.standard_normal(oa.shape_v))
    W = np.asfortranarray(rng.standard_normal(oa.shape_w))
    nt = np.asfortranarray(np.abs(rng.standard_normal(oa.shape_c)) * 0.01)
    oa.set_nu_t(nt)
    ob.set_nu_t(np.asfortranarray(nt[:, ::-1, :]))
    ru, rv, rw = oa.compute_rhs(U, V, W)
    mu, mv, mw = ob.compute_rhs(np.asfortranarray(U[:, ::-1, :]), np.asfortranarray(-V[:, ::-1, :]),
                                np.asfortranarray(W[:, ::-1, :]))
    close(mu, ru[:, ::-1, :], 1e-11, "mirrored rhs_u")
    close(mv, -rv[:, ::-1, :], 1e-11, "mirrored rhs_v")
    close(mw, rw[:, ::-1, :], 1e-11, "mirrored rhs_w")


# ------------------------------------------------------------ statistics
def check_statistics_ranges(path=None):
    """compute_statistics (statistics.f90:43-71) sums U over faces
    1..nx-2 and centres k = 1..nzg-2, W over centres i = 1..nxg-2 and faces
    1..nz-2, V and UV over centres 2..nxg-1 x 2..nzg-1: each range counts
    one periodic image twice (k = 1 is the image of nzg-2, i = 1 of nxg-2,
    i = nxg-1 of 2, k = nzg-1 of 2).  With periodic index patterns
    g_k = cos(2 pi m (k-2)/Nz), h_i = cos(2 pi p (i-2)/Nx) (Nz, Nx unique
    centres), whose sums over one period vanish and whose squares sum to
    N/2, the means and mean squares follow in closed form.  U and W are
    linear in yg, so their face values are exact."""
    o = Oracle(cfg(cells=(12, 10, 8)), path=path)
    nx, nxg, nz, nzg, ny = o.nx, o.nxg, o.nz, o.nzg, o.ny
    Nx, Nz = nxg - 3, nzg - 3
    yg, y = o.grid()["yg"], o.y
    m, p = 1, 2
    kk, ii = np.arange(1, nzg + 1), np.arange(1, nxg + 1)
    g = np.cos(2 * math.pi * m * (kk - 2) / Nz)
    h = np.cos(2 * math.pi * p * (ii - 2) / Nx)
    cu, eu, cw, ew, b, ev, dv = 0.8, 0.3, -0.5, 0.2, 0.6, 0.15, -0.1
    U = np.asfortranarray(np.broadcast_to(cu * yg[None, :, None] + eu * g[None, None, :], o.shape_u))
    W = np.asfortranarray(np.broadcast_to(cw * yg[None, :, None] + ew * h[:nxg, None, None], o.shape_w))
    V = np.asfortranarray(np.broadcast_to(b * y[None, :, None] + ev * h[:, None, None] + dv * g[None, None, :],
                                          o.shape_v))
    o.set_state(U, V, W)
    s = o.compute_statistics()
    c1z = math.cos(2 * math.pi * m / Nz)  # g at k = 1 (the image counted twice)
    c1x = math.cos(2 * math.pi * p / Nx)  # h at i = 1
    dz_, dx_ = nzg - 2, nxg - 2
    uy, wy, vy = cu * y, cw * y, b * y
    close(s["Umean"], uy + eu * c1z / dz_, 1e-13, "Umean")
    close(s["U2mean"], uy ** 2 + (2 * uy * eu * c1z + eu ** 2 * (Nz / 2 + c1z ** 2)) / dz_, 1e-13, "U2mean")
    close(s["Wmean"], wy + ew * c1x / dx_, 1e-13, "Wmean")
    close(s["W2mean"], wy ** 2 + (2 * wy * ew * c1x + ew ** 2 * (Nx / 2 + c1x ** 2)) / dx_, 1e-13, "W2mean")
    # V: sums over 2..nxg-1 and 2..nzg-1 count h_2 = g_2 = 1 twice
    n = dx_ * dz_
    close(s["Vmean"], vy + ev / dx_ + dv / dz_, 1e-13, "Vmean")
    v2 = (vy ** 2 * n + 2 * vy * (ev * dz_ + dv * dx_) + ev ** 2 * dz_ * (Nx / 2 + 1) + dv ** 2 * dx_ * (Nz / 2 + 1)
          + 2 * ev * dv) / n
    close(s["V2mean"], v2, 1e-13, "V2mean")
    # UV: U at the face and x centre is cu y + eu g_k (x-independent)
    uv = (uy * (vy * n + ev * dz_ + dv * dx_) + eu * (vy * dx_ * 1 + ev * 1 * 1 + dv * dx_ * (Nz / 2 + 1))) / n
    close(s["UVmean"], uv, 1e-13, "UVmean")
    assert len(s["Umean"]) == ny


def check_statistics_uv_pairing(path=None):
    """UVmean multiplies V at centre i by U averaged over the faces i-1 and i
    (statistics.f90:58, term(1:nxg-2) against V(2:nxg-1)).  U = cos(k x) on
    the faces, V = cos(k xg) on the centres: the average is
    cos(k xg(i)) cos(k dx/2), and the sum over i = 2..nxg-1 counts i = 2
    twice, so UVmean = cos(k dx/2) (Nx/2 + cos^2(k xg(2))) / (nxg-2)."""
    o = Oracle(cfg(cells=(12, 10, 8)), path=path)
    g = o.grid()
    xg, x = g["xg"], o.x
    k = 2 * math.pi * 2 / (x[-2] - x[0])
    dx = x[2] - x[1]
    U, V, W = _zeros(o)
    U[:] = np.cos(k * x)[:, None, None]
    V[:] = np.cos(k * xg)[:, None, None]
    o.set_state(U, V, W)
    s = o.compute_statistics()
    Nx = o.nxg - 3
    expect = math.cos(k * dx / 2) * (Nx / 2 + math.cos(k * xg[1]) ** 2) / (o.nxg - 2)
    close(s["UVmean"], np.full(o.ny, expect), 1e-13, "UVmean pairing")


def check_statistics_utau_and_bulk(path=None):
    """u_tau (statistics.f90:121-155) from the wall balance
    utau^2 = (UV_t - UV_b - (nu+nut_t) dU/dy_t + (nu+nut_b) dU/dy_b)/Ly.
    (1) U = A yg(2-yg), V = 0, nu_t = c: the one-sided wall differences of
    a parabola are its slopes at the walls y = 0, 2 (the ghost centres are
    mirrored), +2A and -2A, so utau^2 = 4 A (nu + c)/Ly.  (2) U = a,
    V = b y: UVmean(ny) - UVmean(1) = 2ab, so utau^2 = 2ab/Ly.
    Re_tau = utau (y(ny)-y(1))/2/nu.  Bulk flows (mass_flow.f90:25-132) of
    linear profiles are exact under the trapezoid rule: U = cu yg gives
    cu, V = b y gives b (the grid is symmetric about y = 1)."""
    nu, A, c = 0.003, 0.7, 0.002
    o = Oracle(cfg(nu=nu), path=path)
    yg = o.grid()["yg"]
    U, V, W = _zeros(o)
    U[:] = (A * yg * (2 - yg))[None, :, None]
    o.set_state(U, V, W)
    o.set_nu_t(np.full(o.shape_c, c, order="F"))
    s = o.compute_statistics()
    Ly = o.y[-1] - o.y[0]
    ut = math.sqrt(4 * A * (nu + c) / Ly)
    close(s["utau"], ut, 1e-13, "utau of a parabola")
    close(s["Retau"], ut * Ly / 2 / nu, 1e-12, "Retau")
    a, b = 1.1, 0.4
    o = Oracle(cfg(nu=nu), path=path)
    U, V, W = _zeros(o)
    U[:] = a
    V[:] = (b * o.y)[None, :, None]
    o.set_state(U, V, W)
    s = o.compute_statistics()
    close(s["utau"], math.sqrt(2 * a * b / Ly), 1e-13, "utau from the wall Reynolds stresses")
    cu = 0.9
    U[:] = (cu * yg)[None, :, None]
    o.set_state(U, V, W)
    s = o.compute_statistics()
    close(s["Qflow_x"], cu, 1e-13, "bulk U of a linear profile")
    close(s["Qflow_y"], b, 1e-13, "bulk V of a linear profile")


ALL = [check_rhs_linear_profiles, check_rhs_v_stretched_viscous, check_rhs_v_nu_t_on_edges,
       check_rhs_wall_mirror, check_statistics_ranges, check_statistics_uv_pairing,
       check_statistics_utau_and_bulk]


# -------------------------------------------------------------- mass flow
def check_centerline_mean(path=None):
    """compute_dPx_for_constant_Ucenterline (mass_flow.f90:192-209) takes the
    plane mean of U(1:nx-2, jc, 1:nzg-2) with jc = Floor(Real(ny/2))+1, the
    integer division first.  U = f(yg) + e g_k (g_k the periodic pattern of
    check_statistics_ranges, k = 1 counted besides its image nzg-2):
    mean = f(yg(jc)) + e g_1/(nzg-2)."""
    o = Oracle(cfg(cells=(12, 14, 8)), path=path)
    yg = o.grid()["yg"]
    nzg, ny = o.nzg, o.ny
    Nz = nzg - 3
    g = np.cos(2 * math.pi * (np.arange(1, nzg + 1) - 2) / Nz)
    f = lambda t: 1.5 * t * (2 - t)
    U = np.asfortranarray(np.broadcast_to(f(yg)[None, :, None] + 0.2 * g[None, None, :], o.shape_u))
    _, V, W = _zeros(o)
    o.set_state(U, V, W)
    jc = ny // 2 + 1
    close(o.compute_mean_centerline_U(), f(yg[jc - 1]) + 0.2 * math.cos(2 * math.pi / Nz) / (nzg - 2), 1e-14,
          "centreline mean")


def _stepped_pair(kw0, kw1, path):
    """The same projected random field stepped once by two oracles whose
    configurations differ only in the mass-flow options."""
    outs = []
    for kw in (kw0, kw1):
        c = cfg(cells=(12, 14, 8), nu=0.01, cfl=-1e-3, dPdx=0.002, **kw)
        o = Oracle(c, path=path)
        U, V, W = cases.initial_field(c)
        o.initialise(U, V, W)
        o.set_scalars(Qflow_y_0=0.0)
        before = o.get_state()
        o.run(1)
        outs.append((o, before, o.get_state()))
    return outs


def check_mass_flow_corrections(path=None):
    """After the third substep (time_integration.f90:152-171):
    x_mass_cte = 2 adds d = Ucenterline_x - <U(1:nx-2, jc, 1:nzg-2)> to
    every interior U and sets dPdx = d/dt; y_mass_cte = 1 adds
    d = Qflow_y_0 - trapz(V)/norm to the interior rows 2..ny-1 of V (the wall
    rows stay 0) and sets dPdy = d/dt.  Stepping the same field with and
    without the option isolates the correction, whose value the test
    computes from the uncorrected field with the reference's ranges."""
    (o0, _, (U0, V0, W0)), (o2, _, (U2, V2, W2)) = _stepped_pair(dict(x_mass_cte=0),
                                                               dict(x_mass_cte=2, u_centerline=1.2), path)
    nx, ny, nzg = o0.nx, o0.ny, o0.nzg
    jc = ny // 2 + 1
    d = 1.2 - U0[0:nx - 2, jc - 1, 0:nzg - 2].sum() / ((nx - 2) * (nzg - 2))
    close(U2[1:nx - 1, 1:-1, 1:nzg - 1] - U0[1:nx - 1, 1:-1, 1:nzg - 1], d, 1e-13, "centreline correction of U")
    close(V2, V0, 1e-15, "V untouched by the x correction")
    dt = o2.scalars["dt"]
    close(o2.scalars["dPdx"] * dt, d, 1e-13, "dPdx = d/dt")
    (o0, _, (U0, V0, W0)), (o1, _, (U1, V1, W1)) = _stepped_pair(dict(x_mass_cte=0), dict(x_mass_cte=0, y_mass_cte=1),
                                                               path)
    y = o0.y
    nxg = o0.nxg
    col = V0[1:nxg - 1, :, 1:nzg - 1]
    q = ((col[:, 1:, :] + col[:, :-1, :]) * 0.5 * np.diff(y)[None, :, None]).sum()
    nrm = (2 * 0.5 * np.diff(y)).sum() * (nxg - 2) * (nzg - 2)
    d = 0.0 - q / nrm
    assert abs(d) > 1e-6, d  # the random field carries a bulk V to remove
    close(V1[1:nxg - 1, 1:ny - 1, 1:nzg - 1] - V0[1:nxg - 1, 1:ny - 1, 1:nzg - 1], d, 1e-13, "y correction of V")
    close(V1[:, [0, ny - 1], :], 0.0, 1e-15, "wall rows of V")
    close(U1, U0, 1e-15, "U untouched by the y correction")
    close(o1.scalars["dPdy"] * o1.scalars["dt"], d, 1e-13, "dPdy = d/dt")


ALL += [check_centerline_mean, check_mass_flow_corrections]
