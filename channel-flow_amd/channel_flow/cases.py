"""Grids, run configurations and the seeded initial field.

The configuration record follows the reference's input file
(input_output.f90:21-81) and adds the constants the reference hard-codes.
Grid sizes are quoted as unique cells NX x NY x NZ; the reference's face
counts are nx_global = NX+2, ny_global = NY+1, nz_global = NZ+2
(initialization.f90:79-99, 246-247).
"""
import math
from dataclasses import dataclass, field, asdict

import numpy as np


@dataclass
class CaseConfig:
    name: str = "case"
    cells: tuple = (64, 64, 64)          # unique cells NX, NY, NZ
    lengths: tuple = (2 * math.pi, 2.0, math.pi)
    stretch: float = 2.0                 # tanh stretching in y (0 = uniform)
    nu: float = 1.0 / 2800.0
    cfl: float = -2e-3                   # < 0: fixed dt = -cfl
    dPdx: float = 0.0
    dPdz: float = 0.0
    omega_z: float = 0.0
    x_mass_cte: int = 1                  # 1 constant bulk flow, 2 constant centreline velocity
    y_mass_cte: int = 0                  # 1 constant bulk flow in y
    u_centerline: float = 0.0            # x_mass_cte = 2 target (initialization.f90:453-455)
    les_model: int = 0
    iwall_model: int = 0
    dirichlet_nu_t: int = 0
    cs: float = 0.1
    h_wm: float = 0.1                    # wall-model matching height (jref nearest)
    eqwm_ncv: int = 100
    eqwm_stretch: float = 1.10
    eqwm_tol: float = 1e-7
    eqwm_max_iter: int = 100
    steps: int = 100
    seed: int = 2018
    amp: float = 0.1                     # +-10% U_b perturbation
    precision: str = "fp64"
    wall_surrogate: int = 0

    @property
    def faces(self):
        NX, NY, NZ = self.cells
        return NX + 2, NY + 1, NZ + 2

    def grid(self):
        return make_grid(self.faces, self.lengths, self.stretch)

    def jref(self, y):
        """Index j (1-based) of the y centre nearest h_wm above the wall."""
        yg = centres(y)
        j = int(np.argmin(np.abs(yg[1:len(yg) // 2] - (y[0] + self.h_wm)))) + 2
        return j

    def as_dict(self):
        return asdict(self)


def make_grid(faces, lengths, stretch):
    """Face coordinates x(nx), y(ny), z(nz).  x and z are uniform with the
    periodic length Lx spanning faces 1..nx-1; y is tanh-stretched on [0, 2]
    (uniform when stretch == 0)."""
    nx, ny, nz = faces
    Lx, Ly, Lz = lengths
    x = np.arange(nx, dtype=np.float64) * (Lx / (nx - 2))
    z = np.arange(nz, dtype=np.float64) * (Lz / (nz - 2))
    eta = -1.0 + 2.0 * np.arange(ny, dtype=np.float64) / (ny - 1)
    if stretch and stretch > 0:
        y = 1.0 + np.tanh(stretch * eta) / math.tanh(stretch)
    else:
        y = 1.0 + eta
    y = y * (Ly / 2.0)
    return x, y, z


def centres(f):
    """Centres plus mirrored ghosts of a face grid (initialization.f90:169-199)."""
    f = np.asarray(f, dtype=np.float64)
    m = 0.5 * (f[:-1] + f[1:])
    g = np.empty(len(f) + 1)
    g[1:-1] = m
    g[0] = m[0] - 2.0 * (m[0] - f[0])
    g[-1] = m[-1] + 2.0 * (f[-1] - m[-1])
    return g


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


def counter_uniform(seed, comp, shape):
    """Uniform [0,1) numbers keyed by (seed, component, i, j, k), 1-based
    Fortran indices, so any traversal order gives the same field."""
    n1, n2, n3 = shape
    with np.errstate(over="ignore"):
        i = np.arange(1, n1 + 1, dtype=np.uint64)[:, None, None]
        j = np.arange(1, n2 + 1, dtype=np.uint64)[None, :, None]
        k = np.arange(1, n3 + 1, dtype=np.uint64)[None, None, :]
        key = (np.uint64(comp) << np.uint64(56)) | (k << np.uint64(40)) | (j << np.uint64(20)) | i
        base = _splitmix64(np.uint64(seed) & _M64)
        h = _splitmix64(key ^ base)
    r = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.asfortranarray(r)


def initial_field(cfg: CaseConfig):
    """Raw (not yet projected) initial field: Poiseuille 1.5(1-(y-1)^2) in U
    plus a seeded +-amp perturbation in U, V, W.  Fortran-ordered arrays with
    the reference extents U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz)."""
    nx, ny, nz = cfg.faces
    x, y, z = cfg.grid()
    yg = centres(y)
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    a = cfg.amp
    prof = 1.5 * (1.0 - (yg - 1.0) ** 2)
    U = prof[None, :, None] + a * (2.0 * counter_uniform(cfg.seed, 0, (nx, nyg, nzg)) - 1.0)
    V = a * (2.0 * counter_uniform(cfg.seed, 1, (nxg, ny, nzg)) - 1.0)
    W = a * (2.0 * counter_uniform(cfg.seed, 2, (nxg, nyg, nz)) - 1.0)
    return np.asfortranarray(U), np.asfortranarray(V), np.asfortranarray(W)


def large_scale_field(cfg: CaseConfig, amp=0.3, modes=24, seed=None):
    """Poiseuille profile plus a seeded large-scale perturbation, for runs
    that must become turbulent (the grid-scale noise of initial_field decays
    at Re_b of a few thousand): each component is a sum of `modes` waves
    sin(pi n y/Ly) cos(a x + b z + phase) with n in 1..3 and the x and z
    wavenumbers among the four lowest of the box, amplitudes scaled so the
    perturbation rms is `amp` (in units of the bulk velocity).  Vanishes at
    the walls; not divergence-free until projected (initialise projects)."""
    nx, ny, nz = cfg.faces
    x, y, z = cfg.grid()
    xg, yg, zg = centres(x), centres(y), centres(z)
    Lx, Ly, Lz = cfg.lengths
    rng = np.random.default_rng(cfg.seed if seed is None else seed)
    prof = 1.5 * (1.0 - (yg - 1.0) ** 2)
    out = []
    # positions of each component: U on (x faces, y centres, z centres),
    # V on (x centres, y faces, z centres), W on (x centres, y centres, z faces)
    for c, (X, Y, Z) in enumerate(((x, yg, zg), (xg, y, zg), (xg, yg, z))):
        f = np.zeros((len(X), len(Y), len(Z)))
        for _ in range(modes):
            n = rng.integers(1, 4)
            a = 2.0 * np.pi * rng.integers(0, 4) / Lx
            b = 2.0 * np.pi * rng.integers(0, 4) / Lz
            ph = rng.uniform(0.0, 2.0 * np.pi)
            w = rng.standard_normal()
            f += w * (np.sin(np.pi * n * np.clip(Y, 0.0, Ly) / Ly)[None, :, None] *
                      np.cos(a * X[:, None, None] + b * Z[None, None, :] + ph))
        f *= amp / max(np.sqrt(np.mean(f ** 2)), 1e-300)
        if c == 0:
            f += prof[None, :, None]
        out.append(np.asfortranarray(f))
    return tuple(out)


def retau_from_reb(reb):
    """Dean's correlation Re_tau = 0.09 (2 Re_b)^0.88."""
    return 0.09 * (2.0 * reb) ** 0.88


def reb_from_retau(retau):
    """Inverse of Dean's correlation."""
    return 0.5 * (retau / 0.09) ** (1.0 / 0.88)


# ---- the five workloads ---------------------------------------------------
def case_example():
    """The reference's example input (input_parameters_test:1-37): 34x41x34
    faces (32x40x32 cells), nu = 8.888...e-5, constant dPdx = 0.00138, the
    dynamic Smagorinsky model with Neumann nu_t ghosts, no wall model,
    CFL 0.5, 1000 steps.  The grid is this package's tanh 2.0 grid."""
    return CaseConfig(name="example", cells=(32, 40, 32), stretch=2.0, nu=8.888888888888888e-05,
                      cfl=0.5, dPdx=0.001379795599360, x_mass_cte=0, les_model=2, iwall_model=0,
                      dirichlet_nu_t=0, steps=1000)


def case_dns64():
    """64^3 DNS, Re_b = 2800, tanh 2.0, dt = 2e-3, 100 steps."""
    return CaseConfig(name="dns64", cells=(64, 64, 64), stretch=2.0, nu=1 / 2800.0,
                      cfl=-2e-3, dPdx=(retau_from_reb(2800) / 2800.0) ** 2, steps=100)


def case_dns384(precision="fp64"):
    """384x256x384 DNS, Re_b = 10000, tanh 2.2, CFL 0.5, 1000 steps."""
    return CaseConfig(name="dns384", cells=(384, 256, 384), stretch=2.2, nu=1 / 10000.0,
                      cfl=0.5, dPdx=(retau_from_reb(10000) / 10000.0) ** 2, steps=1000,
                      precision=precision)


def case_wmles128():
    """128x64x128 wall-modelled LES at Re_tau = 2000, uniform y, Smagorinsky
    Cs = 0.1, EQWM at h = 0.1 with 30 cells, tol 1e-8."""
    nu = 1.0 / reb_from_retau(2000.0)
    return CaseConfig(name="wmles128", cells=(128, 64, 128), stretch=0.0, nu=nu, cfl=0.5,
                      dPdx=(2000.0 * nu) ** 2, x_mass_cte=1, les_model=1, iwall_model=3, cs=0.1,
                      h_wm=0.1, eqwm_ncv=30, eqwm_stretch=1.10, eqwm_tol=1e-8,
                      eqwm_max_iter=100, steps=5000, precision="mixed")


def case_wmles256():
    """256x96x256 wall-modelled LES at Re_tau = 5200 with the learned wall model."""
    nu = 1.0 / reb_from_retau(5200.0)
    c = case_wmles128()
    c.name = "wmles256_mlp"
    c.cells = (256, 96, 256)
    c.nu = nu
    c.dPdx = (5200.0 * nu) ** 2
    c.wall_surrogate = 1
    return c
