"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test
infrastructure only, used as the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_eqwm.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ["cfl", "nu", "dPdx", "dPdz", "omega_z", "cs", "eqwm_stretch", "eqwm_tol", "tau_mean_eqwm"]] + \
               [(n, ctypes.c_int) for n in
                ["x_mass_cte", "y_mass_cte", "les_model", "iwall_model", "dirichlet_nu_t", "jref_eqwm",
                 "eqwm_ncv", "eqwm_max_iter", "nthreads"]] + [("u_centerline", ctypes.c_double)]


_libs = {}


def lib(path=None):
    """The oracle library (oracle/liboracle.so), or a build of it at `path`
    (the planted-mistake variants of tests/test_oracle_planted.py)."""
    path = path or ORACLE_SO
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(path)
        v = ctypes.c_void_p
        sig = {
            "orc_create": (v, [ctypes.c_int] * 3 + [_dp] * 3 + [ctypes.POINTER(OrcParams)]),
            "orc_destroy": (None, [v]),
            "orc_get_grid": (None, [v] + [_dp] * 8),
            "orc_get_dyy": (None, [v, _dp, _dp, _dp]),
            "orc_set_state": (None, [v, _dp, _dp, _dp]),
            "orc_get_state": (None, [v, _dp, _dp, _dp, _dp]),
            "orc_set_nu_t": (None, [v, _dp]),
            "orc_get_nu_t": (None, [v, _dp, _dp]),
            "orc_get_alpha": (None, [v, _dp, _dp, _dp]),
            "orc_set_tau_init": (None, [v, _dp]),
            "orc_set_scalars": (None, [v, _dp]),
            "orc_get_scalars": (None, [v, _dp]),
            "orc_apply_boundary_conditions": (None, [v]),
            "orc_projection_step": (None, [v, ctypes.c_int]),
            "orc_compute_rhs": (None, [v] + [_dp] * 6),
            "orc_compute_dt": (ctypes.c_double, [v]),
            "orc_rk3_step": (None, [v]),
            "orc_mean_mass_flow_U": (ctypes.c_double, [v, _dp]),
            "orc_mean_mass_flow_V": (ctypes.c_double, [v, _dp]),
            "orc_mean_centerline_U": (ctypes.c_double, [v, _dp]),
            "orc_eddy_viscosity": (None, [v]),
            "orc_wall_model": (None, [v, _ip, _ip]),
            "orc_check_divergence": (None, [v, _dp, _dp]),
            "orc_compute_pressure": (None, [v, _dp, _dp, _dp, _dp]),
            "orc_statistics": (None, [v, _dp]),
            "orc_eqwm": (None, [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                         ctypes.c_int, ctypes.c_double, _dp, _ip]),
            "orc_fft": (None, [ctypes.c_int, _dp, _dp, ctypes.c_int]),
            "orc_fft2d": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]),
            "orc_zgtsv": (ctypes.c_int, [ctypes.c_int] + [_dp] * 5),
            "orc_interp_2nd": (None, [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[path] = L
    return _libs[path]


def P(a):
    return a.ctypes.data_as(_dp)


def _f(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


class Oracle:
    """Same surface as channel_flow.solver.ChannelFlow, on the CPU oracle."""

    def __init__(self, cfg, x=None, y=None, z=None, nthreads=0, jref=None, path=None):
        self.L = lib(path)
        if x is None:
            x, y, z = cfg.grid()
        self.x, self.y, self.z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        nx, ny, nz = len(self.x), len(self.y), len(self.z)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxg, self.nyg, self.nzg = nx + 1, ny + 1, nz + 1
        p = OrcParams()
        p.cfl, p.nu, p.dPdx, p.dPdz, p.omega_z = cfg.cfl, cfg.nu, cfg.dPdx, cfg.dPdz, cfg.omega_z
        p.cs, p.eqwm_stretch, p.eqwm_tol, p.tau_mean_eqwm = cfg.cs, cfg.eqwm_stretch, cfg.eqwm_tol, cfg.dPdx
        p.x_mass_cte, p.y_mass_cte, p.les_model, p.iwall_model = cfg.x_mass_cte, cfg.y_mass_cte, cfg.les_model, cfg.iwall_model
        p.dirichlet_nu_t = cfg.dirichlet_nu_t
        p.jref_eqwm = jref if jref is not None else (cfg.jref(self.y) if cfg.iwall_model == 3 else 2)
        p.eqwm_ncv, p.eqwm_max_iter, p.nthreads = cfg.eqwm_ncv, cfg.eqwm_max_iter, nthreads
        p.u_centerline = cfg.u_centerline
        self.p = p
        self.h = self.L.orc_create(nx, ny, nz, P(self.x), P(self.y), P(self.z), ctypes.byref(p))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    shape_u = property(lambda s: (s.nx, s.nyg, s.nzg))
    shape_v = property(lambda s: (s.nxg, s.ny, s.nzg))
    shape_w = property(lambda s: (s.nxg, s.nyg, s.nz))
    shape_c = property(lambda s: (s.nxg, s.nyg, s.nzg))

    def grid(self):
        xg, yg, zg = np.zeros(self.nxg), np.zeros(self.nyg), np.zeros(self.nzg)
        kxx, kzz = np.zeros(self.nx - 2), np.zeros(self.nz - 2)
        w0, w1, dm = np.zeros(self.ny), np.zeros(self.ny), np.zeros(3)
        self.L.orc_get_grid(self.h, P(xg), P(yg), P(zg), P(kxx), P(kzz), P(w0), P(w1), P(dm))
        return dict(xg=xg, yg=yg, zg=zg, kxx=kxx, kzz=kzz, w0=w0, w1=w1, dmin=dm)

    def dyy(self):
        n = self.nyg - 2
        d, l, u = np.zeros(n), np.zeros(n - 1), np.zeros(n - 1)
        self.L.orc_get_dyy(self.h, P(d), P(l), P(u))
        return d, l, u

    def set_state(self, U, V, W):
        U, V, W = (np.asfortranarray(a, dtype=np.float64) for a in (U, V, W))
        self.L.orc_set_state(self.h, P(U), P(V), P(W))

    def get_state(self):
        U, V, W = _f(self.shape_u), _f(self.shape_v), _f(self.shape_w)
        self.L.orc_get_state(self.h, P(U), P(V), P(W), None)
        return U, V, W

    @property
    def P(self):
        a = _f(self.shape_c)
        self.L.orc_get_state(self.h, None, None, None, P(a))
        return a

    @property
    def nu_t(self):
        a = _f(self.shape_c)
        self.L.orc_get_nu_t(self.h, P(a), None)
        return a

    @property
    def avg_nu_t(self):
        a = np.zeros(self.nyg)
        self.L.orc_get_nu_t(self.h, None, P(a))
        return a

    def set_nu_t(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        self.L.orc_set_nu_t(self.h, P(a))

    def wall_state(self):
        ax, az, ti = _f((self.nx, 2, self.nzg)), _f((self.nxg, 2, self.nz)), _f((self.nx, 2, self.nzg))
        self.L.orc_get_alpha(self.h, P(ax), P(az), P(ti))
        return ax, az, ti

    def set_tau_init(self, ti):
        ti = np.asfortranarray(ti, dtype=np.float64)
        self.L.orc_set_tau_init(self.h, P(ti))

    @property
    def scalars(self):
        s = np.zeros(7)
        self.L.orc_get_scalars(self.h, P(s))
        return dict(zip(["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"], s))

    def set_scalars(self, **kw):
        cur = self.scalars
        cur.update(kw)
        s = np.array([cur[k] for k in ["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"]])
        self.L.orc_set_scalars(self.h, P(s))

    def apply_boundary_conditions(self):
        self.L.orc_apply_boundary_conditions(self.h)

    def compute_projection_step(self, rk_step=0):
        self.L.orc_projection_step(self.h, rk_step)

    def compute_dt(self):
        return self.L.orc_compute_dt(self.h)

    def compute_time_step_RK3(self):
        self.L.orc_rk3_step(self.h)

    def run(self, nsteps):
        for _ in range(nsteps):
            self.compute_dt()
            self.compute_time_step_RK3()

    def compute_rhs(self, U_, V_, W_):
        U_, V_, W_ = (np.asfortranarray(a, dtype=np.float64) for a in (U_, V_, W_))
        ru = _f((self.nx - 2, self.nyg - 2, self.nzg - 2))
        rv = _f((self.nxg - 2, self.ny - 2, self.nzg - 2))
        rw = _f((self.nxg - 2, self.nyg - 2, self.nz - 2))
        self.L.orc_compute_rhs(self.h, P(U_), P(V_), P(W_), P(ru), P(rv), P(rw))
        return ru, rv, rw

    def compute_eddy_viscosity(self):
        self.L.orc_eddy_viscosity(self.h)

    def compute_pressure(self, details=False):
        """pressure.f90 compute_pressure: P, and with details also rhs_p
        (2:nxg-1,2:nyg-1,2:nzg-1) and the wall conditions (2:nxg-1,2:nzg-1)."""
        Pa = _f(self.shape_c)
        rp = _f((self.nxg - 2, self.nyg - 2, self.nzg - 2))
        b1, b2 = _f((self.nxg - 2, self.nzg - 2)), _f((self.nxg - 2, self.nzg - 2))
        self.L.orc_compute_pressure(self.h, P(Pa), P(rp), P(b1), P(b2))
        return (Pa, rp, b1, b2) if details else Pa

    def compute_wall_model(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.L.orc_wall_model(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def compute_mean_mass_flow_U(self):
        U, _, _ = self.get_state()
        return self.L.orc_mean_mass_flow_U(self.h, P(U))

    def compute_mean_centerline_U(self):
        U, _, _ = self.get_state()
        return self.L.orc_mean_centerline_U(self.h, P(U))

    def compute_mean_mass_flow_V(self):
        _, V, _ = self.get_state()
        return self.L.orc_mean_mass_flow_V(self.h, P(V))

    def init_mass_flow(self):
        self.set_scalars(Qflow_x_0=self.compute_mean_mass_flow_U())

    def check_divergence(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self.L.orc_check_divergence(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def compute_statistics(self):
        out = np.zeros(7 * self.ny + 4)
        self.L.orc_statistics(self.h, P(out))
        names = ["Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean"]
        d = {n: out[i * self.ny:(i + 1) * self.ny].copy() for i, n in enumerate(names)}
        d.update(utau=out[7 * self.ny], Retau=out[7 * self.ny + 1], Qflow_x=out[7 * self.ny + 2],
                 Qflow_y=out[7 * self.ny + 3])
        return d

    def initialise(self, U, V, W):
        self.set_state(U, V, W)
        self.apply_boundary_conditions()
        self.compute_projection_step(rk_step=0)
        self.apply_boundary_conditions()
        self.init_mass_flow()


def eqwm(u, nu, h, tau_in, ncv=100, stretch=1.10, tol=1e-7, max_iter=100, rho=1.0):
    L = lib()
    t, it = ctypes.c_double(), ctypes.c_int()
    L.orc_eqwm(u, nu, rho, h, ncv, stretch, tol, max_iter, tau_in, ctypes.byref(t), ctypes.byref(it))
    return t.value, it.value


_ref = None


def ref_eqwm_available():
    return os.path.exists(REF_SO)


def ref_eqwm(u, nu, h, tau_in, ncv=100, stretch=1.10, rho=1.0):
    """The reference's own solve_u_eqwm (flang build in oracle/_ref)."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_SO)
        f = _ref.cf_ref_solve_u_eqwm
        f.restype = None
        f.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _ip]
    t, it = ctypes.c_double(), ctypes.c_int()
    _ref.cf_ref_solve_u_eqwm(u, nu, rho, h, ncv, stretch, tau_in, ctypes.byref(t), ctypes.byref(it))
    return t.value, it.value
