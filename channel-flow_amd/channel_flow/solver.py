"""Host-side mirror of the reference's module interface for the time step.

``ChannelFlow`` holds the module state of global.f90 as host mirrors and
forwards each subroutine of the hot path to the HIP library through the
C-ABI (include/channel_flow.h):

    compute_dt                    time_integration.f90:179
    compute_time_step_RK3         time_integration.f90:89
    apply_boundary_conditions     boundary_conditions.f90:20
    compute_projection_step       projection.f90:20
    compute_rhs_u / _v / _w       equations.f90:24, 186, 351
    compute_eddy_viscosity        subgrid.f90:23
    compute_wall_model            wallmodel.f90:24
    compute_mean_mass_flow_U      mass_flow.f90:25
    check_divergence              projection.f90:234
    compute_statistics            statistics.f90:22

Arrays are numpy Fortran-ordered with the reference extents (1-based index
i in the reference is index i-1 here).
"""
import ctypes

import numpy as np

from . import _lib
from .cases import CaseConfig, centres


def _f(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def case_params(cfg: CaseConfig, x, y, z, device=0):
    """cf_params of a case (include/channel_flow.h) on the face grids x, y, z."""
    p = _lib.CfParams()
    p.nx, p.ny, p.nz = len(x), len(y), len(z)
    p.precision = _lib.precision_code(cfg.precision)
    p.cfl, p.nu, p.dPdx, p.dPdz = cfg.cfl, cfg.nu, cfg.dPdx, cfg.dPdz
    p.omega_z = cfg.omega_z
    p.x_mass_cte, p.y_mass_cte = cfg.x_mass_cte, cfg.y_mass_cte
    p.les_model, p.iwall_model = cfg.les_model, cfg.iwall_model
    p.dirichlet_nu_t = cfg.dirichlet_nu_t
    p.jref_eqwm = cfg.jref(y) if cfg.iwall_model == 3 else 2
    p.cs = cfg.cs
    p.eqwm_ncv, p.eqwm_max_iter = cfg.eqwm_ncv, cfg.eqwm_max_iter
    p.eqwm_stretch, p.eqwm_tol = cfg.eqwm_stretch, cfg.eqwm_tol
    p.tau_mean_eqwm = cfg.dPdx
    p.wall_surrogate = cfg.wall_surrogate
    p.device = device
    p.u_centerline = cfg.u_centerline
    return p


class ChannelFlow:
    def __init__(self, cfg: CaseConfig, x=None, y=None, z=None, device=0, _handle=None):
        self.lib = _lib.load()
        self.cfg = cfg
        if x is None:
            x, y, z = cfg.grid()
        self.x, self.y, self.z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        nx, ny, nz = len(self.x), len(self.y), len(self.z)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxg, self.nyg, self.nzg = nx + 1, ny + 1, nz + 1
        p = case_params(cfg, self.x, self.y, self.z, device)
        self.params = p
        self.jref_eqwm = p.jref_eqwm
        self.yg = centres(self.y)
        # a member of a ChannelFlowBatch borrows its handle from the batch
        self._owned = _handle is None
        if _handle is not None:
            self.h = _handle
            return
        h = ctypes.c_void_p()
        _lib.check(self.lib.cf_create(ctypes.byref(p), _lib.dptr(self.x), _lib.dptr(self.y),
                                      _lib.dptr(self.z), ctypes.byref(h)), "cf_create")
        self.h = h

    # ---- lifetime ----
    def close(self):
        if getattr(self, "h", None) and self.h.value:
            if self._owned:
                self.lib.cf_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shapes ----
    @property
    def shape_u(self):
        return (self.nx, self.nyg, self.nzg)

    @property
    def shape_v(self):
        return (self.nxg, self.ny, self.nzg)

    @property
    def shape_w(self):
        return (self.nxg, self.nyg, self.nz)

    @property
    def shape_c(self):
        return (self.nxg, self.nyg, self.nzg)

    # ---- state ----
    def set_state(self, U, V, W):
        U, V, W = (np.asfortranarray(a, dtype=np.float64) for a in (U, V, W))
        assert U.shape == self.shape_u and V.shape == self.shape_v and W.shape == self.shape_w
        _lib.check(self.lib.cf_set_state(self.h, _lib.dptr(U), _lib.dptr(V), _lib.dptr(W)), "cf_set_state")

    def get_state(self):
        U, V, W = _f(self.shape_u), _f(self.shape_v), _f(self.shape_w)
        _lib.check(self.lib.cf_get_state(self.h, _lib.dptr(U), _lib.dptr(V), _lib.dptr(W)), "cf_get_state")
        return U, V, W

    @property
    def P(self):
        P = _f(self.shape_c)
        _lib.check(self.lib.cf_get_pressure(self.h, _lib.dptr(P)), "cf_get_pressure")
        return P

    def compute_pressure(self):
        """pressure.f90 compute_pressure: the actual pressure of the current
        state (FP64, reference layout with ghosts)."""
        P = _f(self.shape_c)
        _lib.check(self.lib.cf_compute_pressure(self.h, _lib.dptr(P)), "cf_compute_pressure")
        return P

    @property
    def nu_t(self):
        a = _f(self.shape_c)
        _lib.check(self.lib.cf_get_nu_t(self.h, _lib.dptr(a)), "cf_get_nu_t")
        return a

    @property
    def avg_nu_t(self):
        """avg_nu_t(1,nyg,1) of the last eddy-viscosity evaluation."""
        a = np.zeros(self.nyg)
        _lib.check(self.lib.cf_get_avg_nu_t(self.h, _lib.dptr(a)), "cf_get_avg_nu_t")
        return a

    def set_nu_t(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        _lib.check(self.lib.cf_set_nu_t(self.h, _lib.dptr(a)), "cf_set_nu_t")

    def wall_state(self):
        ax = _f((self.nx, 2, self.nzg))
        az = _f((self.nxg, 2, self.nz))
        ti = _f((self.nx, 2, self.nzg))
        _lib.check(self.lib.cf_get_wall(self.h, _lib.dptr(ax), _lib.dptr(az), _lib.dptr(ti)), "cf_get_wall")
        return ax, az, ti

    def set_tau_init(self, ti):
        ti = np.asfortranarray(ti, dtype=np.float64)
        _lib.check(self.lib.cf_set_tau_init(self.h, _lib.dptr(ti)), "cf_set_tau_init")

    @property
    def scalars(self):
        s = np.zeros(7)
        _lib.check(self.lib.cf_get_scalars(self.h, _lib.dptr(s)), "cf_get_scalars")
        return dict(zip(["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"], s))

    def set_scalars(self, **kw):
        cur = self.scalars
        cur.update(kw)
        s = np.array([cur[k] for k in ["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"]])
        _lib.check(self.lib.cf_set_scalars(self.h, _lib.dptr(s)), "cf_set_scalars")

    # ---- reference subroutines ----
    def apply_boundary_conditions(self):
        _lib.check(self.lib.cf_apply_boundary_conditions(self.h), "apply_boundary_conditions")

    def compute_projection_step(self, rk_step=0):
        _lib.check(self.lib.cf_projection_step(self.h, rk_step), "compute_projection_step")

    def compute_dt(self):
        dt = ctypes.c_double()
        _lib.check(self.lib.cf_compute_dt(self.h, ctypes.byref(dt)), "compute_dt")
        return dt.value

    def compute_time_step_RK3(self):
        _lib.check(self.lib.cf_rk3_step(self.h), "compute_time_step_RK3")

    def run(self, nsteps, graph=True):
        """main.f90:67-74 for nsteps (compute_dt + RK3 step), device resident."""
        _lib.check(self.lib.cf_run_steps(self.h, int(nsteps), 1 if graph else 0), "cf_run_steps")
        self.sync()

    def compute_rhs(self, U_, V_, W_):
        U_, V_, W_ = (np.asfortranarray(a, dtype=np.float64) for a in (U_, V_, W_))
        ru = _f((self.nx - 2, self.nyg - 2, self.nzg - 2))
        rv = _f((self.nxg - 2, self.ny - 2, self.nzg - 2))
        rw = _f((self.nxg - 2, self.nyg - 2, self.nz - 2))
        _lib.check(self.lib.cf_compute_rhs(self.h, _lib.dptr(U_), _lib.dptr(V_), _lib.dptr(W_),
                                           _lib.dptr(ru), _lib.dptr(rv), _lib.dptr(rw)), "compute_rhs")
        return ru, rv, rw

    def compute_rhs_u(self, U_, V_, W_):
        return self.compute_rhs(U_, V_, W_)[0]

    def compute_rhs_v(self, U_, V_, W_):
        return self.compute_rhs(U_, V_, W_)[1]

    def compute_rhs_w(self, U_, V_, W_):
        return self.compute_rhs(U_, V_, W_)[2]

    def compute_eddy_viscosity(self):
        _lib.check(self.lib.cf_compute_eddy_viscosity(self.h), "compute_eddy_viscosity")

    def compute_wall_model(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.cf_compute_wall_model(self.h, ctypes.byref(a), ctypes.byref(b)), "compute_wall_model")
        return a.value, b.value

    def set_wall_surrogate(self, model):
        """Use a trained WallMLP in the wall model (None detaches it); points
        outside its training hull are solved by the EQWM ODE."""
        _lib.check(self.lib.cf_set_wall_surrogate(self.h, model.h if model is not None else None),
                   "cf_set_wall_surrogate")
        self._surrogate = model  # keep the model alive while attached

    def wall_fallbacks(self):
        """Wall points that the ODE solved in the last wall-model call."""
        n = self.lib.cf_wall_fallbacks(self.h)
        if n < 0:
            _lib.check(n, "cf_wall_fallbacks")
        return n

    def compute_mean_mass_flow_U(self):
        q = ctypes.c_double()
        _lib.check(self.lib.cf_mean_mass_flow_u(self.h, ctypes.byref(q)), "compute_mean_mass_flow_U")
        return q.value

    def compute_mean_mass_flow_V(self):
        q = ctypes.c_double()
        _lib.check(self.lib.cf_mean_mass_flow_v(self.h, ctypes.byref(q)), "compute_mean_mass_flow_V")
        return q.value

    def compute_mean_centerline_U(self):
        q = ctypes.c_double()
        _lib.check(self.lib.cf_mean_centerline_u(self.h, ctypes.byref(q)), "compute_mean_centerline_U")
        return q.value

    def init_mass_flow(self):
        _lib.check(self.lib.cf_init_mass_flow(self.h), "init_mass_flow")

    def check_divergence(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.cf_check_divergence(self.h, ctypes.byref(a), ctypes.byref(b)), "check_divergence")
        return a.value, b.value

    def compute_statistics(self):
        """Plane means, u_tau, Re_tau and the bulk flows; raises when the
        state holds a NaN (statistics.f90:165-167)."""
        out = np.zeros(7 * self.ny + 4)
        _lib.check(self.lib.cf_statistics(self.h, _lib.dptr(out)), "compute_statistics")
        names = ["Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean"]
        d = {n: out[i * self.ny:(i + 1) * self.ny].copy() for i, n in enumerate(names)}
        t = out[7 * self.ny:]
        d.update(utau=float(t[0]), Retau=float(t[1]), Qflow_x=float(t[2]), Qflow_y=float(t[3]))
        return d

    def stats_accumulate(self):
        """Add the current plane statistics to the device window (async)."""
        _lib.check(self.lib.cf_stats_accumulate(self.h), "cf_stats_accumulate")

    def stats_reset(self):
        _lib.check(self.lib.cf_stats_reset(self.h), "cf_stats_reset")

    def stats_window(self):
        """Window averages of the plane statistics, plus the Reynolds stresses
        uu, vv, ww, uv about the window mean, and the sample count."""
        out = np.zeros(7 * self.ny)
        n = ctypes.c_double()
        _lib.check(self.lib.cf_stats_window(self.h, _lib.dptr(out), ctypes.byref(n)), "cf_stats_window")
        names = ["Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean"]
        d = {k: out[i * self.ny:(i + 1) * self.ny].copy() for i, k in enumerate(names)}
        d["uu"] = d["U2mean"] - d["Umean"] ** 2
        d["vv"] = d["V2mean"] - d["Vmean"] ** 2
        d["ww"] = d["W2mean"] - d["Wmean"] ** 2
        d["uv"] = d["UVmean"] - d["Umean"] * d["Vmean"]
        d["samples"] = int(n.value)
        return d

    # ---- instrumentation ----
    def timing(self, on=True):
        _lib.check(self.lib.cf_timing_enable(self.h, 1 if on else 0), "cf_timing_enable")

    def timing_reset(self):
        _lib.check(self.lib.cf_timing_reset(self.h), "cf_timing_reset")

    def timing_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(self.lib.cf_timing_report(self.h, buf, len(buf)), "cf_timing_report")
        out = {}
        for line in buf.value.decode().strip().splitlines():
            name, ms, cnt = line.split()
            out[name] = (float(ms), int(cnt))
        return out

    def device_bytes(self):
        return int(self.lib.cf_device_bytes(self.h))

    def sync(self):
        _lib.check(self.lib.cf_sync(self.h), "cf_sync")

    # ---- convenience: the initialisation sequence used for every case ----
    def initialise(self, U, V, W):
        """Load a raw field, make it periodic/no-slip, project it once and
        set Qflow_x_0 from it (initialization.f90:156-160, 460)."""
        self.set_state(U, V, W)
        self.apply_boundary_conditions()
        self.compute_projection_step(rk_step=0)
        self.apply_boundary_conditions()
        self.init_mass_flow()


def eqwm_batch(umag, nu, h, tau_in, ncv=100, stretch=1.10, tol=1e-7, max_iter=100):
    """solve_u_eqwm (eqwm.f90:90) for many wall points on the GPU."""
    lib = _lib.load()
    umag, nu, h, tau_in = (np.ascontiguousarray(np.broadcast_to(a, np.shape(umag)), dtype=np.float64)
                           for a in (umag, nu, h, tau_in))
    n = umag.size
    tau = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    _lib.check(lib.cf_eqwm_batch(n, _lib.dptr(umag), _lib.dptr(nu), _lib.dptr(h), _lib.dptr(tau_in),
                                 int(ncv), float(stretch), float(tol), int(max_iter),
                                 _lib.dptr(tau), _lib.iptr(it)), "cf_eqwm_batch")
    return tau, it


class ChannelFlowBatch:
    """Cases of one size class stepped together (cf_batch_*): every option
    but nu, the forcing and the fallback stress shared, one grid.  Each step
    kernel is launched once for all members; ``members[m]`` is a ChannelFlow
    for everything but the step (initialise, get_state, statistics, ...)."""

    def __init__(self, cfgs, x=None, y=None, z=None, device=0):
        self.lib = _lib.load()
        cfgs = list(cfgs)
        if x is None:
            x, y, z = cfgs[0].grid()
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        n = len(cfgs)
        arr = (_lib.CfParams * n)()
        for m, c in enumerate(cfgs):
            arr[m] = case_params(c, x, y, z, device)
        h = ctypes.c_void_p()
        _lib.check(self.lib.cf_batch_create(n, arr, _lib.dptr(x), _lib.dptr(y), _lib.dptr(z), ctypes.byref(h)),
                   "cf_batch_create")
        self.h = h
        self.members = []
        for m, c in enumerate(cfgs):
            mh = ctypes.c_void_p()
            _lib.check(self.lib.cf_batch_member(self.h, m, ctypes.byref(mh)), "cf_batch_member")
            self.members.append(ChannelFlow(c, x, y, z, device, _handle=mh))

    def run(self, nsteps, graph=True):
        """nsteps RK3 steps of every member (main.f90:70-73 per case)."""
        _lib.check(self.lib.cf_batch_run_steps(self.h, int(nsteps), 1 if graph else 0), "cf_batch_run_steps")
        if self.members:
            self.members[0].sync()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            for m in self.members:
                m.h = ctypes.c_void_p()
            self.lib.cf_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
