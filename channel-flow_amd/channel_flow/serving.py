"""Host side of the serving loop (``cf_server_*`` in cf_serve.cpp): submit
channel cases of different sizes, Reynolds numbers and models to one GPU,
let the native scheduler batch them continuously, and read results and
latencies back.  The reference has no serving layer (one case is one run of
main.f90); a case here runs exactly the time step of the single-case path.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .cases import CaseConfig, initial_field, reb_from_retau

STATES = {0: "pending", 1: "preparing", 2: "running", 3: "done", 4: "cancelled", 5: "failed"}


def case_params(cfg: CaseConfig, device=0):
    """cf_params and grids of a case (as ChannelFlow builds them)."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in cfg.grid())
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
    return p, x, y, z


def footprint(cfg: CaseConfig):
    """Device bytes a case takes from the serving arena (no device needed)."""
    lib = _lib.load()
    p, x, y, z = case_params(cfg)
    n = lib.cf_footprint(ctypes.byref(p), _lib.dptr(x), _lib.dptr(y), _lib.dptr(z))
    if n < 0:
        _lib.check(-1, "cf_footprint")
    return int(n)


class Server:
    """A serving loop on one GPU.

    hbm_budget: device bytes reserved for all running cases; max_active:
    cases running at once (one HIP stream each); depth: steps in flight per
    case.  batch > 0: cases of one size class are stepped together, up to
    ``batch`` per group, in stretches of at most ``chunk`` steps (one stream
    per group); when ``batch_max_cells`` is nonzero, only cases of at most
    that many unique cells."""

    def __init__(self, hbm_budget, max_active=8, depth=3, device=0, batch=0, chunk=25, batch_max_cells=0):
        self.lib = _lib.load()
        self.device = device
        h = ctypes.c_void_p()
        _lib.check(self.lib.cf_server_create(device, int(hbm_budget), int(max_active), int(depth),
                                             ctypes.byref(h)), "cf_server_create")
        self.h = h
        self.cases = {}
        if batch:
            _lib.check(self.lib.cf_server_set_batching(self.h, int(batch), int(chunk), int(batch_max_cells)),
                       "cf_server_set_batching")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.cf_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, cfg: CaseConfig, nsteps, arrival=0.0, field=None):
        """Queue a case; ``field`` = raw (U, V, W) or the seeded initial field."""
        p, x, y, z = case_params(cfg, self.device)
        U, V, W = field if field is not None else initial_field(cfg)
        U, V, W = (np.asfortranarray(a, dtype=np.float64) for a in (U, V, W))
        cid = ctypes.c_longlong()
        _lib.check(self.lib.cf_server_submit(self.h, ctypes.byref(p), _lib.dptr(x), _lib.dptr(y), _lib.dptr(z),
                                             _lib.dptr(U), _lib.dptr(V), _lib.dptr(W), int(nsteps),
                                             float(arrival), ctypes.byref(cid)), "cf_server_submit")
        self.cases[cid.value] = (cfg, (len(x), len(y), len(z)))
        return cid.value

    def start(self):
        _lib.check(self.lib.cf_server_start(self.h), "cf_server_start")

    def cancel(self, cid):
        _lib.check(self.lib.cf_server_cancel(self.h, cid), "cf_server_cancel")

    def wait(self, timeout=0.0):
        rc = self.lib.cf_server_wait(self.h, float(timeout))
        if rc < 0:
            _lib.check(rc, "cf_server_wait")
        return rc == 0

    def status(self, cid):
        st, n = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.cf_server_status(self.h, cid, ctypes.byref(st), ctypes.byref(n)), "cf_server_status")
        return STATES[st.value], n.value

    def result(self, cid):
        cfg, (nx, ny, nz) = self.cases[cid]
        U = np.zeros((nx, ny + 1, nz + 1), order="F")
        V = np.zeros((nx + 1, ny, nz + 1), order="F")
        W = np.zeros((nx + 1, ny + 1, nz), order="F")
        sc = np.zeros(7)
        _lib.check(self.lib.cf_server_result(self.h, cid, _lib.dptr(U), _lib.dptr(V), _lib.dptr(W),
                                             _lib.dptr(sc)), "cf_server_result")
        names = ["t", "dt", "dPdx", "dPdz", "Qflow_x_0", "dPdy", "Qflow_y_0"]
        return U, V, W, dict(zip(names, sc))

    def info(self, cid):
        v = np.zeros(14)
        _lib.check(self.lib.cf_server_case_info(self.h, cid, _lib.dptr(v), 14), "cf_server_case_info")
        keys = ["arrival", "admitted", "ready", "finished", "steps_done", "steps", "footprint",
                "step_ms_p50", "step_ms_p99", "step_ms_max", "state", "cells", "precision", "iwall_model"]
        d = dict(zip(keys, v))
        d["state"] = STATES[int(d["state"])]
        return d

    def step_times(self, cid):
        n = int(self.info(cid)["steps_done"])
        ms = np.zeros(max(n, 1), dtype=np.float32)
        m = self.lib.cf_server_step_times(self.h, cid, _lib.fptr(ms), n)
        if m < 0:
            _lib.check(m, "cf_server_step_times")
        return ms[:m].astype(np.float64)

    def summary(self):
        v = np.zeros(8)
        _lib.check(self.lib.cf_server_summary(self.h, _lib.dptr(v), 8), "cf_server_summary")
        return dict(zip(["makespan_s", "peak_bytes", "peak_running", "waited_for_memory", "arena_bytes",
                         "cases", "groups", "peak_members"], v))


# ---------------------------------------------------------------- workloads
GRIDS = ((64, 64, 64), (96, 64, 96), (128, 96, 128), (192, 128, 192))


def ensemble_case(grid, retau, wmles, name="case"):
    """One case of the ensemble: DNS (FP64, tanh 2.0 grid) or wall-modelled
    LES (mixed precision, uniform y, Smagorinsky + EQWM at h = 0.1), bulk
    flow from Dean's correlation, CFL 0.5."""
    nu = 1.0 / reb_from_retau(retau)
    if wmles:
        return CaseConfig(name=name, cells=tuple(grid), stretch=0.0, nu=nu, cfl=0.5, dPdx=(retau * nu) ** 2,
                          x_mass_cte=1, les_model=1, iwall_model=3, cs=0.1, h_wm=0.1, eqwm_ncv=30,
                          eqwm_stretch=1.10, eqwm_tol=1e-8, eqwm_max_iter=100, precision="mixed")
    return CaseConfig(name=name, cells=tuple(grid), stretch=2.0, nu=nu, cfl=0.5, dPdx=(retau * nu) ** 2,
                      x_mass_cte=1, precision="fp64")


def ensemble_workload(n=64, seed=5, steps=(200, 2000), retau=(180.0, 2000.0), mean_gap=0.05,
                      steps_scale=1.0):
    """Seeded Poisson arrivals of n cases: grid uniform over GRIDS, Re_tau
    uniform in ``retau``, DNS or WMLES with equal odds, steps uniform in
    ``steps`` (times steps_scale).  Returns [(cfg, nsteps, arrival)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n):
        t += rng.exponential(mean_gap)
        g = GRIDS[rng.integers(len(GRIDS))]
        rt = rng.uniform(*retau)
        wm = bool(rng.integers(2))
        ns = max(1, int(round(rng.integers(steps[0], steps[1] + 1) * steps_scale)))
        out.append((ensemble_case(g, rt, wm, name="c%02d" % q), ns, t))
    return out


class _FieldCache:
    """The seeded initial field depends only on the grid class and model."""

    def __init__(self):
        self.d = {}

    def __call__(self, cfg):
        key = (cfg.cells, cfg.stretch, cfg.seed, cfg.amp)
        if key not in self.d:
            self.d[key] = initial_field(cfg)
        return self.d[key]


def serve(workload, hbm_budget, max_active, depth=3, timeout=1200.0, fields=None, batch=0, chunk=25,
          batch_max_cells=0):
    """Run a workload through one server; per-case records and the summary.
    batch > 0 steps the cases of one size class together (up to ``batch``
    members per group, stretches of at most ``chunk`` steps, cases of at most
    ``batch_max_cells`` unique cells when nonzero)."""
    fields = fields or _FieldCache()
    with Server(hbm_budget, max_active=max_active, depth=depth, batch=batch, chunk=chunk,
                batch_max_cells=batch_max_cells) as sv:
        ids = [sv.submit(cfg, ns, arrival=t, field=fields(cfg)) for cfg, ns, t in workload]
        sv.start()
        if not sv.wait(timeout):
            raise TimeoutError("serving run did not finish in %.0f s" % timeout)
        recs = []
        for cid in ids:
            inf = sv.info(cid)
            inf["step_ms"] = sv.step_times(cid)
            recs.append(inf)
        summ = sv.summary()
    return recs, summ


def latency_report(recs, summ):
    """p50/p99 of per-step and per-case latency, and aggregate throughput."""
    steps = np.concatenate([r["step_ms"] for r in recs]) if recs else np.zeros(0)
    case_lat = np.array([r["finished"] - r["arrival"] for r in recs])
    cell_steps = sum(r["cells"] * r["steps_done"] for r in recs)
    first = min(r["arrival"] for r in recs)
    span = summ["makespan_s"] - first
    return {
        "cases": len(recs),
        "done": sum(1 for r in recs if r["state"] == "done"),
        "step_ms_p50": float(np.percentile(steps, 50)) if steps.size else math.nan,
        "step_ms_p99": float(np.percentile(steps, 99)) if steps.size else math.nan,
        "case_s_p50": float(np.percentile(case_lat, 50)),
        "case_s_p99": float(np.percentile(case_lat, 99)),
        "makespan_s": float(span),
        "mcell_per_s": float(cell_steps / span / 1e6) if span > 0 else math.nan,
        "peak_running": int(summ["peak_running"]),
        "peak_gb": float(summ["peak_bytes"] / 1e9),
        "waited_for_memory": int(summ["waited_for_memory"]),
        "groups": int(summ["groups"]),
        "peak_members": int(summ["peak_members"]),
    }
