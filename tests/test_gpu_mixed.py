"""Mixed-precision path (FP32 stencil/FFT storage with U held as an FP32
deviation from an FP64 plane profile; FP64 Thomas, reductions, statistics)
against the FP64 path on the same seeded inputs.  Tolerances: per-step field
error 1e-4 relative, max |div u| 1e-5, mean profile 1e-4."""
import numpy as np
import pytest

from channel_flow import cases

pytestmark = pytest.mark.gpu


def solver():
    from channel_flow.solver import ChannelFlow
    return ChannelFlow


def cfg_of(prec, **kw):
    c = cases.case_dns64()
    c.cells = kw.pop("cells", (64, 48, 32))
    c.precision = prec
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nsteps", [1, 10])
def test_mixed_tracks_fp64(nsteps):
    CF = solver()
    g64 = CF(cfg_of("fp64"))
    g32 = CF(cfg_of("mixed"))
    U, V, W = cases.initial_field(g64.cfg)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    g64.run(nsteps, graph=True)
    g32.run(nsteps, graph=True)
    a, b = g32.get_state(), g64.get_state()
    assert relerr(a[0], b[0]) < 1e-4
    # V error measured against the velocity scale (max |U|)
    assert np.max(np.abs(a[1] - b[1])) < 1e-4 * np.max(np.abs(b[0]))
    _, div = g32.check_divergence()
    assert div < 1e-5
    s32, s64 = g32.compute_statistics(), g64.compute_statistics()
    assert relerr(s32["Umean"], s64["Umean"]) < 1e-4
    # bulk flow held at its initial value by the FP64 correction; the bulk is
    # taken from U* before the last projection, which changes it only by the
    # FP32 rounding of the projected values (about 1e-12 per step)
    assert abs(g32.compute_mean_mass_flow_U() - g32.scalars["Qflow_x_0"]) < 1e-10


def test_mixed_lane_solve_and_fixed_plans_track_fp64(monkeypatch):
    """The mixed path with the kernels of the 384x256x384 workload (half
    length 192, z length 64 fixed-plan transforms, lane y solve forced and
    the z transforms as launches of their own): ten steps within the mixed
    tolerances of the FP64 path."""
    monkeypatch.setenv("CF_TH_LANE", "1")
    monkeypatch.setenv("CF_ZSOLVE", "0")
    CF = solver()
    g64 = CF(cfg_of("fp64", cells=(384, 24, 64)))
    g32 = CF(cfg_of("mixed", cells=(384, 24, 64)))
    U, V, W = cases.initial_field(g64.cfg)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    g64.run(10, graph=True)
    g32.run(10, graph=True)
    a, b = g32.get_state(), g64.get_state()
    assert relerr(a[0], b[0]) < 1e-4
    assert np.max(np.abs(a[1] - b[1])) < 1e-4 * np.max(np.abs(b[0]))
    _, div = g32.check_divergence()
    assert div < 1e-5


def test_mixed_actual_pressure_tracks_fp64():
    """compute_pressure in a mixed context (FP32 stencils and spectrum, FP64
    y solve) against the FP64 context on the same state: 1e-3 of max |P|
    (the right-hand sides are differences of FP32 velocities)."""
    CF = solver()
    g64 = CF(cfg_of("fp64", dPdx=2e-3))
    g32 = CF(cfg_of("mixed", dPdx=2e-3))
    U, V, W = cases.initial_field(g64.cfg)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    g64.run(2, graph=True)
    g32.run(2, graph=True)
    p64, p32 = g64.compute_pressure(), g32.compute_pressure()
    assert np.all(np.isfinite(p32))
    assert relerr(p32, p64) < 1e-3, relerr(p32, p64)


def test_mixed_deviation_storage_roundtrip():
    """set_state/get_state through the FP32 deviation storage lose less than
    FP32 rounding of the deviation."""
    CF = solver()
    g = CF(cfg_of("mixed"))
    U, V, W = cases.initial_field(g.cfg)
    g.set_state(U, V, W)
    U2, V2, W2 = g.get_state()
    assert np.max(np.abs(U2 - U)) < 4e-8 * 0.3 * 4
    assert np.max(np.abs(V2 - V)) < 1e-8


def test_mixed_wall_modelled_les_runs():
    CF = solver()
    kw = dict(les_model=1, iwall_model=3, stretch=0.0, eqwm_ncv=30, eqwm_tol=1e-8, nu=1e-4, cfl=-1e-3,
              cells=(48, 32, 32))
    g64, g32 = CF(cfg_of("fp64", **kw)), CF(cfg_of("mixed", **kw))
    U, V, W = cases.initial_field(g64.cfg)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    g64.run(5, graph=True)
    g32.run(5, graph=True)
    a, b = g32.get_state(), g64.get_state()
    assert relerr(a[0], b[0]) < 1e-4
    ax32, _, _ = g32.wall_state()
    ax64, _, _ = g64.wall_state()
    assert relerr(ax32, ax64) < 1e-3


def test_mixed_dynamic_smagorinsky_tracks_fp64():
    """The dynamic model computes in FP64 from the FP32 state: after ten
    steps of the reference's example case the mixed fields and nu_t stay
    within the mixed path's per-step tolerances."""
    CF = solver()
    c64, c32 = cases.case_example(), cases.case_example()
    c32.precision = "mixed"
    g64, g32 = CF(c64), CF(c32)
    U, V, W = cases.initial_field(c64)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    g64.run(10, graph=True)
    g32.run(10, graph=True)
    for a, b in zip(g32.get_state(), g64.get_state()):
        assert relerr(a, b) < 1e-4
    assert relerr(g32.nu_t, g64.nu_t) < 1e-2
    _, dmax = g32.check_divergence()
    assert dmax < 1e-5


def _window(g, steps, start, every):
    g.run(start, graph=True)
    g.stats_reset()
    for _ in range((steps - start) // every):
        g.run(every, graph=True)
        g.stats_accumulate()
    return g.stats_window()


@pytest.mark.parametrize("case,steps,start", [("dns64", 1200, 200), ("wmles128", 800, 200)])
def test_window_statistics_within_one_percent(case, steps, start):
    """Mean profile and Reynolds stresses averaged over a window of steps:
    the mixed path within 1% of the FP64 path (relative to each quantity's
    maximum over y), both accumulated in FP64 on the device."""
    CF = solver()
    if case == "dns64":
        mk = lambda p: cfg_of(p, cells=(64, 64, 64))
    else:
        def mk(p):
            c = cases.case_wmles128()
            c.precision = p
            return c
    g64, g32 = CF(mk("fp64")), CF(mk("mixed"))
    U, V, W = cases.initial_field(g64.cfg)
    g64.initialise(U, V, W)
    g32.initialise(U, V, W)
    w64 = _window(g64, steps, start, 5)
    w32 = _window(g32, steps, start, 5)
    assert w64["samples"] == w32["samples"] == (steps - start) // 5
    errs = {}
    for k in ("Umean", "uu", "vv", "ww", "uv"):
        errs[k] = float(np.max(np.abs(w32[k] - w64[k])) / np.max(np.abs(w64[k])))
    print("\n[mixed window %s] %s" % (case, {k: "%.2e" % v for k, v in errs.items()}))
    for k, e in errs.items():
        assert e < 1e-2, (k, e)
    _, div = g32.check_divergence()
    assert div < 1e-5
