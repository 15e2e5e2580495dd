"""Mixed precision in a turbulent channel: the 64^3 DNS workload (Re_b 2800,
tanh 2.0 grid, dt 2e-3, constant mass flow), started from a seeded
large-scale perturbation (cases.large_scale_field, rms 0.3 U_b), which is
turbulent after 40,000 steps (t = 80; the grid-scale noise of the default
initial field decays back to laminar flow at this Reynolds number).

Turbulence is checked before anything is compared: Re_tau within 10% of
Dean's correlation (178.8) at every sample, u_tau steady over the samples,
and a Reynolds shear stress of order u_tau^2 (max |uv| between 0.5 and 1
u_tau^2; 0 in laminar flow).

Two trajectories of a chaotic flow separate from round-off within some
10,000 steps, after which FP32 and FP64 windows differ by sampling noise
alone.  The comparison at the 1% bound is therefore made where precision is
what it measures: the FP64 turbulent state is handed to the mixed path and
both are averaged over the same window of 4,000 steps (t = 8, several
eddy turnovers near the wall, 400 samples).  The long-window comparison
(independent trajectories, with a second FP64 run from another seed as the
measure of sampling noise) is the second test, run when CF_LONG_TESTS=1."""
import os

import numpy as np
import pytest

from channel_flow import cases

pytestmark = pytest.mark.gpu

TRANSITION = 40000
RETAU_DEAN = cases.retau_from_reb(2800.0)
QUANTITIES = ("Umean", "uu", "vv", "ww", "uv")


def _case(precision, seed=2018):
    c = cases.case_dns64()
    c.precision = precision
    c.seed = seed
    return c


def _turbulent(precision="fp64", seed=2018):
    from channel_flow.solver import ChannelFlow
    c = _case(precision, seed)
    g = ChannelFlow(c)
    g.initialise(*cases.large_scale_field(c, 0.3))
    g.run(TRANSITION, graph=True)
    return g


def _check_turbulent(samples, nu):
    retau = np.array([s["Retau"] for s in samples])
    utau = np.array([s["utau"] for s in samples])
    print("Re_tau samples %s" % np.round(retau, 2), flush=True)
    assert np.all(np.abs(retau / RETAU_DEAN - 1.0) < 0.1), retau
    assert np.allclose(retau, utau / nu, rtol=1e-12)
    assert np.std(utau) / np.mean(utau) < 0.03
    for s in samples:
        uv = s["UVmean"] - s["Umean"] * s["Vmean"]
        r = np.max(np.abs(uv)) / s["utau"] ** 2
        assert 0.5 < r < 1.0, r


def _window(gs, steps, every=10):
    for g in gs:
        g.stats_reset()
    for _ in range(steps // every):
        for g in gs:
            g.run(every, graph=True)
            g.stats_accumulate()
    return [g.stats_window() for g in gs]


def _errs(w, ref):
    return {q: float(np.max(np.abs(w[q] - ref[q])) / np.max(np.abs(ref[q]))) for q in QUANTITIES}


def test_mixed_window_statistics_in_turbulent_flow_within_one_percent():
    from channel_flow.solver import ChannelFlow
    g64 = _turbulent()
    samples = []
    for _ in range(4):
        g64.run(2000, graph=True)
        samples.append(g64.compute_statistics())
    _check_turbulent(samples, g64.cfg.nu)
    U, V, W = g64.get_state()
    g32 = ChannelFlow(_case("mixed"))
    g32.set_state(U, V, W)
    g32.set_scalars(**g64.scalars)
    w64, w32 = _window([g64, g32], 4000)
    assert w64["samples"] == w32["samples"] == 400
    errs = _errs(w32, w64)
    print("[mixed vs fp64, shared turbulent start, 4000 steps] %s" % {k: "%.2e" % v for k, v in errs.items()},
          flush=True)
    for q, e in errs.items():
        assert e < 1e-2, (q, e)
    # the window was turbulent in both runs
    _check_turbulent([g64.compute_statistics(), g32.compute_statistics()], g64.cfg.nu)
    _, div = g32.check_divergence()
    assert div < 1e-5


@pytest.mark.skipif(os.environ.get("CF_LONG_TESTS") != "1", reason="long run (about 3 minutes); CF_LONG_TESTS=1")
def test_mixed_long_window_within_sampling_noise():
    """Independent trajectories from the same seed (FP64, mixed) and from
    another seed (FP64): over a window of 160,000 steps (t = 320) the mixed
    path differs from FP64 by less than a second FP64 run does, i.e. by
    sampling noise; the mean profile within 2%."""
    gs = [_turbulent("fp64"), _turbulent("mixed"), _turbulent("fp64", seed=7)]
    w64, w32, w64b = _window(gs, 160000)
    e32, enoise = _errs(w32, w64), _errs(w64b, w64)
    print("[long window] mixed %s | fp64 other seed %s" % ({k: "%.2e" % v for k, v in e32.items()},
                                                          {k: "%.2e" % v for k, v in enoise.items()}), flush=True)
    assert e32["Umean"] < 2e-2
    for q in QUANTITIES:
        assert e32[q] < max(1.5 * enoise[q], 1e-2), (q, e32[q], enoise[q])
