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
    c = cases.case_c1()
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
    assert relerr(a[1], b[1]) < 1e-3 * 0 + 1e-4 * np.max(np.abs(b[0])) / np.max(np.abs(b[1]))
    _, div = g32.check_divergence()
    assert div < 1e-5
    s32, s64 = g32.compute_statistics(), g64.compute_statistics()
    assert relerr(s32["Umean"], s64["Umean"]) < 1e-4
    # bulk flow held at its initial value to FP64 accuracy by the FP64 reduction
    assert abs(g32.compute_mean_mass_flow_U() - g32.scalars["Qflow_x_0"]) < 1e-12


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
