"""Learned wall model on the GPU: BF16 matrix-core forward/backward kernels
against the CPU oracle (oracle/mlp_oracle.py), training on EQWM-generated
pairs, the hull fallback to the ODE inside the wall model, and the gain over
the ODE alone on the 256x96x256 wall-modelled case.

Tolerances: gradients and outputs use the same bf16 operand rounding as the
oracle, so they differ only by float32 accumulation order and tanhf:
relative L2 error per parameter block <= 2e-3, outputs <= 1e-3 relative."""
import os
import sys
import time

import numpy as np
import pytest

from channel_flow import cases

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _synthetic(n, seed):
    """Raw feature rows in the training ranges and a smooth target."""
    rng = np.random.default_rng(seed)
    feat = np.stack([rng.uniform(0.3, 1.5, n), rng.uniform(0.03, 0.3, n), 10 ** rng.uniform(-5.5, -4.2, n),
                     rng.uniform(-1e-3, 3e-3, n)], axis=1)
    tau = 0.4 * feat[:, 2] * feat[:, 0] / feat[:, 1] + (0.04 * feat[:, 0]) ** 2
    return feat, tau


def _blocks(v):
    out, o = {}, 0
    for name, shp in mo.LAYOUT:
        n = int(np.prod(shp))
        out[name] = v[o:o + n]
        o += n
    return out


@pytest.mark.parametrize("B", [1, 16, 200, 1000])
def test_gradients_match_oracle(B):
    from channel_flow.wall_mlp import WallMLP
    m = WallMLP(seed=7)
    feat, tau = _synthetic(max(B, 64), seed=B)
    nrm, hull = mo.fit_norm(feat, tau)
    m.set_norm(nrm, hull)
    feat, tau = feat[:B], tau[:B]
    g, loss = m.gradients(feat, tau)
    theta = m.params.astype(np.float64)
    X = mo.standardise(feat, nrm)
    Y = mo.target(tau, nrm)
    l_o, g_o = mo.loss_and_grad(theta, X, Y, bf16_ops=True)
    assert abs(loss - l_o) <= 1e-4 * l_o + 1e-7
    gb, ob = _blocks(g.astype(np.float64)), _blocks(g_o)
    for name in gb:
        err = np.linalg.norm(gb[name] - ob[name]) / max(np.linalg.norm(ob[name]), 1e-30)
        assert err <= 2e-3, (name, err)
    # and the bf16 gradient stays near the exact float64 one
    _, g64 = mo.loss_and_grad(theta, X, Y, bf16_ops=False)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 3e-2


def test_forward_matches_oracle_and_hull_flags():
    from channel_flow.wall_mlp import WallMLP
    m = WallMLP(seed=3)
    feat, tau = _synthetic(3000, seed=9)
    nrm, hull = mo.fit_norm(feat, tau)
    m.set_norm(nrm, hull)
    probe = feat.copy()
    probe[::7, 0] *= 50.0   # u far above the training box
    probe[::11, 2] *= 1e-3  # nu far below
    t_gpu, inside = m.predict(probe)
    t_orc = mo.predict_tau(m.params.astype(np.float64), probe, nrm, bf16_ops=True)
    assert np.array_equal(inside, mo.inside_hull(probe, nrm, hull))
    assert not inside[::7].any() and not inside[::11].any()
    assert np.max(np.abs(t_gpu - t_orc) / t_orc) < 1e-3


@pytest.fixture(scope="module")
def trained():
    """The 256x96x256 LES training recipe: 2M EQWM samples over Re_tau 1e3-1e4 (seeded),
    ncv = 30, stretch 1.10, tol 1e-8; Adam, batch 4096."""
    from channel_flow.wall_mlp import WallMLP, eqwm_samples
    t0 = time.time()
    feat, tau = eqwm_samples(2_000_000, seed=2024)
    t1 = time.time()
    m = WallMLP(seed=1)
    hist = m.train(feat, tau, epochs=12, batch=4096, lr=3e-3, lr_final=1e-4)
    t2 = time.time()
    print("\n[mlp] %d samples in %.2f s, 12 epochs in %.2f s, loss %s" %
          (len(tau), t1 - t0, t2 - t1, np.array2string(hist, precision=6)))
    return m, hist


def test_training_fits_eqwm(trained):
    from channel_flow.wall_mlp import eqwm_samples
    m, hist = trained
    # standardised target: an untrained network scores about 1
    assert np.all(np.isfinite(hist)) and hist[-1] < hist[0] and hist[-1] < 1e-3
    feat, tau = eqwm_samples(50_000, seed=77)  # held out
    pred, inside = m.predict(feat)
    assert inside.mean() > 0.99
    rel = np.abs(pred[inside] - tau[inside]) / tau[inside]
    print("\n[mlp] held-out |tau err|: median %.3e p99 %.3e max %.3e" %
          (np.median(rel), np.percentile(rel, 99), rel.max()))
    assert np.median(rel) < 1e-2 and np.percentile(rel, 99) < 5e-2


def _loglaw_state(g, cfg, seed=5):
    """Turbulent-like state for the wall model: a log-law mean profile with
    +-25% seeded noise, spanwise noise of 10% u_tau (not projected)."""
    nu = cfg.nu
    ut = np.sqrt(cfg.dPdx)
    yd = np.minimum(g.yg, 2.0 - g.yg)
    yp = np.maximum(np.abs(yd) * ut / nu, 1.0)
    prof = ut * (np.log(yp) / 0.41 + 5.2)
    r = cases.counter_uniform(seed, 1, g.shape_u)
    U = np.asfortranarray(prof[None, :, None] * (1.0 + 0.5 * (r - 0.5)))
    V = np.zeros(g.shape_v, order="F")
    W = np.asfortranarray(0.2 * ut * (cases.counter_uniform(seed, 3, g.shape_w) - 0.5))
    return U, V, W


def _wall_ctx(cfg):
    from channel_flow.solver import ChannelFlow
    g = ChannelFlow(cfg)
    U, V, W = _loglaw_state(g, cfg)
    g.set_state(U, V, W)
    g.apply_boundary_conditions()
    g.compute_eddy_viscosity()
    return g


def test_hull_fallback_is_exact_ode(trained):
    """All points out of the hull: the wall model with the surrogate attached
    gives exactly the ODE result, and counts every point as a fallback."""
    from channel_flow.wall_mlp import WallMLP
    m, _ = trained
    cfg = cases.case_wmles128()
    cfg.cells = (64, 32, 64)
    cfg.precision = "fp64"
    g = _wall_ctx(cfg)
    ti0 = g.wall_state()[2].copy()
    g.compute_wall_model()
    ref = [a.copy() for a in g.wall_state()]
    empty = WallMLP(seed=4)
    nrm, _ = m.norm
    empty.params = m.params
    empty.set_norm(nrm, np.array([1e30] * 4 + [-1e30] * 4, dtype=np.float32))
    g.set_tau_init(ti0)
    g.set_wall_surrogate(empty)
    g.compute_wall_model()
    got = g.wall_state()
    npts = g.nx * 2 * g.nzg
    assert g.wall_fallbacks() == npts
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    # the trained model: points in its hull take the network's tau
    g.set_tau_init(ti0)
    g.set_wall_surrogate(m)
    g.compute_wall_model()
    nf = g.wall_fallbacks()
    tau_m = g.wall_state()[2]
    rel = np.abs(tau_m - ref[2]) / ref[2]
    print("\n[mlp] wmles128 wall: %d of %d points fall back; |tau err| median %.3e max %.3e" %
          (nf, npts, np.median(rel), rel.max()))
    assert nf < 0.05 * npts
    assert np.median(rel) < 1e-2
    g.set_wall_surrogate(None)
    g.set_tau_init(ti0)
    g.compute_wall_model()
    assert np.array_equal(g.wall_state()[2], ref[2])


def test_gain_on_c4_wall(trained):
    """What the network gains over the ODE on the 256x96x256 case: wall-model
    time per call (same inputs, same kernel launch path) and the stress error."""
    m, _ = trained
    cfg = cases.case_wmles256()
    g = _wall_ctx(cfg)
    ti0 = g.wall_state()[2].copy()
    reps = 20

    def timed(model):
        g.set_wall_surrogate(model)
        g.set_tau_init(ti0)
        g.compute_wall_model()  # warm-up
        g.timing(True)
        g.timing_reset()
        for _ in range(reps):
            g.set_tau_init(ti0)
            g.compute_wall_model()
        g.sync()
        rep = g.timing_report()
        g.timing(False)
        ms = sum(v[0] * v[1] for k, v in rep.items() if k in ("wall_model", "mlp_wall")) / reps
        return ms, g.wall_state()[2].copy(), g.wall_fallbacks() if model is not None else None

    t_ode, tau_ode, _ = timed(None)
    t_mlp, tau_mlp, nf = timed(m)
    rel = np.abs(tau_mlp - tau_ode) / tau_ode
    print("\n[mlp] wmles256 wall (%d points): ODE %.4f ms, MLP %.4f ms (x%.2f), fallbacks %d, "
          "|tau err| median %.3e p99 %.3e" % (tau_ode.size, t_ode, t_mlp, t_ode / t_mlp, nf,
                                              np.median(rel), np.percentile(rel, 99)))
    assert t_mlp < t_ode
    assert np.median(rel) < 1e-2 and np.percentile(rel, 99) < 5e-2
    g.set_wall_surrogate(None)


def test_wmles_runs_with_surrogate_track_eqwm(trained):
    """The 256x96x256 wall-modelled LES run with the learned wall model
    against the same run with the ODE, from the same initial field: mean
    velocity profile over a window of steps within 1%, and the whole RK3 step
    faster with the network."""
    from channel_flow.solver import ChannelFlow
    m, _ = trained
    cfg = cases.case_wmles256()
    steps, start, every = 300, 100, 5
    out = {}
    for name, model in (("ode", None), ("mlp", m)):
        g = ChannelFlow(cfg)
        U, V, W = _loglaw_state(g, cfg)
        g.initialise(U, V, W)
        if model is not None:
            g.set_wall_surrogate(model)
        g.run(start, graph=True)
        g.stats_reset()
        g.sync()
        t0 = time.perf_counter()
        for _ in range((steps - start) // every):
            g.run(every, graph=True)
            g.stats_accumulate()
        g.sync()
        ms = (time.perf_counter() - t0) * 1e3 / (steps - start)
        w = g.stats_window()
        assert np.all(np.isfinite(w["Umean"]))
        out[name] = (w, ms, g.wall_fallbacks() if model is not None else 0)
        g.set_wall_surrogate(None)
        g.close()
    (wo, mso, _), (wm, msm, nf) = out["ode"], out["mlp"]
    err = float(np.max(np.abs(wm["Umean"] - wo["Umean"])) / np.max(np.abs(wo["Umean"])))
    err_uu = float(np.max(np.abs(wm["uu"] - wo["uu"])) / np.max(np.abs(wo["uu"])))
    print("\n[mlp] wmles256 run, steps %d-%d: mean-profile err %.3e, uu err %.3e; step %.3f ms (ODE) "
          "vs %.3f ms (MLP), last-step fallbacks %d" % (start + 1, steps, err, err_uu, mso, msm, nf))
    assert err < 1e-2
    assert msm < mso  # the surrogate makes the whole step faster
