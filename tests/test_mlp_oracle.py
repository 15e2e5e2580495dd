"""Pins for the MLP oracle (oracle/mlp_oracle.py), CPU only.

- bf16 rounding against torch's own float32 -> bfloat16 cast (an independent
  round-to-nearest-even implementation) and hand-made ties;
- the analytic gradient (float64 mode) against central finite differences of
  the loss, so a dropped tanh' factor, a transposed weight or a missing 1/B
  fails;
- the bf16 mode against the float64 mode: close (bf16 operand error only) but
  not identical (the rounding is really applied);
- standardisation/hull identities."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_ties_and_known_values():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # spacing of bfloat16 at 1
    vals = np.array([1.0, 1.0 + ulp / 2, 1.0 + 1.5 * ulp, 1.0 + ulp / 2 + 2 ** -20, -3.0, 0.0], dtype=np.float32)
    got = mo.bf16(vals)
    # ties go to even mantissa: 1+ulp/2 -> 1, 1+1.5ulp -> 1+2ulp; above the tie rounds up
    assert list(got) == [1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -3.0, 0.0]
    assert mo.bf16(one) == 1.0


def test_bf16_matches_torch_cast():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(mo.bf16(x), ref)


def _problem(B=24, seed=3):
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal(mo.NPARAMS) * 0.3
    X = rng.standard_normal((B, 4))
    Y = rng.standard_normal(B)
    return theta, X, Y


def test_gradient_matches_finite_differences():
    theta, X, Y = _problem()
    _, g = mo.loss_and_grad(theta, X, Y, bf16_ops=False)
    B = X.shape[0]
    rng = np.random.default_rng(11)
    # every bias and a sample of every weight block
    offs, o = [], 0
    for name, shp in mo.LAYOUT:
        n = int(np.prod(shp))
        offs += list(o + rng.choice(n, size=min(n, 12), replace=False))
        o += n
    h = 1e-6
    for q in offs:
        tp, tm = theta.copy(), theta.copy()
        tp[q] += h
        tm[q] -= h
        lp, _ = mo.loss_and_grad(tp, X, Y, bf16_ops=False)
        lm, _ = mo.loss_and_grad(tm, X, Y, bf16_ops=False)
        fd = (lp - lm) / (2 * h) / B
        assert abs(fd - g[q]) <= 1e-7 + 1e-6 * abs(fd), (q, fd, g[q])


def test_bf16_mode_close_to_fp64_but_not_equal():
    theta, X, Y = _problem(B=64)
    l64, g64 = mo.loss_and_grad(theta, X, Y, bf16_ops=False)
    l16, g16 = mo.loss_and_grad(theta, X, Y, bf16_ops=True)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < rel < 3e-2
    assert abs(l16 - l64) / l64 < 3e-2


def test_standardise_and_hull():
    rng = np.random.default_rng(2)
    n = 500
    feat = np.stack([rng.uniform(0.5, 2, n), rng.uniform(0.05, 0.2, n), rng.uniform(1e-5, 1e-4, n),
                     rng.uniform(-2e-3, 2e-3, n)], axis=1)
    tau = rng.uniform(1e-3, 1e-2, n)
    nrm, hull = mo.fit_norm(feat, tau)
    X = mo.standardise(feat, nrm)
    assert np.allclose(X.mean(axis=0), 0, atol=1e-5) and np.allclose(X.std(axis=0), 1, atol=1e-4)
    assert np.allclose(mo.target(tau, nrm).std(), 1, atol=1e-4)
    assert mo.inside_hull(feat, nrm, hull).all()
    out = feat.copy()
    out[:, 0] *= 100.0
    assert not mo.inside_hull(out, nrm, hull).any()
