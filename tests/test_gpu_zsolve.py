"""The z transforms fused with the y solve (k_zsolve_fwd / k_zsolve_bwd over
kx slabs, projection.f90:84-113) against the unfused chain (z transform,
lane-form y solve, inverse z transform).  Three forms: one workgroup per slab
eliminating from the bottom wall, which repeats the lane solve's operations
in the same order, so in FP64 the state after several steps is
bit-identical at the fixed-plan z lengths (within 1e-14 with run-time plans,
where the compiler may contract the same operations differently in the two
kernels); two workgroups per slab eliminating from both walls and
joining in the middle (the FP32 default); and both halves in one workgroup,
one thread group each (the FP64 default).  The two-ended forms round
differently and agree to 1e-12 in FP64.  In the mixed path, where the
eliminated values are stored in FP32, the state stays within FP32 round-off
of the unfused chain.  The grids cover the z lengths compiled with fixed
plans (384, 256, 192, 128) and run-time plans (96; 300 = 4*3*25 and the
largest accepted length, 512, on 512 threads), y row counts that leave the
last batches part-filled (down to 5 rows, halves of 3 and 2, and the
smallest accepted, 4), the actual-pressure solve, the oracle at 1e-10,
and the batched step."""
import numpy as np
import pytest

from channel_flow import cases
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _cfg(cells, **kw):
    cfg = cases.case_dns64()
    cfg.cells = cells
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


FORMS = {"one_half": {"CF_ZS_HALVES": "1", "CF_ZS_G": "2"},
         "two_halves": {"CF_ZS_HALVES": "2", "CF_ZS_G": "1"},
         "twin_halves": {"CF_ZS_HALVES": "3", "CF_ZS_G": "2"}}


def _run(cfg, monkeypatch, fused, nsteps=6, graph=True, form=None):
    from channel_flow.solver import ChannelFlow
    monkeypatch.setenv("CF_ZSOLVE", "1" if fused else "0")
    monkeypatch.setenv("CF_TH_LANE", "1")
    for k, v in (FORMS[form] if form else {}).items():
        monkeypatch.setenv(k, v)
    g = ChannelFlow(cfg)
    g.initialise(*cases.initial_field(cfg))
    g.run(nsteps, graph=graph)
    st, sc, P = g.get_state(), g.scalars, g.compute_pressure()
    g.close()
    return st, sc, P


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("form", ["one_half", "two_halves", "twin_halves"])
@pytest.mark.parametrize("cells", [(64, 20, 384), (48, 33, 256), (40, 16, 192), (32, 24, 128),
                                   (32, 17, 96), (24, 12, 300), (32, 5, 128), (32, 4, 128), (16, 8, 512)])
def test_fused_matches_the_unfused_chain_fp64(cells, form, monkeypatch):
    print("cells %s fp64 %s" % (cells, form), flush=True)
    cfg = _cfg(cells)
    (a, sa, pa) = _run(cfg, monkeypatch, False)
    (b, sb, pb) = _run(cfg, monkeypatch, True, form=form)
    if form == "one_half" and cells[2] in (384, 256, 192, 128):
        # fixed-plan lengths: the transforms and sweeps of both chains round alike
        for x, y in zip(a, b):
            assert np.array_equal(x, y), relerr(y, x)
        assert sa["dt"] == sb["dt"] and sa["dPdx"] == sb["dPdx"]
        assert np.array_equal(pa, pb), relerr(pb, pa)
    elif form == "one_half":
        # run-time plans: the same operations compiled into two kernels may
        # contract multiply-adds differently (a length-512 case differs by
        # one ulp)
        for x, y in zip(a, b):
            assert relerr(y, x) < 1e-14, relerr(y, x)
        assert relerr(pb, pa) < 1e-12, relerr(pb, pa)
    else:
        for x, y in zip(a, b):
            assert relerr(y, x) < 1e-12, relerr(y, x)
        assert abs(sa["dt"] - sb["dt"]) <= 1e-12 * sa["dt"]
        assert relerr(pb, pa) < 1e-10, relerr(pb, pa)


@pytest.mark.parametrize("form", ["one_half", "two_halves", "twin_halves"])
@pytest.mark.parametrize("cells", [(64, 20, 384), (32, 17, 96)])
def test_fused_mixed_tracks_unfused(cells, form, monkeypatch):
    print("cells %s mixed %s" % (cells, form), flush=True)
    cfg = _cfg(cells, precision="mixed")
    (a, _, pa) = _run(cfg, monkeypatch, False, nsteps=10)
    (b, _, pb) = _run(cfg, monkeypatch, True, nsteps=10, form=form)
    for x, y in zip(a, b):
        assert relerr(y, x) < 1e-5, relerr(y, x)
    assert relerr(pb, pa) < 1e-4, relerr(pb, pa)


@pytest.mark.parametrize("form", ["one_half", "two_halves", "twin_halves"])
def test_fused_against_oracle(form, monkeypatch):
    """Ten steps with the fused solve at a grid the oracle steps in seconds
    (z length 384 on its fixed plan, 20 rows)."""
    from channel_flow.solver import ChannelFlow
    monkeypatch.setenv("CF_ZSOLVE", "1")
    for k, v in FORMS[form].items():
        monkeypatch.setenv(k, v)
    cfg = _cfg((64, 20, 384))
    g = ChannelFlow(cfg)
    o = Oracle(cfg, jref=g.jref_eqwm)
    U, V, W = cases.initial_field(cfg)
    g.initialise(U, V, W)
    o.initialise(U, V, W)
    g.run(10, graph=True)
    o.run(10)
    for a, b in zip(g.get_state(), o.get_state()):
        assert relerr(a, b) < 1e-10, relerr(a, b)
    _, dmax = g.check_divergence()
    assert dmax < 1e-10
    g.close()


@pytest.mark.parametrize("form", ["one_half", "two_halves", "twin_halves"])
def test_fused_in_the_batched_step(form, monkeypatch):
    """Three cases of one size class through the batched step with the fused
    solve (the checkpoint scratch at the case stride): each member gives the
    bits it gives alone."""
    from channel_flow.serving import ensemble_case
    from channel_flow.solver import ChannelFlow, ChannelFlowBatch
    monkeypatch.setenv("CF_ZSOLVE", "1")
    for k, v in FORMS[form].items():
        monkeypatch.setenv(k, v)
    cfgs = []
    for q, rt in enumerate((250.0, 700.0, 1500.0)):
        c = ensemble_case((32, 20, 128), rt, False, name="z%d" % q)
        c.seed = 40 + q
        cfgs.append(c)
    b = ChannelFlowBatch(cfgs)
    for m, c in zip(b.members, cfgs):
        m.initialise(*cases.initial_field(c))
    b.run(8, graph=True)
    got = [m.get_state() for m in b.members]
    b.close()
    for q, c in enumerate(cfgs):
        g = ChannelFlow(c)
        g.initialise(*cases.initial_field(c))
        g.run(8, graph=True)
        for x, y in zip(got[q], g.get_state()):
            assert np.array_equal(x, y), q
        g.close()
