"""Batched step of one size class (cf_batch_*): each kernel of the RK3 step is
launched once for all members, whose buffers sit at one stride in one
allocation.  Every member computes the same bits as the same case alone
(state, ghosts included, and scalars) after ten steps, for the serving
ensemble's 64^3 DNS class (FP64, Re_tau spread, seeds differ, CFL time
steps, bulk mass flow), its wall-modelled LES class (mixed precision,
Smagorinsky, equilibrium wall model with each case's own nu and fallback
stress, deviation profile), and a small ragged class (partitioned y solve,
fixed dt, centreline mass flow) stepped without graphs.  A member is stepped
only through its batch."""
import numpy as np
import pytest

from channel_flow import cases
from channel_flow.serving import ensemble_case

pytestmark = pytest.mark.gpu


def _alone(cfg, nsteps, graph):
    from channel_flow.solver import ChannelFlow
    g = ChannelFlow(cfg)
    g.initialise(*cases.initial_field(cfg))
    g.run(nsteps, graph=graph)
    out = g.get_state(), g.scalars
    g.close()
    return out


def _check(cfgs, nsteps, graph):
    from channel_flow.solver import ChannelFlowBatch
    b = ChannelFlowBatch(cfgs)
    for m, c in zip(b.members, cfgs):
        m.initialise(*cases.initial_field(c))
    b.run(nsteps, graph=graph)
    got = [(m.get_state(), m.scalars) for m in b.members]
    b.close()
    for q, c in enumerate(cfgs):
        (U, V, W), sc = got[q]
        (Ua, Va, Wa), sca = _alone(c, nsteps, graph)
        for x, y in ((U, Ua), (V, Va), (W, Wa)):
            assert np.array_equal(x, y), (q, np.max(np.abs(x - y)))
        for k in ("t", "dt", "dPdx", "Qflow_x_0"):
            assert sc[k] == sca[k], (q, k, sc[k], sca[k])
    return got


def test_batch_of_ensemble_dns_class_is_bitwise_each_case_alone():
    rng = np.random.default_rng(3)
    cfgs = []
    for q in range(4):
        c = ensemble_case((64, 64, 64), float(rng.uniform(180.0, 2000.0)), False, name="b%d" % q)
        c.seed = 100 + q
        cfgs.append(c)
    print("batch 64^3 x4: nu %s" % [round(c.nu, 8) for c in cfgs], flush=True)
    got = _check(cfgs, 10, True)
    # the members really differ (no cross-talk would hide behind equal inputs)
    assert not np.array_equal(got[0][0][0], got[1][0][0])


def test_batch_of_ensemble_wmles_class_is_bitwise_each_case_alone():
    rng = np.random.default_rng(4)
    cfgs = []
    for q in range(3):
        c = ensemble_case((64, 64, 64), float(rng.uniform(180.0, 2000.0)), True, name="w%d" % q)
        c.seed = 200 + q
        cfgs.append(c)
    print("batch 64^3 WMLES x3: nu %s" % [round(c.nu, 9) for c in cfgs], flush=True)
    _check(cfgs, 10, True)


def test_batch_of_ragged_class_without_graphs():
    cfgs = []
    for q, rt in enumerate((200.0, 550.0, 900.0)):
        c = ensemble_case((48, 32, 40), rt, False, name="r%d" % q)
        c.cfl = -2e-3
        c.x_mass_cte = 2
        c.u_centerline = 1.15
        c.seed = 7 + q
        cfgs.append(c)
    print("batch 48x32x40 x3, fixed dt, centreline flow", flush=True)
    _check(cfgs, 10, False)


def test_member_is_stepped_only_through_its_batch():
    from channel_flow.solver import ChannelFlowBatch
    cfgs = [ensemble_case((32, 24, 16), rt, False) for rt in (300.0, 400.0)]
    b = ChannelFlowBatch(cfgs)
    m = b.members[1]
    m.initialise(*cases.initial_field(cfgs[1]))
    with pytest.raises(RuntimeError, match="batch"):
        m.run(1)
    assert b.lib.cf_destroy(m.h) != 0
    b.close()
