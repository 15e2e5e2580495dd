"""Serving loop on the GPU: each case gives the same bits alone (the
single-case path, its own context and stream) and in a continuously batched
mix; admission respects an imposed HBM budget (cases queue for memory);
cancelling one case leaves the others bit-identical."""
import numpy as np
import pytest

from channel_flow import cases
from channel_flow.serving import Server, ensemble_case, footprint

pytestmark = pytest.mark.gpu

SMALL = [((32, 24, 32), 300.0, False, 40, 0.00),
         ((48, 32, 48), 900.0, True, 30, 0.00),
         ((64, 32, 64), 500.0, False, 25, 0.01),
         ((32, 24, 32), 1500.0, True, 50, 0.02),
         ((48, 32, 48), 200.0, False, 35, 0.03),
         ((64, 48, 64), 1200.0, True, 20, 0.05)]


def _workload():
    return [(ensemble_case(g, rt, wm, name="s%d" % q), n, t) for q, (g, rt, wm, n, t) in enumerate(SMALL)]


@pytest.fixture(scope="module")
def alone():
    """Reference results: every case alone through the single-case API."""
    from channel_flow.solver import ChannelFlow
    out = []
    for cfg, n, _ in _workload():
        g = ChannelFlow(cfg)
        g.initialise(*cases.initial_field(cfg))
        g.run(n, graph=True)
        U, V, W = g.get_state()
        out.append((U, V, W, g.scalars, g.device_bytes()))
        g.close()
    return out


def _check_same(sv, ids, alone, skip=()):
    for q, cid in enumerate(ids):
        if q in skip:
            continue
        try:
            U, V, W, sc = sv.result(cid)
        except RuntimeError as e:
            raise AssertionError("case %d: %s %s" % (q, sv.status(cid), e))
        Ua, Va, Wa, sca, _ = alone[q]
        assert np.all(np.isfinite(U))
        for a, b in ((U, Ua), (V, Va), (W, Wa)):
            assert np.array_equal(a, b), q
        for k in ("t", "dt", "dPdx", "Qflow_x_0"):
            assert sc[k] == sca[k], (q, k, sc[k], sca[k])


def test_same_alone_and_in_mix(alone):
    wl = _workload()
    with Server(8 << 30, max_active=4, depth=3) as sv:
        ids = [sv.submit(cfg, n, arrival=t) for cfg, n, t in wl]
        sv.start()
        assert sv.wait(120.0)
        _check_same(sv, ids, alone)
        s = sv.summary()
        infos = [sv.info(c) for c in ids]
    assert s["peak_running"] >= 2  # cases really overlapped
    for (cfg, n, _), inf, a in zip(wl, infos, alone):
        assert inf["steps_done"] == n and inf["footprint"] >= a[4]
        assert inf["step_ms_p50"] > 0


def test_admission_by_hbm_budget(alone):
    """A budget that holds the largest case and little else: cases queue for
    memory, the bytes in use never exceed the budget, results unchanged."""
    wl = _workload()
    fps = [footprint(cfg) for cfg, _, _ in wl]
    budget = int(max(fps) * 1.3)
    with Server(budget, max_active=4, depth=2) as sv:
        ids = [sv.submit(cfg, n, arrival=0.0) for cfg, n, _ in wl]
        sv.start()
        assert sv.wait(120.0)
        _check_same(sv, ids, alone)
        s = sv.summary()
        infos = [sv.info(c) for c in ids]
    assert s["waited_for_memory"] >= 1
    assert s["peak_bytes"] <= budget
    # replay the admission intervals: resident footprints never exceed the budget
    ev = sorted([(i["admitted"], fp) for i, fp in zip(infos, fps)] +
                [(i["finished"], -fp) for i, fp in zip(infos, fps)], key=lambda e: (e[0], e[1]))
    used = 0
    for _, d in ev:
        used += d
        assert used <= budget


def test_cancel_leaves_others_identical(alone):
    wl = _workload()
    victim = 1
    wl[victim] = (wl[victim][0], 100000, 0.0)  # long enough to be cancelled mid-run
    with Server(8 << 30, max_active=4, depth=3) as sv:
        ids = [sv.submit(cfg, n, arrival=t) for cfg, n, t in wl]
        sv.start()
        for _ in range(2000):
            st, done = sv.status(ids[victim])
            if st == "running" and done >= 5:
                break
            sv.wait(0.005)
        sv.cancel(ids[victim])
        assert sv.wait(120.0)
        assert sv.status(ids[victim])[0] == "cancelled"
        _check_same(sv, ids, alone, skip=(victim,))
        with pytest.raises(RuntimeError):
            sv.result(ids[victim])


def test_dynamic_model_cases_in_the_mix():
    """Cases with the dynamic Smagorinsky model (its work arrays carved from
    the serving arena) beside Smagorinsky and DNS cases: each equals the
    same case run alone."""
    from channel_flow.solver import ChannelFlow
    ex = cases.case_example()
    ex.name = "example"
    dsm = ensemble_case((48, 32, 48), 900.0, True, name="dsm")
    dsm.les_model = 2
    wl = [(ex, 12, 0.0), (dsm, 10, 0.0), _workload()[0]]
    alone = []
    for cfg, n, _ in wl:
        g = ChannelFlow(cfg)
        g.initialise(*cases.initial_field(cfg))
        g.run(n, graph=True)
        alone.append(g.get_state())
        g.close()
    with Server(4 << 30, max_active=3, depth=2) as sv:
        ids = [sv.submit(cfg, n, arrival=t) for cfg, n, t in wl]
        sv.start()
        assert sv.wait(120.0)
        for q, cid in enumerate(ids):
            U, V, W, _ = sv.result(cid)
            for a, b in zip((U, V, W), alone[q]):
                assert np.array_equal(a, b), q
