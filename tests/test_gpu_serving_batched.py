"""The serving loop with batching on: cases of one size class (the same
options but nu and the forcing, the same grid) are stepped together by the
batched step in groups on one stream each, in stretches of a few steps;
cases that arrive while a group runs join it between stretches, and a
finished member leaves without stopping the others.  Every case gives the
same bits as alone through the single-case path.  The workload mixes the
ensemble's DNS (FP64) and WMLES (mixed) classes on two grids with step
counts 6..17 and seeded arrivals, so groups form, fill, drain and refill;
one case is cancelled while its group runs, and a cell limit keeps the
larger class out of the groups."""
import numpy as np
import pytest

from channel_flow import cases
from channel_flow.serving import Server, ensemble_case

pytestmark = pytest.mark.gpu


def _workload():
    rng = np.random.default_rng(21)
    wl, t = [], 0.0
    for q in range(14):
        t += rng.exponential(0.0015)
        g = (64, 64, 64) if q % 3 else (96, 64, 96)
        wm = bool(q % 2)
        c = ensemble_case(g, float(rng.uniform(180.0, 2000.0)), wm, name="s%02d" % q)
        c.seed = 300 + q
        wl.append((c, int(rng.integers(6, 18)), t))
    return wl


@pytest.fixture(scope="module")
def alone():
    from channel_flow.solver import ChannelFlow
    out = []
    for cfg, n, _ in _workload():
        g = ChannelFlow(cfg)
        g.initialise(*cases.initial_field(cfg))
        g.run(n, graph=True)
        out.append((*g.get_state(), g.scalars))
        g.close()
    return out


def test_batched_serving_same_bits_as_alone(alone):
    wl = _workload()
    with Server(int(24e9), max_active=4, depth=3, batch=4, chunk=4) as sv:
        ids = [sv.submit(cfg, n, arrival=t, field=cases.initial_field(cfg)) for cfg, n, t in wl]
        sv.start()
        assert sv.wait(300.0)
        s = sv.summary()
        infos = [sv.info(c) for c in ids]
        res = [sv.result(c) for c in ids]
    print("batched serving: %s" % s, flush=True)
    for q, (U, V, W, sc) in enumerate(res):
        Ua, Va, Wa, sca = alone[q]
        for a, b in ((U, Ua), (V, Va), (W, Wa)):
            assert np.array_equal(a, b), wl[q][0].name
        for k in ("t", "dt", "dPdx", "Qflow_x_0"):
            assert sc[k] == sca[k], (wl[q][0].name, k)
    for (cfg, n, _), inf in zip(wl, infos):
        assert inf["steps_done"] == n and inf["state"] == "done"
    assert s["groups"] >= 4          # four classes at least once each
    assert s["peak_members"] >= 2    # members stepped together
    assert s["peak_running"] >= 3


def test_cancel_inside_a_group():
    wl = _workload()[:6]
    with Server(int(24e9), max_active=4, depth=3, batch=4, chunk=3) as sv:
        ids = [sv.submit(cfg, 40, arrival=0.0, field=cases.initial_field(cfg)) for cfg, _, _ in wl]
        sv.start()
        sv.cancel(ids[1])
        assert sv.wait(300.0)
        states = [sv.status(c)[0] for c in ids]
    assert states[1] == "cancelled"
    assert all(st == "done" for q, st in enumerate(states) if q != 1)


def test_batching_limited_to_small_classes(alone):
    """With a cell limit below the 96x64x96 grid, only the 64^3 classes form
    groups; the 96x64x96 cases run one per stream.  Every case still gives
    the bits it gives alone."""
    wl = _workload()
    n64 = sum(1 for c, _, _ in wl if c.cells == (64, 64, 64))
    with Server(int(24e9), max_active=6, depth=3, batch=4, chunk=4, batch_max_cells=64 * 64 * 64) as sv:
        ids = [sv.submit(cfg, n, arrival=t, field=cases.initial_field(cfg)) for cfg, n, t in wl]
        sv.start()
        assert sv.wait(300.0)
        s = sv.summary()
        res = [sv.result(c) for c in ids]
    print("limited batching: %s" % s, flush=True)
    for q, (U, V, W, sc) in enumerate(res):
        Ua, Va, Wa, sca = alone[q]
        for a, b in ((U, Ua), (V, Va), (W, Wa)):
            assert np.array_equal(a, b), wl[q][0].name
        assert sc["t"] == sca["t"] and sc["dt"] == sca["dt"], wl[q][0].name
    assert 2 <= s["groups"] <= n64  # groups of the two 64^3 classes only
    assert s["peak_members"] >= 2
