"""The serving ensemble on its own grid set: 64^3, 96x64x96, 128x96x128 and
192x128x192, DNS (FP64, tanh grid) and wall-modelled LES (mixed precision)
in equal numbers, Re_tau spread over 180..2000, seeded Poisson arrivals.
An HBM budget of 1.5 times the largest case makes cases queue for memory.
Every case gives the same bits in the mix as alone through the single-case
path, with each case on its own stream and with cases of one size class
stepped together in groups.  Step counts are cut to 8..24 per case: the identity does not depend
on the count, and the throughput of the full counts is measured elsewhere."""
import numpy as np
import pytest

from channel_flow import cases
from channel_flow.serving import GRIDS, Server, ensemble_case, footprint

pytestmark = pytest.mark.gpu


def _workload():
    rng = np.random.default_rng(11)
    wl, t = [], 0.0
    for q, g in enumerate(GRIDS):
        for wm in (False, True):
            t += rng.exponential(0.01)
            rt = float(rng.uniform(180.0, 2000.0))
            n = int(rng.integers(8, 25))
            wl.append((ensemble_case(g, rt, wm, name="c5_%d%s" % (q, "w" if wm else "d")), n, t))
    return wl


@pytest.fixture(scope="module")
def alone():
    from channel_flow.solver import ChannelFlow
    out = []
    for cfg, n, _ in _workload():
        print("alone: %s %s %s, %d steps" % (cfg.name, cfg.cells, cfg.precision, n), flush=True)
        g = ChannelFlow(cfg)
        g.initialise(*cases.initial_field(cfg))
        g.run(n, graph=True)
        out.append((*g.get_state(), g.scalars))
        g.close()
    return out


@pytest.mark.parametrize("batch", [0, 4])
def test_c5_grid_set_same_alone_and_in_mix_under_budget(alone, batch):
    wl = _workload()
    assert sorted({c.cells for c, _, _ in wl}) == sorted(GRIDS)
    assert sum(c.precision == "mixed" for c, _, _ in wl) == len(wl) // 2
    fps = [footprint(cfg) for cfg, _, _ in wl]
    budget = int(max(fps) * 1.5)
    assert budget < sum(fps)
    print("c5 mix: %d cases, batch %d, budget %.2f GB, footprints %s" % (len(wl), batch, budget / 1e9, fps),
          flush=True)
    with Server(budget, max_active=4, depth=3, batch=batch, chunk=5) as sv:
        ids = [sv.submit(cfg, n, arrival=t, field=cases.initial_field(cfg)) for cfg, n, t in wl]
        sv.start()
        assert sv.wait(300.0)
        s = sv.summary()
        infos = [sv.info(c) for c in ids]
        for q, cid in enumerate(ids):
            U, V, W, sc = sv.result(cid)
            Ua, Va, Wa, sca = alone[q]
            for a, b in ((U, Ua), (V, Va), (W, Wa)):
                assert np.array_equal(a, b), wl[q][0].name
            for k in ("t", "dt", "dPdx", "Qflow_x_0"):
                assert sc[k] == sca[k], (wl[q][0].name, k)
    assert s["waited_for_memory"] >= 1
    assert s["peak_bytes"] <= budget
    assert s["peak_running"] >= 2
    for (cfg, n, _), inf in zip(wl, infos):
        assert inf["steps_done"] == n and inf["state"] == "done"
