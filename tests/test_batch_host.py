"""Batch creation checks that run before any device is touched: the members
of a batch share every option but nu, dPdx, dPdz, u_centerline and
tau_mean_eqwm, and the batched step takes FP64 DNS cases only."""
import pytest

from channel_flow.serving import ensemble_case


def _make(cfgs):
    from channel_flow.solver import ChannelFlowBatch
    return ChannelFlowBatch(cfgs)


def test_mixed_precision_and_wall_model_classes_are_refused():
    wm = ensemble_case((64, 64, 64), 1000.0, True)
    with pytest.raises(RuntimeError, match="FP64 DNS"):
        _make([wm, wm])


def test_members_differing_in_a_shared_option_are_refused():
    a = ensemble_case((64, 64, 64), 500.0, False)
    b = ensemble_case((64, 64, 64), 900.0, False)
    b.cfl = 0.4
    with pytest.raises(RuntimeError, match="differ"):
        _make([a, b])
    c = ensemble_case((64, 64, 64), 900.0, False)
    c.x_mass_cte = 2
    with pytest.raises(RuntimeError, match="differ"):
        _make([a, c])
