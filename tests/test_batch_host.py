"""Batch creation checks that run before any device is touched: the members
of a batch share every option but nu, dPdx, dPdz, u_centerline and
tau_mean_eqwm, and the batched step takes neither the dynamic Smagorinsky
model, nor y mass flow, nor the learned wall model."""
import pytest

from channel_flow.serving import ensemble_case


def _make(cfgs):
    from channel_flow.solver import ChannelFlowBatch
    return ChannelFlowBatch(cfgs)


@pytest.mark.parametrize("option", ["les_model", "y_mass_cte", "wall_surrogate"])
def test_unbatched_options_are_refused(option):
    c = ensemble_case((64, 64, 64), 1000.0, True)
    setattr(c, option, 2 if option == "les_model" else 1)
    with pytest.raises(RuntimeError, match="batched step takes no"):
        _make([c, c])


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
