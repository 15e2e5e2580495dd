"""Host package of the MI355X channel-flow time-step library."""
from .cases import (CaseConfig, make_grid, centres, initial_field, counter_uniform,
                    case_example, case_dns64, case_dns384, case_wmles128, case_wmles256)

__all__ = ["CaseConfig", "make_grid", "centres", "initial_field", "counter_uniform",
           "case_example", "case_dns64", "case_dns384", "case_wmles128", "case_wmles256"]


def solver():
    """Import the GPU solver lazily (it loads libcfhip.so)."""
    from .solver import ChannelFlow, eqwm_batch
    return ChannelFlow, eqwm_batch
