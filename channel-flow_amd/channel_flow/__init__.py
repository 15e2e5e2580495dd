"""Host package of the MI355X channel-flow time-step library."""
from .cases import (CaseConfig, make_grid, centres, initial_field, counter_uniform,
                    case_c1, case_c2, case_c3, case_c4)

__all__ = ["CaseConfig", "make_grid", "centres", "initial_field", "counter_uniform",
           "case_c1", "case_c2", "case_c3", "case_c4"]


def solver():
    """Import the GPU solver lazily (it loads libcfhip.so)."""
    from .solver import ChannelFlow, eqwm_batch
    return ChannelFlow, eqwm_batch
