"""Closed-form pins of the oracle's right-hand sides and plane statistics
on a stretched grid (the checks live in pins_closed_form.py, so that
test_oracle_planted.py can run them against builds with planted mistakes)."""
import pytest

import pins_closed_form as pins


@pytest.mark.parametrize("check", pins.ALL, ids=lambda f: f.__name__)
def test_closed_form(check):
    check()
