"""The oracle's EQWM restatement against vectors of the reference's own
solve_u_eqwm (tests/golden/eqwm_ref.json, made by make_eqwm_golden.py from
the flang build of eqwm.f90)."""
import json
import os

import pytest

from oracle_lib import eqwm, ref_eqwm, ref_eqwm_available

HERE = os.path.dirname(os.path.abspath(__file__))
ROWS = json.load(open(os.path.join(HERE, "golden", "eqwm_ref.json")))["rows"]


@pytest.mark.parametrize("row", ROWS, ids=[str(i) for i in range(len(ROWS))])
def test_oracle_eqwm_matches_reference_vectors(row):
    tau, it = eqwm(row["u"], row["nu"], row["h"], row["tau_in"], ncv=row["ncv"], stretch=row["stretch"],
                   tol=1e-7, max_iter=100)
    assert it == row["iter"]
    assert tau == pytest.approx(row["tau"], rel=1e-13, abs=0)


@pytest.mark.skipif(not ref_eqwm_available(), reason="reference eqwm.f90 build absent")
def test_reference_build_reproduces_fixture():
    row = ROWS[0]
    tau, it = ref_eqwm(row["u"], row["nu"], row["h"], row["tau_in"], ncv=row["ncv"], stretch=row["stretch"])
    assert (tau, it) == (row["tau"], row["iter"])
