"""Each planted mistake of the oracle (ORC_PLANT sites in oracle/cf_oracle.c)
makes at least one closed-form pin fail.  The oracle is rebuilt on the CPU
with -DORC_PLANT=n into a temporary directory for every n; the normal build
is untouched."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import pins_closed_form as pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cf_oracle.c")

# planted mistake -> the pins expected to catch it
PLANTED = {
    1: ("rhs_u: nu_t interpolated to the lower y edge with the upper edge's weights",
        [pins.check_rhs_linear_profiles]),
    2: ("rhs_v: outer y difference over a face spacing instead of the centre spacing",
        [pins.check_rhs_v_stretched_viscous]),
    3: ("rhs_v: the two face spacings of dv/dy swapped", [pins.check_rhs_v_stretched_viscous]),
    4: ("rhs_u: -d(uv)/dy with the y weights swapped", [pins.check_rhs_linear_profiles]),
    5: ("rhs_w: -d(vw)/dy, lower edge weights of the wrong row", [pins.check_rhs_linear_profiles]),
    6: ("rhs_v: Coriolis sign", [pins.check_rhs_linear_profiles]),
    7: ("rhs_u: Coriolis V average one row off", [pins.check_rhs_linear_profiles]),
    8: ("rhs_v: -dv^2/dy over a face spacing", [pins.check_rhs_linear_profiles]),
    9: ("rhs_v: nu_t of the wrong centre below the face", [pins.check_rhs_linear_profiles]),
    10: ("statistics: U z range without the duplicated plane", [pins.check_statistics_ranges]),
    11: ("statistics: W2mean without the square", [pins.check_statistics_ranges]),
    12: ("statistics: wall Reynolds stresses subtracted the wrong way", [pins.check_statistics_utau_and_bulk]),
    13: ("statistics: UV pairs V(i) with the U average one column right", [pins.check_statistics_uv_pairing]),
    14: ("rhs_v: nu_t at the x edges with the weights of row j-1", [pins.check_rhs_v_nu_t_on_edges]),
    15: ("statistics: V x range without the duplicated column", [pins.check_statistics_ranges]),
    16: ("centreline row ny/2 instead of ny/2+1", [pins.check_centerline_mean, pins.check_mass_flow_corrections]),
    17: ("centreline mean without the periodic image plane k = 1", [pins.check_centerline_mean]),
    18: ("y mass-flow correction with the wrong sign", [pins.check_mass_flow_corrections]),
}


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")

    def build(n):
        out = str(d / ("liboracle_plant%d.so" % n))
        subprocess.check_call(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=gnu11", "-shared",
                               "-DORC_PLANT=%d" % n, "-o", out, SRC, "-lm"])
        return n, out

    with ThreadPoolExecutor(max_workers=8) as ex:
        return dict(ex.map(build, sorted(PLANTED)))


@pytest.mark.parametrize("n", sorted(PLANTED))
def test_planted_mistake_is_caught(n, variants):
    what, checks = PLANTED[n]
    caught = []
    for check in checks:
        try:
            check(path=variants[n])
        except AssertionError:
            caught.append(check.__name__)
    assert caught, "planted mistake %d (%s) passed every pin" % (n, what)
