"""Serving host logic without a GPU: the exact footprint that admission
uses (a dry allocation pass of the context) and the seeded ensemble."""
import numpy as np

from channel_flow import cases
from channel_flow.serving import ensemble_workload, footprint, GRIDS


def _field_bytes(cfg, esize):
    NX, NY, NZ = cfg.cells
    nx, ny, nz = NX + 2, NY + 1, NZ + 2
    return esize * (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz)


def test_footprint_covers_the_state_and_scales():
    c = cases.case_dns64()
    fp = footprint(c)
    # two ping-pong copies of (U, V, W) plus the RK3 accumulators: at least 3x one state
    assert fp >= 3 * _field_bytes(c, 8)
    assert fp % 256 == 0
    big = cases.case_dns64()
    big.cells = (128, 128, 128)
    assert 7.0 < footprint(big) / fp < 9.0  # 8x the cells
    m = cases.case_dns64()
    m.precision = "mixed"
    assert footprint(m) < 0.7 * fp


def test_footprint_of_large_cases():
    c = cases.case_dns384()
    fp = footprint(c)
    assert 3 * _field_bytes(c, 8) < fp < 20 * _field_bytes(c, 8) / 3
    assert fp < 6e9


def test_ensemble_is_seeded_and_mixed():
    a = ensemble_workload(n=64, seed=5)
    b = ensemble_workload(n=64, seed=5)
    assert [(x.cells, x.nu, x.iwall_model, n, t) for x, n, t in a] == \
           [(x.cells, x.nu, x.iwall_model, n, t) for x, n, t in b]
    arr = np.array([t for _, _, t in a])
    assert np.all(np.diff(arr) > 0)
    assert {x.cells for x, _, _ in a} == set(GRIDS)
    wm = sum(x.iwall_model == 3 for x, _, _ in a)
    assert 20 <= wm <= 44
    assert all(200 <= n <= 2000 for _, n, _ in a)
    s = ensemble_workload(n=8, seed=5, steps_scale=0.1)
    assert all(20 <= n <= 200 for _, n, _ in s)


def test_footprint_of_the_dynamic_model():
    """les_model 2 adds the model's work arrays: 10 centre fields, the
    filtered velocities and the plane partials."""
    c = cases.case_example()
    s = cases.case_example()
    s.les_model = 1
    NX, NY, NZ = c.cells
    ncent = (NX + 3) * (NY + 2) * (NZ + 3)
    assert footprint(c) - footprint(s) >= 8 * (10 * ncent) + _field_bytes(c, 8)


def test_invalid_parameters_are_rejected():
    """cf_create's checks, exercised through the dry footprint pass: each
    bad parameter gives an error naming it."""
    import pytest

    def err(**kw):
        c = cases.case_dns64()
        c.cells = (32, 24, 16)
        for k, v in kw.items():
            setattr(c, k, v)
        with pytest.raises(RuntimeError) as e:
            footprint(c)
        return str(e.value)

    assert "les_model" in err(les_model=3)
    assert "iwall_model" in err(iwall_model=2)
    assert "even" in err(cells=(35, 24, 16))
    assert "FFT" in err(cells=(30, 24, 2 * 11))   # z length 22: prime factor 11
    assert "FFT" in err(cells=(2 * 13, 24, 16))  # half x length 13
    for cells in ((40, 24, 28), (70, 16, 50), (30, 20, 42)):  # factors 5 and 7 are planned
        c = cases.case_dns64()
        c.cells = cells
        assert footprint(c) > 0
    assert "nu" in err(nu=0.0)
    assert "eqwm_ncv" in err(iwall_model=3, eqwm_ncv=500)
    c = cases.case_dns64()
    c.precision = "fp16"
    with pytest.raises(ValueError):
        footprint(c)
