"""The ten closed-form cases of validate_analytical.py on the HIP solver, at the reference's thresholds.

Same set-ups, step counts and error metrics as the reference (restated in
tests/analytical_cases.py, which the CPU oracle also passes); thresholds from
validate_analytical.py:1174-1223.
"""
import pytest

import analytical_cases as ac

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(ac.CASES))
def test_case(gpu, name):
    from hip_backend import HipBackend
    print(f"[analytical] {name}", flush=True)
    value = ac.CASES[name](HipBackend())
    print(f"[analytical] {name}: {value:.4f} (threshold {ac.THRESHOLDS[name]})", flush=True)
    assert ac.passes(name, value), f"{name}: {value:.4f} vs threshold {ac.THRESHOLDS[name]}"
