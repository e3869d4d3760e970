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


@pytest.mark.parametrize("name", list(ac.CASES))
def test_case_fp32_error_against_fp64(gpu, name):
    """The same case on the FP64 build of the same kernels: both pass, and the FP32 result's error against
    FP64 is stated (printed) and small against the case's threshold."""
    from hip_backend import HipBackend
    v32 = ac.CASES[name](HipBackend(32))
    v64 = ac.CASES[name](HipBackend(64))
    print(f"[analytical fp64] {name}: fp32 {v32:.6f}, fp64 {v64:.6f}, |fp32 - fp64| {abs(v32 - v64):.2e} "
          f"(threshold {ac.THRESHOLDS[name]})", flush=True)
    assert ac.passes(name, v64), f"{name} (FP64): {v64:.4f} vs threshold {ac.THRESHOLDS[name]}"
    assert abs(v32 - v64) <= 0.05 * abs(ac.THRESHOLDS[name]), (v32, v64)
