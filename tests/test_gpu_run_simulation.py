"""The drop-in run_simulation against the oracle's restatement of src/simulation.py:14-162.

Same mask, parameters and stop rule on both sides: the stop step must agree
and the returned FP32 fields must be identical (the passes are bit-exact;
only the order of the stat sums differs, which cannot move a stop decision
that is not within rounding of the tolerance).
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def small_design(seed=0):
    from xadvcool_amd import geometry as G
    rng = np.random.default_rng(seed)
    core = (24, 18, 12)
    mask = G.tpms_mask(3, period=float(rng.uniform(8, 12)), threshold=float(rng.uniform(-0.3, 0.1)), core=core,
                       buf=4)
    heat = G.heat_source_3d(mask.shape, "quad_core", core=core[:2], buf=4)
    return mask, heat


@pytest.mark.parametrize("pump", ["constant_flow", "constant_power"])
def test_run_simulation_matches_oracle(gpu, pump):
    print(f"[run_simulation] pump={pump}", flush=True)
    from xadvcool_amd.simulation import run_simulation
    mask, heat = small_design()
    kw = dict(max_steps=2600, tau_fluid=0.55, heat_power=0.1, t_inlet=25.0, u_inlet_val=0.03, stats_every=100,
              tol_ke=1e-5, tol_T=1e-5, pump_mode=pump, target_power=0.05, silent=True, heat_source=heat, Pr=7.0,
              k_ratio=628.0, min_steps_rule=False)
    got = run_simulation(mask, device="cuda", **kw)
    ref = orc.run_simulation(mask, **kw)
    assert got["steps"] == ref["steps"], (got["steps"], ref["steps"])
    assert got["converged"] == ref["converged"]
    for k in ("rho", "u", "T"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("d_ke", "dT_conv", "dQ_conv"):
        assert abs(got[k] - ref[k]) <= 1e-9 + 1e-6 * abs(ref[k]), (k, got[k], ref[k])


def test_run_simulation_reference_defaults(gpu):
    """Default heat map (z = 0 plane), default floor, a short max_steps: the shape of the returned dict."""
    from xadvcool_amd.simulation import run_simulation
    mask, _ = small_design(1)
    got = run_simulation(mask, max_steps=300, silent=True)
    assert set(got) >= {"u", "rho", "T", "mask", "steps", "d_ke", "dT_conv", "dQ_conv", "converged"}
    assert got["u"].shape == mask.shape + (3,) and got["u"].dtype == np.float32
    assert got["rho"].shape == mask.shape and got["T"].shape == mask.shape
    ref = orc.run_simulation(mask, max_steps=300)
    assert got["steps"] == ref["steps"] and got["converged"] == ref["converged"]
    assert np.array_equal(got["T"], ref["T"]) and np.array_equal(got["u"], ref["u"])
