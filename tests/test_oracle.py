"""Pins of the CPU oracle (oracle/lbm_oracle.cpp, oracle/oracle.py) — no GPU.

What fixes each oracle function (none of them can be checked against the
reference itself: Warp is not installed and the reference ships no stored
vectors, only closed-form tests and its published dataset):

* equilibrium, BGK, streaming with inlet / outlet / bounce-back, the six test
  streaming and collision kernels, thermal BGK + source, thermal streaming:
  the ten closed-form cases of validate_analytical.py at the reference's
  thresholds (test_closed_form_case);
* the same cases also catch plausible mistakes planted in the oracle — a
  dropped (1 - omega/2) source factor, bounce-back of the wrong direction, a
  sign flip of the moving-wall term (test_planted_mistakes_are_caught);
* exact invariants: the rest state is a fixed point, mass and heat are
  conserved in periodic and closed boxes to FP64 rounding
  (test_rest_state_fixed_point, test_mass_conservation_*);
* FP32 against FP64 on the production schedule (test_fp32_tracks_fp64);
* the stats slots against direct numpy sums (test_stats_slots);
* run_simulation's loop: the converged scalars of the published dataset are
  checked on the GPU (tests/test_gpu_parquet.py) with the same loop logic,
  and test_gpu_run_simulation.py pins this restatement against the HIP
  run_simulation step for step.
"""
import numpy as np
import pytest

import analytical_cases as ac
from oracle import oracle as orc

FAST = [n for n in ac.CASES if n != "volumetric_heat_source"]


@pytest.mark.parametrize("name", FAST)
def test_closed_form_case(name):
    v = ac.CASES[name](ac.OracleBackend())
    assert ac.passes(name, v), f"{name}: {v:.4f} vs {ac.THRESHOLDS[name]}"


@pytest.mark.slow
def test_closed_form_volumetric_heat_source():
    v = ac.case_volumetric_heat_source(ac.OracleBackend())
    assert ac.passes("volumetric_heat_source", v), v


@pytest.mark.parametrize("mutation,case", [(1, "volumetric_heat_source"), (2, "kuwabara_pin_fin"),
                                           (3, "couette_flow")])
def test_planted_mistakes_are_caught(mutation, case):
    v = ac.CASES[case](ac.OracleBackend(mutation=mutation))
    assert not ac.passes(case, v), f"mutation {mutation} not caught by {case} ({v:.4f})"


def test_rest_state_fixed_point():
    """rho = 1, u = 0, no forcing: every population stays at its weight (periodic, with solids).

    Exact up to the rounding of sum(w_i) in FP64 (the weights do not sum to 1 exactly)."""
    m = np.zeros((6, 5, 4), np.int32)
    m[2, 2, :] = 1
    st = orc.State(m.shape, np.float64)
    st.init_fluid()
    f0 = st.f_old.copy()
    orc.run(st, m, 20, fluid=2, fluid_stream=2, tau=0.7)
    np.testing.assert_allclose(st.f_old, f0, rtol=1e-14, atol=0)
    assert np.abs(st.u).max() < 1e-15


def test_mass_conservation_periodic_fp64():
    rng = np.random.default_rng(0)
    m = np.zeros((10, 9, 8), np.int32)
    m[3:5, 3:6, :] = 1
    rho = 1.0 + 0.01 * rng.standard_normal(m.shape)
    u = 0.02 * rng.standard_normal(m.shape + (3,))
    u[m == 1] = 0.0
    rho[m == 1] = 1.0
    st = orc.State(m.shape, np.float64)
    st.init_fluid(rho, u)
    mass0 = st.f_old.sum()
    orc.run(st, m, 200, fluid=2, fluid_stream=2, tau=0.62)
    assert abs(st.f_old.sum() - mass0) / mass0 < 1e-13


def test_mass_conservation_closed_box_bounce_back():
    """A closed solid box (channel streaming with walls all around): bounce-back conserves mass."""
    rng = np.random.default_rng(1)
    m = np.ones((9, 8, 7), np.int32)
    m[1:-1, 1:-1, 1:-1] = 0
    u = 0.02 * rng.standard_normal(m.shape + (3,))
    u[m == 1] = 0.0
    st = orc.State(m.shape, np.float64)
    st.init_fluid(np.ones(m.shape), u)
    fluid = m == 0
    mass0 = st.f_old[:, fluid].sum()
    orc.run(st, m, 300, fluid=1, tau=0.6)
    assert abs(st.f_old[:, fluid].sum() - mass0) / mass0 < 1e-13


def test_heat_conservation_periodic_thermal():
    rng = np.random.default_rng(2)
    m = np.zeros((12, 6, 5), np.int32)
    m[4:7] = 1
    u = 0.03 * rng.standard_normal(m.shape + (3,))
    T0 = 20.0 + rng.uniform(0, 5, m.shape)
    st = orc.State(m.shape, np.float64)
    st.u[...] = u
    st.init_thermal(T0, m)
    h0 = st.g_old.sum()
    orc.run(st, m, 150, thermal=2, thermal_stream=1, tau_th_fluid=0.55, tau_th_solid=0.9)
    assert abs(st.g_old.sum() - h0) / h0 < 1e-13


def _coupled(dtype, steps=200):
    from test_gpu_parity import coupled_params, oracle_coupled
    p = coupled_params(seed=4)
    return oracle_coupled(p, steps, dtype)


def test_fp32_tracks_fp64():
    a, b = _coupled(np.float32), _coupled(np.float64)
    for name in ("rho", "u", "T"):
        x, y = getattr(a, name).astype(np.float64), getattr(b, name)
        if name == "rho":
            x, y = x - 1.0, y - 1.0
        if name == "T":
            x, y = x - 25.0, y - 25.0
        err = np.linalg.norm(x - y) / np.linalg.norm(y)
        assert err < 1e-4, (name, err)


def test_stats_slots():
    st = _coupled(np.float32, 60)
    from test_gpu_parity import coupled_params
    m = coupled_params(seed=4)["mask"]
    s = orc.stats(st.rho, st.u, st.T, m)
    v2 = (st.u.astype(np.float32) ** 2).sum(-1)
    assert s[2] == pytest.approx(float(v2.astype(np.float64).sum()), rel=1e-6)
    assert s[3] == np.sum(m != 1)
    assert s[8] == m.shape[1] * m.shape[2] and s[9] == m.shape[1] * m.shape[2]
    assert s[10] == np.sum(m == 3) and s[15] == np.sum(m == 3)
    assert s[11] == pytest.approx(float(st.T.max()))
    assert s[0] == pytest.approx(float(np.sqrt(v2[m != 1]).max()), rel=1e-7)
