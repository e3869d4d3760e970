"""The stop-rule bookkeeping of the drivers (host logic, no GPU).

``rule="reference"`` must be src/simulation.py:134-136 verbatim; ``rule="rise"`` applies the same formulas
to the temperature rise max T - T_in and to the absorbed heat sum(u_x T) - T_in sum(u_x).  The last test
shows why the option exists: a slowly relaxing temperature rise passes the reference rule at 1e-6 while
it is still percents from its final value, and does not pass the rise rule.
"""
import math

import pytest

from xadvcool_amd.simulation import _Residuals


def stats(ke, max_T, Q_out, ux_out=10.0, max_v=0.05):
    return {"max_v": max_v, "sum_ke": ke, "max_T": max_T, "Q_out": Q_out, "sum_ux_out": ux_out}


def test_reference_rule_is_the_reference_formulas():
    r = _Residuals(100, 1e-6, 1e-6, 0)
    r.update(100, stats(2.0, 30.0, 260.0))
    assert (r.d_ke, r.dT_conv, r.dQ_conv) == (0.0, 0.0, 1.0)  # first check (simulation.py:134-136)
    r.update(200, stats(2.2, 30.5, 262.0))
    assert r.d_ke == abs(2.2 - 2.0) / (2.0 + 1e-9) / 100
    assert r.dT_conv == abs(30.5 - 30.0) / (30.0 + 1e-9) / 100
    assert r.dQ_conv == abs(262.0 - 260.0) / (abs(260.0) + 1e-12) / 100


def test_rise_rule_uses_rise_and_absorbed_heat():
    T_in = 25.0
    r = _Residuals(100, 1e-6, 1e-6, 0, rule="rise", t_inlet=T_in)
    r.update(100, stats(2.0, 30.0, 260.0, ux_out=10.0))
    r.update(200, stats(2.0, 30.5, 262.0, ux_out=10.0))
    assert r.dT_conv == pytest.approx(abs(5.5 - 5.0) / 5.0 / 100, rel=1e-12)
    assert r.dQ_conv == pytest.approx(abs((262.0 - 250.0) - (260.0 - 250.0)) / 10.0 / 100, rel=1e-12)


def test_unknown_rule_is_refused():
    with pytest.raises(ValueError):
        _Residuals(100, 1e-6, 1e-6, 0, rule="field")


def test_slow_rise_passes_the_reference_rule_too_early():
    """Rise R(t) = R (1 - exp(-t / tau)) with tau = 15,000 steps and R = 4 over T_in = 25; the outlet heat
    follows the same law.  Stop at the first passing check (tol 1e-6) and compare with the final rise."""
    T_in, R, tau, flow = 25.0, 4.0, 15000.0, 10.0

    def stop_error(rule):
        res = _Residuals(100, 1e-6, 1e-6, 0, rule=rule, t_inlet=T_in)
        for step in range(100, 400001, 100):
            rise = R * (1.0 - math.exp(-step / tau))
            res.update(step, stats(1.0, T_in + rise, flow * (T_in + rise), ux_out=flow))
            if step > 100 and res.converged(step):
                return 1.0 - rise / R
        return None

    ref_err, rise_err = stop_error("reference"), stop_error("rise")
    assert ref_err is not None and ref_err > 0.05       # stops with > 5 % of the rise still to come
    assert rise_err is not None and rise_err < 0.02     # the rise rule stops near tol * tau = 1.5 %
