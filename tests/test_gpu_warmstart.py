"""Warm start on the GPU: the cold LBM is the fallback and the correctness reference.

* a warm start that blows up (temperatures far above the divergence limit) is abandoned at
  its first residual check and the design is re-run cold: the result equals
  the cold run of the same design bit for bit (same steps, same metrics);
* the trained surrogate (the shipped pair of networks, predictions averaged), on
  designs it never saw, reaches the same stop rule in fewer steps than the cold
  start, with converged pressure drop and thermal resistance within 0.5 % of the
  cold result.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _designs(n, seed_off):
    from xadvcool_amd import warmstart as W
    return W.make_designs(W.COMPARE_SEED0 + seed_off, n)


def test_blown_up_warm_start_falls_back_to_cold(gpu):
    from xadvcool_amd import warmstart as W

    designs = _designs(2, 7000)

    def bad(d):  # a thermal explosion: max T far above the divergence limit at the first check
        nx, ny, nz = d["mask"].shape
        u = torch.zeros(nx, ny, nz, 3, device="cuda")
        rho = torch.ones(nx, ny, nz, device="cuda")
        T = torch.full((nx, ny, nz), 1.0e6, device="cuda")
        return rho, u, T

    kw = dict(slots=2, tol=1e-5, max_steps=6000)
    cold = W.solve_designs(designs, None, **kw)
    warm = W.solve_designs(designs, None, predictor=bad, **kw)
    assert len(warm["fallbacks"]) == len(designs)
    for k, status, steps in warm["fallbacks"]:
        assert status == "diverged" and steps == 100
    for d in designs:
        c, w = cold["results"][d["key"]], warm["results"][d["key"]]
        assert w["abandoned_steps"] == 100
        assert (w["steps"], w["status"]) == (c["steps"], c["status"])
        for m in ("pressure_drop", "R_th", "Nu", "Q_vol"):
            assert w["metrics"][m] == c["metrics"][m], m


def test_unphysical_converged_warm_result_is_rerun_cold(gpu):
    """The physics gate: a warm run that meets the stop rule with fields failing the acceptance checks
    (here a prediction with the flow reversed: at the first check the flow through the plane one cell
    before the outlet still runs backwards, so analyze_dataset.py's mass balance fails) is not accepted;
    the design is re-run cold and the cold result is the one reported."""
    from xadvcool_amd import surrogate as S
    from xadvcool_amd import warmstart as W

    designs = _designs(2, 7100)
    net = S.UNet3D.load_file(W.DEFAULT_WEIGHTS)

    def reversed_flow(d):
        rho, u, T = W.predict_fields(net, d)
        return rho, (-u).contiguous(), T

    # A loose tolerance stops every run at its first check (step 100), well inside one acoustic crossing.
    kw = dict(slots=2, tol=10.0, max_steps=2000, confirm=1)
    cold = W.solve_designs(designs, None, **kw)
    warm = W.solve_designs(designs, None, predictor=reversed_flow, **kw)
    print(f"[gate] fallbacks {warm['fallbacks']}", flush=True)
    assert len(warm["fallbacks"]) == len(designs)
    for k, why, steps in warm["fallbacks"]:
        assert why.startswith("checks:") and "mass_warn" in why and steps == 100, why
    for d in designs:
        c, w = cold["results"][d["key"]], warm["results"][d["key"]]
        assert c["converged"] and c["steps"] == 100
        assert w["abandoned_steps"] == 100 and (w["steps"], w["status"]) == (c["steps"], c["status"])
        assert w["checks"] == c["checks"]
        for m in ("pressure_drop", "R_th", "Nu", "Q_vol"):
            assert w["metrics"][m] == c["metrics"][m], m


def test_trained_surrogate_cuts_steps_on_unseen_designs(gpu):
    """Eight unseen designs (all families), rise residuals at 3e-7 on five consecutive checks, 150,000-step cap.

    Both arms stop on the rise residuals (the reference's formulas on the temperature rise and the absorbed
    heat).  A rate-based rule stops an exponential approach with time constant tau about tol * tau from its
    steady state; on these lattices tau reaches ~15,000 steps, so at 1e-6 a cold run can stop 0.7 % away and
    could not judge the warm one at 0.5 %; at 3e-7 it stops within ~0.2 %."""
    from xadvcool_amd import warmstart as W

    for p in W.DEFAULT_ENSEMBLE:
        assert os.path.exists(p), f"trained weights missing from the package: {p}"
    net = W.load_surrogate()
    assert isinstance(net, W.Ensemble) and len(net.nets) == 2
    n = 8
    designs = _designs(n, 5000)
    print("[warmstart] cold vs warm, 8 designs, rise residuals at 3e-7", flush=True)
    summ, rows = W.compare(net, designs, slots=n, tol=3e-7, max_steps=150000, stop_rule="rise")
    print(summ, flush=True)
    both = [r for r in rows if r["cold_status"] == "converged" and r["warm_status"] == "converged"]
    assert len(both) >= 7
    for r in both:  # the cold LBM is the reference for the warm result
        assert r["d_pressure_drop"] <= 5e-3, r
        assert r["d_R_th"] <= 5e-3, r
    cold_steps = np.mean([r["cold_steps"] for r in both])
    warm_steps = np.mean([r["warm_steps"] for r in both])
    assert warm_steps < 0.5 * cold_steps, (warm_steps, cold_steps)
    assert summ["speedup_solves_per_hr"] >= 2.0, summ


def test_ensemble_prediction_is_the_mean_of_its_networks(gpu):
    """The shipped surrogate averages its two networks' de-normalised fields; solids and INLET cells keep the
    cold-start state in every member and so in the mean."""
    from xadvcool_amd import warmstart as W
    ens = W.load_surrogate()
    d = _designs(1, 7200)[0]
    print("[warmstart] ensemble mean", flush=True)
    a, b = (W.predict_fields(n, d) for n in ens.nets)
    m = W.predict_fields(ens, d)
    for x, y, z in zip(a, b, m):
        assert torch.equal(z, (x + y) / 2)
    solid = torch.from_numpy(d["mask"] == 1).cuda()
    assert torch.all(m[0][solid] == 1.0) and torch.all(m[1][solid] == 0.0)
    assert not torch.equal(a[2], b[2])  # two different networks
