"""Host-side rules of the warm-started batched solver (no GPU needed)."""
from types import SimpleNamespace

import numpy as np

from xadvcool_amd import warmstart as W

TOL = 1e-6


def _feed(stall, design, dT, every=100):
    """Residual series with d_ke at FP32-noise level and dQ = dT."""
    rng = np.random.default_rng(0)
    out = []
    for i, v in enumerate(dT):
        r = SimpleNamespace(d_ke=float(rng.uniform(1e-11, 1e-9)), dT_conv=float(v), dQ_conv=float(v))
        out.append(stall(design, (i + 1) * every, r))
    return out


def test_decaying_residual_with_dips_is_never_abandoned():
    st = W._Stall(TOL, TOL)
    dT = 3e-5 * np.exp(-np.arange(600) / 150.0)
    dT[[10, 11, 57, 200, 201, 202]] = 1e-9  # rate crossing zero at a few checks
    assert not any(_feed(st, {"key": 1, "warm": True}, dT))


def test_growing_residual_is_abandoned_after_the_start_window():
    st = W._Stall(TOL, TOL, after=5000)
    dT = np.concatenate([3e-6 * np.exp(-np.arange(60) / 40.0), 2e-6 * 1.05 ** np.arange(200)])
    out = _feed(st, {"key": 2, "warm": True}, dT)
    first = out.index(True)
    assert (first + 1) * 100 > 5000
    assert dT[first] > 4 * dT[:first].min()


def test_stalled_residual_is_abandoned_after_patience():
    st = W._Stall(TOL, TOL, after=5000, patience=100)
    dT = np.full(400, 3e-6)
    out = _feed(st, {"key": 3, "warm": True}, dT)
    first = out.index(True)
    # the median window fills at check 20, the best is set there, then 100 checks without improvement
    assert first + 1 == 20 + 100


def test_cold_runs_are_never_abandoned():
    st = W._Stall(TOL, TOL)
    out = _feed(st, {"key": 4}, 1e-6 * 1.1 ** np.arange(300))
    assert not any(out)


def test_compare_designs_are_unseen():
    train = {d["key"] for d in W.make_designs(0, 8)}
    test = {d["key"] for d in W.make_designs(W.COMPARE_SEED0, 8)}
    assert not train & test
    assert W.COMPARE_SEED0 > 100_000  # far above any training set size used


def test_warm_start_dip_then_slow_decay_is_not_abandoned():
    """A warm start's residual often dips right after the start, rises while the slow modes develop, then
    decays with the thermal time constant (~15,000 steps); the early dip is not the level to beat."""
    st = W._Stall(TOL, TOL)
    i = np.arange(400)
    dip = 1e-5 * np.exp(-i / 5.0) + 1e-8
    rise = 3e-6 * np.clip((i - 30) / 30.0, 0.0, 1.0) * np.exp(-np.clip(i - 60, 0, None) / 150.0)
    dT = np.maximum(dip, rise)
    assert not any(_feed(st, {"key": 5, "warm": True}, dT))


def test_decay_slower_than_the_stall_bound_is_abandoned():
    """Decaying by less than 10 % over 100 checks counts as stalled."""
    st = W._Stall(TOL, TOL, after=5000, patience=100)
    dT = 3e-6 * np.exp(-np.arange(400) / 2000.0)
    out = _feed(st, {"key": 6, "warm": True}, dT)
    assert True in out


def test_compare_summary_counts_and_ratios():
    """The benchmark's summary from synthetic results: converged counts, solves per hour, agreement only where
    both arms converged, fallback reasons, per-family ratios."""
    from xadvcool_amd.checks import GATE
    ok = {k: True for k in GATE}

    def res(key, steps, status, R, dP):
        return {"key": key, "steps": steps, "status": status, "converged": status == "converged",
                "metrics": {"R_th": R, "pressure_drop": dP, "Nu": 1.0, "P_pump": 1.0}, "checks": dict(ok)}

    designs = [{"key": k, "family": f, "heat_name": "uniform", "u_inlet_val": 0.02}
               for k, f in ((1, "gyroid"), (2, "gyroid"), (3, "straight"))]
    cold = {"results": {1: res(1, 60000, "converged", 1.0, 0.01), 2: res(2, 120000, "max_steps", 1.0, 0.01),
                        3: res(3, 20000, "converged", 2.0, 0.02)}, "wall": 100.0, "fallbacks": [],
            "inference_s": 0.0}
    w1 = dict(res(1, 20000, "converged", 1.004, 0.01), abandoned_steps=0)
    w2 = dict(res(2, 30000, "converged", 1.0, 0.01), abandoned_steps=10000)
    w3 = dict(res(3, 5000, "converged", 2.02, 0.02), abandoned_steps=0)
    warm = {"results": {1: w1, 2: w2, 3: w3}, "wall": 25.0, "fallbacks": [(2, "stalled", 10000)],
            "inference_s": 0.1}
    rows = W.compare_rows(designs, cold, warm)
    assert [r["warm_steps"] for r in rows] == [20000, 40000, 5000]
    s = W.summarize(rows, cold, warm)
    assert s["cold"]["converged"] == 2 and s["warm"]["converged"] == 3
    assert s["cold"]["solves_per_hr"] == 72.0 and s["warm"]["solves_per_hr"] == 432.0
    assert s["speedup_solves_per_hr"] == 6.0
    assert s["both_converged"] == 2 and s["within_0.5pct"] == 1 and s["warm_only_converged"] == 1
    assert abs(s["max_d_R_th"] - 0.01) < 1e-12
    assert s["warm"]["fallback_reasons"] == {"stalled": 1}
    g = s["by_family"]["gyroid"]
    assert (g["cold_converged"], g["warm_converged"], g["both"], g["within_0.5pct"]) == (1, 2, 1, 1)
    assert g["converged_per_step_ratio"] == round((2 / 60000) / (1 / 180000), 3)


def test_shipped_ensemble_members_share_the_architecture():
    """The two shipped weight files hold the same arrays (names and shapes) and the same normalisation."""
    import os
    a, b = (np.load(p, allow_pickle=False) for p in W.DEFAULT_ENSEMBLE)
    assert all(os.path.exists(p) for p in W.DEFAULT_ENSEMBLE)
    assert sorted(a.files) == sorted(b.files)
    for k in a.files:
        assert a[k].shape == b[k].shape, k
    norm = [k for k in a.files if "norm" in k or "mean" in k or "std" in k]
    assert norm, a.files
    for k in norm:
        assert np.array_equal(a[k], b[k]), k
    assert any(not np.array_equal(a[k], b[k]) for k in a.files)  # two different networks


def test_load_surrogate_forms_one_network_or_an_ensemble(monkeypatch):
    """A path gives one network; several (a sequence, or paths joined by commas) an Ensemble in that order."""
    monkeypatch.setattr(W.S.UNet3D, "load_file", classmethod(lambda cls, p, device="cuda": ("net", p)))
    assert W.load_surrogate("a.npz") == ("net", "a.npz")
    e = W.load_surrogate("a.npz,b.npz")
    assert isinstance(e, W.Ensemble) and e.nets == [("net", "a.npz"), ("net", "b.npz")]
    e = W.load_surrogate(W.DEFAULT_ENSEMBLE)
    assert [p for _, p in e.nets] == list(W.DEFAULT_ENSEMBLE)
    try:
        W.Ensemble([])
    except ValueError:
        pass
    else:
        raise AssertionError("an empty ensemble must be refused")
