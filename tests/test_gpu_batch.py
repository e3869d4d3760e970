"""Batched solving: every slot of a mixed batch reproduces its own single solve bit for bit.

Designs of different lattice sizes share one fused launch per step; slots
retire at their own convergence and take queued designs (continuous
batching).  Each result must match run_simulation on the same design (same
stop step, identical FP32 fields), which the oracle pins separately.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def designs():
    from xadvcool_amd import geometry as G
    out = []
    for i, (fam, core) in enumerate([("straight", (20, 14, 10)), ("gyroid", (26, 16, 12)), ("pinfin", (18, 12, 9)),
                                     ("schwarz_p", (22, 18, 10)), ("schwarz_d", (24, 12, 11))]):
        d = G.family_design(100 + i, core=core, family=fam, heat_name="quad_core")
        mask = d["mask"]
        d["mask"] = mask
        d["key"] = i
        out.append(d)
    return out


@pytest.mark.parametrize("alternate_sweep", [False, True])
def test_batch_matches_single_solves(gpu, alternate_sweep):
    from xadvcool_amd.batch import BatchedSolver
    from xadvcool_amd.simulation import run_simulation
    ds = designs()
    cap = tuple(max(d["mask"].shape[k] for d in ds) for k in range(3))
    print(f"[batch] {len(ds)} designs in 3 slots, cap {cap}, alternate_sweep={alternate_sweep}", flush=True)
    bs = BatchedSolver(3, cap, tol_ke=1e-5, tol_T=1e-5, max_steps=3000, collect_fields=True, metrics=False,
                       alternate_sweep=alternate_sweep)
    try:
        res = {r["key"]: r for r in bs.run(ds)}
    finally:
        bs.close()
    assert len(res) == len(ds)
    for d in ds:
        r = res[d["key"]]
        ref = run_simulation(d["mask"], max_steps=3000, tau_fluid=d["tau_fluid"], heat_power=d["heat_power"],
                             u_inlet_val=d["u_inlet_val"], tol_ke=1e-5, tol_T=1e-5, silent=True,
                             heat_source=d["heat"], min_steps_rule=False)
        print(f"[batch] design {d['key']}: steps {r['steps']} ({r['status']}) vs {ref['steps']}", flush=True)
        assert r["steps"] == ref["steps"]
        rho, u, T = r["fields"]
        assert np.array_equal(rho, ref["rho"]) and np.array_equal(u, ref["u"]) and np.array_equal(T, ref["T"])
