"""The plain small cold plate (74 x 64 x 16, 2,000 steps) through the drop-in run_simulation on the GPU:
fields identical to the oracle's run of the reference schedule and to the committed fixture."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_config1_run_simulation_bit_exact(gpu):
    from xadvcool_amd import geometry as G
    from xadvcool_amd.simulation import run_simulation
    print("[config1] 74x64x16, 2000 steps", flush=True)
    c = G.CONFIG1
    m = G.config1_mask()
    kw = dict(max_steps=c["steps"], tau_fluid=c["tau_fluid"], heat_power=c["heat_power"], t_inlet=c["t_inlet"],
              u_inlet_val=c["u_inlet_val"], tol_ke=0.0, tol_T=0.0, Pr=c["Pr"], k_ratio=c["k_ratio"])
    got = run_simulation(m, device="cuda", silent=True, **kw)
    assert got["steps"] == c["steps"] and not got["converged"]
    with open(os.path.join(HERE, "golden", "config1.json")) as f:
        want = json.load(f)
    md5 = hashlib.md5(got["u"].tobytes() + got["rho"].tobytes() + got["T"].tobytes()).hexdigest()
    if md5 != want["fields_md5"]:
        orc.set_threads(orc.max_threads())
        ref = orc.run_simulation(m, **kw)
        for k in ("rho", "u", "T"):
            d = np.abs(got[k].astype(np.float64) - ref[k])
            print(f"[config1] {k}: max |diff| {d.max():.3e}", flush=True)
        pytest.fail("config-1 fields differ from the oracle's")
