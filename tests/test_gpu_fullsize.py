"""Full-size parity of the benchmark's own launch configuration.

test_gpu_parity.py pins every launch shape on small lattices.  This test pins the configuration
bench.py measures: three 148x128x32 gyroid designs per launch (seeds 1000-1002, the benchmark's
first three), compact fluid storage, graph-captured chunks, 256 threads x 2 cells, linear order,
alternating sweep, thermal loads issued right after the fluid gather (and the XCD-contiguous order
without that prefetch).  After 24 steps every slot's FP32 rho, u and T equal the oracle's bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 24
SHAPE = (148, 128, 32)


@pytest.fixture(scope="module")
def cases():
    from oracle import oracle as orc
    from xadvcool_amd.geometry import config2_design
    orc.set_threads(orc.max_threads())
    out = []
    for sd in (1000, 1001, 1002):
        d = config2_design(sd)
        ttf, tts = orc.thermal_taus(d["tau_fluid"], 7.0, 628.0)
        heat = orc.normalized_heat(d["heat"], SHAPE)
        p = dict(mask=d["mask"], heat=heat, tau=d["tau_fluid"], ttf=ttf, tts=tts, P=d["heat_power"], T_in=25.0,
                 u_in=d["u_inlet_val"])
        st = orc.State(SHAPE, np.float32)
        st.T[...] = p["T_in"]
        st.init_thermal(mask=p["mask"])
        st.init_fluid()
        orc.run(st, p["mask"], STEPS, fluid=1, thermal=1, tau=p["tau"], u_in=p["u_in"], rho_out=1.0,
                tau_th_fluid=ttf, tau_th_solid=tts, heat_power=p["P"], T_in=p["T_in"], heat=heat)
        out.append((sd, p, (st.rho.copy(), st.u.copy(), st.T.copy())))
    return out


@pytest.mark.parametrize("shape", [dict(block=256, cells_per_thread=2, xcd_remap=False, alternate_sweep=True,
                                        thermal_prefetch=True),
                                   dict(block=256, cells_per_thread=1, xcd_remap=True, alternate_sweep=True)],
                         ids=["bench-shape", "xcd-order"])
def test_three_fullsize_designs_per_launch_match_oracle(gpu, cases, shape):
    from xadvcool_amd.engine import LatticeBatch
    print(f"[fullsize] 3 x {SHAPE}, {STEPS} steps, {shape}", flush=True)
    b = LatticeBatch(slots=3, cap_shape=SHAPE, graph_chunk=12, **shape)
    try:
        for s, (_, p, _) in enumerate(cases):
            b.load(s, p["mask"], heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"], tau_th_solid=p["tts"],
                   heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"])
            b.init(s)
        b.step(STEPS, want_fields=True)
        b.readout(with_stats=False)
        for s, (sd, _, ref) in enumerate(cases):
            got = [t.cpu().numpy() for t in b.fields(s)]
            for name, a, r in zip(("rho", "u", "T"), got, ref):
                assert np.array_equal(a, r), f"seed {sd}: {name} differs from the oracle"
    finally:
        b.close()
