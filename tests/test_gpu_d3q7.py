"""The D3Q7 thermal variant on the GPU (the fused pass with thermal lattice 7) against the oracle's D3Q7
schedule (thermal schedule 3 of oracle/lbm_oracle.cpp, pinned by tests/test_oracle_d3q7.py): rho, u and T
bit for bit in FP32 and FP64, cold and warm starts, every launch shape, graph chunks, two slots of
different lattices with one refilled late, and config 1 through the drop-in run_simulation."""
import numpy as np
import pytest

from oracle import oracle as orc
from test_gpu_parity import assert_match, coupled_params

pytestmark = pytest.mark.gpu


def oracle7(p, steps, dtype, init=None):
    st = orc.State(p["mask"].shape, dtype)
    st.T[...] = p["T_in"]
    if init is not None:
        st.rho[...], st.u[...], st.T[...] = init
    orc.init_thermal7(st, mask=p["mask"])
    st.init_fluid()
    orc.run(st, p["mask"], steps, fluid=1, thermal=3, tau=p["tau"], u_in=p["u_in"], rho_out=1.0,
            tau_th_fluid=p["ttf"], tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], heat=p["heat"])
    return st


def hip7(p, steps, precision, chunks=None, init=None, **kw):
    import torch
    from xadvcool_amd.engine import LatticeSolver
    s = LatticeSolver(p["mask"], fluid=True, thermal=True, heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"],
                      tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"], precision=precision,
                      thermal_lattice="D3Q7", **kw)
    try:
        if init is None:
            s.init_equilibrium()
        else:
            s.init_equilibrium(*(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in init))
        for n in (chunks or [steps]):
            s.step(n)
        return tuple(t.cpu().numpy() for t in s.fields_torch())
    finally:
        s.close()


@pytest.mark.parametrize("precision,seed", [(32, 0), (32, 1), (64, 2)])
def test_d3q7_matches_oracle(gpu, precision, seed):
    dtype = np.float32 if precision == 32 else np.float64
    p = coupled_params(seed=seed)
    steps = 60
    print(f"[d3q7] parity FP{precision} seed {seed}, {steps} steps", flush=True)
    r, u, T = hip7(p, steps, precision, graph_chunk=0)
    st = oracle7(p, steps, dtype)
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)
    assert_match("T", T, st.T)
    assert float(T.max()) > p["T_in"] + 1e-3  # the heat reached the fields


def test_d3q7_warm_start_matches_oracle(gpu):
    p = coupled_params(seed=3)
    rng = np.random.default_rng(7)
    shp = p["mask"].shape
    init = (1.0 + 0.01 * rng.standard_normal(shp).astype(np.float32),
            (0.01 * rng.standard_normal(shp + (3,))).astype(np.float32),
            (25.0 + rng.random(shp)).astype(np.float32))
    print("[d3q7] warm start", flush=True)
    r, u, T = hip7(p, 31, 32, init=init, graph_chunk=0)
    st = oracle7(p, 31, np.float32, init=init)
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)
    assert_match("T", T, st.T)


@pytest.mark.parametrize("shape", [dict(cells_per_thread=2, xcd_remap=True, graph_chunk=10),
                                   dict(cells_per_thread=1, block=128, fluid_storage="dense", graph_chunk=20,
                                        alternate_sweep=False)])
def test_d3q7_launch_shapes_and_graphs(gpu, shape):
    """Graph-captured chunks hold the D3Q7 kernel after every fluid pass; launch shapes and storages give the
    same bits."""
    p = coupled_params(seed=4)
    print(f"[d3q7] launch shape {shape}", flush=True)
    r, u, T = hip7(p, 60, 32, chunks=[20, 20, 20] if shape["graph_chunk"] == 20 else [30, 30], **shape)
    st = oracle7(p, 60, np.float32)
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)
    assert_match("T", T, st.T)


def test_d3q7_batch_slots_and_late_start(gpu):
    """Two slots of different lattices; slot 1 starts 20 steps late (at a chunk boundary).  Each equals its
    own oracle run."""
    from xadvcool_amd.engine import LatticeBatch
    pa, pb = coupled_params(seed=5), coupled_params(seed=6, shape=(18, 10, 9))
    b = LatticeBatch(slots=2, cap_shape=pa["mask"].shape, fluid=True, thermal=True, graph_chunk=10,
                     thermal_lattice="D3Q7")
    try:
        for s, p in ((0, pa), (1, pb)):
            b.load(s, p["mask"], heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"], tau_th_solid=p["tts"],
                   heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"])
        b.init(0)
        b.set_active(1, False)
        b.step(20)
        b.init(1)
        b.step(30)
        b.readout(with_stats=False)
        got = [b.fields(s) for s in (0, 1)]
    finally:
        b.close()
    for (r, u, T), p, n in ((got[0], pa, 50), (got[1], pb, 30)):
        st = oracle7(p, n, np.float32)
        assert_match("rho", r.cpu().numpy(), st.rho)
        assert_match("u", u.cpu().numpy(), st.u)
        assert_match("T", T.cpu().numpy(), st.T)


def test_d3q7_config1_run_simulation(gpu):
    """Config 1 (the 74 x 64 x 16 cold plate, 2,000 steps, k_ratio 10) through the drop-in run_simulation with
    the D3Q7 thermal lattice: bit for bit the oracle's run_simulation restatement on the D3Q7 schedule; the
    temperature rise within 3 % of the D3Q19 run's."""
    from xadvcool_amd.geometry import config1_mask
    from xadvcool_amd.simulation import run_simulation
    mask = config1_mask()
    kw = dict(max_steps=2000, tau_fluid=0.6, heat_power=0.1, t_inlet=25.0, u_inlet_val=0.02, tol_ke=0.0, tol_T=0.0,
              silent=True, k_ratio=10.0)
    print("[d3q7] config 1 run_simulation", flush=True)
    g7 = run_simulation(mask, thermal_lattice="D3Q7", **kw)
    g19 = run_simulation(mask, **kw)
    o7 = orc.run_simulation(mask, thermal_lattice="D3Q7", **kw)
    assert g7["steps"] == o7["steps"] == 2000
    for k in ("rho", "u", "T"):
        assert_match(k, g7[k], o7[k])
    r7, r19 = g7["T"] - 25.0, g19["T"] - 25.0
    assert abs(r7.max() / r19.max() - 1) < 0.03
    assert abs(r7.mean() / r19.mean() - 1) < 0.03
