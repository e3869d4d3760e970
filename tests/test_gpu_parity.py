"""Parity of the HIP AA-pattern passes with the CPU oracle, field by field.

The oracle (oracle/lbm_oracle.cpp) runs the reference's six-kernel,
two-buffer schedule; the HIP library runs one fused in-place pass per step.
Both use the reference's operations in the reference's order with no
contraction, so in FP32 the fields must agree bit for bit; the bar written
into each test is a rel-L2 bound of 1e-5 plus an exact-match check
on rho, u and T.  FP64 runs are checked the same way against the FP64 oracle.

Lattices are small and seeded so the oracle finishes in well under a second;
they include every cell type, solid pins touching the inlet and outlet
planes, heat on the base and (for the general kernels) every test boundary
mode of validate_analytical.py.
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

BC_FLUID, BC_SOLID, BC_INLET, BC_OUTLET = 0, 1, 2, 3


def channel_mask(nx=22, ny=12, nz=10, seed=0, base=2, pins=True):
    """A cold-plate-like mask: side walls, base, lid, inlet/outlet planes, random pins."""
    rng = np.random.default_rng(seed)
    m = np.zeros((nx, ny, nz), np.int32)
    if pins:
        for _ in range(6):
            x0, y0 = rng.integers(2, nx - 3), rng.integers(1, ny - 2)
            m[x0:x0 + 2, y0:y0 + 2, :] = BC_SOLID
        m[1, 3, :] = BC_SOLID      # solid right behind the inlet
        m[nx - 2, 5, :] = BC_SOLID  # solid left of the outlet
    m[:, 0, :] = m[:, -1, :] = BC_SOLID
    m[:, :, :base] = BC_SOLID
    m[:, :, -1:] = BC_SOLID
    m[0] = np.where(m[0] == BC_SOLID, BC_SOLID, BC_FLUID)
    m[0, 1:-1, base:-1] = BC_INLET
    m[-1, 1:-1, base:-1] = BC_OUTLET
    return m


def relL2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def assert_match(name, got, ref, tol=1e-5, exact=True):
    e = relL2(got, ref)
    assert e <= tol, f"{name}: rel-L2 {e:.3e} > {tol:g}"
    if exact:
        bad = np.argwhere(got != ref)
        assert bad.size == 0, (f"{name}: {len(bad)} values differ, first at {tuple(bad[0])}: "
                               f"{got[tuple(bad[0])]!r} vs {ref[tuple(bad[0])]!r} (rel-L2 {e:.2e})")


def coupled_params(seed=0, shape=(22, 12, 10)):
    m = channel_mask(*shape, seed=seed)
    rng = np.random.default_rng(seed + 100)
    hs = np.zeros(m.shape, np.float32)
    hs[2:-2, :, 0] = rng.uniform(0.0, 1.0, size=hs[2:-2, :, 0].shape).astype(np.float32)
    hs[2:-2, 3:5, 0] = 0.0
    hs /= hs.max()
    tau, Pr, kr = 0.56, 7.0, 628.0
    ttf, tts = orc.thermal_taus(tau, Pr, kr)
    return dict(mask=m, heat=hs, tau=tau, ttf=ttf, tts=tts, P=0.1, T_in=25.0, u_in=0.03)


def oracle_coupled(p, steps, dtype):
    st = orc.State(p["mask"].shape, dtype)
    st.T[...] = p["T_in"]
    st.init_thermal(mask=p["mask"])
    st.init_fluid()
    orc.run(st, p["mask"], steps, fluid=1, thermal=1, tau=p["tau"], u_in=p["u_in"], rho_out=1.0,
            tau_th_fluid=p["ttf"], tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], heat=p["heat"])
    return st


def hip_coupled(p, steps, precision, chunks=None, graph_chunk=100, **shape):
    from xadvcool_amd.engine import LatticeSolver
    s = LatticeSolver(p["mask"], fluid=True, thermal=True, heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"],
                      tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"], precision=precision,
                      graph_chunk=graph_chunk, **shape)
    try:
        s.init_equilibrium()
        for n in (chunks or [steps]):
            s.step(n)
        r, u, T = (t.cpu().numpy() for t in s.fields_torch())
        return r, u, T
    finally:
        s.close()


@pytest.mark.parametrize("storage", ["compact", "dense"])
@pytest.mark.parametrize("steps", [1, 2, 3, 4, 7, 50])
def test_coupled_fp32_matches_oracle(gpu, steps, storage):
    """Both fluid storages (non-solid cells only, or every cell) give the oracle's bits."""
    print(f"[parity] coupled fp32 steps={steps} storage={storage}", flush=True)
    p = coupled_params()
    ref = oracle_coupled(p, steps, np.float32)
    r, u, T = hip_coupled(p, steps, 32, fluid_storage=storage)
    assert_match("rho", r, ref.rho)
    assert_match("u", u, ref.u)
    assert_match("T", T, ref.T)


@pytest.mark.parametrize("block,cpt,xcd,alt,storage,pre", [(64, 1, False, False, "compact", False),
                                                           (128, 2, True, False, "compact", False),
                                                           (256, 1, True, False, "compact", False),
                                                           (64, 2, False, False, "compact", False),
                                                           (128, 2, True, False, "dense", False),
                                                           (256, 1, False, False, "dense", False),
                                                           (256, 1, False, True, "compact", False),
                                                           (64, 2, True, True, "compact", False),
                                                           (128, 1, False, True, "dense", False),
                                                           (256, 2, False, True, "compact", True),
                                                           (64, 1, True, False, "compact", True)])
def test_coupled_launch_shapes(gpu, block, cpt, xcd, alt, storage, pre):
    """Every launch shape the autotuner may pick computes the same bits, in either block order, with the thermal
    loads issued after the fluid gather or with the fluid stores."""
    print(f"[parity] coupled shape block={block} cpt={cpt} xcd={xcd} alt={alt} storage={storage} prefetch={pre}",
          flush=True)
    p = coupled_params(seed=2, shape=(23, 13, 11))
    ref = oracle_coupled(p, 13, np.float32)
    r, u, T = hip_coupled(p, 13, 32, block=block, cells_per_thread=cpt, xcd_remap=xcd, alternate_sweep=alt,
                          fluid_storage=storage, thermal_prefetch=pre)
    assert_match("rho", r, ref.rho)
    assert_match("u", u, ref.u)
    assert_match("T", T, ref.T)


@pytest.mark.parametrize("storage,pre", [("compact", False), ("dense", False), ("compact", True)])
def test_coupled_fp64_matches_oracle(gpu, storage, pre):
    print(f"[parity] coupled fp64 steps=20 storage={storage} prefetch={pre}", flush=True)
    p = coupled_params(seed=3)
    ref = oracle_coupled(p, 20, np.float64)
    r, u, T = hip_coupled(p, 20, 64, fluid_storage=storage, thermal_prefetch=pre)
    assert_match("rho", r, ref.rho, tol=1e-12)
    assert_match("u", u, ref.u, tol=1e-12)
    assert_match("T", T, ref.T, tol=1e-12)


def test_coupled_graph_and_split_steps(gpu):
    """Graph replays (chunks of 10) and odd-sized step calls give the same state."""
    print("[parity] coupled graph/split", flush=True)
    p = coupled_params(seed=5)
    ref = oracle_coupled(p, 40, np.float32)
    r, u, T = hip_coupled(p, 40, 32, chunks=[10, 10, 3, 7, 10], graph_chunk=10)
    assert_match("rho", r, ref.rho)
    assert_match("u", u, ref.u)
    assert_match("T", T, ref.T)


@pytest.mark.parametrize("storage", ["compact", "dense"])
@pytest.mark.parametrize("steps", [1, 2, 5, 30])
def test_fluid_only_channel(gpu, steps, storage):
    print(f"[parity] fluid-only channel steps={steps} storage={storage}", flush=True)
    from xadvcool_amd.engine import FluidEngine3D
    m = channel_mask(18, 11, 9, seed=1)
    st = orc.State(m.shape, np.float32)
    st.init_fluid()
    orc.run(st, m, steps, fluid=2, fluid_stream=0, tau=0.6, u_in=0.02, rho_out=1.0)
    s = FluidEngine3D(m, tau=0.6, u_inlet_val=0.02, fluid_storage=storage)
    try:
        s.step(steps)
        r, u, _ = (t.cpu().numpy() for t in s.fields_torch())
    finally:
        s.close()
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)


@pytest.mark.parametrize("steps", [1, 2, 9])
def test_thermal_only_channel(gpu, steps):
    """Thermal lattice with a frozen velocity, INLET planes at both ends, solids and a heat source."""
    print(f"[parity] thermal-only channel steps={steps}", flush=True)
    from xadvcool_amd.engine import ThermalEngine3D
    rng = np.random.default_rng(7)
    m = channel_mask(16, 9, 8, seed=2)
    m[-1] = np.where(m[-1] == BC_OUTLET, BC_OUTLET, m[-1])
    u = (rng.uniform(-0.03, 0.03, size=m.shape + (3,))).astype(np.float32)
    T0 = rng.uniform(20.0, 30.0, size=m.shape).astype(np.float32)
    hs = np.zeros(m.shape, np.float32)
    hs[:, :, 0] = rng.uniform(0, 1, size=hs[:, :, 0].shape)
    st = orc.State(m.shape, np.float32)
    st.u[...] = u
    st.init_thermal(T0, m)
    orc.run(st, m, steps, thermal=2, thermal_stream=0, tau_th_fluid=0.51, tau_th_solid=0.9, heat_power=0.05,
            T_in=25.0, heat=hs)
    s = ThermalEngine3D(m, tau_fluid=0.51, tau_solid=0.9, t_inlet=25.0, heat=hs, heat_power=0.05)
    try:
        s.init_equilibrium(None, u, T0)
        s.step(steps)
        T = s.fields_torch()[2].cpu().numpy()
    finally:
        s.close()
    assert_match("T", T, st.T)


@pytest.mark.parametrize("storage", ["compact", "dense"])
def test_coupled_from_given_fields(gpu, storage):
    """Initial rho, u, T given everywhere (solids included, as a warm start may): both storages match."""
    print(f"[parity] coupled from given fields storage={storage}", flush=True)
    from xadvcool_amd.engine import LatticeSolver
    p = coupled_params(seed=4, shape=(21, 14, 12))
    rng = np.random.default_rng(44)
    shape = p["mask"].shape
    rho0 = (1.0 + rng.uniform(-0.01, 0.01, size=shape)).astype(np.float32)
    u0 = rng.uniform(-0.02, 0.02, size=shape + (3,)).astype(np.float32)
    T0 = rng.uniform(25.0, 27.0, size=shape).astype(np.float32)
    st = orc.State(shape, np.float32)
    st.init_fluid(rho0, u0)
    st.init_thermal(T0, p["mask"])
    orc.run(st, p["mask"], 9, fluid=1, thermal=1, tau=p["tau"], u_in=p["u_in"], rho_out=1.0, tau_th_fluid=p["ttf"],
            tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], heat=p["heat"])
    s = LatticeSolver(p["mask"], fluid=True, thermal=True, heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"],
                      tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"],
                      fluid_storage=storage)
    try:
        s.init_equilibrium(rho0, u0, T0)
        s.step(9)
        r, u, T = (t.cpu().numpy() for t in s.fields_torch())
    finally:
        s.close()
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)
    assert_match("T", T, st.T)


def _tgv(L=16, nz=3, U0=0.05):
    k = 2 * np.pi / L
    X, Y = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    u = np.zeros((L, L, nz, 3), np.float32)
    u[..., 0] = (U0 * np.sin(k * X) * np.cos(k * Y))[..., None]
    u[..., 1] = (-U0 * np.cos(k * X) * np.sin(k * Y))[..., None]
    rho = np.ones((L, L, nz), np.float32)
    rho[...] = (1.0 - 0.75 * U0 ** 2 * (np.cos(2 * k * X) + np.cos(2 * k * Y)))[..., None]
    return rho, u


@pytest.mark.parametrize("mode", ["periodic", "couette", "forced_solids"])
def test_general_fluid_modes(gpu, mode):
    print(f"[parity] general fluid {mode}", flush=True)
    from xadvcool_amd.engine import LatticeSolver
    if mode == "periodic":
        rho0, u0 = _tgv()
        m = np.zeros(rho0.shape, np.int32)
        kw_o = dict(fluid_stream=1)
        kw_h = dict(periodic=(1, 1, 1))
    elif mode == "couette":
        m = np.zeros((5, 6, 11), np.int32)
        m[2, 2, 3] = BC_SOLID
        rho0, u0 = np.ones(m.shape, np.float32), np.zeros(m.shape + (3,), np.float32)
        kw_o = dict(fluid_stream=3, u_wall=0.05)
        kw_h = dict(periodic=(1, 1, 0), u_wall=0.05)
    else:
        m = np.zeros((12, 10, 3), np.int32)
        m[4:7, 3:6, :] = BC_SOLID
        m[:, 0, :] = BC_SOLID
        rho0, u0 = np.ones(m.shape, np.float32), np.zeros(m.shape + (3,), np.float32)
        kw_o = dict(fluid_stream=2, forced=1, force_x=1e-4)
        kw_h = dict(periodic=(1, 1, 1), force_x=1e-4)
    steps = 25
    st = orc.State(m.shape, np.float32)
    st.init_fluid(rho0, u0)
    orc.run(st, m, steps, fluid=2, tau=0.6, **kw_o)
    s = LatticeSolver(m, fluid=True, thermal=False, tau=0.6, **kw_h)
    try:
        s.init_equilibrium(rho0, u0, None)
        s.step(steps)
        r, u, _ = (t.cpu().numpy() for t in s.fields_torch())
    finally:
        s.close()
    assert_match("rho", r, st.rho)
    assert_match("u", u, st.u)


@pytest.mark.parametrize("mode", ["periodic", "dirichlet"])
def test_general_thermal_modes(gpu, mode):
    print(f"[parity] general thermal {mode}", flush=True)
    from xadvcool_amd.engine import LatticeSolver
    rng = np.random.default_rng(11)
    if mode == "periodic":
        m = np.zeros((20, 4, 5), np.int32)
        m[5:8] = BC_SOLID
        kw_o = dict(thermal_stream=1)
        kw_h = dict(periodic=(1, 1, 1))
    else:
        m = np.zeros((20, 4, 5), np.int32)
        m[:10] = BC_SOLID
        m[0] = BC_INLET
        m[-1] = BC_INLET
        kw_o = dict(thermal_stream=0, dirichlet_x=0, dirichlet_T=100.0)
        kw_h = dict(dirichlet=(0, 100.0))
    u = rng.uniform(-0.04, 0.04, size=m.shape + (3,)).astype(np.float32)
    T0 = rng.uniform(0.0, 10.0, size=m.shape).astype(np.float32)
    steps = 17
    st = orc.State(m.shape, np.float32)
    st.u[...] = u
    st.init_thermal(T0, m)
    orc.run(st, m, steps, thermal=2, tau_th_fluid=0.6, tau_th_solid=0.8, T_in=0.0, **kw_o)
    s = LatticeSolver(m, fluid=False, thermal=True, tau_th_fluid=0.6, tau_th_solid=0.8, T_in=0.0,
                      frozen_velocity=True, **kw_h)
    try:
        s.init_equilibrium(None, u, T0)
        s.step(steps)
        T = s.fields_torch()[2].cpu().numpy()
    finally:
        s.close()
    assert_match("T", T, st.T)


def test_stats_match_oracle(gpu):
    """The 16 on-device reduction slots against the oracle's ordered FP64 sums."""
    print("[parity] stats", flush=True)
    from xadvcool_amd.engine import LatticeSolver
    p = coupled_params(seed=9)
    ref = oracle_coupled(p, 30, np.float32)
    want = orc.stats(ref.rho, ref.u, ref.T, p["mask"])
    s = LatticeSolver(p["mask"], fluid=True, thermal=True, heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"],
                      tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"])
    try:
        s.init_equilibrium()
        s.step(30)
        got = np.array(list(s.stats().values()))
    finally:
        s.close()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_compact_storage_limits(gpu):
    """Compact fluid storage needs nz <= 32: a taller lattice is refused with a message, and the
    automatic choice of a batch sized for it is dense storage (which then solves it)."""
    print("[parity] compact storage limits", flush=True)
    from xadvcool_amd.engine import LatticeBatch
    m = np.zeros((6, 4, 40), np.int32)
    m[:, :, 0] = BC_SOLID
    b = LatticeBatch(slots=1, cap_shape=(6, 4, 40), fluid_storage="compact")
    try:
        with pytest.raises(ValueError, match="nz <= 32"):
            b.load(0, m, tau=0.6)
    finally:
        b.close()
    b = LatticeBatch(slots=1, cap_shape=(6, 4, 40))
    try:
        b.load(0, m, tau=0.6)
        b.init(0)
        b.step(4)
    finally:
        b.close()
