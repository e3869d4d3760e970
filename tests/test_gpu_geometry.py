"""Designs built on the device: masks, heat maps and slot loading.

* Masks from xadv_geometry_mask are compared with the host restatement of
  src/geometry.py (``design_arrays`` / ``generate_mask``) for every family,
  random core sizes, and the reference's feature_size/spacing mapping: they
  must be the same integers.
* Heat fields from xadv_heat_map are compared with the host generators
  (``heat_source_3d``, which calls scipy.ndimage.gaussian_filter): maps
  without Gaussian hot spots use only + and * on the device and must be the
  same FP32 values; gaussian_hotspots also evaluates exp on the device and is
  held to 1e-6 relative (a last-bit difference of exp).
* A slot loaded by xadv_instance_load_device must step to the same bits as
  the same design loaded from host arrays by xadv_instance_load, in compact
  and dense storage and in the general (periodic / Dirichlet) kernels; a
  rejected device load leaves the slot as it was.
"""
import numpy as np
import pytest
import torch

from xadvcool_amd import geometry as G
from xadvcool_amd.engine import LatticeBatch

pytestmark = pytest.mark.gpu

CAP = (148, 128, 32)


@pytest.fixture(scope="module")
def gen_batch():
    b = LatticeBatch(slots=1, cap_shape=CAP)
    yield b
    b.close()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_device_masks_match_host(gen_batch):
    n = 0
    for family in G.FAMILIES:
        for seed in range(5):
            core = None if seed < 4 else (128, 128, 32)
            p = G.family_params(1000 + seed, core=core, family=family, heat_name="uniform")
            print(f"mask {family} seed {1000 + seed} core {p['core']}", flush=True)
            ref, _ = G.design_arrays(p)
            got, _ = G.device_design(gen_batch, p)
            got = _np(got)
            assert got.shape == ref.shape
            bad = np.argwhere(got != ref)
            assert bad.size == 0, f"{family} seed {1000 + seed}: {len(bad)} cells differ, first {tuple(bad[0])}"
            n += 1
    # The reference's design_idx mapping (feature_size, spacing -> channels, pins, TPMS level and cell size).
    rng = np.random.default_rng(7)
    for idx in range(6):
        name = G.DESIGN_NAMES[idx]
        ranges = G.DESIGN_CONFIGS[name]["ranges"] if name in G.DESIGN_CONFIGS else {"feature_size": (2, 8),
                                                                                   "spacing": (2, 7)}
        for _ in range(3):
            fs = float(rng.uniform(*ranges["feature_size"]))
            sp = float(rng.uniform(*ranges["spacing"]))
            print(f"mask design_idx {idx} feature_size {fs:.4f} spacing {sp:.4f}", flush=True)
            ref = G.generate_mask({"design_type": idx, "feature_size": fs, "spacing": sp, "buffer_size": G.BUFFER})
            got = _np(gen_batch.geometry_mask(ref.shape, idx, buffer=G.BUFFER, base=G.BASE_THICKNESS,
                                              top=G.TOP_THICKNESS, feature_size=fs, spacing=sp, period=0.0,
                                              manifold=5))
            assert np.array_equal(got, ref), f"design_idx {idx}: {int((got != ref).sum())} cells differ"
            n += 1
    assert n == 5 * 5 + 6 * 3


@pytest.mark.parametrize("mx,my", [(128, 128), (76, 64), (100, 97), (128, 80)])
def test_device_heat_maps_match_host(gen_batch, mx, my):
    shape = (mx + 2 * G.BUFFER, my, 5)
    for name in G.HEAT_SOURCE_NAMES:
        for seed in ([0, 7, 12345] if name == "gaussian_hotspots" else [0]):
            print(f"heat {name} {mx}x{my} seed {seed}", flush=True)
            ref = G.heat_source_3d(shape, name, core=(mx, my), seed=seed)
            got = _np(gen_batch.heat_map(G.heat_program(name, mx, my, seed), shape, G.BUFFER))
            if name == "gaussian_hotspots":
                rel = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)
                assert float(rel.max()) <= 1e-6, f"{name} seed {seed}: max rel diff {rel.max():.3e}"
                assert np.array_equal(got == 0, ref == 0)
            else:
                bad = np.argwhere(got != ref)
                assert bad.size == 0, (f"{name} {mx}x{my}: {len(bad)} values differ, first {tuple(bad[0])}: "
                                       f"{got[tuple(bad[0])]!r} vs {ref[tuple(bad[0])]!r}")


def _solve_params(p):
    tau = p["tau_fluid"]
    ttf = 0.5 + (tau - 0.5) / 7.0
    return dict(tau=tau, tau_th_fluid=ttf, tau_th_solid=0.5 + 628.0 * (ttf - 0.5), heat_power=p["heat_power"],
                T_in=25.0, rho_out=1.0, u_in=p["u_inlet_val"])


def _assert_same_slots(b, s0, s1, what):
    b.readout(True)
    st = b.stats()
    f0, f1 = b.fields(s0), b.fields(s1)
    for name, a, c in zip(("rho", "u", "T"), f0, f1):
        a, c = _np(a), _np(c)
        bad = np.argwhere(a != c)
        assert bad.size == 0, f"{what}: {name} differs at {len(bad)} cells, first {tuple(bad[0])}"
    assert np.array_equal(st[s0], st[s1]), f"{what}: stats differ"


@pytest.mark.parametrize("storage", ["compact", "dense"])
def test_device_load_matches_host_load(storage):
    b = LatticeBatch(slots=3, cap_shape=CAP, fluid_storage=storage)
    try:
        for seed, family, heat in [(11, "gyroid", "quad_core"), (12, "pinfin", "gaussian_hotspots"),
                                   (13, "schwarz_d", "peripheral_ring"), (14, "straight", "uniform")]:
            p = G.family_params(seed, family=family, heat_name=heat)
            print(f"load {storage} {family} {heat} core {p['core']}", flush=True)
            mask, hs = G.design_arrays(p)
            kw = _solve_params(p)
            b.load(0, mask, heat=hs, **kw)
            # Same host arrays uploaded: isolates the device loader.
            b.load_device(1, torch.from_numpy(mask.astype(np.int8)).cuda(), heat=torch.from_numpy(hs).cuda(), **kw)
            # Fully device-built design.
            dm, dh = G.device_design(b, p)
            if heat != "gaussian_hotspots":
                assert np.array_equal(_np(dh), hs)
            b.load_device(2, dm, heat=dh if heat != "gaussian_hotspots" else torch.from_numpy(hs).cuda(), **kw)
            for s in range(3):
                b.init(s)
            b.step(200)
            _assert_same_slots(b, 0, 1, f"{storage} {family} device load")
            _assert_same_slots(b, 0, 2, f"{storage} {family} device-built design")
            b.load_device(1, torch.from_numpy(mask.astype(np.int8)).cuda(), heat=None, **kw)
            b.load(0, mask, heat=None, **kw)
            b.init(0)
            b.init(1)
            b.set_active(2, 0)
            b.step(100)
            _assert_same_slots(b, 0, 1, f"{storage} {family} no heat")
    finally:
        b.close()


def test_device_load_general_kernels():
    rng = np.random.default_rng(3)
    m = np.zeros((20, 12, 9), np.int8)
    for _ in range(5):
        x, y = rng.integers(2, 17), rng.integers(1, 10)
        m[x:x + 2, y:y + 2, 2:6] = 1
    heat = np.zeros(m.shape, np.float32)
    heat[4:12, 3:9, 0] = 1.0
    b = LatticeBatch(slots=2, cap_shape=m.shape, general=True)
    try:
        for kw in [dict(periodic=(1, 1, 0), force_x=1e-5, u_wall=0.01),
                   dict(periodic=(0, 1, 1), dirichlet=(0, 1.0), heat_power=0.01)]:
            print(f"general {kw}", flush=True)
            b.load(0, m, heat=heat, tau=0.7, tau_th_fluid=0.8, tau_th_solid=0.9, **kw)
            b.load_device(1, torch.from_numpy(m).cuda(), heat=torch.from_numpy(heat).cuda(), tau=0.7,
                          tau_th_fluid=0.8, tau_th_solid=0.9, **kw)
            b.init(0)
            b.init(1)
            b.step(60)
            _assert_same_slots(b, 0, 1, f"general {kw}")
    finally:
        b.close()


def test_device_load_rejects_bad_masks_and_keeps_the_slot():
    p = G.family_params(21, family="pinfin", heat_name="dual_core")
    mask, hs = G.design_arrays(p)
    kw = _solve_params(p)
    b = LatticeBatch(slots=2, cap_shape=CAP)
    try:
        dm = torch.from_numpy(mask.astype(np.int8)).cuda()
        dh = torch.from_numpy(hs).cuda()
        b.load_device(0, dm, heat=dh, **kw)
        b.load(1, mask, heat=hs, **kw)
        b.init(0)
        b.init(1)
        b.step(100)
        bad_code = mask.copy()
        bad_code[30, 40, 10] = 7
        bad_inlet = mask.copy()
        bad_inlet[5, 40, 10] = 2
        bad_outlet = mask.copy()
        bad_outlet[50, 40, 10] = 3
        for bad, msg in [(bad_code, "code other than 0..3"), (bad_inlet, "INLET cells must lie on x = 0"),
                         (bad_outlet, "OUTLET cells must lie on x = nx-1")]:
            print(f"rejects: {msg}", flush=True)
            with pytest.raises(ValueError, match=msg):
                b.load_device(0, torch.from_numpy(bad.astype(np.int8)).cuda(), heat=dh, **kw)
        b.step(100)
        _assert_same_slots(b, 0, 1, "after rejected loads")
    finally:
        b.close()
    with pytest.raises(ValueError, match="non-periodic x axis"):
        g = LatticeBatch(slots=1, cap_shape=(20, 12, 9), general=True)
        try:
            m = np.zeros((20, 12, 9), np.int8)
            m[0] = 2
            m[-1] = 3
            g.load_device(0, torch.from_numpy(m).cuda(), periodic=(1, 0, 0))
        finally:
            g.close()


def test_device_features_equal_host_features():
    from xadvcool_amd import surrogate as S
    from xadvcool_amd.batch import device_inputs
    for seed, family, heat in [(31, "gyroid", "quad_core"), (32, "straight", "gaussian_hotspots"),
                               (33, "schwarz_p", "uniform")]:
        p = G.family_params(seed, family=family, heat_name=heat)
        mask, hs = G.design_arrays(p)
        print(f"features {family} {heat} core {p['core']}", flush=True)
        ref = S.features(mask, hs, p["u_inlet_val"], p["tau_fluid"], p["heat_power"], (160, 128, 32))
        m, h = device_inputs(dict(mask=mask, heat=hs))
        got = _np(S.features_device(m, h, p["u_inlet_val"], p["tau_fluid"], p["heat_power"], (160, 128, 32)))
        bad = np.argwhere(got != ref)
        assert bad.size == 0, f"{family}: {len(bad)} feature values differ, first {tuple(bad[0])}"


def test_solver_on_device_built_designs_matches_host_designs():
    """BatchedSolver on parameter-only designs (built on the GPU when a slot starts) gives the same steps,
    metrics and fields as on the same designs' host arrays."""
    from xadvcool_amd.batch import BatchedSolver
    params = [G.config2_params(1000 + i) for i in range(3)] + [G.family_params(41, core=(128, 128, 32),
                                                                               family="pinfin", heat_name="dual_core")]
    host, dev = [], []
    for i, p in enumerate(params):
        mask, hs = G.design_arrays(p)
        common = dict(key=i, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"], heat_power=p["heat_power"])
        host.append(dict(mask=mask, heat=hs, **common))
        dev.append(dict(params=p, **common))
    out = []
    for designs in (host, dev):
        bs = BatchedSolver(2, CAP, max_steps=600, tol_ke=1e-6, tol_T=1e-6, collect_fields=True)
        try:
            out.append({r["key"]: r for r in bs.run(designs)})
        finally:
            bs.close()
    for k in range(len(params)):
        a, b = out[0][k], out[1][k]
        print(f"solver design {k}: {a['status']} {a['steps']}", flush=True)
        assert (a["status"], a["steps"]) == (b["status"], b["steps"])
        assert a["metrics"] == b["metrics"]
        for fa, fb in zip(a["fields"], b["fields"]):
            assert np.array_equal(fa, fb)
    assert "mask" not in dev[0] and isinstance(host[0]["mask"], np.ndarray)
