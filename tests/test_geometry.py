"""The pyvista-free geometry and heat-map restatement against the reference's dataset files (no GPU).

Golden data (tests/golden/geometry_rows.json, from dataset/metadata.parquet
and dataset/failed_samples.json): design ids regenerated by the restated
Latin-hypercube sampler, porosity and D_h of accepted designs (exact), the
printed porosity of rejected designs (to the 0.01 % the file holds).
"""
import json
import os

import numpy as np
import pytest

from xadvcool_amd import geometry as G

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "geometry_rows.json")))


def _mask(r):
    return G.generate_mask({"design_type": r["design_idx"], "feature_size": r["feature_size"],
                            "spacing": r["spacing"], "buffer_size": 10})


def test_sampler_regenerates_dataset_ids():
    ids = {G.sample_id(p) for name in G.DESIGN_CONFIGS for p in G.generate_samples(name, 200, seed=42)}
    assert len(ids) == 10000
    for group in ("accepted", "rejected", "diverged"):
        for r in GOLD[group]:
            assert r["sample_id"] in ids


@pytest.mark.parametrize("idx", range(0, 120, 4))
def test_accepted_porosity_and_dh_exact(idx):
    for r in GOLD["accepted"][idx:idx + 4]:
        m = _mask(r)
        assert G.compute_porosity(m) == r["porosity"]
        assert G.compute_hydraulic_diameter(m)[0] == r["Dh"]
        assert G.is_valid_geometry(m)[0]


def test_rejected_designs_are_rejected():
    for r in GOLD["rejected"]:
        m = _mask(r)
        ok, por = G.is_valid_geometry(m)
        assert not ok
        assert round(por * 100.0, 2) == pytest.approx(r["porosity_pct_2dp"] * 100.0, abs=5e-3)


def test_mask_frame():
    m = _mask(GOLD["accepted"][0])
    assert m.shape == (148, 128, 32)
    assert np.all(m[:, 0, :] == 1) and np.all(m[:, -1, :] == 1)
    assert np.all(m[:, :, :4] == 1) and np.all(m[:, :, -1] == 1)
    assert np.all(m[0, 1:-1, 4:-1] == 2) and np.all(m[-1, 1:-1, 4:-1] == 3)
    assert set(np.unique(m[0])) <= {1, 2} and set(np.unique(m[-1])) <= {1, 3}


@pytest.mark.parametrize("name", G.HEAT_SOURCE_NAMES)
def test_heat_maps_normalised(name):
    hm = G.HEAT_SOURCE_REGISTRY[name](128, 128, seed=3)
    assert hm.shape == (128, 128) and hm.dtype == np.float32
    assert hm.max() == pytest.approx(1.0) and hm.min() >= 0.0


def test_workload_generators():
    d = G.config2_design(0)
    assert d["mask"].shape == (148, 128, 32)
    assert 16 <= d["period"] <= 32 and -0.4 <= d["threshold"] <= 0.4 and 0.01 <= d["u_inlet_val"] <= 0.05
    for s in range(6):
        d = G.family_design(s)
        nx, ny, nz = d["mask"].shape
        assert 96 <= nx <= 148 and 64 <= ny <= 128 and 16 <= nz <= 32
        assert d["heat"].shape == d["mask"].shape


@pytest.mark.parametrize("mx,my", [(128, 128), (76, 64), (100, 97), (128, 80), (64, 128)])
def test_heat_programs_reproduce_the_generators(mx, my):
    """The paint programs the device heat-map builder takes, rendered on the host in the device's order of
    operations (FP64 blur sums: centre, then the symmetric pairs from the outermost in), give the generators'
    exact FP32 values, scipy.ndimage.gaussian_filter included."""
    for name in G.HEAT_SOURCE_NAMES:
        for seed in ([0, 7, 12345] if name == "gaussian_hotspots" else [0]):
            prog = G.heat_program(name, mx, my, seed)
            assert len(prog["prims"]) <= 16
            if prog["sigma"] is not None:
                assert G.gauss_weights(prog["sigma"])[0] <= 63
            ref = G.HEAT_SOURCE_REGISTRY[name](mx, my, seed=seed)
            got = G.render_heat_program(prog)
            assert got.dtype == np.float32
            assert np.array_equal(got, ref), f"{name} {mx}x{my} seed {seed}: {int((got != ref).sum())} differ"


def test_gauss_weights_match_scipy():
    from scipy.ndimage import gaussian_filter1d
    for sigma in (1.28, 1.6, 128 / 70.0, 128 / 60.0, 0.7, 3.3):
        r, w = G.gauss_weights(sigma)
        impulse = np.zeros(4 * r + 3)
        impulse[2 * r + 1] = 1.0
        resp = gaussian_filter1d(impulse, sigma)
        assert np.array_equal(resp[2 * r + 1:3 * r + 2], w)
        assert np.array_equal(resp[2 * r + 1::-1][:r + 1], w)
        assert abs(w[0] + 2 * w[1:].sum() - 1.0) < 1e-15


def test_family_params_drive_family_design():
    for seed in range(12):
        d = G.family_design(seed)
        p = G.family_params(seed)
        mask, heat = G.design_arrays(p)
        assert np.array_equal(mask, d["mask"]) and np.array_equal(heat, d["heat"])
        assert (p["family"], p["heat_name"], p["tau_fluid"], p["u_inlet_val"]) == \
            (d["family"], d["heat_name"], d["tau_fluid"], d["u_inlet_val"])
        assert mask.shape == (p["core"][0] + 2 * G.BUFFER, p["core"][1], p["core"][2])
