"""Dataset writer: the reference's metadata schema and sample layout from a small batched run."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# Columns of the reference's metadata.parquet (generate_dataset.py:538-579).
COLUMNS = ['sample_id', 'Pr', 'k_ratio', 'design_name', 'design_idx', 'heat_source_type', 'feature_size', 'spacing',
           'tau_fluid', 'u_inlet_val', 'heat_power', 'porosity', 'Dh', 'Re', 'max_velocity', 'max_temperature',
           'mean_temperature', 'pressure_drop', 'elapsed_s', 'sim_steps', 'converged', 'grid_nx', 'grid_ny', 'grid_nz',
           'grid_cells', 'total_lattice_updates', 'mlups', 'gpu_name', 'R_th', 'Nu', 'f_friction', 'P_pump', 'COP',
           'Q_vol', 'T_max_surface', 'T_avg_surface', 'delta_T_max', 'sigma_T', 'fluid_temp_rise']


def test_small_dataset(gpu, tmp_path):
    import pandas as pd
    from xadvcool_amd.dataset import generate_dataset
    print("[dataset] small run", flush=True)
    kw = dict(samples_per_design=1, max_steps=6000, slots=8, core=(32, 24, 12))
    rows, failed = generate_dataset(str(tmp_path), limit=6, **kw)
    assert len(rows) + len(failed) >= 6 and rows
    df = pd.read_parquet(tmp_path / "metadata.parquet")
    assert list(df.columns) == COLUMNS
    for sid in df.sample_id:
        z = np.load(tmp_path / f"{sid}.npz")
        assert z["mask"].dtype == np.int8 and z["velocity"].shape == z["mask"].shape + (3,)
        assert z["pressure"].dtype == np.float32 and z["temperature"].shape == z["mask"].shape
        attrs = json.loads(str(z["attrs"]))
        assert attrs["sim_steps"] == int(df[df.sample_id == sid].sim_steps.iloc[0])
        assert np.isclose(attrs["R_th"], float(df[df.sample_id == sid].R_th.iloc[0]))
    # resume: solved and failed ids are skipped, new ones are appended
    print("[dataset] resume", flush=True)
    _, failed2 = generate_dataset(str(tmp_path), limit=4, **kw)
    df2 = pd.read_parquet(tmp_path / "metadata.parquet")
    assert set(df.sample_id) <= set(df2.sample_id)
    assert len(df2) == len(set(df2.sample_id))
    new = (set(df2.sample_id) - set(df.sample_id)) | (set(failed2) - set(failed))
    assert len(new) == 4


def test_warm_started_dataset_matches_the_cold_one(gpu, tmp_path):
    """generate_dataset with the shipped surrogate against the cold run of the same design points (full-size
    lattices, the reference's protocol with its floor): the same schema; every cold sample is also a warm one,
    and a design only the warm run keeps was rejected cold as not converged within the step cap; the warm runs
    stop no later than the cold ones on average; the pressure drop and R_th of the two runs are close (the
    reference's residuals at 1e-5 stop a run short of its steady state, from either side, so the bound is
    loose: 5 %)."""
    import pandas as pd
    from xadvcool_amd import warmstart as W
    from xadvcool_amd.dataset import generate_dataset
    kw = dict(samples_per_design=1, max_steps=20000, slots=3, limit=3, log=lambda m: print(m, flush=True))
    print("[dataset] cold, 3 full-size designs", flush=True)
    rows_c, failed_c = generate_dataset(str(tmp_path / "cold"), **kw)
    print("[dataset] warm-started, the same designs", flush=True)
    rows_w, failed_w = generate_dataset(str(tmp_path / "warm"), surrogate=W.load_surrogate(), **kw)
    dc = pd.read_parquet(tmp_path / "cold" / "metadata.parquet").set_index("sample_id")
    dw = pd.read_parquet(tmp_path / "warm" / "metadata.parquet").set_index("sample_id")
    assert ["sample_id"] + list(dw.columns) == COLUMNS
    assert set(dc.index) | set(failed_c) == set(dw.index) | set(failed_w) and len(dc) >= 2
    assert set(dc.index) <= set(dw.index) and set(failed_w) <= set(failed_c)
    for sid in set(dw.index) - set(dc.index):
        assert "not converged" in failed_c[sid], failed_c[sid]
    common = sorted(dc.index)
    assert dw.sim_steps[common].mean() <= dc.sim_steps[common].mean()
    for sid in common:
        if dc.converged[sid] and dw.converged[sid]:
            assert abs(dw.pressure_drop[sid] / dc.pressure_drop[sid] - 1) < 0.05, sid
            assert abs(dw.R_th[sid] / dc.R_th[sid] - 1) < 0.05, sid
    with pytest.raises(ValueError):
        generate_dataset(str(tmp_path / "bad"), surrogate=W.load_surrogate(), core=(32, 24, 12), limit=1)
