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
