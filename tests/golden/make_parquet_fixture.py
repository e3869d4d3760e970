"""Extract a stratified set of rows of the reference's dataset/metadata.parquet as a JSON fixture.

The parquet holds converged scalars of the reference solver (Warp FP32 on an
RTX 3090) for 8,071 designs whose parameters regenerate exactly from the
Latin-hypercube sampler (xadvcool_amd.geometry.generate_samples, seed 42,
200 points per design family).  gaussian_hotspots rows are skipped: their
heat maps used a salted Python hash as seed.  Run in a checkout that has the
reference at /root/reference; the GPU tests read only the JSON.
"""
import json
import os
import sys

import pandas as pd

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

PARQUET = "/root/reference/dataset/metadata.parquet"
COLS = ["sample_id", "design_name", "design_idx", "heat_source_type", "feature_size", "spacing", "tau_fluid",
        "u_inlet_val", "heat_power", "porosity", "Dh", "Re", "sim_steps", "converged", "pressure_drop", "P_pump",
        "Q_vol", "R_th", "Nu", "f_friction", "COP", "T_max_surface", "T_avg_surface", "delta_T_max", "sigma_T",
        "fluid_temp_rise", "max_velocity", "max_temperature", "mean_temperature", "elapsed_s", "mlups"]


def main(per_family=3):
    df = pd.read_parquet(PARQUET)
    df = df[df.heat_source_type != "gaussian_hotspots"]
    rows = []
    for fam, g in df.groupby("design_name"):
        g = g.sort_values("sim_steps").reset_index(drop=True)
        for q in [0.2, 0.5, 0.8][:per_family]:
            rows.append(g.iloc[int(q * (len(g) - 1))][COLS].to_dict())
    for r in rows:
        for k, v in r.items():
            if hasattr(v, "item"):
                r[k] = v.item()
    out = os.path.join(HERE, "parquet_rows.json")
    with open(out, "w") as f:
        json.dump({"source": "reference dataset/metadata.parquet (Warp FP32, RTX 3090)",
                   "protocol": {"max_steps": 30000, "tol_ke": 1e-5, "tol_T": 1e-5, "stats_every": 100, "Pr": 7.0,
                                "k_ratio": 628.0, "floor": True},
                   "rows": rows}, f, indent=1)
    print(out, len(rows))


if __name__ == "__main__":
    main()
