"""Dataset generation on the batched solver, with the reference's schema.

Restates generate_dataset.py:305-597 around ``BatchedSolver``: the same
Latin-hypercube design points (shuffled per design family and interleaved),
md5 sample ids, geometry validity gate, solve protocol (run_simulation's
residual rule with the flow-through floor), divergence and plausibility gates,
metrics and the 39 metadata columns.  Storage: h5py is not available here, so
each accepted sample is one ``<sample_id>.npz`` holding the reference's HDF5
datasets (mask int8, heat_source, velocity, pressure, temperature; FP32) plus
its attributes as JSON, next to ``metadata.parquet`` and
``failed_samples.json``; resume skips ids already present in either.

The gaussian_hotspots maps of the reference used Python's salted string hash as
seed (generate_dataset.py:406), which no run can reproduce; here the seed is
the first 8 hex digits of the sample id's md5.

With ``surrogate`` (not a reference option) every design starts from the
surrogate's predicted fields under the same protocol (the reference's
residuals at 1e-5 on every check, the flow-through floor, the step cap); a
warm run that stalls, diverges or converges to fields failing the physics
gate is re-run cold (``warmstart.solve_designs``).  Samples, metadata and
attributes have the cold protocol's form; their stop step, wall time and
MLUPS are the warm run's.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import geometry as G
from .batch import BatchedSolver

# compute_macroscopic_metrics's keys (generate_dataset.py:167-179)
METRIC_KEYS = ("R_th", "Nu", "f_friction", "P_pump", "COP", "Q_vol", "T_max_surface", "T_avg_surface",
               "delta_T_max", "sigma_T", "fluid_temp_rise")
COOLANT_PR = 7.0
COOLANT_K_RATIO = 628.0
BUF = 10


def ordered_samples(samples_per_design, seed=42):
    """Design points in the reference's order (generate_dataset.py:345-356)."""
    rng = np.random.default_rng(seed)
    per = {}
    for name in G.DESIGN_CONFIGS:
        s = G.generate_samples(name, samples_per_design, seed=seed)
        rng.shuffle(s)
        per[name] = s
    out = []
    for i in range(max(len(v) for v in per.values())):
        for name in G.DESIGN_CONFIGS:
            if i < len(per[name]):
                out.append(per[name][i])
    return out


def _metadata_row(sid, p, d, r, met, elapsed, gpu_name, Pr, k_ratio, max_v, attrs):
    n = int(np.prod(d["mask"].shape))
    return {
        "sample_id": sid, "Pr": Pr, "k_ratio": k_ratio, "design_name": p["design_name"],
        "design_idx": p["design_idx"], "heat_source_type": p["heat_source_type"],
        "feature_size": p["feature_size"], "spacing": p["spacing"], "tau_fluid": p["tau_fluid"],
        "u_inlet_val": p["u_inlet_val"], "heat_power": p["heat_power"], "porosity": d["porosity"], "Dh": d["Dh"],
        "Re": d["Re"], "max_velocity": max_v, "max_temperature": attrs["max_temperature"],
        "mean_temperature": attrs["mean_temperature"], "pressure_drop": met["pressure_drop"], "elapsed_s": elapsed,
        "sim_steps": r["steps"], "converged": r["converged"], "grid_nx": d["mask"].shape[0],
        "grid_ny": d["mask"].shape[1], "grid_nz": d["mask"].shape[2], "grid_cells": n,
        "total_lattice_updates": n * r["steps"], "mlups": n * r["steps"] / (elapsed * 1e6) if elapsed > 0 else 0.0,
        "gpu_name": gpu_name, "R_th": met["R_th"], "Nu": met["Nu"], "f_friction": met["f_friction"],
        "P_pump": met["P_pump"], "COP": met["COP"], "Q_vol": met["Q_vol"], "T_max_surface": met["T_max_surface"],
        "T_avg_surface": met["T_avg_surface"], "delta_T_max": met["delta_T_max"], "sigma_T": met["sigma_T"],
        "fluid_temp_rise": met["fluid_temp_rise"],
    }


def plausibility(met, converged):
    """generate_dataset.py:458-485: reasons a converged sample is rejected."""
    reasons = []
    Nu, R_th, COP, rise = met["Nu"], met["R_th"], met["COP"], met["fluid_temp_rise"]
    if Nu < 0 or Nu > 2000 or np.isnan(Nu):
        reasons.append(f"Nu={Nu:.1f}")
    if R_th < 0 or R_th > 1000 or np.isnan(R_th):
        reasons.append(f"R_th={R_th:.1f}")
    if COP < 0 or np.isnan(COP):
        reasons.append(f"COP={COP:.2e}")
    if rise < 0 or np.isnan(rise):
        reasons.append(f"fluid_temp_rise={rise:.2e}")
    if not converged:
        reasons.append("not converged")
    return reasons


def generate_dataset(output_dir, samples_per_design=50, max_steps=30000, seed=42, resume=True, slots=64,
                     core=(G.CORE_NX, G.CORE_NY, G.CORE_NZ), limit=None, log=print, surrogate=None, warm_chunk=256):
    """Solve the design points (up to ``limit``) and write samples, metadata.parquet and failed_samples.json.

    ``surrogate`` (``warmstart.load_surrogate()``): warm-start every design, ``warm_chunk`` designs per
    batched run (their predicted fields are held on the GPU until their slot starts); the lattice must be the
    surrogate's (the default core).  Returns (rows, failed) and, with a surrogate, logs the fallbacks."""
    import pandas as pd
    import torch
    if surrogate is not None and tuple(core) != (G.CORE_NX, G.CORE_NY, G.CORE_NZ):
        raise ValueError("the surrogate predicts fields on the default core only")
    Pr, k_ratio = COOLANT_PR, COOLANT_K_RATIO
    os.makedirs(output_dir, exist_ok=True)
    meta_path = os.path.join(output_dir, "metadata.parquet")
    failed_path = os.path.join(output_dir, "failed_samples.json")
    existing = {f[:-4] for f in os.listdir(output_dir) if f.endswith(".npz")} if resume else set()
    failed = json.load(open(failed_path)) if (resume and os.path.exists(failed_path)) else {}
    rows = pd.read_parquet(meta_path).to_dict("records") if (resume and os.path.exists(meta_path)) else []
    gpu_name = torch.cuda.get_device_name(0)

    designs = []
    for p in ordered_samples(samples_per_design, seed):
        sid = G.sample_id(p)
        if sid in existing or sid in failed:
            continue
        mask = G.generate_mask({"design_type": p["design_idx"], "feature_size": p["feature_size"],
                                "spacing": p["spacing"], "buffer_size": BUF}, core=core)
        ok, porosity = G.is_valid_geometry(mask)
        if not ok:
            failed[sid] = f"invalid geometry (porosity={porosity:.2%})"
            continue
        Dh, _, _ = G.compute_hydraulic_diameter(mask)
        Re = p["u_inlet_val"] * Dh / ((p["tau_fluid"] - 0.5) / 3.0)
        hs_seed = int(sid.rsplit("_", 1)[1][:8], 16)
        heat = G.heat_source_3d(mask.shape, p["heat_source_type"], core=core[:2], seed=hs_seed)
        designs.append(dict(mask=mask, heat=heat, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"],
                            heat_power=p["heat_power"], key=sid, params=p, porosity=porosity, Dh=Dh, Re=Re))
        if limit is not None and len(designs) >= limit:
            break
    by_key = {d["key"]: d for d in designs}
    log(f"[dataset] {len(designs)} designs to solve, {len(existing)} present, {len(failed)} failed before")
    bs = None
    if surrogate is None:
        cap = (tuple(max(d["mask"].shape[k] for d in designs) for k in range(3)) if designs else
               (core[0] + 2 * BUF, core[1], core[2]))
        bs = BatchedSolver(slots, cap, tol_ke=1e-5, tol_T=1e-5, max_steps=max_steps, min_steps_rule=True, Pr=Pr,
                           k_ratio=k_ratio, collect_fields=True, metrics=True)

    def on_result(r):
        d = by_key[r["key"]]
        sid, p = r["key"], d["params"]
        rho, u, T = r.pop("fields")
        met = r["metrics"]
        max_v = float(np.max(np.linalg.norm(u, axis=-1)))
        max_T = float(np.max(T))
        Ma = max_v * 3.0 ** 0.5
        if np.isnan(max_v) or Ma > 0.3 or np.isnan(max_T) or max_T > 1e4:
            failed[sid] = f"diverged (max_v={max_v:.2e}, Ma={Ma:.2f}, max_T={max_T:.2e})"
            return
        reasons = plausibility(met, r["converged"])
        if reasons:
            failed[sid] = f"unphysical ({', '.join(reasons)})"
            return
        # attribute order and precedence of generate_dataset.py:495-517 (the metrics are written last)
        n_cells = int(d["mask"].size)
        attrs = dict(p)
        attrs.update(coolant="water", Pr=Pr, k_ratio=k_ratio, porosity=d["porosity"], Dh=d["Dh"], Re=d["Re"],
                     max_velocity=max_v, max_temperature=max_T, mean_temperature=float(np.mean(T)),
                     grid_shape=list(d["mask"].shape), elapsed_s=r["wall_s"], sim_steps=r["steps"],
                     converged=r["converged"], grid_cells=n_cells, total_lattice_updates=n_cells * r["steps"],
                     mlups=n_cells * r["steps"] / (r["wall_s"] * 1e6) if r["wall_s"] > 0 else 0.0,
                     gpu_name=gpu_name)
        attrs.update({k: met[k] for k in METRIC_KEYS})
        np.savez_compressed(os.path.join(output_dir, sid + ".npz"), mask=d["mask"].astype(np.int8),
                            heat_source=d["heat"].astype(np.float32), velocity=u.astype(np.float32),
                            pressure=(rho / 3.0).astype(np.float32), temperature=T.astype(np.float32),
                            attrs=json.dumps(attrs, default=float))
        rows.append(_metadata_row(sid, p, d, r, met, r["wall_s"], gpu_name, Pr, k_ratio, max_v, attrs))

    t0 = time.perf_counter()
    if surrogate is not None:
        from . import warmstart as W
        nfb = 0
        for i in range(0, len(designs), warm_chunk):
            out = W.solve_designs(designs[i:i + warm_chunk], surrogate, slots=slots, tol=1e-5, max_steps=max_steps,
                                  log=None, confirm=1, stop_rule="reference", min_steps_rule=True,
                                  collect_fields=True, on_final=on_result)
            nfb += len(out["fallbacks"])
            for k, why, steps in out["fallbacks"]:
                log(f"[dataset] {k}: warm run abandoned after {steps} steps ({why}), solved cold")
        log(f"[dataset] warm-started {len(designs)} designs, {nfb} re-run cold")
    else:
        try:
            if designs:
                bs.run(designs, on_result=on_result)
        finally:
            bs.close()
    if rows:
        df = pd.DataFrame(rows).drop_duplicates(subset=["sample_id"], keep="last")
        df.to_parquet(meta_path, index=False)
    with open(failed_path, "w") as f:
        json.dump(failed, f, indent=1)
    log(f"[dataset] wrote {len(rows)} samples, {len(failed)} failed, {time.perf_counter() - t0:.1f} s")
    return rows, failed
