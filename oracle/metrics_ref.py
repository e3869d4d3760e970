"""Host numpy restatement of the reference's post-solve metrics (test infrastructure only).

compute_macroscopic_metrics follows generate_dataset.py:80-179 line for line,
and the metadata reductions follow generate_dataset.py:438-441,553-556.  The
checker for xadvcool_amd.metrics.device_metrics.
"""
from __future__ import annotations

import numpy as np

BC_SOLID, BC_INLET, BC_OUTLET = 1, 2, 3
CORE_NX, CORE_NY = 128, 128


def compute_macroscopic_metrics(results, mask, params, Dh, A_wetted, Pr=7.0, core=(CORE_NX, CORE_NY)):
    u_np, rho_np, T_np = results["u"], results["rho"], results["T"]
    nx, ny, nz = mask.shape
    t_inlet = 25.0
    outlet_bc = (mask[-1, :, :] == BC_OUTLET)
    inlet_fluid = (mask[1, :, :] != BC_SOLID)
    outlet_fluid = (mask[-2, :, :] != BC_SOLID)
    Q_vol = float(np.sum(u_np[-1, :, :, 0][outlet_bc]))
    rho_in_avg = float(np.mean(rho_np[1, :, :][inlet_fluid])) if np.any(inlet_fluid) else 1.0
    rho_out_avg = float(np.mean(rho_np[-2, :, :][outlet_fluid])) if np.any(outlet_fluid) else 1.0
    delta_P = (rho_in_avg - rho_out_avg) / 3.0
    P_pump = abs(delta_P * Q_vol)
    buf = 10
    T_surface = T_np[buf:buf + core[0], :core[1], 0]
    T_max_surface = float(np.max(T_surface))
    T_avg_surface = float(np.mean(T_surface))
    T_min_surface = float(np.min(T_surface))
    delta_T_max = T_max_surface - T_min_surface
    sigma_T = float(np.std(T_surface))
    ux_out = u_np[-1, :, :, 0][outlet_bc]
    T_out = T_np[-1, :, :][outlet_bc]
    total_flux = float(np.sum(ux_out))
    T_outlet_avg = float(np.sum(ux_out * T_out)) / total_flux if abs(total_flux) > 1e-12 else t_inlet
    fluid_temp_rise = T_outlet_avg - t_inlet
    Q_out_flux = float(np.sum(ux_out * T_out))
    Q_in_flux = t_inlet * total_flux
    Q_absorbed = abs(Q_out_flux - Q_in_flux) if abs(Q_out_flux - Q_in_flux) > 1e-12 else 1e-12
    R_th = (T_max_surface - t_inlet) / Q_absorbed
    core_fluid = (mask[buf:-buf, 1:-1, 4:-2] != BC_SOLID)
    T_fluid_avg = float(np.mean(T_np[buf:-buf, 1:-1, 4:-2][core_fluid])) if np.any(core_fluid) else t_inlet
    dT_conv = T_avg_surface - T_fluid_avg
    h_avg = Q_absorbed / (A_wetted * abs(dT_conv)) if (abs(dT_conv) > 1e-6 and A_wetted > 0) else 0.0
    tau_th_fluid = 0.5 + (params["tau_fluid"] - 0.5) / Pr
    k_fluid = (1.0 / 3.0) * (tau_th_fluid - 0.5)
    Nu = (h_avg * Dh / k_fluid) if k_fluid > 1e-12 else 0.0
    L = nx - 2 * buf
    u_inlet = params["u_inlet_val"]
    rho_avg = (rho_in_avg + rho_out_avg) / 2.0
    f_friction = (abs(delta_P) * Dh) / (2.0 * L * rho_avg * u_inlet ** 2) if u_inlet > 1e-12 else 0.0
    COP = Q_absorbed / P_pump if P_pump > 1e-12 else 0.0
    return {"R_th": float(R_th), "Nu": float(Nu), "f_friction": float(f_friction), "P_pump": float(P_pump),
            "COP": float(COP), "Q_vol": float(Q_vol), "T_max_surface": float(T_max_surface),
            "T_avg_surface": float(T_avg_surface), "delta_T_max": float(delta_T_max), "sigma_T": float(sigma_T),
            "fluid_temp_rise": float(fluid_temp_rise)}


def metadata_reductions(results):
    """max_velocity, max_temperature, mean_temperature and pressure_drop (generate_dataset.py:438-441,553-556)."""
    u_np, rho_np, T_np = results["u"], results["rho"], results["T"]
    return {"max_velocity": float(np.max(np.linalg.norm(u_np, axis=-1))),
            "max_temperature": float(np.max(T_np)), "mean_temperature": float(np.mean(T_np)),
            "pressure_drop": float(np.mean(rho_np[1, :, :]) - np.mean(rho_np[-2, :, :])) / 3.0}
