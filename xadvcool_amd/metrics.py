"""Post-solve performance metrics from the on-device reductions.

``device_metrics`` gives the dict of compute_macroscopic_metrics
(generate_dataset.py:80-179): R_th, Nu, f_friction, P_pump, COP, Q_vol,
T_max_surface, T_avg_surface, delta_T_max, sigma_T, fluid_temp_rise, plus the
metadata scalars of generate_dataset.py:538-579 that are reductions of the
fields (pressure_drop, max_velocity, max_temperature, mean_temperature).  The
sums, maxima and minima come from one wavefront/LDS reduction kernel over the
fields of the last step (xadv_metrics); only the closing scalar arithmetic
runs on the host, in the reference's formulas.
"""
from __future__ import annotations

import math

from . import _lib  # noqa: F401  (the reductions are the library's)

T_INLET = 25.0  # generate_dataset.py:96


def from_reductions(m, *, tau_fluid, u_inlet_val, nx, Dh, A_wetted, buf=10, Pr=7.0, t_inlet=T_INLET):
    """Close the metrics from the XADV_M_* reduction vector ``m`` (generate_dataset.py:100-179)."""
    Q_vol = m[0]
    rho_in_avg = m[1] / m[2] if m[2] > 0 else 1.0
    rho_out_avg = m[3] / m[4] if m[4] > 0 else 1.0
    delta_P = (rho_in_avg - rho_out_avg) / 3.0
    P_pump = abs(delta_P * Q_vol)
    T_max_s, T_min_s = m[8], m[9]
    ns = m[12]
    T_avg_s = m[10] / ns
    sigma_T = math.sqrt(max(m[11] / ns - T_avg_s * T_avg_s, 0.0))
    total_flux = Q_vol
    Q_out_flux = m[13]
    T_outlet_avg = Q_out_flux / total_flux if abs(total_flux) > 1e-12 else t_inlet
    fluid_temp_rise = T_outlet_avg - t_inlet
    Q_in_flux = t_inlet * total_flux
    Q_absorbed = abs(Q_out_flux - Q_in_flux) if abs(Q_out_flux - Q_in_flux) > 1e-12 else 1e-12
    R_th = (T_max_s - t_inlet) / Q_absorbed
    T_fluid_avg = m[14] / m[15] if m[15] > 0 else t_inlet
    dT_conv = T_avg_s - T_fluid_avg
    h_avg = Q_absorbed / (A_wetted * abs(dT_conv)) if (abs(dT_conv) > 1e-6 and A_wetted > 0) else 0.0
    tau_th_fluid = 0.5 + (tau_fluid - 0.5) / Pr
    k_fluid = (1.0 / 3.0) * (tau_th_fluid - 0.5)
    Nu = (h_avg * Dh / k_fluid) if k_fluid > 1e-12 else 0.0
    L = nx - 2 * buf
    rho_avg = (rho_in_avg + rho_out_avg) / 2.0
    f_friction = (abs(delta_P) * Dh) / (2.0 * L * rho_avg * u_inlet_val ** 2) if u_inlet_val > 1e-12 else 0.0
    COP = Q_absorbed / P_pump if P_pump > 1e-12 else 0.0
    plane_n = m[7]
    return {
        "R_th": float(R_th), "Nu": float(Nu), "f_friction": float(f_friction), "P_pump": float(P_pump),
        "COP": float(COP), "Q_vol": float(Q_vol), "T_max_surface": float(T_max_s), "T_avg_surface": float(T_avg_s),
        "delta_T_max": float(T_max_s - T_min_s), "sigma_T": float(sigma_T), "fluid_temp_rise": float(fluid_temp_rise),
        # metadata columns (generate_dataset.py:553-556)
        "pressure_drop": float((m[5] / plane_n - m[6] / plane_n) / 3.0),
        "max_velocity": float(m[16]), "max_temperature": float(m[17]), "mean_temperature": float(m[18] / m[19]),
        "nonfinite_cells": int(m[20]),
    }


def device_metrics(batch, slot, mask, *, tau_fluid, u_inlet_val, buf=10, core=(128, 128), Pr=7.0):
    """Metrics of the slot's current fields (call after the solve's last readout)."""
    from .geometry import compute_hydraulic_diameter
    Dh, _, A = compute_hydraulic_diameter(mask, buf)
    m = batch.metrics(slot, buf, core[0], core[1])
    return from_reductions(m, tau_fluid=tau_fluid, u_inlet_val=u_inlet_val, nx=mask.shape[0], Dh=Dh, A_wetted=A,
                           buf=buf, Pr=Pr)
