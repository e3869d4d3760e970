"""ctypes front end of the CPU oracle (lbm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker for the HIP solver, never the product.

Arrays use the reference layout (populations ``(19, nx, ny, nz)``, fields
``(nx, ny, nz)``, velocity ``(nx, ny, nz, 3)``, mask int32).  ``run_simulation``
restates ``src/simulation.py:14-162`` line for line on top of the oracle
kernels: same initialisation (:60-90), same six-launch step (:94-99), same
stats cadence, residual formulas, pump controller and stop rule (:102-160).

Parity pins for every function here are listed in the header of
``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")
_lib = None

BC_FLUID, BC_SOLID, BC_INLET, BC_OUTLET = 0, 1, 2, 3


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("fluid", ctypes.c_int), ("fluid_stream", ctypes.c_int), ("forced", ctypes.c_int),
        ("thermal", ctypes.c_int), ("thermal_stream", ctypes.c_int), ("dirichlet_x", ctypes.c_int),
        ("tau", ctypes.c_double), ("force_x", ctypes.c_double), ("u_wall", ctypes.c_double),
        ("u_in", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("tau_th_fluid", ctypes.c_double), ("tau_th_solid", ctypes.c_double),
        ("heat_power", ctypes.c_double), ("T_in", ctypes.c_double), ("dirichlet_T", ctypes.c_double),
        ("mutation", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
        _lib.orc_max_threads.restype = ctypes.c_int
        global _max0
        _max0 = int(_lib.orc_max_threads())
    return _lib


_max0 = None


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    """OpenMP's thread count for this process (OMP_NUM_THREADS or the host's cores), as it was when the library
    loaded: set_threads() changes what OpenMP reports afterwards, not this."""
    lib()
    return _max0


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


class State:
    """Two-buffer lattice state exactly as the reference allocates it (simulation.py:60-75)."""

    def __init__(self, shape, dtype=np.float32):
        nx, ny, nz = shape
        self.shape = (nx, ny, nz)
        self.dtype = np.dtype(dtype)
        z4 = lambda: np.zeros((19, nx, ny, nz), self.dtype)
        self.f_old, self.f_new, self.f_eq = z4(), z4(), z4()
        self.g_old, self.g_new, self.g_eq = z4(), z4(), z4()
        self.rho = np.ones((nx, ny, nz), self.dtype)
        self.u = np.zeros((nx, ny, nz, 3), self.dtype)
        self.T = np.zeros((nx, ny, nz), self.dtype)

    def init_fluid(self, rho=None, u=None):
        """compute_equilibrium_kernel into f_old (simulation.py:90; validate_analytical.py:151-152)."""
        if rho is not None:
            self.rho[...] = rho
        if u is not None:
            self.u[...] = u
        nx, ny, nz = self.shape
        getattr(lib(), "orc_equilibrium_" + _sfx(self.dtype))(
            nx, ny, nz, _p(self.rho), _p(self.u), _p(self.f_eq))
        self.f_old[...] = self.f_eq

    def init_thermal(self, T=None, mask=None):
        """compute_thermal_equilibrium_kernel into g_old (simulation.py:89)."""
        if T is not None:
            self.T[...] = T
        nx, ny, nz = self.shape
        m = np.ascontiguousarray(mask, dtype=np.int32)
        getattr(lib(), "orc_thermal_equilibrium_" + _sfx(self.dtype))(
            nx, ny, nz, _p(self.u), _p(self.T), _p(self.g_eq), _p(m))
        self.g_old[...] = self.g_eq


def init_thermal7(state: State, T=None, mask=None):
    """D3Q7 variant: g_eq7(T, u) into planes 0..6 of g_old (the initial state of the D3Q7 schedules,
    ``thermal=3`` and ``4`` of ``run``)."""
    if T is not None:
        state.T[...] = T
    nx, ny, nz = state.shape
    m = np.ascontiguousarray(mask if mask is not None else np.zeros(state.shape, np.int32), dtype=np.int32)
    getattr(lib(), "orc_thermal_equilibrium7_" + _sfx(state.dtype))(
        nx, ny, nz, _p(state.u), _p(state.T), _p(state.g_eq), _p(m))
    state.g_old[...] = state.g_eq


def run(state: State, mask, nsteps, *, fluid=0, fluid_stream=0, forced=0, thermal=0, thermal_stream=0,
        dirichlet_x=-1, tau=0.6, force_x=0.0, u_wall=0.0, u_in=0.0, rho_out=1.0,
        tau_th_fluid=0.6, tau_th_solid=0.6, heat_power=0.0, T_in=0.0, dirichlet_T=0.0, heat=None, mutation=0):
    """Advance ``nsteps`` with the chosen reference schedule (see orc_cfg in lbm_oracle.cpp)."""
    nx, ny, nz = state.shape
    cfg = OrcCfg(nx, ny, nz, fluid, fluid_stream, forced, thermal, thermal_stream, dirichlet_x,
                 tau, force_x, u_wall, u_in, rho_out, tau_th_fluid, tau_th_solid, heat_power, T_in, dirichlet_T,
                 mutation)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    hs = np.zeros(state.shape, state.dtype) if heat is None else np.ascontiguousarray(heat, dtype=state.dtype)
    fn = getattr(lib(), "orc_run_" + _sfx(state.dtype))
    rc = fn(ctypes.byref(cfg), int(nsteps), _p(m), _p(hs), _p(state.f_old), _p(state.f_new), _p(state.f_eq),
            _p(state.rho), _p(state.u), _p(state.g_old), _p(state.g_new), _p(state.g_eq), _p(state.T))
    if rc != 0:
        raise ValueError("oracle: bad configuration")
    return state


def stats(rho, u, T, mask):
    """Fluid slots 0-10 and thermal slots 0-4 of the reference stats kernels, as 16 doubles."""
    nx, ny, nz = mask.shape
    out = np.zeros(16, np.float64)
    lib().orc_stats_f32(nx, ny, nz, _p(np.ascontiguousarray(rho, np.float32)),
                        _p(np.ascontiguousarray(u, np.float32)), _p(np.ascontiguousarray(T, np.float32)),
                        _p(np.ascontiguousarray(mask, np.int32)), _p(out))
    return out


def thermal_taus(tau_fluid, Pr, k_ratio):
    """simulation.py:70-71 (Python double arithmetic)."""
    tau_th_fluid = 0.5 + ((tau_fluid - 0.5) / Pr)
    tau_th_solid = 0.5 + k_ratio * (tau_th_fluid - 0.5)
    return tau_th_fluid, tau_th_solid


def normalized_heat(heat_source, shape):
    """simulation.py:77-85: default z=0 plane, normalised by its max in FP32."""
    nx, ny, nz = shape
    if heat_source is not None:
        hm = np.array(heat_source, dtype=np.float32)
        if hm.shape != (nx, ny, nz):
            raise ValueError(f"heat_source shape {hm.shape} != domain shape ({nx}, {ny}, {nz})")
    else:
        hm = np.zeros((nx, ny, nz), dtype=np.float32)
        hm[:, :, 0] = 1.0
    if hm.max() > 0:
        hm /= hm.max()
    return hm


def pump_update(u_in, sum_rho_in, sum_rho_out, c_in, c_out, target_power, mutation=0):
    """One constant-power pump update (src/simulation.py:139-146).

    The reference's operands are float32: the stats come back as a float32
    array (stats_warp.numpy(), :128), u_in is read back from the float32
    device array (:140), and NumPy's promotion keeps float32 when a float32
    meets a Python float (0.005, 3.0, target_power, the clip bounds).  So
    every operation here rounds to float32, in the reference's order; the
    stats arrive as doubles and are rounded to float32 first, as the
    reference's FP32 sums are.  Returns the new inlet velocity (np.float32).

    ``mutation`` plants a mistake for the tests that must catch it:
    1 = the arithmetic in float64 (rounded to float32 only at the end),
    2 = no clip, 3 = the correction's sign flipped, 4 = the /3 of dP dropped.
    """
    if mutation == 1:
        s8, s9 = float(np.float32(c_in)), float(np.float32(c_out))
        rin = float(np.float32(sum_rho_in)) / s8 if s8 > 0 else 1.0
        rout = float(np.float32(sum_rho_out)) / s9 if s9 > 0 else 1.0
        power = (rin - rout) / 3.0 * (float(np.float32(u_in)) * s8)
        return np.float32(np.clip(float(np.float32(u_in)) + 0.005 * (target_power - power), 0.0001, 0.15))
    f = np.float32
    u = f(u_in)
    s5, s6, s8, s9 = f(sum_rho_in), f(sum_rho_out), f(c_in), f(c_out)
    rho_in_avg = f(s5 / s8) if s8 > 0 else f(1.0)
    rho_out_avg = f(s6 / s9) if s9 > 0 else f(1.0)
    dP = f(f(rho_in_avg - rho_out_avg) / f(3.0)) if mutation != 4 else f(rho_in_avg - rho_out_avg)
    power = f(dP * f(u * s8))
    step = f(f(0.005) * f(f(target_power) - power))
    if mutation == 3:
        step = -step
    if mutation == 2:
        return f(u + step)
    return f(min(max(f(u + step), f(0.0001)), f(0.15)))


def run_simulation(mask, max_steps=40000, tau_fluid=0.52, heat_power=0.1, t_inlet=25.0, u_inlet_val=0.03,
                   stats_every=100, tol_ke=1e-7, tol_T=1e-7, pump_mode="constant_flow", target_power=1.0,
                   device="cpu", silent=True, heat_source=None, Pr=7.0, k_ratio=628.0, dtype=np.float32,
                   min_steps_rule=True, pump_mutation=0, thermal_lattice="D3Q19"):
    """src/simulation.py:14-162 on the oracle kernels.

    ``thermal_lattice="D3Q7"`` runs the oracle's D3Q7 thermal variant (schedule 3 of ``run``) in the same loop.

    ``min_steps_rule=False`` drops the flow-through floor (simulation.py:104);
    everything else is the reference loop.  ``pump_mutation`` plants a pump
    mistake for the tests: 1-4 as in pump_update, 5 = the controller gated at
    step >= 2000 instead of > 2000.
    """
    mask = np.asarray(mask)
    nx, ny, nz = mask.shape
    st = State((nx, ny, nz), dtype)
    st.rho[...] = 1.0
    st.u[...] = 0.0
    st.T[...] = t_inlet
    ttf, tts = thermal_taus(tau_fluid, Pr, k_ratio)
    hm = normalized_heat(heat_source, (nx, ny, nz))
    th = 3 if thermal_lattice == "D3Q7" else 1
    if th == 3:
        init_thermal7(st, mask=mask)
    else:
        st.init_thermal(mask=mask)
    st.init_fluid()
    min_steps = min(int(0.6 * max_steps), max(5000, int(2.0 * nx / max(u_inlet_val, 1e-6))))
    if not min_steps_rule:
        min_steps = 0
    prev_ke = prev_max_T = prev_Q_out = 0.0
    d_ke = dT_conv = 0.0
    dQ_conv = 1.0
    u_in = np.float32(u_inlet_val)
    step = 0
    while step < max_steps:
        n = stats_every - (step % stats_every)
        n = min(n, max_steps - step)
        run(st, mask, n, fluid=1, thermal=th, tau=tau_fluid, u_in=float(u_in), rho_out=1.0,
            tau_th_fluid=ttf, tau_th_solid=tts, heat_power=heat_power, T_in=t_inlet, heat=hm)
        step += n
        if step % stats_every == 0:
            s = stats(st.rho, st.u, st.T, mask)
            max_v, ke, max_T, Q_out = s[0], s[2], s[11], s[14]
            d_ke = abs(ke - prev_ke) / (prev_ke + 1e-9) / stats_every if step > stats_every else 0.0
            dT_conv = abs(max_T - prev_max_T) / (prev_max_T + 1e-9) / stats_every if step > stats_every else 0.0
            dQ_conv = abs(Q_out - prev_Q_out) / (abs(prev_Q_out) + 1e-12) / stats_every if step > stats_every else 1.0
            prev_ke, prev_max_T, prev_Q_out = ke, max_T, Q_out
            gate = step >= 2000 if pump_mutation == 5 else step > 2000
            if pump_mode == "constant_power" and gate:
                u_in = pump_update(u_in, s[5], s[6], s[8], s[9], target_power,
                                   mutation=pump_mutation if pump_mutation < 5 else 0)
            if step > min_steps and d_ke < tol_ke and dT_conv < tol_T and dQ_conv < tol_T:
                break
            if np.isnan(max_v) or max_v > 0.17 or np.isnan(max_T) or max_T > 1e4:
                break
    converged = step > min_steps and d_ke < tol_ke and dT_conv < tol_T and dQ_conv < tol_T
    return {"u": st.u.copy(), "rho": st.rho.copy(), "T": st.T.copy(), "mask": mask, "steps": step,
            "d_ke": d_ke, "dT_conv": dT_conv, "dQ_conv": dQ_conv, "converged": bool(converged),
            "u_inlet": float(u_in)}
