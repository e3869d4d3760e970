"""Drop-in ``run_simulation`` (same signature and return dict as src/simulation.py:14-162).

The loop logic is the reference's: equilibrium start at rho = 1, u = 0,
T = t_inlet; stats every ``stats_every`` steps; the residuals d_ke, dT_conv,
dQ_conv; the optional constant-power pump; the flow-through floor; the stop
and divergence rules.  Each ``stats_every`` block of steps is one replay of a
captured hipGraph of fused AA passes; the stats are deterministic on-device
reductions read back once per block.

Two keyword-only extras (defaults keep the reference behaviour):
``min_steps_rule`` (False drops the floor of src/simulation.py:104, for
residual-only comparisons) and ``init_fields`` (rho, u, T to start from
instead of rest, for warm starts).
"""
from __future__ import annotations

import numpy as np

from .constants import BC_SOLID
from .engine import STAT_NAMES, CoupledSolver


def hydraulic_diameter(mask, buf=10):
    """D_h = 4 V_fluid / A_wetted over the core (src/simulation.py:47-58, generate_dataset.py:57-68)."""
    core = mask[buf:-buf, 1:-1, 4:-2]
    fluid = core != BC_SOLID
    solid = core == BC_SOLID
    V = int(np.sum(fluid))
    A = int(np.sum(fluid[1:] & solid[:-1]) + np.sum(solid[1:] & fluid[:-1]) +
            np.sum(fluid[:, 1:] & solid[:, :-1]) + np.sum(solid[:, 1:] & fluid[:, :-1]) +
            np.sum(fluid[:, :, 1:] & solid[:, :, :-1]) + np.sum(solid[:, :, 1:] & fluid[:, :, :-1]))
    Dh = 4.0 * V / A if A > 0 else float(mask.shape[1])
    return Dh, V, A


STOP_RULES = ("reference", "rise")


class _Residuals:
    """Residual bookkeeping of src/simulation.py:107-137 and the stop rules of :150-156.

    ``rule="reference"`` is the reference's: the thermal residuals are the
    relative changes of max T and of the outlet flux sum(u_x T), per step.
    ``rule="rise"`` applies the same formulas to the temperature rise
    (max T - T_in) and the absorbed heat (sum(u_x T) - T_in sum(u_x)) at the
    outlet.  The reference's quantities carry the inlet temperature (T_in = 25
    against rises of a few units), which dilutes their relative changes by
    T_in / rise, so a run can pass at 1e-6 while R_th = rise / absorbed heat is
    still percents from its steady value; the rise rule measures the changes of
    the quantities R_th is made of.  Not a reference option."""

    def __init__(self, stats_every, tol_ke, tol_T, min_steps, rule="reference", t_inlet=25.0):
        if rule not in STOP_RULES:
            raise ValueError(f"stop rule must be one of {STOP_RULES}")
        self.every, self.tol_ke, self.tol_T, self.min_steps = stats_every, tol_ke, tol_T, min_steps
        self.rule, self.t_inlet = rule, t_inlet
        self.prev_ke = self.prev_max_T = self.prev_Q_out = 0.0
        self.d_ke = self.dT_conv = 0.0
        self.dQ_conv = 1.0
        self.max_v = self.max_T = 0.0

    def update(self, step, s):
        self.max_v, ke, self.max_T, Q_out = s["max_v"], s["sum_ke"], s["max_T"], s["Q_out"]
        T_meas = self.max_T
        if self.rule == "rise":
            T_meas = self.max_T - self.t_inlet
            Q_out = Q_out - self.t_inlet * s["sum_ux_out"]
        first = step <= self.every
        self.d_ke = 0.0 if first else abs(ke - self.prev_ke) / (self.prev_ke + 1e-9) / self.every
        den_T = abs(self.prev_max_T) if self.rule == "rise" else self.prev_max_T
        self.dT_conv = 0.0 if first else abs(T_meas - self.prev_max_T) / (den_T + 1e-9) / self.every
        self.dQ_conv = 1.0 if first else abs(Q_out - self.prev_Q_out) / (abs(self.prev_Q_out) + 1e-12) / self.every
        self.prev_ke, self.prev_max_T, self.prev_Q_out = ke, T_meas, Q_out

    def converged(self, step):
        return (step > self.min_steps and self.d_ke < self.tol_ke and self.dT_conv < self.tol_T
                and self.dQ_conv < self.tol_T)

    def diverged(self):
        return bool(np.isnan(self.max_v) or self.max_v > 0.17 or np.isnan(self.max_T) or self.max_T > 1e4)


def min_steps_floor(nx, max_steps, u_inlet_val):
    """Flow-through floor of src/simulation.py:104."""
    return min(int(0.6 * max_steps), max(5000, int(2.0 * nx / max(u_inlet_val, 1e-6))))


def pump_update(u_in, sum_rho_in, sum_rho_out, c_in, c_out, target_power):
    """One update of the constant-power pump (src/simulation.py:139-146), in float32 as the reference does it.

    In the reference every operand is float32: the stats are read back as a
    float32 array, u_in from the float32 device scalar, and NumPy keeps
    float32 when it meets a Python float (0.005, 3.0, target_power, the clip
    bounds).  Here the stats are FP64 sums, rounded to float32 first; then
    each operation rounds to float32 in the reference's order:
    u += 0.005 (P_target - dP u c_in) with dP = (mean rho(x=1) - mean rho(x=nx-2)) / 3,
    clipped to [1e-4, 0.15].  Returns np.float32."""
    f32 = np.float32
    u = f32(u_in)
    rin, rout, cin, cout = f32(sum_rho_in), f32(sum_rho_out), f32(c_in), f32(c_out)
    rho_in_avg = f32(rin / cin) if cin > 0 else f32(1.0)
    rho_out_avg = f32(rout / cout) if cout > 0 else f32(1.0)
    dP = f32(f32(rho_in_avg - rho_out_avg) / f32(3.0))
    power = f32(dP * f32(u * cin))
    u_new = f32(u + f32(f32(0.005) * f32(f32(target_power) - power)))
    return f32(min(max(u_new, f32(0.0001)), f32(0.15)))


def solve_on_batch(batch, slot, *, nx, max_steps, tau_fluid, u_inlet_val, stats_every=100, tol_ke=1e-7, tol_T=1e-7,
                   pump_mode="constant_flow", target_power=1.0, min_steps_rule=True, pbar=None, Dh=0.0,
                   timer=None):
    """The step loop of src/simulation.py:102-160 on slot ``slot`` of a one-slot, initialised batch.

    ``timer`` (optional) gets ``begin()`` before and ``end(n)`` after every
    block of n steps (events on the batch stream).  Returns the loop's outcome.
    """
    min_steps = min_steps_floor(nx, max_steps, u_inlet_val) if min_steps_rule else 0
    res = _Residuals(stats_every, tol_ke, tol_T, min_steps)
    u_in_eff = float(np.float32(u_inlet_val))
    step = 0
    diverged = False
    while step < max_steps:
        n = min(stats_every - (step % stats_every), max_steps - step)
        if timer is not None:
            timer.begin()
        batch.step(n, want_fields=True)
        if timer is not None:
            timer.end(n)
        step += n
        if pbar is not None:
            pbar.update(n)
        if step % stats_every != 0:
            continue
        batch.readout(with_stats=True)
        s = dict(zip(STAT_NAMES, batch.stats()[slot].tolist()))
        res.update(step, s)
        if pump_mode == "constant_power" and step > 2000:
            u_in_eff = float(pump_update(u_in_eff, s["sum_rho_in"], s["sum_rho_out"], s["c_in"], s["c_out"],
                                         target_power))
            batch.set_inlet_velocity(slot, u_in_eff)
        if pbar is not None:
            pbar.set_postfix({"Re": f"{(u_in_eff * Dh) / ((tau_fluid - 0.5) / 3.0):.1f}", "MaxV": f"{res.max_v:.4f}",
                              "MaxT": f"{res.max_T:.2f}", "dKE": f"{res.d_ke:.1e}", "dQ": f"{res.dQ_conv:.1e}"})
        if res.converged(step):
            break
        if res.diverged():
            diverged = True
            break
    if step % stats_every != 0:
        batch.readout(with_stats=False)
    return {"steps": step, "d_ke": res.d_ke, "dT_conv": res.dT_conv, "dQ_conv": res.dQ_conv,
            "converged": bool(res.converged(step)), "diverged": diverged, "max_v": res.max_v, "max_T": res.max_T,
            "u_inlet": u_in_eff}


def run_simulation(mask,
                   max_steps=40000,
                   tau_fluid=0.52,
                   heat_power=0.1,
                   t_inlet=25.0,
                   u_inlet_val=0.03,
                   stats_every=100,
                   tol_ke=1e-7,
                   tol_T=1e-7,
                   pump_mode="constant_flow",
                   target_power=1.0,
                   device="cuda",
                   silent=False,
                   heat_source=None,
                   Pr=7.0,
                   k_ratio=628.0,
                   *,
                   min_steps_rule=True,
                   init_fields=None,
                   precision=32,
                   thermal_lattice="D3Q19"):
    """Run one coupled solve; returns {"u","rho","T","mask","steps","d_ke","dT_conv","dQ_conv","converged"}.

    ``thermal_lattice="D3Q7"`` runs the seven-velocity thermal lattice (no reference counterpart; the same
    relaxation times, so the reference's diffusivities and heat input per step)."""
    if not str(device).startswith("cuda"):
        raise ValueError(f"device={device!r}: this solver runs on the GPU only (use device='cuda')")
    mask = np.asarray(mask)
    nx, ny, nz = mask.shape
    if not silent:
        print(f"Simulation Domain: {nx}x{ny}x{nz}")
    Dh = hydraulic_diameter(mask)[0] if not silent else 0.0

    graph_chunk = stats_every if stats_every % 2 == 0 else 0
    solver = CoupledSolver(mask, heat_source, tau_fluid=tau_fluid, heat_power=heat_power, t_inlet=t_inlet,
                           u_inlet_val=u_inlet_val, Pr=Pr, k_ratio=k_ratio, precision=precision,
                           graph_chunk=graph_chunk, thermal_lattice=thermal_lattice)
    try:
        if init_fields is not None:
            solver.init_equilibrium(*init_fields)
        else:
            solver.init_equilibrium()
        pbar = None
        if not silent:
            from tqdm import tqdm
            pbar = tqdm(total=max_steps)
        out = solve_on_batch(solver.batch, 0, nx=nx, max_steps=max_steps, tau_fluid=tau_fluid,
                             u_inlet_val=u_inlet_val, stats_every=stats_every, tol_ke=tol_ke, tol_T=tol_T,
                             pump_mode=pump_mode, target_power=target_power, min_steps_rule=min_steps_rule,
                             pbar=pbar, Dh=Dh)
        if pbar is not None:
            pbar.close()
            if out["diverged"]:
                print(f"\nSimulation diverged at step {out['steps']}: MaxV={out['max_v']:.2e}, "
                      f"Ma={out['max_v'] * 3 ** 0.5:.2f}, MaxT={out['max_T']:.2e}")
        rho, u, T = solver.batch.fields(0)
        return {"u": u.cpu().numpy(), "rho": rho.cpu().numpy(), "T": T.cpu().numpy(), "mask": mask,
                "steps": out["steps"], "d_ke": out["d_ke"], "dT_conv": out["dT_conv"], "dQ_conv": out["dQ_conv"],
                "converged": out["converged"]}
    finally:
        solver.close()
