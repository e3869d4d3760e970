"""Batched solving: many designs of uneven lattice sizes advanced by one fused launch per step.

``BatchedSolver`` keeps ``slots`` lattice slots (each sized for the largest
lattice) in one C-ABI batch.  Every step is one launch of the fused AA pass
over all active slots; blocks of retired or empty slots exit at once.  Steps
run in graph-captured chunks of ``stats_every``; after each chunk one
reduction launch forms T and the 16 stats of every slot, and the host applies
the reference's residual rule (src/simulation.py:130-156) per slot:

* converged or diverged (or at ``max_steps``) slots retire: their fields and
  on-device metrics are collected and the slot takes the next queued design
  (continuous batching), so the launch stays full until the queue drains;
* all slots share the AA parity, so a new design starts at a chunk boundary.

The queue holds design dicts (``mask``, ``heat``, ``tau_fluid``,
``u_inlet_val``, ``heat_power``; optional ``init_fields`` = (rho, u, T) for a
warm start, and ``key``).  ``mask`` and ``heat`` are host arrays or device
tensors (int8 / FP32 CUDA); a design with ``params`` (``geometry.family_params``
or ``config2_params``) and no ``mask`` is built on the GPU when it starts.
Every slot is loaded from device-resident inputs (xadv_instance_load_device):
host arrays are uploaded once and their flags, heat source and column table
are built on the GPU.
"""
from __future__ import annotations

import time
from collections import deque

import numpy as np
import torch

from .engine import STAT_NAMES, LatticeBatch, normalized_heat, thermal_taus

__all__ = ["BatchedSolver", "device_inputs", "run_grouped"]
from .simulation import _Residuals, min_steps_floor


def device_inputs(d, batch=None):
    """(mask, heat) of a design as device tensors (int8, FP32 divided by its maximum), uploading host arrays and
    building ``params``-only designs on ``batch``; the tensors are stored back into ``d``."""
    if d.get("mask") is None:
        from .geometry import device_design
        if batch is None:
            raise ValueError("a design without a mask needs a batch to be built on")
        d["mask"], d["heat"] = device_design(batch, d["params"])
    mask = d["mask"]
    if not torch.is_tensor(mask):
        mask = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.int8)).cuda()
    heat = d.get("heat")
    if heat is None:
        heat = torch.from_numpy(normalized_heat(None, tuple(mask.shape))).cuda()
    else:
        if not torch.is_tensor(heat):
            heat = torch.from_numpy(np.ascontiguousarray(heat, dtype=np.float32)).cuda()
        # Division by the maximum as src/simulation.py:84-85 does it; by a device tensor, so each cell is an
        # IEEE division (as numpy's), not a multiplication by a reciprocal.
        mx = heat.max()
        heat = torch.where(mx > 0, heat / mx, heat).contiguous()
    d["mask"], d["heat"] = mask, heat
    return mask, heat


class BatchedSolver:
    def __init__(self, slots, cap_shape=(148, 128, 32), *, stats_every=100, tol_ke=1e-6, tol_T=1e-6,
                 max_steps=30000, min_steps_rule=False, Pr=7.0, k_ratio=628.0, t_inlet=25.0, precision=32,
                 block=256, cells_per_thread=1, xcd_remap=False, alternate_sweep=True, fluid_storage="auto",
                 collect_fields=False,
                 metrics=True, stall_check=None, confirm=1, stop_rule="reference", checks=False, slab_planes=0,
                 thermal_prefetch=False):
        self.slots = int(slots)
        self.cap_shape = tuple(cap_shape)
        self.batch = LatticeBatch(slots=self.slots, cap_shape=self.cap_shape, precision=precision, block=block,
                                  cells_per_thread=cells_per_thread, xcd_remap=xcd_remap, graph_chunk=stats_every,
                                  fluid_storage=fluid_storage, alternate_sweep=alternate_sweep,
                                  slab_planes=slab_planes, thermal_prefetch=thermal_prefetch)
        self.every = int(stats_every)
        self.tol_ke, self.tol_T = tol_ke, tol_T
        self.max_steps = int(max_steps)
        self.min_steps_rule = min_steps_rule
        self.Pr, self.k_ratio, self.t_inlet = Pr, k_ratio, t_inlet
        self.collect_fields = collect_fields
        self.metrics = metrics
        # Physics acceptance checks (analyze_dataset.py:184-325) of each retired slot's fields, on the device.
        self.checks = checks
        # Optional callable(design, step, residuals) -> True to abandon a run
        # (used by the warm-start fallback to detect a stalled or growing residual).
        self.stall_check = stall_check
        # A slot converges when the reference's residual test passes on `confirm` consecutive checks
        # (1 = the reference's rule; more keeps a run from stopping where a transient's rate crosses zero).
        self.confirm = max(1, int(confirm))
        # "reference": the reference's residuals; "rise": the same on the temperature rise and absorbed heat
        # (see simulation._Residuals).
        self.stop_rule = stop_rule
        self.state = [None] * self.slots
        self.lups = 0
        self.passes = 0

    def close(self):
        self.batch.close()

    # -- slots ----------------------------------------------------------------
    def _start(self, s, d):
        d = dict(d)  # the slot's device inputs stay with the slot, not in the caller's design dict
        mask, heat = device_inputs(d, self.batch)
        ttf, tts = thermal_taus(d["tau_fluid"], self.Pr, self.k_ratio)
        self.batch.load_device(s, mask, heat=heat, tau=d["tau_fluid"], tau_th_fluid=ttf, tau_th_solid=tts,
                               heat_power=d["heat_power"], T_in=self.t_inlet, rho_out=1.0, u_in=d["u_inlet_val"])
        init = d.get("init_fields")
        if init is not None:
            self.batch.init(s, *init)
        else:
            self.batch.init(s)
        nx = mask.shape[0]
        min_steps = min_steps_floor(nx, self.max_steps, d["u_inlet_val"]) if self.min_steps_rule else 0
        self.state[s] = dict(design=d, steps=0, res=_Residuals(self.every, self.tol_ke, self.tol_T, min_steps, self.stop_rule,
                                                    self.t_inlet),
                             cells=int(mask.numel()), t0=time.perf_counter(), history=[], ok=0,
                             bytes=19 * 4 * 2 * (int((mask != 1).sum()) + int(mask.numel())))

    def _finish(self, s, status):
        st = self.state[s]
        d = st["design"]
        out = dict(key=d.get("key"), steps=st["steps"], status=status, converged=status == "converged",
                   d_ke=st["res"].d_ke, dT_conv=st["res"].dT_conv, dQ_conv=st["res"].dQ_conv,
                   max_v=st["res"].max_v, max_T=st["res"].max_T, history=st["history"],
                   wall_s=time.perf_counter() - st["t0"])
        if self.metrics:
            from .metrics import device_metrics
            mask = d["mask"]
            core = (mask.shape[0] - 20, mask.shape[1])
            out["metrics"] = device_metrics(self.batch, s, mask, tau_fluid=d["tau_fluid"],
                                            u_inlet_val=d["u_inlet_val"], core=core, Pr=self.Pr)
        if self.checks:
            from .checks import device_checks
            m = out.get("metrics", {})
            out["checks"] = device_checks(self.batch, s, d["mask"], d["heat"], tau_fluid=d["tau_fluid"],
                                          heat_power=d["heat_power"], Nu=m.get("Nu", 0.0),
                                          f_friction=m.get("f_friction", 0.0), Pr=self.Pr, k_ratio=self.k_ratio)
        if self.collect_fields:
            rho, u, T = self.batch.fields(s)
            out["fields"] = (rho.cpu().numpy(), u.cpu().numpy(), T.cpu().numpy())
            out["inputs"] = (d["mask"], d["heat"])  # the design's device mask and normalised heat map
        self.batch.set_active(s, False)
        self.state[s] = None
        return out

    # -- driver ---------------------------------------------------------------
    def active_bytes_per_pass(self):
        """Algorithmic bytes of one FP32 pass over the active slots: 152 B per non-solid cell (fluid lattice)
        plus 152 B per cell (thermal lattice)."""
        return sum(st["bytes"] for st in self.state if st is not None)

    def busy(self):
        return any(st is not None for st in self.state)

    def fill(self, queue):
        """Load queued designs into the empty slots."""
        for s in range(self.slots):
            if self.state[s] is None and queue:
                self._start(s, queue.popleft())

    def chunk(self, queue, on_result=None, timer=None):
        """One graph chunk of ``stats_every`` passes over the active slots, the residual check of each, and
        retirement + refill from ``queue``; returns the results of the slots that retired."""
        results = []
        if timer is not None:
            timer.begin()
        self.batch.step(self.every, want_fields=True)
        if timer is not None:
            timer.end(self.every, self.active_bytes_per_pass())
        self.passes += self.every
        self.batch.readout(with_stats=True)
        stats = self.batch.stats()
        for s, st in enumerate(self.state):
            if st is None:
                continue
            st["steps"] += self.every
            self.lups += st["cells"] * self.every
            sd = dict(zip(STAT_NAMES, stats[s].tolist()))
            res = st["res"]
            res.update(st["steps"], sd)
            st["history"].append((st["steps"], res.d_ke, res.dT_conv, res.dQ_conv))
            status = None
            st["ok"] = st["ok"] + 1 if res.converged(st["steps"]) else 0
            if st["ok"] >= self.confirm:
                status = "converged"
            elif res.diverged() or not (np.isfinite(sd["sum_ke"]) and np.isfinite(sd["sum_T"])):
                # non-finite sums: the fields hold NaN/inf (the max reductions skip NaN), retire now
                status = "diverged"
            elif self.stall_check is not None and self.stall_check(st["design"], st["steps"], res):
                status = "stalled"
            elif st["steps"] >= self.max_steps:
                status = "max_steps"
            if status is not None:
                r = self._finish(s, status)
                results.append(r)
                if on_result is not None:
                    more = on_result(r)
                    if more:
                        queue.extend(more)
                if queue:
                    self._start(s, queue.popleft())
        return results

    def run(self, designs, on_result=None, timer=None, progress=None):
        """Solve every design in ``designs``; returns results in completion order.

        ``on_result(r)`` (optional) sees each result as its slot retires and may return designs to append to
        the queue.  ``timer`` (optional) gets ``begin()`` / ``end(n, bytes_per_pass)`` around every chunk of
        passes."""
        return run_grouped([self], designs, on_result=on_result, timer=timer, progress=progress)


def run_grouped(solvers, designs, on_result=None, timer=None, progress=None):
    """Solve ``designs`` on several batched solvers that share one queue, one graph chunk per solver in turn.

    With small batches (a few full-size lattices each), a batch's state stays partly in the 256 MB Infinity
    Cache across the passes of its chunk, which a single large batch (all designs in one launch, its state in
    HBM) cannot do; the groups take turns chunk by chunk, and a slot that retires takes the next queued
    design as in one batch.  Results in completion order."""
    queue = deque(designs)
    results = []
    for bs in solvers:
        bs.fill(queue)
    t_last = time.perf_counter()
    while True:
        active = [bs for bs in solvers if bs.busy()]
        if not active:
            return results
        if progress is not None and time.perf_counter() - t_last > 60.0:  # state of a long run, once a minute
            t_last = time.perf_counter()
            progress(sum(sum(st is not None for st in b.state) for b in solvers), len(queue), len(results))
        for bs in active:
            if timer is not None and len(solvers) > 1:
                timer.stream = bs.batch.stream()  # each group's passes run on its own batch's stream
            results += bs.chunk(queue, on_result=on_result, timer=timer)
            for other in solvers:  # designs re-queued by on_result may go to a group that fell idle
                if queue and not other.busy():
                    other.fill(queue)
