"""Benchmark: converged cold-plate solves per hour on one MI355X.

Workload (one "step" = one converged solve of one synthetic design):
full-size single solve on a 148 x 128 x 32 lattice — gyroid TPMS core
(sin x cos y + sin y cos z + sin z cos x < t, cell period in [16, 32], t in
[-0.4, 0.4]), quad_core heat map on the base, inlet u in [0.01, 0.05], tau in
[0.52, 0.56], heat power in [0.05, 0.20], all drawn from the design seed;
FP32 D3Q19 fluid + D3Q19 thermal lattices; run to the residual tolerance
(the reference's d_ke / dT / dQ residuals, floor off) or 30,000 steps.

Printed (one JSON line): converged solves/hour, ms per solve, MLUPS (all
cells x steps / wall time), the fused pass's achieved bandwidth against the
MI355X HBM roofline (HIP events on the solver's own stream over the timed
region; algorithmic bytes 152 B per non-solid cell for the fluid lattice plus
152 B per cell for the thermal lattice, per pass), and the CPU oracle timed on
this host as the baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LATTICE = (148, 128, 32)


class PassTimer:
    """HIP events around each block of passes on the solver's stream."""

    def __init__(self, stream):
        import torch
        self.torch = torch
        self.stream = stream
        self.pairs = []
        self.passes = 0

    def begin(self):
        self._e0 = self.torch.cuda.Event(enable_timing=True)
        self._e0.record(self.stream)

    def end(self, n, bytes_per_pass=None):
        e1 = self.torch.cuda.Event(enable_timing=True)
        e1.record(self.stream)
        self.pairs.append((self._e0, e1, n, bytes_per_pass))
        self.passes += n

    def total_ms(self):
        self.torch.cuda.synchronize()
        return sum(p[0].elapsed_time(p[1]) for p in self.pairs)

    def first(self, k):
        """(ms, passes, bytes) over the first k chunks."""
        self.torch.cuda.synchronize()
        sel = self.pairs[:k]
        return (sum(p[0].elapsed_time(p[1]) for p in sel), sum(p[2] for p in sel),
                sum(p[2] * (p[3] or 0) for p in sel))


def algorithmic_bytes_per_pass(mask, rsize=4):
    n_all = mask.size
    n_fluid = int(np.sum(mask != 1))
    return 19 * rsize * 2 * (n_fluid + n_all)


def load_design(batch, d, Pr=7.0, k_ratio=628.0, T_in=25.0):
    from xadvcool_amd.engine import normalized_heat, thermal_taus
    ttf, tts = thermal_taus(d["tau_fluid"], Pr, k_ratio)
    batch.load(0, d["mask"], heat=normalized_heat(d["heat"], d["mask"].shape), tau=d["tau_fluid"],
               tau_th_fluid=ttf, tau_th_solid=tts, heat_power=d["heat_power"], T_in=T_in, rho_out=1.0,
               u_in=d["u_inlet_val"])
    batch.init(0)


def gpu_solves(designs, max_steps, tol, timer_stream=True):
    import torch
    from xadvcool_amd.engine import LatticeBatch
    from xadvcool_amd.simulation import solve_on_batch
    batch = LatticeBatch(slots=1, cap_shape=LATTICE, graph_chunk=100)
    timer = PassTimer(batch.stream()) if timer_stream else None
    results = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in designs:
        load_design(batch, d)
        r = solve_on_batch(batch, 0, nx=d["mask"].shape[0], max_steps=max_steps, tau_fluid=d["tau_fluid"],
                           u_inlet_val=d["u_inlet_val"], stats_every=100, tol_ke=tol, tol_T=tol,
                           min_steps_rule=False, timer=timer)
        r["bytes_per_pass"] = algorithmic_bytes_per_pass(d["mask"])
        r["cells"] = d["mask"].size
        results.append(r)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    batch.close()
    return results, wall, timer


def cpu_baseline(design, seconds, threads):
    """The oracle (reference six-kernel schedule, FP32) on the same lattice, timed on this host."""
    from oracle import oracle as orc
    orc.set_threads(threads)
    mask = design["mask"]
    st = orc.State(mask.shape, np.float32)
    st.T[...] = 25.0
    st.init_thermal(mask=mask)
    st.init_fluid()
    ttf, tts = orc.thermal_taus(design["tau_fluid"], 7.0, 628.0)
    hm = orc.normalized_heat(design["heat"], mask.shape)
    kw = dict(fluid=1, thermal=1, tau=design["tau_fluid"], u_in=design["u_inlet_val"], rho_out=1.0,
              tau_th_fluid=ttf, tau_th_solid=tts, heat_power=design["heat_power"], T_in=25.0, heat=hm)
    orc.run(st, mask, 1, **kw)  # warm caches / page in
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        orc.run(st, mask, 1, **kw)
        steps += 1
    dt = time.perf_counter() - t0
    return mask.size * steps / dt / 1e6, steps, dt


def profile_passes(npasses, seed=1000):
    """Only fused passes on one config-2 design (graph chunks of 100), for kernel traces and counters."""
    import torch
    from xadvcool_amd.engine import LatticeBatch
    from xadvcool_amd.geometry import config2_design
    d = config2_design(seed)
    batch = LatticeBatch(slots=1, cap_shape=LATTICE, graph_chunk=100)
    load_design(batch, d)
    batch.step(200, want_fields=False)  # past the start-up transient
    torch.cuda.synchronize()
    timer = PassTimer(batch.stream())
    done = 0
    while done < npasses:
        n = min(100, npasses - done)
        timer.begin()
        batch.step(n, want_fields=False)
        timer.end(n)
        done += n
    ms = timer.total_ms() / max(done, 1)
    nb = algorithmic_bytes_per_pass(d["mask"])
    print(json.dumps({"mode": "passes", "passes": done, "avg_pass_us": round(1000 * ms, 2),
                      "alg_bytes_per_pass": nb, "achieved_GBs": round(nb / (ms * 1e-3) / 1e9, 1)}), flush=True)
    batch.close()


def profile_copy(mbytes, reps=20):
    """A plain device copy of known size (torch): calibrates the FETCH_SIZE / WRITE_SIZE counters and gives the
    copy bandwidth for that working set (read + write bytes / time)."""
    import torch
    n = int(mbytes * (1 << 20)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"mode": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": reps + 3,
                      "copy_GBs": round(8 * n / (ms * 1e-3) / 1e9, 1), "us": round(1000 * ms, 2)}), flush=True)


def batch_bench(n_designs, slots, max_steps, tol, seed0=2000, **shape):
    """Config-3 style batch: mixed families and heat maps, uneven lattices, one launch per step."""
    import torch
    from xadvcool_amd.batch import BatchedSolver
    from xadvcool_amd.geometry import family_design
    designs = [dict(family_design(seed0 + i), key=seed0 + i) for i in range(n_designs)]
    bs = BatchedSolver(slots, (148, 128, 32), tol_ke=tol, tol_T=tol, max_steps=max_steps, **shape)
    timer = PassTimer(bs.batch.stream())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = bs.run(designs, timer=timer)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, npass, nbytes = timer.first(3)
    total_ms = timer.total_ms()
    out = {"designs": n_designs, "slots": slots, "wall_s": round(wall, 2),
           "converged": sum(r["converged"] for r in res),
           "status": {k: sum(r["status"] == k for r in res) for k in ("converged", "diverged", "max_steps", "stalled")},
           "solves_per_hr": round(3600.0 * sum(r["converged"] for r in res) / wall, 1),
           "mlups": round(bs.lups / wall / 1e6, 1),
           "full_batch_pass_us": round(1000 * ms / max(npass, 1), 2),
           "full_batch_alg_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "all_passes_ms": round(total_ms, 1), "passes": bs.passes,
           "cells_total": int(sum(d["mask"].size for d in designs)),
           "steps": sorted(r["steps"] for r in res)}
    bs.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed solves (distinct seeded designs)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves before the timed ones")
    ap.add_argument("--max-steps", type=int, default=30000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["solves", "passes", "copy", "batch"], default="solves",
                    help="solves: the benchmark; passes / copy: profiling drivers (pure fused passes; a plain copy)")
    ap.add_argument("--passes", type=int, default=400)
    ap.add_argument("--copy-mb", type=float, default=512.0)
    ap.add_argument("--designs", type=int, default=64)
    ap.add_argument("--slots", type=int, default=64)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver runs independent designs per GPU and this build has no "
              "multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.mode == "passes":
        return profile_passes(args.passes)
    if args.mode == "copy":
        return profile_copy(args.copy_mb)
    if args.mode == "batch":
        print(json.dumps({"mode": "batch", **batch_bench(args.designs, args.slots, args.max_steps, args.tol)}),
              flush=True)
        return

    from xadvcool_amd.geometry import config2_design

    warm = [config2_design(900 + i) for i in range(args.warmup)]
    timed = [config2_design(1000 + i) for i in range(args.steps)]
    if warm:
        gpu_solves(warm, args.max_steps, args.tol, timer_stream=False)
    results, wall, timer = gpu_solves(timed, args.max_steps, args.tol)

    n_conv = sum(1 for r in results if r["converged"])
    total_steps = sum(r["steps"] for r in results)
    lups = sum(r["steps"] * r["cells"] for r in results)
    pass_ms_total = timer.total_ms()
    alg_bytes = sum(r["steps"] * r["bytes_per_pass"] for r in results)
    avg_pass_ms = pass_ms_total / max(timer.passes, 1)
    avg_bytes = alg_bytes / max(total_steps, 1)
    achieved = alg_bytes / (pass_ms_total * 1e-3) / 1e9
    value = 3600.0 * n_conv / wall

    line = {
        "metric": "converged solves/hr, warm-start vs plain LBM; MLUPS and % of HBM peak",
        "value": round(value, 2),
        "unit": "converged solves/hr (plain cold-start LBM)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / max(args.steps, 1), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gyroid designs, quad_core heat map)",
        "config": {"workload": "full-size single solve, 148x128x32 gyroid TPMS, to residual tol or 30000 steps",
                   "lattice": list(LATTICE), "tol": args.tol, "max_steps": args.max_steps,
                   "stop_rule": "reference d_ke/dT/dQ residuals every 100 steps, flow-through floor off",
                   "seeds": [1000 + i for i in range(args.steps)]},
        "mlups": round(lups / wall / 1e6, 1),
        "converged": n_conv,
        "steps_to_converge": [r["steps"] for r in results],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_pass (fused AA fluid+thermal pass)",
                     "avg_launch_us": round(1000.0 * avg_pass_ms, 2),
                     "alg_bytes_per_launch": int(avg_bytes),
                     "note": "one 148x128x32 instance (~95 MB of AA state) fits the 256 MB Infinity Cache: "
                             "LLC-resident, not a pure-HBM figure"},
    }
    if not args.no_cpu:
        mlups1, n1, t1 = cpu_baseline(timed[0], args.cpu_seconds, 1)
        mean_steps = total_steps / max(len(results), 1)
        cells = timed[0]["mask"].size
        line["cpu_baseline"] = {
            "value": round(3600.0 / (mean_steps * cells / (mlups1 * 1e6)), 3),
            "unit": "converged solves/hr (projected from MLUPS at the GPU's mean steps to converge)",
            "cores": 1, "kind": "port", "mlups": round(mlups1, 3),
            "sample": f"oracle six-kernel FP32 schedule, 148x128x32 design seed 1000, {n1} steps in {t1:.1f} s, "
                      f"1 thread",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
