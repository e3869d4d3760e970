"""Benchmark: converged cold-plate solves per hour on one MI355X, warm-started vs plain LBM.

Workload (default): surrogate-warm-started batched solving of 256 unseen
seeded designs on 148 x 128 x 32 lattices — all five core families (straight
channels, staggered pin fins, gyroid, Schwarz P and Schwarz D TPMS) and all
ten heat maps, tau, heat power and inlet velocity drawn from each design's
seed (seeds from 1,000,000; the surrogate never saw them).  FP32 D3Q19 fluid
+ D3Q19 thermal lattices.  Every design is solved twice, one arm after the
other, three lattices per launch with retirement and refill: cold
(equilibrium at rest) and warm (surrogate-predicted rho, u, T; a warm run
that stalls, diverges, or converges to fields failing the physics
acceptance checks is re-run cold).  Stop rule of both arms: the reference's
residual formulas applied to the temperature rise and the absorbed heat
(``--stop-rule rise``; the reference's own residuals with ``reference``),
checked every 100 steps, below 1e-6 on five consecutive checks, flow-through
floor off, at most 30,000 steps.  Each arm's wall time covers building its
designs on the GPU from their parameters, the solves, metrics, checks and,
for the warm arm, the surrogate inference and any fallback.  ``--steps``
does not change the workload (a step is one design solved to the stop
rule); ``--designs`` does.  The fused pass runs 256 threads x 2 cells,
linear block order, alternating sweep, thermal loads issued right after the
fluid gather (``--launch tune`` runs the project's autotuner first instead).

Printed (one JSON line): warm-start converged solves/hour (``value``), the
cold-start figure and their ratio, the agreement of the converged pressure
drop and thermal resistance where both arms converged (count within 0.5 %,
median and maximum differences), per-family counts, MLUPS (all cells x
steps / wall time), the fused pass's achieved bandwidth against the MI355X
HBM roofline (HIP events on the solver's own stream over every pass of both
arms; algorithmic bytes 152 B per non-solid cell for the fluid lattice plus
152 B per cell for the thermal lattice, per pass), and the CPU oracle timed
on this host as the baseline.
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

    def totals(self):
        """(ms, passes, bytes) over every chunk."""
        return self.first(len(self.pairs))

    def first(self, k):
        """(ms, passes, bytes) over the first k chunks."""
        self.torch.cuda.synchronize()
        sel = self.pairs[:k]
        return (sum(p[0].elapsed_time(p[1]) for p in sel), sum(p[2] for p in sel),
                sum(p[2] * (p[3] or 0) for p in sel))


COUNTERS_FILE = os.path.join("profiles", "k_pass_counters.json")


def pass_counters():
    """Measured HBM traffic of the fused pass relative to its algorithmic bytes (rocprofv3 FETCH_SIZE and
    WRITE_SIZE, calibrated on a plain copy), as committed with the profiles; None if absent."""
    try:
        with open(os.path.join(ROOT, COUNTERS_FILE)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def algorithmic_bytes_per_pass(mask, rsize=4, thermal_q=19):
    """Populations read and written once per pass: 19 per non-solid cell (compact fluid lattice) and
    ``thermal_q`` per cell (dense thermal lattice)."""
    n_all = mask.size
    n_fluid = int(np.sum(mask != 1))
    return rsize * 2 * (19 * n_fluid + thermal_q * n_all)


def load_design(batch, d, Pr=7.0, k_ratio=628.0, T_in=25.0, slot=0):
    from xadvcool_amd.engine import normalized_heat, thermal_taus
    ttf, tts = thermal_taus(d["tau_fluid"], Pr, k_ratio)
    batch.load(slot, d["mask"], heat=normalized_heat(d["heat"], d["mask"].shape), tau=d["tau_fluid"],
               tau_th_fluid=ttf, tau_th_solid=tts, heat_power=d["heat_power"], T_in=T_in, rho_out=1.0,
               u_in=d["u_inlet_val"])
    batch.init(slot)


def host_cpu():
    """The host CPU the oracle ran on: model name, logical CPUs in the machine and in this process's affinity
    mask, and the oracle's OpenMP thread count."""
    from oracle import oracle as orc
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_in_affinity": len(os.sched_getaffinity(0)),
            "oracle_threads": orc.max_threads()}


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


def profile_passes(npasses, slots=1, seed=1000, thermal_lattice="D3Q19", precision=32, **launch):
    """Only fused passes on ``slots`` config-2 designs (seeds from ``seed``; graph chunks of 100, fewer when
    npasses is smaller), for kernel traces and counters.  One slot (~140 MB touched per pass) stays in the
    256 MB Infinity Cache; many slots are HBM-resident.  With the D3Q7 thermal lattice the fused pass moves
    56 B per cell of thermal populations; precision 64 moves 8-byte populations."""
    import torch
    from xadvcool_amd.engine import LatticeBatch
    from xadvcool_amd.geometry import config2_design
    designs = [config2_design(seed + i) for i in range(slots)]
    chunk = 100 if npasses >= 100 else max(2, npasses - npasses % 2)
    batch = LatticeBatch(slots=slots, cap_shape=LATTICE, graph_chunk=chunk, thermal_lattice=thermal_lattice,
                         precision=precision, **launch)
    for s, d in enumerate(designs):
        load_design(batch, d, slot=s)
    batch.step(200, want_fields=False)  # past the start-up transient
    torch.cuda.synchronize()
    timer = PassTimer(batch.stream())
    done = 0
    while done < npasses:
        n = min(chunk, npasses - done)
        timer.begin()
        batch.step(n, want_fields=False)
        timer.end(n)
        done += n
    ms = timer.total_ms() / max(done, 1)
    q = 7 if thermal_lattice == "D3Q7" else 19
    nb = sum(algorithmic_bytes_per_pass(d["mask"], rsize=precision // 8, thermal_q=q) for d in designs)
    print(json.dumps({"mode": "passes", "slots": slots, "thermal_lattice": thermal_lattice, "precision": precision,
                      **launch,
                      "passes": done, "avg_pass_us": round(1000 * ms, 2),
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


def designs_bench(n=64, seed0=2_000_000):
    """Design parameters to a loaded slot: host path (numpy generators + host flag build) against the device path
    (xadv_geometry_mask, xadv_heat_map, xadv_instance_load_device), over the same n full-size designs of every
    family and batch heat map (148x128x32)."""
    import torch
    from xadvcool_amd import geometry as G
    from xadvcool_amd.engine import LatticeBatch
    core = (128, 128, 32)
    params = [G.family_params(seed0 + i, core=core) for i in range(n)]
    b = LatticeBatch(slots=1, cap_shape=(core[0] + 2 * G.BUFFER, core[1], core[2]))

    def kw(p):
        tau = p["tau_fluid"]
        ttf = 0.5 + (tau - 0.5) / 7.0
        return dict(tau=tau, tau_th_fluid=ttf, tau_th_solid=0.5 + 628.0 * (ttf - 0.5), heat_power=p["heat_power"],
                    T_in=25.0, u_in=p["u_inlet_val"])

    def host(p):
        t0 = time.perf_counter()
        mask, heat = G.design_arrays(p)
        t1 = time.perf_counter()
        b.load(0, mask, heat=heat, **kw(p))
        return t1 - t0, time.perf_counter() - t1

    def dev(p):
        t0 = time.perf_counter()
        mask, heat = G.device_design(b, p)
        t1 = time.perf_counter()
        b.load_device(0, mask, heat=heat, **kw(p))
        return t1 - t0, time.perf_counter() - t1

    out = {"mode": "designs", "designs": n, "lattice": [core[0] + 2 * G.BUFFER, core[1], core[2]],
           "families": {f: sum(p["family"] == f for p in params) for f in G.FAMILIES}}
    for name, fn in (("host", host), ("device", dev)):
        fn(params[0])  # first call: allocations, module loads
        torch.cuda.synchronize()
        t = time.perf_counter()
        parts = np.array([fn(p) for p in params])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        out[name] = {"designs_per_s": round(n / wall, 1), "ms_per_design": round(1e3 * wall / n, 3),
                     "build_ms": round(1e3 * parts[:, 0].mean(), 3), "load_ms": round(1e3 * parts[:, 1].mean(), 3)}
        print(f"[designs] {name}: {out[name]}", file=sys.stderr, flush=True)
    out["speedup"] = round(out["host"]["ms_per_design"] / out["device"]["ms_per_design"], 1)
    b.close()
    print(json.dumps(out), flush=True)


def train_step_bench(steps=4, batch=8, forwards=3):
    """Surrogate training steps (forward, weighted MSE, backward, Adam) at batch 8 on 152x128x32 with random inputs
    of the real shape (no data generation): the kernels of the warm-start network, for traces."""
    import torch
    from xadvcool_amd import surrogate as S
    torch.manual_seed(0)
    net = S.UNet3D(seed=0)
    net.norm = (np.zeros(S.NOUT, np.float32), np.ones(S.NOUT, np.float32))
    X, Y, Z = 152, 128, 32
    x = torch.randn(batch, X, Y, Z, S.CIN, device="cuda").to(torch.bfloat16)
    t = torch.randn(batch, X, Y, Z, S.NOUT, device="cuda")
    w = torch.ones(batch, X, Y, Z, S.NOUT, device="cuda")
    sc = torch.from_numpy(np.stack([S.scalars(0.52 + 0.005 * i, 0.1, 0.02) for i in range(batch)])).cuda()
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = net.forward(x, sc, keep=True)
        _, g = S.mse_loss(out, t, w)
        net.zero_grad()
        net.backward(g)
        net.adam_step(1e-4)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(forwards):
        net.forward(x, sc)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / max(forwards, 1)
    step = min(times[1:] or times)
    flop_fwd = 0.0  # 2 * 27 * cin * cout FLOP per output voxel of each layer, at its level's resolution
    shapes = {"c1a": 0, "c1b": 0, "d1a": 0, "d1b": 0, "head": 0, "c2a": 1, "c2b": 1, "d2a": 1, "d2b": 1,
              "c3a": 2, "c3b": 2, "d3a": 2, "d3b": 2, "c4a": 3, "c4b": 3, "c4c": 3, "c4d": 3}
    for k, c in net.conv.items():
        flop_fwd += 2.0 * 27 * c.cin * c.cout * batch * (X * Y * Z) / 8 ** shapes[k]
    print(json.dumps({"mode": "train-step", "batch": batch, "lattice": [X, Y, Z], "ms_per_step": round(1e3 * step, 2),
                      "forward_ms": round(1e3 * fwd, 2), "forward_GFLOP": round(flop_fwd / 1e9, 1),
                      "forward_TFLOPs": round(flop_fwd / fwd / 1e12, 1),
                      "step_TFLOPs": round(3 * flop_fwd / step / 1e12, 1)}), flush=True)


def batch_bench(n_designs, slots, max_steps, tol, seed0=2000, groups=1, **shape):
    """Config-3 style batch: mixed families and heat maps, uneven lattices.  ``groups`` = 1: every design in
    one launch per step; more: that many batches of ``slots`` lattices share the queue and take turns chunk
    by chunk (each batch's state partly resident in the Infinity Cache during its chunk)."""
    import torch
    from xadvcool_amd.batch import BatchedSolver, run_grouped
    from xadvcool_amd.geometry import family_design
    designs = [dict(family_design(seed0 + i), key=seed0 + i) for i in range(n_designs)]
    solvers = [BatchedSolver(slots, (148, 128, 32), tol_ke=tol, tol_T=tol, max_steps=max_steps, **shape)
               for _ in range(groups)]
    timer = PassTimer(solvers[0].batch.stream())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run_grouped(solvers, designs, timer=timer)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, npass, nbytes = timer.first(3 * groups)
    total_ms, _, total_bytes = timer.totals()
    lups, passes = sum(b.lups for b in solvers), sum(b.passes for b in solvers)
    for b in solvers:
        b.close()
    out = {"designs": n_designs, "slots": slots, "groups": groups, "wall_s": round(wall, 2),
           "converged": sum(r["converged"] for r in res),
           "status": {k: sum(r["status"] == k for r in res) for k in ("converged", "diverged", "max_steps", "stalled")},
           "solves_per_hr": round(3600.0 * sum(r["converged"] for r in res) / wall, 1),
           "mlups": round(lups / wall / 1e6, 1),
           "first_chunks_pass_us": round(1000 * ms / max(npass, 1), 2),
           "first_chunks_alg_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "all_passes_ms": round(total_ms, 1), "passes": passes,
           "all_passes_alg_GBs": round(total_bytes / (total_ms * 1e-3) / 1e9, 1),
           "all_passes_hbm_frac": round(total_bytes / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "cells_total": int(sum(d["mask"].size for d in designs)),
           "steps": sorted(r["steps"] for r in res)}
    return out


def config1_bench(cpu_steps_1=100, cpu_steps_all=500):
    """Plain small cold plate (74x64x16, 2,000 steps): the drop-in run_simulation on the GPU and the oracle
    on this host's cores at 1 thread and at all threads (bounded step counts, MLUPS), with the reference's
    D3Q19 thermal lattice and, reported separately, the D3Q7 variant (same relaxation times and heat input)."""
    import torch
    from oracle import oracle as orc
    from xadvcool_amd import geometry as G
    from xadvcool_amd.simulation import run_simulation
    c = G.CONFIG1
    m = G.config1_mask()
    kw = dict(tau_fluid=c["tau_fluid"], heat_power=c["heat_power"], t_inlet=c["t_inlet"], u_inlet_val=c["u_inlet_val"],
              tol_ke=0.0, tol_T=0.0, Pr=c["Pr"], k_ratio=c["k_ratio"])
    fluid = m != 1

    def summary(r):  # pressure drop over the fluid cells of x = 1 and x = nx-2, temperature rise
        dp = (float(r["rho"][1][fluid[1]].mean()) - float(r["rho"][-2][fluid[-2]].mean())) / 3.0
        rise = r["T"] - c["t_inlet"]
        return {"pressure_drop": dp, "T_rise_max": float(rise.max()), "T_rise_mean": float(rise.mean())}

    out = {"mode": "config1", "lattice": list(m.shape)}
    for lat in ("D3Q19", "D3Q7"):
        run_simulation(m, max_steps=200, device="cuda", silent=True, thermal_lattice=lat, **kw)  # kernels, graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = run_simulation(m, max_steps=c["steps"], device="cuda", silent=True, thermal_lattice=lat, **kw)
        gpu_s = time.perf_counter() - t0
        key = "" if lat == "D3Q19" else "_d3q7"
        out["steps"] = r["steps"]
        out["gpu" + key] = {"wall_s": round(gpu_s, 4), "mlups": round(m.size * r["steps"] / gpu_s / 1e6, 1),
                            **{k: round(v, 7) for k, v in summary(r).items()},
                            "note": "run_simulation end to end: init, graph replays, stats every 100 steps, field "
                                    "copies" + ("" if lat == "D3Q19" else "; thermal lattice D3Q7 (no reference "
                                                "counterpart; reported beside the D3Q19 result)")}
        for threads, n in ((1, cpu_steps_1), (orc.max_threads(), cpu_steps_all)):
            orc.set_threads(threads)
            t0 = time.perf_counter()
            orc.run_simulation(m, max_steps=n, thermal_lattice=lat, **kw)
            dt = time.perf_counter() - t0
            out[f"oracle{key}_{threads}t"] = {"steps": n, "wall_s": round(dt, 3), "mlups": round(m.size * n / dt / 1e6, 2)}
    g19, g7 = out["gpu"], out["gpu_d3q7"]
    out["d3q7_vs_d3q19"] = {"T_rise_max": round(g7["T_rise_max"] / g19["T_rise_max"] - 1, 5),
                            "T_rise_mean": round(g7["T_rise_mean"] / g19["T_rise_mean"] - 1, 5),
                            "pressure_drop": round(g7["pressure_drop"] / g19["pressure_drop"] - 1, 7),
                            "speed": round(g7["mlups"] / g19["mlups"], 3)}
    out["host"] = host_cpu()
    print(json.dumps(out), flush=True)


def config2_bench(seed=0, tol=1e-6, max_steps=30000, cpu_seconds=12.0):
    """Full-size single solve (148x128x32 gyroid TPMS, quad_core heat map, parameters from the seed) through the
    drop-in run_simulation, to the residual tolerance (reference residuals, flow-through floor off) or max_steps;
    the oracle's per-step rate on the same design at 1 thread and at all threads (bounded samples)."""
    import torch
    from oracle import oracle as orc
    from xadvcool_amd.geometry import config2_design
    from xadvcool_amd.simulation import run_simulation
    d = config2_design(seed)
    kw = dict(tau_fluid=d["tau_fluid"], heat_power=d["heat_power"], u_inlet_val=d["u_inlet_val"],
              heat_source=d["heat"], tol_ke=tol, tol_T=tol, silent=True, min_steps_rule=False)
    run_simulation(d["mask"], max_steps=200, **kw)  # load kernels, build graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = run_simulation(d["mask"], max_steps=max_steps, **kw)
    gpu_s = time.perf_counter() - t0
    cells = int(d["mask"].size)
    out = {"mode": "config2", "lattice": list(d["mask"].shape), "seed": seed, "family": d["family"],
           "period": d["period"], "threshold": d["threshold"], "tau_fluid": d["tau_fluid"],
           "u_inlet_val": d["u_inlet_val"], "heat_power": d["heat_power"], "tol": tol, "max_steps": max_steps,
           "steps": r["steps"], "converged": bool(r["converged"]),
           "gpu": {"wall_s": round(gpu_s, 3), "mlups": round(cells * r["steps"] / gpu_s / 1e6, 1),
                   "us_per_step": round(1e6 * gpu_s / max(r["steps"], 1), 2),
                   "note": "run_simulation end to end: init, graph replays, residual checks every 100 steps, "
                           "field copies"}}
    for threads, sec in ((1, cpu_seconds), (orc.max_threads(), cpu_seconds / 2)):
        mlups, n, t = cpu_baseline(d, sec, threads)
        out[f"oracle_{threads}t"] = {"steps": n, "wall_s": round(t, 2), "mlups": round(mlups, 3),
                                     "s_per_step": round(t / max(n, 1), 4),
                                     "solve_s_at_this_rate": round(r["steps"] * t / max(n, 1), 1)}
    out["host"] = host_cpu()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256,
                    help="reported as given; the workload is --designs designs (a step = one design solved)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed designs (1,000 steps each arm) before the timed ones")
    ap.add_argument("--max-steps", type=int, default=30000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["solves", "passes", "copy", "ceilings", "batch", "config1", "config2",
                                       "train-step", "designs"],
                    default="solves",
                    help="solves: the benchmark; passes / copy: profiling drivers (pure fused passes; a plain copy); "
                         "ceilings: bandwidth of plain copies and of the pass's access shapes")
    ap.add_argument("--passes", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0, help="config2 mode: design seed")
    ap.add_argument("--copy-mb", type=float, default=512.0)
    ap.add_argument("--designs", type=int, default=None,
                    help="solves: timed designs (default 256); batch / designs modes: design count (default 64)")
    ap.add_argument("--stop-rule", choices=["rise", "reference"], default="rise",
                    help="solves: residuals of both arms (rise: the reference's formulas on the temperature rise "
                         "and the absorbed heat)")
    ap.add_argument("--rows", default=None, help="solves: also write the per-design rows to this JSON file")
    ap.add_argument("--groups", type=int, default=1,
                    help="batch: batches of --slots lattices sharing the queue, taking turns chunk by chunk")
    ap.add_argument("--slots", type=int, default=None, help="batch slots (default: 1 for passes, 64 for batch)")
    ap.add_argument("--launch", choices=["tune", "fixed"], default="fixed",
                    help="solves mode: autotune the fused pass's launch shape before the timed arms, or use the "
                         "--block/--cpt/--xcd/--sweep flags")
    ap.add_argument("--sweep", choices=["alternate", "forward"], default="alternate",
                    help="block order of the fused pass: alternate reverses every even pass (Infinity-Cache reuse)")
    ap.add_argument("--block", type=int, default=256, help="threads per block of the fused pass")
    ap.add_argument("--cpt", type=int, default=2, help="lattice cells per thread of the fused pass")
    ap.add_argument("--xcd", action="store_true", help="map consecutive lattice blocks to one XCD")
    ap.add_argument("--slab", type=int, default=0,
                    help="passes/batch modes: x-planes per slab of the wavefront pass pairs (0: whole passes)")
    ap.add_argument("--thermal-prefetch", choices=["auto", "on", "off"], default="auto",
                    help="fused D3Q19 pass: issue the thermal loads right after the fluid gather (faster at a few "
                         "lattices per launch, slower when the state is HBM-resident); auto: on in the solves "
                         "mode (three lattices per launch), off elsewhere")
    ap.add_argument("--precision", type=int, choices=[32, 64], default=32, help="passes mode: FP32 or FP64 lattices")
    ap.add_argument("--thermal-lattice", choices=["D3Q19", "D3Q7"], default="D3Q19",
                    help="passes mode: thermal lattice (D3Q7: the seven-velocity variant in the same fused pass)")
    ap.add_argument("--weights", default=None,
                    help="surrogate weights: one .npz or several joined by commas, whose predictions are averaged "
                         "(default: the shipped pair)")
    args = ap.parse_args()
    if args.mode == "solves" and args.designs is None:
        args.designs = 256

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver runs independent designs per GPU and has no multi-GPU "
              "path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.mode == "passes":
        return profile_passes(args.passes, args.slots or 1, thermal_lattice=args.thermal_lattice,
                              precision=args.precision,
                              thermal_prefetch=args.thermal_prefetch == "on", block=args.block,
                              cells_per_thread=args.cpt, xcd_remap=args.xcd,
                              alternate_sweep=args.sweep == "alternate", slab_planes=args.slab)
    if args.mode == "copy":
        return profile_copy(args.copy_mb)
    if args.mode == "ceilings":
        # The bandwidth probe is a measurement aid, not part of the library: compiled here on demand.
        import subprocess
        from xadvcool_amd.build import HIPCC_FLAGS
        src, exe = os.path.join(ROOT, "bw_ceiling.hip"), os.path.join(ROOT, "bw_ceiling")
        if not os.path.exists(exe) or os.path.getmtime(exe) < os.path.getmtime(src):
            flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
            subprocess.run([os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"), *flags, "-o", exe, src], check=True)
        sys.exit(subprocess.run([exe]).returncode)
    if args.mode == "config1":
        return config1_bench()
    if args.mode == "config2":
        return config2_bench(seed=args.seed, tol=args.tol, max_steps=args.max_steps, cpu_seconds=args.cpu_seconds)
    if args.mode == "train-step":
        return train_step_bench()
    if args.mode == "designs":
        return designs_bench(args.designs or 64)
    if args.mode == "batch":
        print(json.dumps({"mode": "batch", **batch_bench(args.designs or 64, args.slots or 64, args.max_steps,
                                                         args.tol, groups=args.groups, block=args.block,
                                                         cells_per_thread=args.cpt,
                                                         alternate_sweep=args.sweep == "alternate",
                                                         slab_planes=args.slab,
                                                         thermal_prefetch=args.thermal_prefetch == "on")}),
              flush=True)
        return

    from xadvcool_amd import warmstart as W
    from xadvcool_amd.geometry import config2_design

    # The surrogate-warm-started batched workload: unseen seeded designs of all five families and all ten heat
    # maps on 148x128x32, parameters only (each arm builds its designs on the GPU inside its wall time).
    # --steps does not change it: a "step" here is one design solved to the stop rule.
    timed = W.make_design_params(W.COMPARE_SEED0, args.designs)
    net = W.load_surrogate(args.weights or W.DEFAULT_ENSEMBLE)
    nslots = args.slots or 3
    launch = dict(block=args.block, cells_per_thread=args.cpt, xcd_remap=args.xcd,
                  alternate_sweep=args.sweep == "alternate", thermal_prefetch=args.thermal_prefetch != "off")
    tuned_us = None
    if args.launch == "tune":
        # Untimed set-up, like a compile step: the project's autotuner times each launch shape on ``nslots``
        # full-size designs outside the timed set.  Every shape computes the same bits.
        from xadvcool_amd import autotune
        tune_designs = [config2_design(sd) for sd in range(900, 900 + nslots)]
        best, times = autotune.tune(tune_designs, LATTICE, passes=300, save=False,
                                    candidates=[c for c in autotune.CANDIDATES if c["block"] >= 128])
        launch = dict({k: best[k] for k in ("block", "cells_per_thread", "xcd_remap", "alternate_sweep")},
                      thermal_prefetch=args.thermal_prefetch != "off")
        tuned_us = round(times[json.dumps(best, sort_keys=True)], 2)
        print(f"[bench] autotuned launch {launch}: {tuned_us} us/pass", file=sys.stderr, flush=True)
    kw = dict(slots=nslots, tol=args.tol, max_steps=args.max_steps, launch=launch, stop_rule=args.stop_rule,
              confirm=W.CONFIRM, log=lambda m: print(m, file=sys.stderr, flush=True))
    if args.warmup:  # untimed: kernel loads, graph captures, the surrogate's first run
        w = W.make_design_params(900_000, args.warmup)
        W.solve_designs(w, None, **dict(kw, max_steps=1000))
        W.solve_designs(w, net, **dict(kw, max_steps=1000))
    timer = PassTimer(None)  # every pass of both arms (the population a kernel trace of this command averages)
    t_start = time.perf_counter()
    cold = W.solve_designs(timed, None, timer=timer, **kw)
    warm = W.solve_designs(timed, net, timer=timer, **kw)
    t_arms = time.perf_counter() - t_start
    rows = W.compare_rows(timed, cold, warm)
    summ = W.summarize(rows, cold, warm)
    cells = int(np.prod(LATTICE))
    c_steps = [r["cold_steps"] for r in rows]
    w_steps = [r["warm_steps"] for r in rows]
    pass_ms, npass, alg_bytes = timer.totals()
    avg_pass_ms = pass_ms / max(npass, 1)
    achieved = alg_bytes / (pass_ms * 1e-3) / 1e9
    counters = pass_counters()
    c_rate, w_rate = summ["cold"]["solves_per_hr"], summ["warm"]["solves_per_hr"]
    rule = ("rise residuals (the reference's d_ke / dT / dQ formulas on the temperature rise max T - T_in and the "
            "absorbed heat sum(u_x T) - T_in sum(u_x))" if args.stop_rule == "rise"
            else "the reference's d_ke / dT / dQ residuals")
    line = {
        "metric": "converged solves/hr, warm-start vs plain LBM; MLUPS and % of HBM peak",
        "value": w_rate,
        "unit": "converged solves/hr (surrogate warm start, physics gate, cold fallback, inference included)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * warm["wall"] / max(len(timed), 1), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded unseen designs (seeds from 1,000,000), all five families and all ten heat maps; "
                "surrogate (two networks, predictions averaged) trained only on the solver's own converged "
                "solutions of other seeds",
        "config": {"workload": f"surrogate-warm-started batched solving: {len(timed)} unseen 148x128x32 designs "
                               "(straight, pin fin, gyroid, Schwarz P, Schwarz D; ten heat maps), each solved cold "
                               "and warm to the stop rule",
                   "designs": len(timed), "lattice": list(LATTICE), "tol": args.tol, "max_steps": args.max_steps,
                   "stop_rule": f"{rule} every 100 steps below tol on {W.CONFIRM} consecutive checks, "
                                "flow-through floor off, both arms",
                   "step_definition": "one design solved to the stop rule (--steps does not change the workload)",
                   "designs_per_launch": nslots,
                   "families": {f: v["designs"] for f, v in summ["by_family"].items()},
                   "launch": dict(launch, autotuned=args.launch == "tune", tuned_us_per_pass=tuned_us)},
        "cold_start": dict(summ["cold"], mlups=round(cells * sum(c_steps) / cold["wall"] / 1e6, 1)),
        "warm_start": dict(summ["warm"], mlups=round(cells * sum(w_steps) / warm["wall"] / 1e6, 1)),
        "warm_vs_cold": summ["speedup_solves_per_hr"],
        "acceptance": {k: summ[k] for k in ("both_converged", "within_0.5pct", "warm_only_converged",
                                            "max_d_pressure_drop", "median_d_pressure_drop", "max_d_R_th",
                                            "median_d_R_th", "cold_converged_failing_gate")},
        "by_family": summ["by_family"],
        "mlups": round(cells * (sum(c_steps) + sum(w_steps)) / (cold["wall"] + warm["wall"]) / 1e6, 1),
        "timed_s": round(t_arms, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (int(counters["traffic_per_alg_byte"] * alg_bytes / max(npass, 1))
                                 if counters else None),
                     "kernel": "k_pass (fused AA fluid+thermal pass)",
                     "avg_launch_us": round(1000.0 * avg_pass_ms, 2),
                     "alg_bytes_per_launch": int(alg_bytes / max(npass, 1)),
                     "traffic_source": (f"rocprofv3 FETCH_SIZE x{counters['fetch_scale']} + WRITE_SIZE "
                                        f"x{counters['write_scale']} per pass / algorithmic bytes = "
                                        f"{counters['traffic_per_alg_byte']:.3f} on {counters['workload']} "
                                        f"({COUNTERS_FILE})" if counters else None),
                     "passes_timed": npass,
                     "note": f"HIP events on the solver's stream around every graph chunk of both arms; up to "
                             f"{nslots} lattices per launch (~140 MB touched per lattice per pass): partly resident "
                             "in the 256 MB Infinity Cache; the all-HBM figure is bench.py --mode passes --slots 64"},
    }
    if not args.no_cpu:
        d0 = dict(config2_design(0), key=0)  # a full-size gyroid design: the oracle's per-step rate
        mlups1, n1, t1 = cpu_baseline(d0, args.cpu_seconds, 1)
        cpu_s = sum(c_steps) * cells / (mlups1 * 1e6)  # the cold arm's lattice updates at the oracle's rate
        line["cpu_baseline"] = {
            "value": round(3600.0 * summ["cold"]["converged"] / cpu_s, 4),
            "unit": "converged solves/hr (the cold arm's designs and steps at the oracle's measured MLUPS)",
            "cores": 1, "kind": "port", "mlups": round(mlups1, 3),
            "sample": f"oracle six-kernel FP32 schedule on a 148x128x32 gyroid design, {n1} steps in {t1:.1f} s, "
                      f"1 thread",
        }
        from oracle import oracle as orc
        nt = orc.max_threads()
        mlups_n, n_n, t_n = cpu_baseline(d0, args.cpu_seconds / 2, nt)
        line["cpu_baseline"]["all_threads"] = {
            "threads": nt, "mlups": round(mlups_n, 3),
            "solves_per_hr": round(3600.0 * summ["cold"]["converged"] / (sum(c_steps) * cells / (mlups_n * 1e6)), 4),
            "sample": f"{n_n} steps in {t_n:.1f} s"}
        line["cpu_baseline"]["host"] = host_cpu()
    if args.rows:
        with open(args.rows, "w") as f:
            json.dump({"line": line, "rows": rows}, f)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
