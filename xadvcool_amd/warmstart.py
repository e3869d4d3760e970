"""Warm-started batched solving: data generation, surrogate training, and the cold/warm comparison.

* ``generate``: designs are solved cold by the batched solver; their converged
  fields become training pairs (features, standardised targets, loss weights),
  kept on the GPU.
* ``train``: minibatch Adam on the project's HIP kernels (surrogate.py).
* ``solve_designs``: one batched run over a design list, either cold or warm.
  Warm designs start from the surrogate's (rho, u, T) (solids at rest);
  a warm run whose residual grows for ``stall_checks`` consecutive checks
  after ``stall_after`` steps, or that diverges, is abandoned and re-queued
  cold (the plain LBM is the fallback and the correctness reference).
* ``compare``: cold vs warm on the same designs with the same stop rule:
  steps, converged solves per hour (inference and fallbacks included in the
  warm arm's wall time), and the relative differences of the converged
  pressure drop and thermal resistance.

Command line: ``python -m xadvcool_amd.warmstart train|compare ...`` (see main()).
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from . import surrogate as S
from .batch import BatchedSolver, device_inputs
from .checks import gate_failures
from .geometry import HEAT_SOURCE_NAMES, family_design, family_params

CORE = (128, 128, 32)  # 148 x 128 x 32 lattices
PAD = (S.round8(CORE[0] + 20), S.round8(CORE[1]), S.round8(CORE[2]))
CONFIRM = 5  # consecutive passing residual checks before a run counts as converged (both arms)
COMPARE_SEED0 = 1_000_000  # evaluation designs: seeds far from every training seed
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")
# The shipped surrogate: two networks of the same architecture (the second fine-tuned from the first on more
# solver solutions) whose predicted fields are averaged.
DEFAULT_ENSEMBLE = (DEFAULT_WEIGHTS,
                    os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights_2.npz"))


class Ensemble:
    """Surrogates whose predicted (rho, u, T) are averaged, each de-normalised with its own statistics."""

    def __init__(self, nets):
        self.nets = list(nets)
        if not self.nets:
            raise ValueError("an ensemble needs at least one network")


def load_surrogate(paths=DEFAULT_ENSEMBLE):
    """One network (a path) or the average of several (a sequence of paths, or paths joined by commas)."""
    if isinstance(paths, str):
        paths = [p for p in paths.split(",") if p]
    nets = [S.UNet3D.load_file(p) for p in paths]
    return nets[0] if len(nets) == 1 else Ensemble(nets)


def make_designs(seed0, n, heats=HEAT_SOURCE_NAMES, families=None):
    """n seeded 148x128x32 designs from seed0 on; with ``families``, the first n of those families."""
    out = []
    for p in make_design_params(seed0, n, heats, families):
        d = family_design(p["key"], core=CORE, heat_name=p["heat_name"])
        d["key"] = p["key"]
        out.append(d)
    return out


def make_design_params(seed0, n, heats=HEAT_SOURCE_NAMES, families=None):
    """The designs of ``make_designs`` as parameters only (``params`` plus the solver's scalars): the
    batched solver builds each one on the GPU when its slot starts."""
    out = []
    s = seed0
    while len(out) < n:
        rng = np.random.default_rng(10_000_019 * (s + 1))
        heat = heats[int(rng.integers(0, len(heats)))]
        p = family_params(s, core=CORE, heat_name=heat)
        if families is None or p["family"] in families:
            out.append(dict(params=p, key=s, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"],
                            heat_power=p["heat_power"], family=p["family"], heat_name=p["heat_name"]))
        s += 1
    return out


def config2_design_params(seed0, n):
    """n full-size single-solve designs (gyroid, quad_core heat map, u_in up to 0.05) as parameters only."""
    from .geometry import config2_params
    out = []
    for s in range(seed0, seed0 + n):
        p = config2_params(s)
        out.append(dict(params=p, key=s, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"],
                        heat_power=p["heat_power"], family=p["family"], heat_name=p["heat_name"]))
    return out


def _sample(d, fields, inputs=None, t_weight=1.0):
    """One training pair on the GPU: (features bf16, raw targets, loss weights, design scalars).
    ``inputs``: the design's device (mask, normalised heat) when the solver built it."""
    rho, u, T = fields
    mask, heat = inputs if inputs is not None else device_inputs(dict(d))
    f = S.features_device(mask, heat, d["u_inlet_val"], d["tau_fluid"], d["heat_power"], PAD)
    t, w = S.targets(mask.cpu().numpy(), rho, u, T, d["u_inlet_val"], d["heat_power"], PAD, t_weight=t_weight)
    sc = torch.from_numpy(S.scalars(d["tau_fluid"], d["heat_power"], d["u_inlet_val"])).cuda()
    return f.to(torch.bfloat16), torch.from_numpy(t).cuda(), torch.from_numpy(w).cuda(), sc


def _stack(samples, idx, m_t, s_t):
    """Minibatch tensors (x, standardised targets, weights, scalars) of samples[idx]."""
    x = torch.stack([samples[i][0] for i in idx])
    t = ((torch.stack([samples[i][1] for i in idx]) - m_t) / s_t).contiguous()
    w = torch.stack([samples[i][2] for i in idx])
    sc = torch.stack([samples[i][3] for i in idx]).contiguous()
    return x, t, w, sc


def generate(designs, net=None, slots=3, tol=1e-6, max_steps=120000, log=print, confirm=CONFIRM,
             stop_rule="rise", keep_unconverged=False, t_weight=1.0):
    """Training pairs from the batched solver: returns (samples, results) for the designs that converged
    (and, with ``keep_unconverged``, those stopped at ``max_steps``).

    With ``net`` the solves start from its predictions (the warm-start path, fallback and physics gate
    included): they reach the same steady state as cold solves, in fewer steps.  The targets must be the
    converged state the warm runs will be stopped at, so the stop rule and tolerance are the evaluation's:
    the thermal field relaxes with a time constant of ~15,000 steps on these lattices, and fields stopped at
    residual 1e-5 still differ from the 1e-6 ones by 10-25 % in R_th."""
    by_key = {d["key"]: d for d in designs}
    samples, results = [], []
    t0 = time.perf_counter()

    def keep(r):
        fields, inputs = r.pop("fields"), r.pop("inputs")
        if r["status"] == "converged" or (keep_unconverged and r["status"] == "max_steps"):
            samples.append(_sample(by_key[r["key"]], fields, inputs, t_weight=t_weight))
        results.append(r)
        if len(results) % 32 == 0:
            log(f"[datagen] {len(results)}/{len(designs)} solved ({len(samples)} kept as samples), "
                f"{time.perf_counter() - t0:.0f} s")

    out = solve_designs(designs, net, slots=slots, tol=tol, max_steps=max_steps, log=None, confirm=confirm,
                        stop_rule=stop_rule, collect_fields=True, on_final=keep)
    log(f"[datagen] {len(designs)} designs in {out['wall']:.0f} s: {len(samples)} samples, "
        f"{len(out['fallbacks'])} warm fallbacks")
    return samples, results


def fit_norm(samples):
    """Per-channel mean/std of the targets over their weighted entries."""
    s1 = torch.zeros(S.NOUT, dtype=torch.float64, device="cuda")
    s2 = torch.zeros_like(s1)
    n = torch.zeros_like(s1)
    for _, t, w, _ in samples:
        tw = t.double() * w
        s1 += tw.sum(dim=(0, 1, 2))
        s2 += (tw * t.double()).sum(dim=(0, 1, 2))
        n += w.double().sum(dim=(0, 1, 2))
    mean = s1 / n
    std = torch.sqrt(torch.clamp(s2 / n - mean * mean, min=1e-12))
    return mean.float().cpu().numpy(), std.float().cpu().numpy()


def train(net, samples, epochs=40, batch=8, lr=5e-4, seed=0, val=None, log=print, ckpt=None, ckpt_every=10,
          warmup=100, deadline=None):
    """Adam, linear warm-up over ``warmup`` steps then cosine decay.  (Peak rates of 2e-3 and above blow up the
    decoder activations of this un-normalised network within ten steps and leave it dead; 5e-4 trains.)
    ``deadline`` (time.perf_counter() value): stop after the epoch that passes it."""
    mean, std = net.norm
    m_t, s_t = torch.tensor(mean, device="cuda"), torch.tensor(std, device="cuda")
    rng = np.random.default_rng(seed)
    nsteps = epochs * max(1, len(samples) // batch)
    step = 0
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot, cnt = 0.0, 0
        for b0 in range(0, len(order) - batch + 1, batch):
            idx = order[b0:b0 + batch]
            x, t, w, sc = _stack(samples, idx, m_t, s_t)
            out = net.forward(x, sc, keep=True)
            loss, g = S.mse_loss(out, t, w)
            net.zero_grad()
            net.backward(g)
            cur_lr = lr * min(1.0, (step + 1) / warmup) * 0.5 * (1.0 + np.cos(np.pi * step / max(nsteps, 1)))
            net.adam_step(float(cur_lr))
            step += 1
            tot += loss.item()
            cnt += 1
        entry = {"epoch": ep + 1, "train_loss": tot / max(cnt, 1), "s": round(time.perf_counter() - t0, 1)}
        if val is not None and (ep % 5 == 4 or ep == epochs - 1):
            entry["val_loss"] = evaluate(net, val)
        hist.append(entry)
        log(f"[train] {json.dumps(entry)}")
        if ckpt is not None and (ep + 1) % ckpt_every == 0:
            net.save(ckpt)
        if deadline is not None and time.perf_counter() > deadline:
            log(f"[train] time budget reached after epoch {ep + 1}")
            break
    return hist


def time_train_step(net, samples, batch=8, reps=3):
    """Seconds per forward + backward pass of one batch (no weight update; gradients cleared after)."""
    mean, std = net.norm
    m_t, s_t = torch.tensor(mean, device="cuda"), torch.tensor(std, device="cuda")
    idx = list(range(min(batch, len(samples))))
    x, t, w, sc = _stack(samples, idx, m_t, s_t)
    dt = 0.0
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = net.forward(x, sc, keep=True)
        _, g = S.mse_loss(out, t, w)
        net.backward(g)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    net.zero_grad()
    return dt


@torch.no_grad()
def evaluate_channels(net, samples, batch=8):
    """Weighted MSE of each standardised target channel (u_x, u_y, u_z, rho, T) over ``samples``."""
    mean, std = net.norm
    m_t, s_t = torch.tensor(mean, device="cuda"), torch.tensor(std, device="cuda")
    num = torch.zeros(S.NOUT, dtype=torch.float64, device="cuda")
    den = torch.zeros_like(num)
    for b0 in range(0, len(samples), batch):
        x, t, w, sc = _stack(samples, range(b0, min(b0 + batch, len(samples))), m_t, s_t)
        out = net.forward(x, sc)[..., :S.NOUT]
        num += ((out - t) ** 2 * w).double().sum(dim=(0, 1, 2, 3))
        den += w.double().sum(dim=(0, 1, 2, 3))
    return (num / den).cpu().tolist()


@torch.no_grad()
def evaluate(net, samples, batch=8):
    mean, std = net.norm
    m_t, s_t = torch.tensor(mean, device="cuda"), torch.tensor(std, device="cuda")
    tot, n = 0.0, 0
    for b0 in range(0, len(samples), batch):
        idx = range(b0, min(b0 + batch, len(samples)))
        x, t, w, sc = _stack(samples, idx, m_t, s_t)
        loss, _ = S.mse_loss(net.forward(x, sc), t, w)
        tot += loss.item() * len(idx)
        n += len(idx)
    return tot / max(n, 1)


@torch.no_grad()
def predict_fields(net, d):
    """Surrogate (rho, u, T) for a design, on the GPU (FP32, lattice shape); an Ensemble gives the mean of its
    networks' fields.

    Solid and INLET cells keep the cold-start values (rho 1, u 0).  Solid
    populations never change after the init.  INLET cells never read a
    neighbour (src/fluid_engine3d.py:196-244 rebuilds them from their own
    post-collision set), so their state is a memory of the initial one: the
    inlet density, hence the pressure level at x = 1 and the converged pressure
    drop, would otherwise depend on the prediction (measured: up to 13 %)."""
    if isinstance(net, Ensemble):
        outs = [predict_fields(n, d) for n in net.nets]
        return tuple((sum(o[i] for o in outs) / len(outs)).contiguous() for i in range(3))
    mask, heat = device_inputs(dict(d))
    x = S.features_device(mask, heat, d["u_inlet_val"], d["tau_fluid"], d["heat_power"], PAD).to(torch.bfloat16)[None]
    sc = torch.from_numpy(S.scalars(d["tau_fluid"], d["heat_power"], d["u_inlet_val"])).cuda()[None]
    out = net.forward(x, sc)[0]
    nx, ny, nz = mask.shape
    mean, std = net.norm
    y = out[:nx, :ny, :nz, :S.NOUT] * torch.tensor(std, device="cuda") + torch.tensor(mean, device="cuda")
    free = (mask != S.BC_SOLID) & (mask != S.BC_INLET)
    u = torch.where(free[..., None], y[..., 0:3] * d["u_inlet_val"], torch.zeros_like(y[..., 0:3]))
    rho = torch.where(free, 1.0 + y[..., 3] * d["u_inlet_val"], torch.ones_like(y[..., 3]))
    T = S.T_IN + y[..., 4] * d["heat_power"]
    return rho.contiguous(), u.contiguous(), T.contiguous()


class _Stall:
    """Decides when a warm run is abandoned for a cold re-run.

    The run's residual is the worst of its three stop-rule ratios
    (d_ke / tol_ke, dT / tol_T, dQ / tol_T), smoothed by a running median over
    ``window`` checks so that a transient's rate crossing zero or FP32 noise in
    a converged quantity does not count.  After ``after`` steps (about one
    flow-through time: a warm start's residual typically dips right after the
    start, rises while the slow modes develop, then decays), the run is
    abandoned when the smoothed residual
      * grows: it is more than ``growth`` times its value ``grow_span`` checks
        earlier, or
      * stalls: over the last ``patience`` checks it stayed within a factor
        ``plateau`` (neither decaying nor growing).  A warm run on a TPMS core can
        hold a plateau for 10,000-15,000 steps while two thermal modes cross and
        then decay again, still far ahead of where a cold run is (measured: a
        run abandoned this way at 26,000 steps had the residual a cold run only
        reaches at ~80,000 steps), so the window is 300 checks.
    Cold runs are never abandoned.
    """

    def __init__(self, tol_ke, tol_T, after=10000, window=20, growth=2.0, grow_span=30, patience=300, plateau=1.25):
        self.tol_ke, self.tol_T = tol_ke, tol_T
        self.after, self.window, self.patience = after, window, patience
        self.growth, self.grow_span, self.plateau = growth, grow_span, plateau
        self.runs = {}

    def __call__(self, design, step, res):
        if not design.get("warm"):
            return False
        worst = max(res.d_ke / self.tol_ke, res.dT_conv / self.tol_T, res.dQ_conv / self.tol_T)
        st = self.runs.setdefault(design["key"], {"hist": [], "smooth": []})
        st["hist"].append(worst)
        if len(st["hist"]) > self.window:
            st["hist"].pop(0)
        if len(st["hist"]) < self.window:
            return False
        smooth = float(np.median(st["hist"]))
        sm = st["smooth"]
        sm.append(smooth)
        keep = max(self.patience, self.grow_span) + 1
        if len(sm) > keep:
            sm.pop(0)
        if step <= self.after:
            return False
        grows = len(sm) > self.grow_span and smooth > self.growth * sm[-self.grow_span - 1]
        recent = sm[-self.patience - 1:]
        stalls = len(sm) > self.patience and max(recent) < self.plateau * min(recent)
        return grows or stalls


def solve_designs(designs, net=None, slots=64, tol=1e-6, max_steps=30000, log=print, predictor=None, timer=None,
                  confirm=CONFIRM, stop_rule="reference", launch=None, min_steps_rule=False, gate=True,
                  collect_fields=False, on_final=None):
    """One batched run; with ``net`` (or any ``predictor(design) -> (rho, u, T)``) designs start warm.

    Fallback: a warm design that stalls or diverges is re-queued cold in the same run, and so is a warm design
    that converges to fields failing the physics acceptance gate (``gate``: checks.GATE, the field checks of
    analyze_dataset.py:184-325 reduced on the device; every result carries its ``checks``).  Both arms stop
    when the reference's residual test passes on ``confirm`` consecutive checks.  ``timer`` (optional):
    ``begin()`` / ``end(n, bytes_per_pass)`` around every chunk of passes; its ``stream`` is set to the solver's.
    ``launch``: launch-shape options of the fused pass (``block``, ``cells_per_thread``, ``xcd_remap``,
    ``alternate_sweep``); they change timing only.  ``min_steps_rule``: the reference's flow-through floor
    (src/simulation.py:104; off by default here, on in the reference).  ``collect_fields``: results carry
    host copies of (rho, u, T) and the device inputs; ``on_final(r)`` sees each final result as its slot
    retires (and may drop the fields)."""
    if predictor is None and net is not None:
        predictor = lambda d: predict_fields(net, d)  # noqa: E731
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs = BatchedSolver(slots, (CORE[0] + 20, CORE[1], CORE[2]), tol_ke=tol, tol_T=tol, max_steps=max_steps,
                       metrics=True, stall_check=_Stall(tol, tol), confirm=confirm, stop_rule=stop_rule,
                       min_steps_rule=min_steps_rule, checks=True, collect_fields=collect_fields,
                       **(launch or {}))
    queue = []
    t_inf = 0.0
    for d in designs:
        dd = dict(d)
        if predictor is not None:
            if dd.get("mask") is None:  # parameters only: build the design on the GPU for the prediction
                device_inputs(dd, bs.batch)
            ti = time.perf_counter()
            dd["init_fields"] = predictor(dd)
            torch.cuda.synchronize()
            t_inf += time.perf_counter() - ti
            dd["warm"] = True
        queue.append(dd)
    if timer is not None:  # events go on the solver's own stream
        timer.stream = bs.batch.stream()
    final = {}
    fallbacks = []
    by_key = {q["key"]: q for q in queue}
    fell_back = set()
    done = [0]

    def handle(r):
        done[0] += 1
        if log is not None and done[0] % 32 == 0:
            log(f"[solve] {'warm' if predictor is not None else 'cold'}: {done[0]} runs finished, "
                f"{time.perf_counter() - t0:.0f} s")
        k = r["key"]
        why = None
        if predictor is not None and k not in fell_back:
            if r["status"] in ("stalled", "diverged"):
                why = r["status"]  # the warm start did not lead to convergence
            elif gate and r["converged"]:
                failed = gate_failures(r["checks"])
                if failed:  # converged to fields that fail the physics checks
                    why = "checks:" + ",".join(failed)
        if why is not None:  # rerun this design cold, at the end of the queue
            fell_back.add(k)
            fallbacks.append((k, why, r["steps"]))
            final[k] = {"abandoned_steps": r["steps"]}
            r.pop("fields", None)
            return [{kk: v for kk, v in by_key[k].items() if kk not in ("init_fields", "warm")}]
        r["abandoned_steps"] = final.get(k, {}).get("abandoned_steps", 0)
        if on_final is not None:
            on_final(r)
        final[k] = r
        return None

    def progress_line(active, queued, finished):
        if log is not None:
            log(f"[solve] {'warm' if predictor is not None else 'cold'}: {active} running, {queued} queued, "
                f"{len(final)} finished, {time.perf_counter() - t0:.0f} s")

    try:
        bs.run(queue, timer=timer, on_result=handle, progress=progress_line)
    finally:
        lups = bs.lups
        bs.close()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return dict(results=final, wall=wall, inference_s=t_inf, fallbacks=fallbacks, lups=lups)


def compare_rows(designs, cold, warm):
    """Per-design comparison of a cold and a warm ``solve_designs`` run over the same designs."""
    rows = []
    for d in designs:
        c, w = cold["results"][d["key"]], warm["results"][d["key"]]
        row = dict(key=d["key"], family=d["family"], heat=d["heat_name"], u_in=d["u_inlet_val"],
                   cold_steps=c["steps"], warm_steps=w["steps"] + w.get("abandoned_steps", 0),
                   cold_status=c["status"], warm_status=w["status"],
                   cold_gate=gate_failures(c["checks"]) if "checks" in c else None,
                   warm_gate=gate_failures(w["checks"]) if "checks" in w else None)
        for arm, r in (("cold", c), ("warm", w)):
            if r.get("converged"):
                row[arm + "_R_th"] = r["metrics"]["R_th"]
                row[arm + "_pressure_drop"] = r["metrics"]["pressure_drop"]
        if c["converged"] and w["converged"]:
            for m in ("pressure_drop", "R_th", "Nu", "P_pump"):
                a, b = w["metrics"][m], c["metrics"][m]
                row["d_" + m] = abs(a - b) / max(abs(b), 1e-30)
        rows.append(row)
    return rows


def summarize(rows, cold, warm):
    """Solves per hour of both arms, their ratio, and the agreement of the converged results.

    Converged results are compared where both arms converged ("both"); a warm result whose cold run hit
    the step cap has no cold result to judge it and is counted apart ("warm_only": accepted by the physics
    gate alone)."""
    def arm(a, key):
        n = sum(r[key] == "converged" for r in rows)
        return {"converged": n, "wall_s": round(a["wall"], 2), "solves_per_hr": round(3600 * n / a["wall"], 1),
                "mean_steps": float(np.mean([r[key.replace("status", "steps")] for r in rows]))}

    both = [r for r in rows if "d_R_th" in r]
    within = [r for r in both if r["d_pressure_drop"] <= 5e-3 and r["d_R_th"] <= 5e-3]
    reasons = {}
    for _, why, _ in warm["fallbacks"]:
        key = why.split(":")[0]
        reasons[key] = reasons.get(key, 0) + 1
    summ = {
        "designs": len(rows),
        "cold": arm(cold, "cold_status"),
        "warm": dict(arm(warm, "warm_status"), inference_s=round(warm["inference_s"], 2),
                     fallbacks=len(warm["fallbacks"]), fallback_reasons=reasons),
        "both_converged": len(both),
        "within_0.5pct": len(within),
        "warm_only_converged": sum(1 for r in rows if r["warm_status"] == "converged"
                                   and r["cold_status"] != "converged"),
        "max_d_pressure_drop": max((r["d_pressure_drop"] for r in both), default=None),
        "median_d_pressure_drop": float(np.median([r["d_pressure_drop"] for r in both])) if both else None,
        "max_d_R_th": max((r["d_R_th"] for r in both), default=None),
        "median_d_R_th": float(np.median([r["d_R_th"] for r in both])) if both else None,
        "cold_converged_failing_gate": sum(1 for r in rows if r["cold_status"] == "converged" and r["cold_gate"]),
    }
    summ["speedup_solves_per_hr"] = round(summ["warm"]["solves_per_hr"] / max(summ["cold"]["solves_per_hr"], 1e-9),
                                          3)
    by_family = {}
    for r in rows:
        fam = by_family.setdefault(r["family"], {"designs": 0, "cold_converged": 0, "warm_converged": 0,
                                                 "both": 0, "within_0.5pct": 0, "cold_steps": 0, "warm_steps": 0})
        fam["designs"] += 1
        fam["cold_converged"] += r["cold_status"] == "converged"
        fam["warm_converged"] += r["warm_status"] == "converged"
        fam["cold_steps"] += r["cold_steps"]
        fam["warm_steps"] += r["warm_steps"]
        if "d_R_th" in r:
            fam["both"] += 1
            fam["within_0.5pct"] += r["d_R_th"] <= 5e-3 and r["d_pressure_drop"] <= 5e-3
    for fam in by_family.values():  # steps-weighted: each arm's share of the GPU time goes with its steps
        fam["steps_ratio_cold_over_warm"] = round(fam["cold_steps"] / max(fam["warm_steps"], 1), 3)
        # converged designs per lattice step of each arm (every design of the table has the same lattice size,
        # so this is the ratio of converged solves per hour the family would see alone)
        if fam["cold_converged"]:
            fam["converged_per_step_ratio"] = round((fam["warm_converged"] / max(fam["warm_steps"], 1)) /
                                                    (fam["cold_converged"] / fam["cold_steps"]), 3)
    summ["by_family"] = by_family
    return summ


def compare(net, designs, slots=64, tol=1e-6, max_steps=30000, log=print, confirm=CONFIRM, stop_rule="reference",
            min_steps_rule=False):
    kw = dict(confirm=confirm, stop_rule=stop_rule, min_steps_rule=min_steps_rule)
    cold = solve_designs(designs, None, slots, tol, max_steps, log, **kw)
    warm = solve_designs(designs, net, slots, tol, max_steps, log, **kw)
    rows = compare_rows(designs, cold, warm)
    return summarize(rows, cold, warm), rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "compare"])
    ap.add_argument("--n-train", type=int, default=512)
    ap.add_argument("--n-val", type=int, default=32)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed0", type=int, default=None,
                    help="first design seed (default: 0 for train, %d for compare, disjoint from training)"
                    % COMPARE_SEED0)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--weights", default=None,
                    help="train: where the trained weights go (default: weights.npz in --out, so the shipped files "
                         "are never overwritten by accident); compare: the surrogate, one .npz or several joined "
                         "by commas (default: the shipped pair)")
    ap.add_argument("--out", default="warmstart_out")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=30000)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--gen-tol", type=float, default=1e-6, help="residual of the solves that make the data")
    ap.add_argument("--gen-max-steps", type=int, default=120000)
    ap.add_argument("--gen-slots", type=int, default=3, help="lattices per launch of the data-making solves")
    ap.add_argument("--t-weight", type=float, default=1.0,
                    help="train: loss weight of the temperature channel relative to u and rho")
    ap.add_argument("--config2-frac", type=float, default=0.0,
                    help="train: share of the designs drawn as full-size single solves (gyroid, quad_core, "
                         "u_in up to 0.05) beside the all-family designs")
    ap.add_argument("--gen-cold", action="store_true",
                    help="make the data with cold solves (default: warm-started from --init-weights when given)")
    ap.add_argument("--stop-rule", choices=["reference", "rise"], default="reference",
                    help="residuals of the reference, or the same on the temperature rise and absorbed heat")
    ap.add_argument("--gen-stop-rule", choices=["reference", "rise"], default="rise",
                    help="stop rule of the solves that make the training data")
    ap.add_argument("--keep-unconverged", action="store_true",
                    help="also train on designs stopped at --gen-max-steps (not on diverged ones)")
    ap.add_argument("--init-weights", default=None, help="continue training from these weights (and their norm)")
    ap.add_argument("--warmup", type=int, default=100, help="linear learning-rate warm-up steps")
    ap.add_argument("--floor", action="store_true",
                    help="compare: the reference's flow-through floor on (no stop before it; src/simulation.py:104)")
    ap.add_argument("--families", default=None,
                    help="compare: comma-separated families (straight, pinfin, gyroid, schwarz_p, schwarz_d) to keep")
    ap.add_argument("--deadline-s", type=float, default=None,
                    help="stop training after the epoch that ends past this many seconds from the start")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    log = lambda s: print(s, flush=True)
    if args.cmd == "train":
        t_start = time.perf_counter()
        seed0 = 0 if args.seed0 is None else args.seed0
        n = args.n_train + args.n_val
        n2 = int(round(args.config2_frac * n))
        designs = make_design_params(seed0, n - n2)
        if n2:  # interleaved, so the validation designs (the first samples) hold both kinds
            c2 = config2_design_params(seed0 + 500_000, n2)
            step = max(1, (n - n2) // n2)
            mixed = []
            for i, d in enumerate(designs):
                mixed.append(d)
                if i % step == step - 1 and c2:
                    mixed.append(c2.pop(0))
            designs = mixed + c2
        net = S.UNet3D.load_file(args.init_weights) if args.init_weights else None
        samples, results = generate(designs, None if args.gen_cold else net, slots=args.gen_slots,
                                    tol=args.gen_tol, max_steps=args.gen_max_steps, log=log,
                                    stop_rule=args.gen_stop_rule, keep_unconverged=args.keep_unconverged,
                                    t_weight=args.t_weight)
        nval = min(args.n_val, len(samples) // 8)
        val, tr = samples[:nval], samples[nval:]
        if net is None:
            net = S.UNet3D(seed=0)
            net.norm = fit_norm(tr)
        log(f"[train] {len(tr)} train / {len(val)} val samples, {net.n_params()} parameters, norm {net.norm}")
        deadline = t_start + args.deadline_s if args.deadline_s else None
        epochs = args.epochs
        if deadline is not None:  # fit the whole schedule (warm-up + cosine) into the time left
            step_s = time_train_step(net, tr)
            per_epoch = step_s * max(1, len(tr) // 8) * 1.1 + 0.5
            epochs = max(1, min(epochs, int((deadline - time.perf_counter()) / per_epoch)))
            log(f"[train] {step_s * 1e3:.1f} ms per step: {epochs} epochs fit the time budget")
        out_weights = args.weights or os.path.join(args.out, "weights.npz")
        hist = train(net, tr, epochs=epochs, lr=args.lr, val=val, log=log, ckpt=out_weights,
                     warmup=args.warmup, deadline=deadline)
        net.save(out_weights)
        log(f"[train] validation MSE per channel (u_x, u_y, u_z, rho, T): {evaluate_channels(net, val)}")
        with open(f"{args.out}/train_history.json", "w") as f:
            json.dump({"history": hist, "datagen": [{k: r[k] for k in ("key", "steps", "status")} for r in results]},
                      f)
        log(f"[train] saved {out_weights}")
    else:
        net = load_surrogate(args.weights or DEFAULT_ENSEMBLE)
        fams = set(args.families.split(",")) if args.families else None
        designs = make_designs(COMPARE_SEED0 if args.seed0 is None else args.seed0, args.n, families=fams)
        summ, rows = compare(net, designs, slots=args.slots, tol=args.tol, max_steps=args.max_steps, log=log,
                             stop_rule=args.stop_rule, min_steps_rule=args.floor)
        summ["stop_rule"] = args.stop_rule
        summ["flow_through_floor"] = bool(args.floor)
        with open(f"{args.out}/warmstart_compare.json", "w") as f:
            json.dump({"summary": summ, "rows": rows}, f)
        log(json.dumps(summ))


if __name__ == "__main__":
    main()
