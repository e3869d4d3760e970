"""Launch-shape autotuner for the fused pass.

Candidates: threads per block (64, 128, 256) x lattice cells per thread
(1, 2) x XCD-contiguous block mapping (on, off) x alternating sweep
direction (even passes visit the blocks in reverse order: on, off).  Each
candidate is timed with HIP events on the solver's stream over
graph-captured passes of the real lattice (or batch); the fastest shape is
cached per (lattice shapes, device) in a JSON file next to this module
(XADV_TUNE_CACHE overrides the path).  Each candidate is scored with every
slot busy and with one slot idle (time_shape), as in a batched run where
slots retire.  The shapes change only performance: every candidate computes
the same bits (the parity tests run them).

The alternating sweep starts each pass on the lines the previous pass
touched last (still in the Infinity Cache when the batch's state is larger
than it); on 148x128x32 gyroid designs it was never slower.  The
XCD-contiguous order gives each XCD one eighth of every slot, so the
heavier eighths (the all-fluid inlet and outlet buffers are in the first and
last) always land on the same two XCDs: with three such designs per launch
it measured 79.0 us per pass against 75.3 us in linear order.
"""
from __future__ import annotations

import itertools
import json
import os

CANDIDATES = [dict(block=b, cells_per_thread=c, xcd_remap=x, alternate_sweep=a)
              for b, c, x, a in itertools.product((64, 128, 256), (1, 2), (True, False), (True, False))]
DEFAULT = dict(block=256, cells_per_thread=1, xcd_remap=False, alternate_sweep=True)  # the engine's defaults
_CACHE = os.environ.get("XADV_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json"))


def _key(shapes, precision, fluid, thermal):
    import torch
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{dev}|p{precision}|f{int(fluid)}t{int(thermal)}|" + ";".join("x".join(map(str, s)) for s in shapes)


def load_cache():
    try:
        with open(_CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def cached(shapes, precision=32, fluid=True, thermal=True):
    return load_cache().get(_key(shapes, precision, fluid, thermal))


def time_shape(designs, cap_shape, shape, passes=200, precision=32, idle_phase=True):
    """Average pass time (us) of one launch shape on a batch holding ``designs`` (list of dicts with mask/heat).

    With ``idle_phase`` (and more than one design) the shape is timed twice, with every slot busy and with the
    last slot idle, and the mean is returned: in a real batched run slots retire and wait for their next
    design, so an order that parks a slot's work on some XCDs only (and leaves them idle with it) must lose."""
    import torch
    from .engine import LatticeBatch
    b = LatticeBatch(slots=len(designs), cap_shape=cap_shape, graph_chunk=100, precision=precision, **shape)
    try:
        for s, d in enumerate(designs):
            b.load(s, d["mask"], heat=d.get("heat"), tau=d.get("tau_fluid", 0.55), tau_th_fluid=0.51,
                   tau_th_solid=3.0, heat_power=d.get("heat_power", 0.1), T_in=25.0, u_in=d.get("u_inlet_val", 0.02))
            b.init(s)
        b.step(100, want_fields=False)
        st = b.stream()
        times = []
        for phase in range(2 if idle_phase and len(designs) > 1 else 1):
            if phase == 1:
                b.set_active(len(designs) - 1, False)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            b.step(passes, want_fields=False)
            e1.record(st)
            torch.cuda.synchronize()
            times.append(1000.0 * e0.elapsed_time(e1) / passes)
        return sum(times) / len(times)
    finally:
        b.close()


def tune(designs, cap_shape, passes=200, precision=32, candidates=None, save=True, verbose=False):
    """Time every candidate shape; return (best shape, {shape: us})."""
    shapes = [tuple(d["mask"].shape) for d in designs]
    times = {}
    for c in (candidates or CANDIDATES):
        t = time_shape(designs, cap_shape, c, passes, precision)
        times[json.dumps(c, sort_keys=True)] = t
        if verbose:
            print(f"[autotune] {c}: {t:.2f} us/pass", flush=True)
    best = json.loads(min(times, key=times.get))
    if save:
        cache = load_cache()
        cache[_key(shapes, precision, True, True)] = dict(best, us=times[json.dumps(best, sort_keys=True)])
        try:
            with open(_CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best, times
