"""Launch-shape autotuner for the fused pass.

Candidates: threads per block (64, 128, 256) x lattice cells per thread
(1, 2) x XCD-contiguous block mapping (on, off) x alternating sweep
direction (even passes visit the blocks in reverse order: on, off).  Each
candidate is timed with HIP events on the solver's stream over
graph-captured passes of the real lattice (or batch); the fastest shape is
cached per (lattice shapes, device) in a JSON file next to this module
(XADV_TUNE_CACHE overrides the path).  The shapes change only performance:
every candidate computes the same bits (the parity tests run them).

Which shape wins depends on the working set: with a batch state of a few
hundred MB (partly resident in the 256 MB Infinity Cache) the XCD mapping
keeps neighbour lines in one XCD's L2 and wins; with many GB in HBM the
linear order does.  The alternating sweep starts each pass on the lines the
previous pass touched last, so it never loses.
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


def time_shape(designs, cap_shape, shape, passes=200, precision=32):
    """Average pass time (us) of one launch shape on a batch holding ``designs`` (list of dicts with mask/heat)."""
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
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        b.step(passes, want_fields=False)
        e1.record(st)
        torch.cuda.synchronize()
        return 1000.0 * e0.elapsed_time(e1) / passes
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
