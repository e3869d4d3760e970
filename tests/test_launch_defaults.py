"""Launch-shape defaults agree across the layers (no GPU needed).

The engine's constructors (LatticeBatch, LatticeSolver, BatchedSolver) and the
autotuner's DEFAULT name the same launch shape, and every candidate the
autotuner may pick is one the C ABI accepts (block 64/128/256, 1 or 2 cells
per thread, 0/1 flags).
"""
import inspect

KEYS = ("block", "cells_per_thread", "xcd_remap", "alternate_sweep")


def _defaults(fn):
    sig = inspect.signature(fn)
    return {k: sig.parameters[k].default for k in KEYS}


def test_engine_and_autotuner_defaults_agree():
    from xadvcool_amd import autotune
    from xadvcool_amd.batch import BatchedSolver
    from xadvcool_amd.engine import LatticeBatch, LatticeSolver
    want = autotune.DEFAULT
    for fn in (LatticeBatch.__init__, LatticeSolver.__init__, BatchedSolver.__init__):
        assert _defaults(fn) == want, fn.__qualname__


def test_autotune_candidates_are_valid_shapes():
    from xadvcool_amd import autotune
    seen = set()
    for c in autotune.CANDIDATES:
        assert set(c) == set(KEYS)
        assert c["block"] in (64, 128, 256) and c["cells_per_thread"] in (1, 2)
        assert c["xcd_remap"] in (True, False) and c["alternate_sweep"] in (True, False)
        seen.add(tuple(c[k] for k in KEYS))
    assert len(seen) == len(autotune.CANDIDATES) == 24
    assert tuple(autotune.DEFAULT[k] for k in KEYS) in seen
