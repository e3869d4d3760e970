"""Launch-shape autotuner on the GPU (xadvcool_amd/autotune.py).

Every candidate shape runs on a real batch (all slots busy, then one idle),
the fastest one is returned and cached under the batch's shape set.  That
every shape computes the same bits is the parity tests' job
(test_gpu_parity.py::test_coupled_launch_shapes, test_gpu_batch.py).
"""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_autotune_returns_and_caches_the_fastest_shape(gpu, tmp_path, monkeypatch):
    from xadvcool_amd import autotune as A
    from xadvcool_amd import geometry as G

    monkeypatch.setattr(A, "_CACHE", str(tmp_path / "tune_cache.json"))
    ds = [G.family_design(300 + i, core=(24, 16, 12), family=fam, heat_name="quad_core")
          for i, fam in enumerate(("straight", "gyroid"))]
    cap = tuple(max(d["mask"].shape[k] for d in ds) for k in range(3))
    cands = [c for c in A.CANDIDATES if c["block"] == 256]
    print(f"[autotune] {len(cands)} candidates on 2 designs, cap {cap}", flush=True)
    best, times = A.tune(ds, cap, passes=100, candidates=cands, save=True)
    assert len(times) == len(cands)
    assert all(t > 0 for t in times.values())
    assert times[json.dumps(best, sort_keys=True)] == min(times.values())
    hit = A.cached([tuple(d["mask"].shape) for d in ds])
    assert hit is not None and {k: hit[k] for k in best} == best
