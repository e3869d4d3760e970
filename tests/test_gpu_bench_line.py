"""The benchmark's JSON line on a tiny workload: three designs, a loose tolerance and a 2,000-step cap, so
both arms finish in seconds.  Checks the fields the line promises (metric, value, unit, config, roofline
with its traffic, cpu_baseline with its host) and that they are consistent with each other."""
import json
import math
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_bench_line_contract(gpu, capsys, monkeypatch):
    import bench
    monkeypatch.setattr(sys, "argv", ["bench.py", "--designs", "3", "--tol", "1e-2", "--max-steps", "2000",
                                      "--warmup", "0", "--cpu-seconds", "0.6"])
    print("[bench-line] 3 designs, tol 1e-2, 2,000 steps", file=sys.stderr, flush=True)
    bench.main()
    out = capsys.readouterr().out.strip().splitlines()
    line = json.loads(out[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline", "cold_start", "warm_start",
              "acceptance", "by_family"):
        assert k in line, k
    assert line["n_gpus"] == 1 and line["higher_is_better"] is True and line["dtype"] == "f32"
    assert line["config"]["designs"] == 3 and "workload" in line["config"]
    assert line["config"]["launch"]["thermal_prefetch"] is True
    assert line["value"] == line["warm_start"]["solves_per_hr"] >= 0.0
    rf = line["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0.0 < rf["frac"] < 1.2 and math.isclose(rf["frac"], rf["achieved"] / rf["peak"], rel_tol=2e-3)
    assert rf["passes_timed"] > 0 and rf["alg_bytes_per_launch"] > 0
    assert rf["traffic"] is not None and 1.0 <= rf["traffic"] / rf["alg_bytes_per_launch"] < 1.5
    cb = line["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] == 1 and cb["mlups"] > 0.0
    assert cb["host"]["logical_cpus"] >= 1 and cb["all_threads"]["threads"] >= 1
