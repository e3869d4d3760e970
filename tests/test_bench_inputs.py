"""Inputs bench.py reads from the committed profiles (no GPU needed).

The benchmark line's roofline.traffic is the launch's algorithmic bytes times the measured
traffic per algorithmic byte in profiles/k_pass_counters.json; a missing key there would turn the
field into null without an error.  The ratio must follow from the file's own counter values and
calibration scales.
"""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pass_counters_file_has_what_bench_reads():
    import bench
    c = bench.pass_counters()
    assert c is not None, bench.COUNTERS_FILE
    for k in ("workload", "fetch_scale", "write_scale", "traffic_per_alg_byte", "fetch_kib_per_dispatch",
              "write_kib_per_dispatch", "alg_bytes_per_dispatch"):
        assert k in c, k
    traffic = 1024 * (c["fetch_kib_per_dispatch"] * c["fetch_scale"] + c["write_kib_per_dispatch"] * c["write_scale"])
    assert abs(traffic / c["alg_bytes_per_dispatch"] - c["traffic_per_alg_byte"]) < 1e-3
    assert 1.0 <= c["traffic_per_alg_byte"] < 1.5


def test_bench_help_and_gpu_count_guard():
    """bench.py parses its options without a GPU and refuses more than one GPU with a message."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--launch" in r.stdout and "config2" in r.stdout
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
