"""Golden geometry facts from the reference's dataset files, as a small JSON fixture.

* accepted designs: porosity and D_h columns of dataset/metadata.parquet for a
  seeded sample of distinct geometries (one row each);
* rejected designs: the porosity printed in dataset/failed_samples.json for
  "invalid geometry" entries (generate_dataset.py:387-392);
* diverged designs: a few "diverged" entries (generate_dataset.py:444-449).
Design parameters come from the Latin-hypercube sampler restated in
xadvcool_amd.geometry (seed 42, 200 points per family): the md5 sample id of
each regenerated point is matched against the reference files.
"""
import json
import os
import re
import sys

import numpy as np
import pandas as pd

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from xadvcool_amd import geometry as G  # noqa: E402

REF = "/root/reference/dataset"


def main(n_ok=120, n_bad=40, n_div=4, seed=0):
    df = pd.read_parquet(os.path.join(REF, "metadata.parquet")).set_index("sample_id")
    failed = json.load(open(os.path.join(REF, "failed_samples.json")))
    points = {}
    for name in G.DESIGN_CONFIGS:
        for p in G.generate_samples(name, 200, seed=42):
            points[G.sample_id(p)] = p
    ok, bad, div, seen = [], [], [], set()
    for sid, p in points.items():
        key = (p["design_idx"], p["feature_size"], p["spacing"])
        if sid in df.index and key not in seen:
            seen.add(key)
            r = df.loc[sid]
            ok.append(dict(sample_id=sid, design_idx=p["design_idx"], feature_size=p["feature_size"],
                           spacing=p["spacing"], porosity=float(r.porosity), Dh=float(r.Dh)))
        elif sid in failed and failed[sid].startswith("invalid geometry"):
            por = float(re.search(r"porosity=([0-9.]+)%", failed[sid]).group(1)) / 100.0
            bad.append(dict(sample_id=sid, design_idx=p["design_idx"], feature_size=p["feature_size"],
                            spacing=p["spacing"], porosity_pct_2dp=por))
        elif sid in failed and failed[sid].startswith("diverged"):
            div.append(dict(p, sample_id=sid, reason=failed[sid]))
    rng = np.random.default_rng(seed)
    pick = lambda xs, n: [xs[i] for i in sorted(rng.choice(len(xs), size=min(n, len(xs)), replace=False))]
    out = {"accepted": pick(ok, n_ok), "rejected": pick(bad, n_bad), "diverged": pick(div, n_div),
           "counts": {"accepted_geometries": len(ok), "rejected": len(bad), "diverged": len(div)}}
    path = os.path.join(HERE, "geometry_rows.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path, {k: len(v) for k, v in out.items() if isinstance(v, list)}, out["counts"])


if __name__ == "__main__":
    main()
