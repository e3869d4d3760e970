"""The plain small cold plate (74 x 64 x 16, six straight channels, tau 0.6, u_in 0.02, k_ratio 10, 2,000
steps): the mask builder and the oracle's results against the committed regression fixture."""
import json
import os

import numpy as np

from tests.golden.make_config1_fixture import run_case, summary

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    with open(os.path.join(HERE, "golden", "config1.json")) as f:
        return json.load(f)


def test_config1_mask_layout():
    from xadvcool_amd import geometry as G
    m = G.config1_mask()
    assert m.shape == (74, 64, 16)
    # six 4-wide channels with 2-cell fins, open over z in [3, 15) through the core
    core = m[10:64, :, 8]
    open_y = np.flatnonzero(np.all(core == 0, axis=0))
    runs = np.split(open_y, np.flatnonzero(np.diff(open_y) > 1) + 1)
    assert [len(r) for r in runs] == [4] * 6
    assert all(b[0] - a[-1] == 3 for a, b in zip(runs, runs[1:]))  # 2 solid cells between channels
    assert np.all(m[:, :, :3] == 1) and np.all(m[:, :, 15] == 1)
    assert np.all(m[0, 1:-1, 3:15] == 2) and np.all(m[73, 1:-1, 3:15] == 3)


def test_config1_oracle_matches_fixture():
    m, r = run_case()
    got, want = summary(m, r), fixture()
    assert got == want, (got, want)
