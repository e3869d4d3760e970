"""Host-side rules of the warm-started batched solver (no GPU needed)."""
from types import SimpleNamespace

from xadvcool_amd import warmstart as W


def _feed(stall, design, seq, every=100):
    return [stall(design, (i + 1) * every, SimpleNamespace(d_ke=v)) for i, v in enumerate(seq)]


def test_stall_fires_after_five_growing_checks_past_step_1000():
    st = W._Stall(after=1000, checks=5)
    seq = [1e-3 / (i + 1) for i in range(12)] + [1e-5 * 1.5 ** i for i in range(8)]
    out = _feed(st, {"key": 1, "warm": True}, seq)
    first = out.index(True)
    assert (first + 1) * 100 == 1800  # drop at 1300, growth 1400..1800 = 5 checks
    assert not any(out[:first])


def test_stall_needs_consecutive_growth_and_ignores_early_steps():
    st = W._Stall(after=1000, checks=5)
    # grows for 9 checks but all before step 1000: not a stall at that point
    early = _feed(st, {"key": 2, "warm": True}, [1e-6 * 2 ** i for i in range(9)])
    assert not any(early)
    # alternating growth and decay never reaches 5 in a row
    st = W._Stall(after=1000, checks=5)
    zig = _feed(st, {"key": 3, "warm": True}, [1e-5 * (2 if i % 2 else 1) for i in range(40)])
    assert not any(zig)


def test_stall_never_abandons_cold_runs():
    st = W._Stall(after=1000, checks=5)
    out = _feed(st, {"key": 4}, [1e-6 * 2 ** i for i in range(40)])
    assert not any(out)


def test_compare_designs_are_unseen():
    train = {d["key"] for d in W.make_designs(0, 8)}
    test = {d["key"] for d in W.make_designs(W.COMPARE_SEED0, 8)}
    assert not train & test
    assert W.COMPARE_SEED0 > 100_000  # far above any training set size used
