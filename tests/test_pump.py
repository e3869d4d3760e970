"""Pins of the constant-power pump controller (src/simulation.py:139-146) — no GPU.

The controller, once per stats check after step 2000:
  dP = (sum_rho_in / c_in - sum_rho_out / c_out) / 3,  P = dP * u * c_in,
  u <- clip(u + 0.005 (P_target - P), 1e-4, 0.15),
every operand float32 (the reference's stats and u_in are float32 arrays).
It is restated in the oracle (oracle.pump_update, inside oracle.run_simulation)
and in the product's host loop (xadvcool_amd.simulation.pump_update, inside
solve_on_batch).  What fixes them here, independently of the formula's text:

* closed forms: with constant stats the update is the linear recurrence
  u_n = u* + (u_0 - u*) (1 - 0.005 dP c_in)^n with u* = P_target / (dP c_in)
  (checked over 200 updates), an exactly representable fixed point, the clip
  bounds reached exactly, and c_in = c_out = 0 giving u += 0.005 P_target;
* float32 rounding: one golden update where float32 and float64 arithmetic
  give different float32 results (the value below was derived with exact
  rational arithmetic rounded to float32 after each operation), and the same
  exact-rational evaluation on random cases;
* the step > 2000 gate: the oracle loop on a small channel leaves u_in
  untouched through step 2000 and applies exactly one update at step 2100;
  the product's loop is driven by a scripted stats stream on a stand-in
  batch (host logic only);
* planted mistakes in the oracle (float64 arithmetic, no clip, flipped sign,
  dropped /3, a >= 2000 gate) are each caught by one of the checks above.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc
from xadvcool_amd import simulation as sim
from xadvcool_amd.engine import STAT_NAMES

f32 = np.float32


def _r32(q):
    """Exact rational -> nearest float32, ties to even (independent of NumPy's float32 arithmetic)."""
    q = Fraction(q)
    if q == 0:
        return f32(0.0)
    sign, a, e = (-1 if q < 0 else 1), abs(q), 0
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    m = a * (1 << 23)
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return f32(sign * float(Fraction(fl) * Fraction(2) ** (e - 23)))


def _exact_update(u, s5, s6, c_in, c_out, target):
    """The update evaluated with exact rationals, rounded to float32 after every operation."""
    q = lambda x: Fraction(float(f32(x)))  # noqa: E731
    u, s5, s6, c_in, c_out = q(u), q(s5), q(s6), q(c_in), q(c_out)
    ri = Fraction(float(_r32(s5 / c_in))) if c_in > 0 else Fraction(1)
    ro = Fraction(float(_r32(s6 / c_out))) if c_out > 0 else Fraction(1)
    dP = Fraction(float(_r32(Fraction(float(_r32(ri - ro))) / 3)))
    power = Fraction(float(_r32(dP * Fraction(float(_r32(u * c_in))))))
    corr = Fraction(float(_r32(q(0.005) * Fraction(float(_r32(q(target) - power))))))
    un = _r32(u + corr)
    return f32(min(max(un, f32(1e-4)), f32(0.15)))


PUMPS = {"oracle": orc.pump_update, "product": sim.pump_update}


# --------------------------------------------------------------- checks ----
# Each returns None when the controller behaves, else a message; the planted-mistake test reuses them.

def check_fixed_point(pump):
    # rho_in = 70/64 = 1.09375, rho_out = 1, dP = 0.09375 / 3 = 1/32, P = dP * 0.125 * 64 = 0.25 = target
    u = pump(0.125, 70.0, 64.0, 64, 64, 0.25)
    return None if u == f32(0.125) and isinstance(u, np.float32) else f"fixed point moved to {u!r}"


def check_geometric_approach(pump):
    c, dP, target, u0 = 100.0, 0.1, 0.5, 0.01  # u* = 0.05, ratio 1 - 0.005 * 0.1 * 100 = 0.95
    u = f32(u0)
    ustar, r = target / (dP * c), 1.0 - 0.005 * dP * c
    for n in range(1, 201):
        u = pump(u, c * (1.0 + 3.0 * dP), c, c, c, target)
        want = ustar + (u0 - ustar) * r ** n
        if abs(float(u) - want) > 2e-7:
            return f"update {n}: {float(u)!r} vs closed form {want!r}"
    return None


def check_clip(pump):
    hi = pump(0.14, 110.0, 100.0, 100, 100, 1e6)
    lo = pump(0.001, 110.0, 100.0, 100, 100, -1e6)
    if hi != f32(0.15) or lo != f32(0.0001):
        return f"clip bounds {hi!r}, {lo!r}"
    return None


def check_zero_counts(pump):
    u = pump(0.02, 0.0, 0.0, 0, 0, 1.0)  # no cells on the planes: both averages are 1, dP = 0
    want = _r32(Fraction(float(f32(0.02))) + Fraction(float(_r32(Fraction(float(f32(0.005)))))))
    return None if u == want else f"{u!r} vs {want!r}"


# A case where float32 arithmetic and float64 arithmetic round to different float32 results.
GOLDEN_IN = (0.03224639222025871, 152.1195526123047, 150.95083618164062, 151.0, 151.0, 1.0)
GOLDEN_OUT = float.fromhex("0x1.309ba0p-5")


def check_float32_golden(pump):
    u = pump(*GOLDEN_IN)
    return None if float(u) == GOLDEN_OUT else f"{float(u).hex()} vs {GOLDEN_OUT.hex()}"


def check_gate_oracle_loop(pump_mutation=0):
    m = np.zeros((10, 6, 6), np.int32)
    m[:, 0, :] = m[:, -1, :] = 1
    m[:, :, 0] = m[:, :, -1] = 1
    m[0, 1:-1, 1:-1] = 2
    m[-1, 1:-1, 1:-1] = 3
    kw = dict(tau_fluid=0.6, u_inlet_val=0.03, pump_mode="constant_power", target_power=0.5, tol_ke=0.0,
              tol_T=0.0, heat_power=0.05, pump_mutation=pump_mutation)
    r20 = orc.run_simulation(m, max_steps=2000, **kw)
    if r20["u_inlet"] != float(f32(0.03)):
        return f"u_in changed by step 2000: {r20['u_inlet']!r}"
    r21 = orc.run_simulation(m, max_steps=2100, **kw)
    s = orc.stats(r21["rho"], r21["u"], r21["T"], m)
    want = _exact_update(0.03, s[5], s[6], s[8], s[9], 0.5)
    if r21["u_inlet"] != float(want) or want == f32(0.03):
        return f"step 2100: {r21['u_inlet']!r} vs one update {float(want)!r}"
    return None


# ---------------------------------------------------------------- tests ----

@pytest.mark.parametrize("which", sorted(PUMPS))
@pytest.mark.parametrize("check", [check_fixed_point, check_geometric_approach, check_clip, check_zero_counts,
                                   check_float32_golden], ids=lambda c: c.__name__)
def test_pump_update(which, check):
    msg = check(PUMPS[which])
    assert msg is None, msg


@pytest.mark.parametrize("which", sorted(PUMPS))
def test_pump_matches_exact_rational_arithmetic(which):
    rng = np.random.default_rng(11)
    for _ in range(300):
        c = float(rng.integers(1, 4000))
        args = (f32(rng.uniform(1e-4, 0.15)), f32(c * rng.uniform(0.98, 1.02)), f32(c * rng.uniform(0.98, 1.02)),
                c, float(rng.integers(1, 4000)), float(rng.uniform(-2.0, 2.0)))
        assert PUMPS[which](*args) == _exact_update(*args), args


def test_gate_in_oracle_loop():
    msg = check_gate_oracle_loop()
    assert msg is None, msg


class _ScriptedBatch:
    """Stand-in for a one-slot LatticeBatch: stats from a script, inlet-velocity writes recorded."""

    def __init__(self, stats_fn):
        self.stats_fn, self.steps, self.u_writes = stats_fn, 0, []

    def step(self, n, want_fields=True):
        self.steps += n

    def readout(self, with_stats=True):
        pass

    def stats(self):
        return np.array([[self.stats_fn(self.steps)[k] for k in STAT_NAMES]], np.float64)

    def set_inlet_velocity(self, slot, u):
        self.u_writes.append((self.steps, u))


def test_product_loop_gate_and_updates():
    c, dP, target, u0 = 100.0, 0.1, 0.5, 0.01

    def stats(step):
        s = dict.fromkeys(STAT_NAMES, 1.0)
        s.update(sum_ke=1.0 + step, max_T=26.0 + step, Q_out=1.0 + step, max_v=0.05,  # never converges
                 sum_rho_in=c * (1.0 + 3.0 * dP), sum_rho_out=c, c_in=c, c_out=c)
        return s

    b = _ScriptedBatch(stats)
    out = sim.solve_on_batch(b, 0, nx=10, max_steps=4000, tau_fluid=0.6, u_inlet_val=u0, tol_ke=0.0, tol_T=0.0,
                             pump_mode="constant_power", target_power=target, min_steps_rule=False)
    assert out["steps"] == 4000
    steps = [s for s, _ in b.u_writes]
    assert steps == list(range(2100, 4001, 100))  # the first update at the first check after step 2000
    ustar, r = target / (dP * c), 1.0 - 0.005 * dP * c
    for n, (_, u) in enumerate(b.u_writes, 1):
        assert abs(u - (ustar + (u0 - ustar) * r ** n)) < 2e-7
    assert out["u_inlet"] == b.u_writes[-1][1]


@pytest.mark.parametrize("mutation,check", [
    (1, check_float32_golden), (2, check_clip), (3, check_geometric_approach), (4, check_geometric_approach)])
def test_planted_pump_mistakes_are_caught(mutation, check):
    def mutant(*a):
        return orc.pump_update(*a, mutation=mutation)
    assert check(mutant) is not None, f"mutation {mutation} not caught by {check.__name__}"


def test_planted_gate_mistake_is_caught():
    assert check_gate_oracle_loop(pump_mutation=5) is not None
