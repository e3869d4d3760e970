"""Pins of the physics-acceptance oracle (oracle/checks_ref.py, analyze_dataset.py:184-325) — no GPU.

Each check is evaluated on fields whose answer the mathematics fixes, not by
re-running the formula:

* uniform plug flow u = (U, 0, 0) in a straight duct: no mass imbalance, zero
  divergence, Mach = U sqrt(3); with rho falling linearly along x the mass
  imbalance and the pressure drop follow from the two planes one cell inside;
* u_x = s x: the central-difference divergence is s at every interior cell;
* T(x) = T0 + G x^2 / 2 in a solid block: the diffusive fluxes through the
  two x faces are -k G / 2 and -k G (2 nx - 3) / 2 per cell (k = (tau - 1/2) / 3,
  the lattice diffusivity), so the heat removed is -k n G (nx - 2); a heat
  map sized to inject exactly that much ((1 - omega / 2) q per heated cell,
  the source term's net gain) gives zero energy error;
* plug flow carrying T = 25 + a x: the heat removed is U N a (nx - 1);
* the pass / warn thresholds of analyze_dataset.py:111-120 at chosen values,
  and the reference's rho = 3 * pressure round trip.

Planted mistakes in the oracle (the source factor's 1/2 dropped, the
diffusive flux's sign flipped, undivided central differences, the mass
balance on the boundary planes) are each caught.
"""
import numpy as np
import pytest

from oracle import checks_ref as C

TAU, PR, KR, P = 0.52, 7.0, 628.0, 0.1
TAU_TH_F = 0.5 + (TAU - 0.5) / PR
TAU_TH_S = 0.5 + KR * (TAU_TH_F - 0.5)
K_F, K_S = (TAU_TH_F - 0.5) / 3.0, (TAU_TH_S - 0.5) / 3.0
ATTRS = {"tau_fluid": TAU, "heat_power": P, "Pr": PR, "k_ratio": KR}


def duct(nx=12, ny=8, nz=7):
    m = np.zeros((nx, ny, nz), np.int32)
    m[:, 0, :] = m[:, -1, :] = 1
    m[:, :, 0] = m[:, :, -1] = 1
    m[0][m[0] == 0] = 2
    m[-1][m[-1] == 0] = 3
    return m


def plug(m, U):
    u = np.zeros(m.shape + (3,))
    u[m != 1, 0] = U
    return u


def test_plug_flow_is_balanced_and_divergence_free():
    m = duct()
    c = C.validate_fields(m, plug(m, 0.04), np.ones(m.shape), np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS)
    assert c["mass_err"] == 0.0 and c["mean_div"] == 0.0 and c["max_div"] == 0.0
    assert c["mach"] == pytest.approx(0.04 * np.sqrt(3.0), rel=1e-15)
    assert c["rho_dev"] == 0.0 and c["delta_P"] == 0.0 and c["T_lower_bound"] == 0.0


def test_linear_density_mass_imbalance_and_pressure_drop():
    m = duct()
    nx = m.shape[0]
    g = 1e-3
    rho = np.broadcast_to((1.0 + g * (nx - 1 - np.arange(nx)))[:, None, None], m.shape).copy()
    c = C.validate_fields(m, plug(m, 0.03), rho, np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS)
    assert c["mass_err"] == pytest.approx(g * (nx - 3) / (1.0 + g * (nx - 2)), rel=1e-12)
    assert c["delta_P"] == pytest.approx(g * (nx - 3) / 3.0, rel=1e-12)
    assert c["rho_dev"] == pytest.approx(g * (nx - 1), rel=1e-12)  # the inlet plane x = 0 is non-solid


def test_divergence_of_a_linear_field():
    m = duct()
    s = 2e-4
    u = np.zeros(m.shape + (3,))
    u[..., 0] = s * np.arange(m.shape[0])[:, None, None]
    c = C.validate_fields(m, u, np.ones(m.shape), np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS)
    assert c["mean_div"] == pytest.approx(s, rel=1e-12) and c["max_div"] == pytest.approx(s, rel=1e-12)


def conduction_case(mutation=0):
    """A solid block with T = T0 + G x^2 / 2 and a base heat map injecting exactly the heat conducted out."""
    nx, ny, nz = 10, 6, 5
    m = np.ones((nx, ny, nz), np.int32)
    G = -0.01
    x = np.arange(nx, dtype=np.float64)
    T = np.broadcast_to((40.0 + G * x * x / 2.0)[:, None, None], m.shape).copy()
    n = ny * nz
    removed = -K_S * n * G * (nx - 2)
    hs = np.zeros(m.shape)
    gain = 1.0 - 0.5 / TAU_TH_S  # (1 - omega_s / 2): net heat per unit source in a solid cell
    hs[:, :, 0] = removed / (gain * P * nx * ny)
    return C.validate_fields(m, np.zeros(m.shape + (3,)), np.ones(m.shape), T, hs, ATTRS, mutation=mutation), removed


def test_conduction_energy_balance():
    c, removed = conduction_case()
    assert c["Q_conv"] == pytest.approx(removed, rel=1e-12)
    assert c["Q_source_rate"] == pytest.approx(removed, rel=1e-12)
    assert c["energy_err"] < 1e-12 and c["energy_pass"]


def test_advected_heat():
    m = duct()
    nx = m.shape[0]
    U, a = 0.02, 0.05
    T = np.broadcast_to((25.0 + a * np.arange(nx))[:, None, None], m.shape).copy()
    c = C.validate_fields(m, plug(m, U), np.ones(m.shape), T, np.zeros(m.shape), ATTRS)
    N = int(np.sum(m[-1] == 3))
    assert c["Q_conv"] == pytest.approx(U * N * a * (nx - 1), rel=1e-12)
    assert c["Q_source_rate"] == 0.0 and c["energy_err"] == 0.0  # no source: the reference reports 0


def test_thresholds_and_flags():
    m = duct()
    U = 0.2 / np.sqrt(3.0)  # Mach 0.2: fails 0.15, passes the 0.30 warning level
    T = np.full(m.shape, 25.0)
    T[5, 3, 3] = 24.7  # fluid minimum 0.3 below the inlet: fails >= 0, passes >= -0.5
    rho = np.ones(m.shape)
    rho[5, 4, 3] = 1.07  # 0.07 < max(0.05, 2 Ma^2 = 0.08): passes
    c = C.validate_fields(m, plug(m, U), rho, T, np.zeros(m.shape), {**ATTRS, "Nu": 0.5, "f_friction": 2.0,
                                                                        "Re": 50.0})
    assert (c["mach_pass"], c["mach_warn"]) == (False, True)
    assert c["T_lower_bound"] == pytest.approx(-0.3) and (c["T_lb_pass"], c["T_lb_warn"]) == (False, True)
    assert c["rho_dev"] == pytest.approx(0.07) and c["rho_dev_pass"]
    assert not c["Nu_pass"] and c["f_Re"] == 100.0 and c["f_Re_pass"]
    c2 = C.validate_fields(m, plug(m, U), rho, T, np.zeros(m.shape), {**ATTRS, "f_friction": 5.0, "Re": 50.0})
    assert not c2["f_Re_pass"]  # 250 > 200


def test_pressure_round_trip():
    m = duct()
    rho = np.float32(1.0) + np.random.default_rng(0).uniform(0, 1e-3, m.shape).astype(np.float32)
    pr = (rho / np.float32(3.0)).astype(np.float32)
    a = C.validate_fields(m, plug(m, 0.03), rho, np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS, pressure=pr)
    b = C.validate_fields(m, plug(m, 0.03), pr * 3.0, np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS)
    assert a == b


def _caught(mutation):
    if mutation in (1, 2):
        c, removed = conduction_case(mutation)
        return not (c["energy_err"] < 1e-9)
    m = duct()
    if mutation == 3:
        u = np.zeros(m.shape + (3,))
        u[..., 0] = 2e-4 * np.arange(m.shape[0])[:, None, None]
        c = C.validate_fields(m, u, np.ones(m.shape), np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS,
                              mutation=3)
        return c["mean_div"] != pytest.approx(2e-4, rel=1e-12)
    nx, g = m.shape[0], 1e-3
    rho = np.broadcast_to((1.0 + g * (nx - 1 - np.arange(nx)))[:, None, None], m.shape).copy()
    c = C.validate_fields(m, plug(m, 0.03), rho, np.full(m.shape, 25.0), np.zeros(m.shape), ATTRS, mutation=4)
    return c["mass_err"] != pytest.approx(g * (nx - 3) / (1.0 + g * (nx - 2)), rel=1e-12)


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(mutation):
    assert _caught(mutation)
