"""Cold-plate masks and heat maps (host-side inputs of the solver), without pyvista.

``generate_mask(state)`` and the ten heat-map generators restate
src/geometry.py:69-526 (same parameter mapping, carving, walls, base and
inlet/outlet planes); ``compute_porosity``, ``is_valid_geometry`` and
``compute_hydraulic_diameter`` restate generate_dataset.py:57-74,280-293;
``DESIGN_CONFIGS`` / ``generate_samples`` restate the Latin-hypercube design
sampler of generate_dataset.py:193-274.  Every function takes the core size
as an argument (default 128 x 128 x 32, the reference's CORE_NX/NY/NZ), so
lattices of other sizes can be built the same way; ``tpms_mask`` and
``family_mask`` add the direct (period, threshold) and seeded-family
parameterisations used for the benchmark workloads.
"""
from __future__ import annotations

import numpy as np

from .constants import BC_FLUID, BC_INLET, BC_OUTLET, BC_SOLID

CORE_NX, CORE_NY, CORE_NZ = 128, 128, 32
BUFFER = 10
BASE_THICKNESS = 4
TOP_THICKNESS = 1
DESIGN_NAMES = ["Empty Channel", "Straight Channels", "Pin-Fin Staggered", "Gyroid TPMS", "Schwarz P TPMS",
                "Schwarz D TPMS"]


def _tpms_field(kind, nx, ny, nz, omega):
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    wx, wy, wz = X * omega, Y * omega, Z * omega
    if kind == 3:  # gyroid
        return np.sin(wx) * np.cos(wy) + np.sin(wy) * np.cos(wz) + np.sin(wz) * np.cos(wx)
    if kind == 4:  # Schwarz P
        return np.cos(wx) + np.cos(wy) + np.cos(wz)
    # Schwarz D
    return (np.sin(wx) * np.sin(wy) * np.sin(wz) + np.sin(wx) * np.cos(wy) * np.cos(wz) +
            np.cos(wx) * np.sin(wy) * np.cos(wz) + np.cos(wx) * np.cos(wy) * np.sin(wz))


def _frame(is_solid, buf, base=BASE_THICKNESS, top=TOP_THICKNESS):
    """Buffers, walls, base, lid and the inlet/outlet planes (src/geometry.py:506-524)."""
    nx, ny, nz = is_solid.shape
    mask = np.full((nx + 2 * buf, ny, nz), BC_FLUID, dtype=np.int32)
    mask[buf:buf + nx, :, :][is_solid] = BC_SOLID
    mask[:, 0, :] = BC_SOLID
    mask[:, -1, :] = BC_SOLID
    mask[:, :, :base] = BC_SOLID
    mask[:, :, -top:] = BC_SOLID
    mask[0, 1:-1, base:-top] = BC_INLET
    mask[-1, 1:-1, base:-top] = BC_OUTLET
    return mask


def core_solid(design_idx, feature_size=3.0, spacing=3.0, core=(CORE_NX, CORE_NY, CORE_NZ)):
    """Solid cells of the core for one design (src/geometry.py:365-503)."""
    nx, ny, nz = core
    design_idx = int(min(max(int(design_idx), 0), len(DESIGN_NAMES) - 1))
    is_solid = np.zeros((nx, ny, nz), dtype=bool)
    raw_width = feature_size
    raw_count = spacing
    channel_width = max(2, int(raw_width * 2.0))
    target_channels = max(4, int(raw_count * 4.0))
    max_n_possible = int((ny - 1) / (channel_width + 1))
    n_channels = max(2, min(target_channels, max_n_possible))
    if design_idx == 1:  # straight channels
        total_channel_width = n_channels * channel_width
        if total_channel_width < ny:
            gap = float(ny - total_channel_width) / (n_channels + 1)
            y_end = int(gap)
            if y_end > 0:
                is_solid[:, :y_end, :] = True
            current_y = gap
            for i in range(n_channels):
                current_y += channel_width
                y_start = int(current_y)
                y_end = int(current_y + gap)
                if i == n_channels - 1:
                    y_end = ny
                if y_start < ny:
                    is_solid[:, y_start:y_end, :] = True
                current_y += gap
    elif design_idx == 2:  # staggered pin fins
        d_pin = max(2, int(raw_width * 2.0))
        pitch = max(d_pin + 2, int(raw_count * 3.0))
        radius = d_pin / 2.0
        X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        dx_a = (X % pitch) - pitch / 2
        dy_a = (Y % pitch) - pitch / 2
        dx_b = ((X - pitch / 2) % pitch) - pitch / 2
        dy_b = ((Y - pitch / 2) % pitch) - pitch / 2
        is_solid[:, :, :] = (np.minimum(dx_a ** 2 + dy_a ** 2, dx_b ** 2 + dy_b ** 2) < radius ** 2)[:, :, None]
    elif design_idx in (3, 4, 5):  # TPMS
        C = 0.8 - ((raw_width - 1.0) / 9.0) * 1.6
        cell_size = 20.0 + ((raw_count - 1.0) / 9.0) * 80.0
        is_solid[:, :, :] = _tpms_field(design_idx, nx, ny, nz, 2.0 * np.pi / cell_size) < C
        is_solid[:5, :, :] = False
        is_solid[-5:, :, :] = False
    return is_solid


def generate_mask(state, core=(CORE_NX, CORE_NY, CORE_NZ)):
    """Full mask for a design state dict (src/geometry.py:339-526)."""
    np.random.seed(42)  # the reference seeds the global RNG here (src/geometry.py:342)
    solid = core_solid(state["design_type"], state.get("feature_size", 3.0), state.get("spacing", 3.0), core)
    return _frame(solid, int(state["buffer_size"]))


def tpms_mask(kind, period, threshold, core=(CORE_NX, CORE_NY, CORE_NZ), buf=BUFFER):
    """TPMS core with an explicit cell period and threshold (solid where f < threshold), 5-cell open manifolds."""
    solid = _tpms_field(kind, *core, 2.0 * np.pi / float(period)) < threshold
    solid[:5] = False
    solid[-5:] = False
    return _frame(solid, buf)


# ------------------------------------------------------------- heat maps ---

def _blur(hm, sigma):
    import scipy.ndimage as ndimage
    return ndimage.gaussian_filter(hm, sigma=sigma)


def _norm(hm):
    if hm.max() > 0:
        hm /= hm.max()
    return hm


def generate_dual_core_heatmap(nx, ny):
    hm = np.zeros((nx, ny), dtype=np.float32)
    hm[:, :] = 0.1
    hm[int(nx * 0.25):int(nx * 0.75), int(ny * 0.55):int(ny * 0.85)] = 1.0
    hm[int(nx * 0.25):int(nx * 0.75), int(ny * 0.15):int(ny * 0.45)] = 1.0
    hm[int(nx * 0.30):int(nx * 0.70), int(ny * 0.45):int(ny * 0.55)] = 0.5
    return _norm(_blur(hm, nx / 80.0))


def generate_quad_core_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.08, dtype=np.float32)
    for x0f, y0f, x1f, y1f in [(0.15, 0.55, 0.45, 0.85), (0.15, 0.15, 0.45, 0.45), (0.55, 0.55, 0.85, 0.85),
                               (0.55, 0.15, 0.85, 0.45)]:
        hm[int(nx * x0f):int(nx * x1f), int(ny * y0f):int(ny * y1f)] = 1.0
    hm[int(nx * 0.45):int(nx * 0.55), int(ny * 0.10):int(ny * 0.90)] = 0.45
    hm[int(nx * 0.10):int(nx * 0.90), int(ny * 0.45):int(ny * 0.55)] = 0.45
    return _norm(_blur(hm, nx / 80.0))


def generate_gpu_die_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.05, dtype=np.float32)
    hm[int(nx * 0.15):int(nx * 0.85), int(ny * 0.15):int(ny * 0.85)] = 0.75
    hm[int(nx * 0.10):int(nx * 0.90), int(ny * 0.02):int(ny * 0.12)] = 1.0
    hm[int(nx * 0.10):int(nx * 0.90), int(ny * 0.88):int(ny * 0.98)] = 1.0
    for cx_f in [0.30, 0.50, 0.70]:
        cx, cy = int(nx * cx_f), int(ny * 0.50)
        r = int(min(nx, ny) * 0.04)
        xx, yy = np.ogrid[:nx, :ny]
        hm[(xx - cx) ** 2 + (yy - cy) ** 2 < r ** 2] = 1.0
    return _norm(_blur(hm, nx / 60.0))


def generate_chiplet_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.03, dtype=np.float32)
    cw, ch = 0.18, 0.22
    for x0f, y0f in [(0.12, 0.08), (0.12, 0.40), (0.12, 0.72), (0.58, 0.08), (0.58, 0.40), (0.58, 0.72)]:
        hm[int(nx * x0f):int(nx * (x0f + cw)), int(ny * y0f):int(ny * (y0f + ch))] = 1.0
    hm[int(nx * 0.38):int(nx * 0.52), int(ny * 0.10):int(ny * 0.90)] = 0.35
    return _norm(_blur(hm, nx / 70.0))


def generate_igbt_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.02, dtype=np.float32)
    die_h, igbt_w, diode_w = 0.10, 0.50, 0.25
    for i, y0f in enumerate([0.05, 0.20, 0.38, 0.55, 0.72, 0.88]):
        if i % 2 == 0:
            hm[int(nx * 0.20):int(nx * (0.20 + igbt_w)), int(ny * y0f):int(ny * (y0f + die_h))] = 1.0
        else:
            hm[int(nx * 0.30):int(nx * (0.30 + diode_w)), int(ny * y0f):int(ny * (y0f + die_h))] = 0.55
    return _norm(_blur(hm, nx / 60.0))


def generate_soc_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.06, dtype=np.float32)
    hm[int(nx * 0.10):int(nx * 0.35), int(ny * 0.60):int(ny * 0.90)] = 1.0
    for dy in [0.08, 0.22]:
        for dx in [0.10, 0.28]:
            hm[int(nx * dx):int(nx * (dx + 0.12)), int(ny * (0.35 + dy)):int(ny * (0.35 + dy + 0.10))] = 0.30
    hm[int(nx * 0.50):int(nx * 0.90), int(ny * 0.35):int(ny * 0.90)] = 0.65
    hm[int(nx * 0.05):int(nx * 0.95), int(ny * 0.02):int(ny * 0.12)] = 0.40
    hm[int(nx * 0.45):int(nx * 0.60), int(ny * 0.15):int(ny * 0.30)] = 0.80
    return _norm(_blur(hm, nx / 70.0))


def generate_gaussian_hotspots_heatmap(nx, ny, seed=0):
    rng = np.random.RandomState(seed)
    hm = np.full((nx, ny), 0.05, dtype=np.float32)
    n_spots = rng.randint(1, 4)
    xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for _ in range(n_spots):
        cx = rng.uniform(0.2, 0.8) * nx
        cy = rng.uniform(0.2, 0.8) * ny
        sigma = rng.uniform(0.06, 0.15) * min(nx, ny)
        amplitude = rng.uniform(0.5, 1.0)
        hm += amplitude * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * sigma ** 2))
    return _norm(hm)


def generate_hotspot_on_background_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.20, dtype=np.float32)
    xx, yy = np.ogrid[:nx, :ny]
    cx1, cy1, r1 = int(nx * 0.40), int(ny * 0.60), int(min(nx, ny) * 0.07)
    hm[(xx - cx1) ** 2 + (yy - cy1) ** 2 < r1 ** 2] = 1.0
    cx2, cy2, r2 = int(nx * 0.65), int(ny * 0.35), int(min(nx, ny) * 0.05)
    hm[(xx - cx2) ** 2 + (yy - cy2) ** 2 < r2 ** 2] = 0.85
    return _norm(_blur(hm, nx / 100.0))


def generate_peripheral_ring_heatmap(nx, ny):
    hm = np.full((nx, ny), 0.05, dtype=np.float32)
    rw = 0.10
    hm[:int(nx * rw), :] = 0.85
    hm[int(nx * (1 - rw)):, :] = 0.85
    hm[:, :int(ny * rw)] = 0.85
    hm[:, int(ny * (1 - rw)):] = 0.85
    corner_r = int(min(nx, ny) * 0.12)
    xx, yy = np.ogrid[:nx, :ny]
    for cx_f, cy_f in [(0.0, 0.0), (0.0, 1.0), (1.0, 0.0), (1.0, 1.0)]:
        cx, cy = int(nx * cx_f), int(ny * cy_f)
        hm[(xx - cx) ** 2 + (yy - cy) ** 2 < corner_r ** 2] = 1.0
    hm[int(nx * 0.20):int(nx * 0.80), int(ny * 0.20):int(ny * 0.80)] = 0.15
    return _norm(_blur(hm, nx / 60.0))


HEAT_SOURCE_REGISTRY = {
    "uniform": lambda nx, ny, **kw: np.ones((nx, ny), dtype=np.float32),
    "dual_core": lambda nx, ny, **kw: generate_dual_core_heatmap(nx, ny),
    "quad_core": lambda nx, ny, **kw: generate_quad_core_heatmap(nx, ny),
    "gpu_die": lambda nx, ny, **kw: generate_gpu_die_heatmap(nx, ny),
    "chiplet": lambda nx, ny, **kw: generate_chiplet_heatmap(nx, ny),
    "igbt_half_bridge": lambda nx, ny, **kw: generate_igbt_heatmap(nx, ny),
    "soc_heterogeneous": lambda nx, ny, **kw: generate_soc_heatmap(nx, ny),
    "gaussian_hotspots": lambda nx, ny, **kw: generate_gaussian_hotspots_heatmap(nx, ny, seed=kw.get("seed", 0)),
    "hotspot_on_background": lambda nx, ny, **kw: generate_hotspot_on_background_heatmap(nx, ny),
    "peripheral_ring": lambda nx, ny, **kw: generate_peripheral_ring_heatmap(nx, ny),
}
HEAT_SOURCE_NAMES = list(HEAT_SOURCE_REGISTRY.keys())


def heat_source_3d(mask_shape, name, core=(CORE_NX, CORE_NY), buf=BUFFER, seed=0):
    """2-D map placed on the base of the core, x in [buf, buf+core_nx), z = 0 (generate_dataset.py:402-410)."""
    nx, ny, nz = mask_shape
    hs2 = HEAT_SOURCE_REGISTRY[name](core[0], core[1], seed=seed)
    hs = np.zeros((nx, ny, nz), dtype=np.float32)
    hs[buf:buf + core[0], :core[1], 0] = hs2
    return hs



# ------------------------------------------- heat maps as paint programs ---
#
# Each generator above, restated as the data the device heat-map builder
# (xadv_heat_map) takes: a base value, primitives painted in order (RECT /
# DISK assign a value, GAUSS adds a Gaussian in FP64), an optional Gaussian
# blur (scipy.ndimage.gaussian_filter's weights, truncate 4.0, reflecting
# edges) and the division by the maximum.  Integer bounds are computed here
# exactly as the generators compute their slices.

HEAT_RECT, HEAT_DISK, HEAT_GAUSS = 0, 1, 2


def _rect(x0, x1, y0, y1, value):
    return dict(kind=HEAT_RECT, x0=int(x0), x1=int(x1), y0=int(y0), y1=int(y1), value=float(value))


def _disk(cx, cy, r, value):
    return dict(kind=HEAT_DISK, cx=int(cx), cy=int(cy), r2=int(r) ** 2, value=float(value))


def gauss_weights(sigma, truncate=4.0):
    """(radius, w[0..radius]) of scipy.ndimage.gaussian_filter1d for ``sigma`` (weights of offsets 0..radius)."""
    sd = float(sigma)
    lw = int(truncate * sd + 0.5)
    x = np.arange(-lw, lw + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    phi = phi / phi.sum()
    return lw, np.ascontiguousarray(phi[lw:], dtype=np.float64)


def heat_program(name, nx, ny, seed=0):
    """The paint program of heat map ``name`` on an (nx, ny) map (see the generators above)."""
    P = []
    sigma = None
    normalise = True
    if name == "uniform":
        return dict(mx=nx, my=ny, base=1.0, prims=[], sigma=None, normalise=False)
    if name == "dual_core":
        base = 0.1
        P += [_rect(nx * 0.25, nx * 0.75, ny * 0.55, ny * 0.85, 1.0),
              _rect(nx * 0.25, nx * 0.75, ny * 0.15, ny * 0.45, 1.0),
              _rect(nx * 0.30, nx * 0.70, ny * 0.45, ny * 0.55, 0.5)]
        sigma = nx / 80.0
    elif name == "quad_core":
        base = 0.08
        for x0f, y0f, x1f, y1f in [(0.15, 0.55, 0.45, 0.85), (0.15, 0.15, 0.45, 0.45), (0.55, 0.55, 0.85, 0.85),
                                   (0.55, 0.15, 0.85, 0.45)]:
            P.append(_rect(nx * x0f, nx * x1f, ny * y0f, ny * y1f, 1.0))
        P += [_rect(nx * 0.45, nx * 0.55, ny * 0.10, ny * 0.90, 0.45),
              _rect(nx * 0.10, nx * 0.90, ny * 0.45, ny * 0.55, 0.45)]
        sigma = nx / 80.0
    elif name == "gpu_die":
        base = 0.05
        P += [_rect(nx * 0.15, nx * 0.85, ny * 0.15, ny * 0.85, 0.75),
              _rect(nx * 0.10, nx * 0.90, ny * 0.02, ny * 0.12, 1.0),
              _rect(nx * 0.10, nx * 0.90, ny * 0.88, ny * 0.98, 1.0)]
        for cx_f in [0.30, 0.50, 0.70]:
            P.append(_disk(int(nx * cx_f), int(ny * 0.50), int(min(nx, ny) * 0.04), 1.0))
        sigma = nx / 60.0
    elif name == "chiplet":
        base = 0.03
        cw, ch = 0.18, 0.22
        for x0f, y0f in [(0.12, 0.08), (0.12, 0.40), (0.12, 0.72), (0.58, 0.08), (0.58, 0.40), (0.58, 0.72)]:
            P.append(_rect(nx * x0f, nx * (x0f + cw), ny * y0f, ny * (y0f + ch), 1.0))
        P.append(_rect(nx * 0.38, nx * 0.52, ny * 0.10, ny * 0.90, 0.35))
        sigma = nx / 70.0
    elif name == "igbt_half_bridge":
        base = 0.02
        die_h, igbt_w, diode_w = 0.10, 0.50, 0.25
        for i, y0f in enumerate([0.05, 0.20, 0.38, 0.55, 0.72, 0.88]):
            if i % 2 == 0:
                P.append(_rect(nx * 0.20, nx * (0.20 + igbt_w), ny * y0f, ny * (y0f + die_h), 1.0))
            else:
                P.append(_rect(nx * 0.30, nx * (0.30 + diode_w), ny * y0f, ny * (y0f + die_h), 0.55))
        sigma = nx / 60.0
    elif name == "soc_heterogeneous":
        base = 0.06
        P.append(_rect(nx * 0.10, nx * 0.35, ny * 0.60, ny * 0.90, 1.0))
        for dy in [0.08, 0.22]:
            for dx in [0.10, 0.28]:
                P.append(_rect(nx * dx, nx * (dx + 0.12), ny * (0.35 + dy), ny * (0.35 + dy + 0.10), 0.30))
        P += [_rect(nx * 0.50, nx * 0.90, ny * 0.35, ny * 0.90, 0.65),
              _rect(nx * 0.05, nx * 0.95, ny * 0.02, ny * 0.12, 0.40),
              _rect(nx * 0.45, nx * 0.60, ny * 0.15, ny * 0.30, 0.80)]
        sigma = nx / 70.0
    elif name == "gaussian_hotspots":
        base = 0.05
        rng = np.random.RandomState(seed)
        for _ in range(rng.randint(1, 4)):
            cx = rng.uniform(0.2, 0.8) * nx
            cy = rng.uniform(0.2, 0.8) * ny
            sg = rng.uniform(0.06, 0.15) * min(nx, ny)
            amp = rng.uniform(0.5, 1.0)
            P.append(dict(kind=HEAT_GAUSS, gx=float(cx), gy=float(cy), den=float(2 * sg ** 2), amplitude=float(amp)))
    elif name == "hotspot_on_background":
        base = 0.20
        P += [_disk(int(nx * 0.40), int(ny * 0.60), int(min(nx, ny) * 0.07), 1.0),
              _disk(int(nx * 0.65), int(ny * 0.35), int(min(nx, ny) * 0.05), 0.85)]
        sigma = nx / 100.0
    elif name == "peripheral_ring":
        base = 0.05
        rw = 0.10
        P += [_rect(0, nx * rw, 0, ny, 0.85), _rect(nx * (1 - rw), nx, 0, ny, 0.85),
              _rect(0, nx, 0, ny * rw, 0.85), _rect(0, nx, ny * (1 - rw), ny, 0.85)]
        corner_r = int(min(nx, ny) * 0.12)
        for cx_f, cy_f in [(0.0, 0.0), (0.0, 1.0), (1.0, 0.0), (1.0, 1.0)]:
            P.append(_disk(int(nx * cx_f), int(ny * cy_f), corner_r, 1.0))
        P.append(_rect(nx * 0.20, nx * 0.80, ny * 0.20, ny * 0.80, 0.15))
        sigma = nx / 60.0
    else:
        raise KeyError(f"unknown heat map {name!r}")
    return dict(mx=nx, my=ny, base=base, prims=P, sigma=sigma, normalise=normalise)


def _reflect(i, n):
    return np.where(i < 0, -i - 1, np.where(i >= n, 2 * n - i - 1, i))


def render_heat_program(prog):
    """Host rendering of a paint program, in the device builder's order of operations (test restatement)."""
    nx, ny = prog["mx"], prog["my"]
    hm = np.full((nx, ny), prog["base"], dtype=np.float32)
    xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for q in prog["prims"]:
        if q["kind"] == HEAT_RECT:
            hm[max(q["x0"], 0):max(q["x1"], 0), max(q["y0"], 0):max(q["y1"], 0)] = np.float32(q["value"])
        elif q["kind"] == HEAT_DISK:
            hm[(xx - q["cx"]) ** 2 + (yy - q["cy"]) ** 2 < q["r2"]] = np.float32(q["value"])
        else:
            dx, dy = xx - q["gx"], yy - q["gy"]
            hm = (hm.astype(np.float64) + q["amplitude"] * np.exp(-(dx * dx + dy * dy) / q["den"])).astype(np.float32)
    if prog["sigma"] is not None:
        r, w = gauss_weights(prog["sigma"])
        for axis in (0, 1):
            n = hm.shape[axis]
            idx = np.arange(n)
            src = hm.astype(np.float64)
            take = (lambda j: src[_reflect(idx + j, n), :]) if axis == 0 else (lambda j: src[:, _reflect(idx + j, n)])
            acc = take(0) * w[0]
            for j in range(r, 0, -1):
                acc = acc + (take(-j) + take(j)) * w[j]
            hm = acc.astype(np.float32)
    if prog["normalise"] and hm.max() > 0:
        hm = hm / hm.max()
    return hm.astype(np.float32)

# ------------------------------------------------- design-level helpers ---

def compute_porosity(mask, buf=BUFFER):
    """generate_dataset.py:71-74 (numpy array or torch tensor)."""
    core = mask[buf:-buf, :, :]
    return 1.0 - int((core == BC_SOLID).sum()) / _numel(core)


def _numel(a):
    return a.numel() if hasattr(a, "numel") else a.size


def compute_hydraulic_diameter(mask, buf=BUFFER):
    """generate_dataset.py:57-68: (D_h, V_fluid, A_wetted); ``mask`` a numpy array or a (device) torch tensor."""
    core = mask[buf:-buf, 1:-1, 4:-2]
    fluid = core != BC_SOLID
    solid = core == BC_SOLID
    V = int(fluid.sum())
    A = int((fluid[1:] & solid[:-1]).sum()) + int((solid[1:] & fluid[:-1]).sum()) + \
        int((fluid[:, 1:] & solid[:, :-1]).sum()) + int((solid[:, 1:] & fluid[:, :-1]).sum()) + \
        int((fluid[:, :, 1:] & solid[:, :, :-1]).sum()) + int((solid[:, :, 1:] & fluid[:, :, :-1]).sum())
    Dh = 4.0 * V / A if A > 0 else float(mask.shape[1])
    return Dh, V, A


MIN_POROSITY, MAX_POROSITY = 0.15, 0.90


def is_valid_geometry(mask, buf=BUFFER):
    """generate_dataset.py:280-293."""
    porosity = compute_porosity(mask, buf)
    if porosity < MIN_POROSITY or porosity > MAX_POROSITY:
        return False, porosity
    if np.all(mask[buf, 1:-1, 4:-2] == BC_SOLID) or np.all(mask[-(buf + 1), 1:-1, 4:-2] == BC_SOLID):
        return False, porosity
    return True, porosity


DESIGN_CONFIGS = {
    "Straight Channels": {"design_idx": 1, "ranges": {"feature_size": (2.0, 8.0), "spacing": (1.5, 7.0),
                                                      "tau_fluid": (0.52, 0.56), "u_inlet_val": (0.02, 0.05),
                                                      "heat_power": (0.05, 0.20)}},
    "Pin-Fin Staggered": {"design_idx": 2, "ranges": {"feature_size": (1.5, 5.0), "spacing": (2.0, 7.0),
                                                      "tau_fluid": (0.52, 0.56), "u_inlet_val": (0.02, 0.05),
                                                      "heat_power": (0.05, 0.20)}},
    "Gyroid TPMS": {"design_idx": 3, "ranges": {"feature_size": (4.0, 8.5), "spacing": (2.0, 6.0),
                                                "tau_fluid": (0.52, 0.56), "u_inlet_val": (0.01, 0.025),
                                                "heat_power": (0.05, 0.20)}},
    "Schwarz P TPMS": {"design_idx": 4, "ranges": {"feature_size": (4.0, 8.5), "spacing": (2.0, 6.0),
                                                   "tau_fluid": (0.52, 0.56), "u_inlet_val": (0.01, 0.025),
                                                   "heat_power": (0.05, 0.20)}},
    "Schwarz D TPMS": {"design_idx": 5, "ranges": {"feature_size": (4.0, 8.5), "spacing": (2.0, 6.0),
                                                   "tau_fluid": (0.52, 0.56), "u_inlet_val": (0.01, 0.025),
                                                   "heat_power": (0.05, 0.20)}},
}


def generate_samples(design_name, n_samples, seed=42):
    """Latin-hypercube design points paired with every heat map (generate_dataset.py:249-274)."""
    from scipy.stats import qmc
    cfg = DESIGN_CONFIGS[design_name]
    names = list(cfg["ranges"].keys())
    lows = np.array([cfg["ranges"][p][0] for p in names])
    highs = np.array([cfg["ranges"][p][1] for p in names])
    raw = qmc.LatinHypercube(d=len(names), seed=seed).random(n=n_samples)
    scaled = qmc.scale(raw, lows, highs)
    out = []
    for i in range(n_samples):
        base = {name: float(scaled[i, j]) for j, name in enumerate(names)}
        base["design_idx"] = cfg["design_idx"]
        base["design_name"] = design_name
        for hs_name in HEAT_SOURCE_NAMES:
            p = dict(base)
            p["heat_source_type"] = hs_name
            out.append(p)
    return out


def sample_id(params):
    """Deterministic id of a design point (generate_dataset.py:366-369)."""
    import hashlib
    import json
    s = json.dumps(params, sort_keys=True) + "|coolant=water"
    return f"{params['design_name'].replace(' ', '_')}_water_{hashlib.md5(s.encode()).hexdigest()[:8]}"


# ------------------------------------------------ benchmark workloads -----

CONFIG1 = dict(shape=(74, 64, 16), tau_fluid=0.6, u_inlet_val=0.02, heat_power=0.1, t_inlet=25.0, Pr=7.0,
               k_ratio=10.0, steps=2000)


def config1_mask():
    """Plain small cold plate, 74 x 64 x 16: six straight channels of width 4 separated by 2-cell fins
    (centred in the 62 open cells across y), a 3-cell solid base, a 1-cell lid, side walls, 10-cell open
    buffers with the inlet at x = 0 and the outlet at x = 73 over the open y, z cells.  The reference's
    generator cannot produce this layout (its base is 4 cells and its fin gap is derived from the channel
    count, src/geometry.py:402-403,516), so it is built here and pinned by a committed fixture."""
    nx, ny, nz = CONFIG1["shape"]
    buf = BUFFER
    solid = np.ones((nx - 2 * buf, ny, nz), dtype=bool)
    n_ch, width, fin = 6, 4, 2
    span = n_ch * width + (n_ch - 1) * fin
    y0 = 1 + (ny - 2 - span) // 2
    for k in range(n_ch):
        a = y0 + k * (width + fin)
        solid[:, a:a + width, :] = False
    return _frame(solid, buf, base=3, top=1)


def config2_params(seed, core=(CORE_NX, CORE_NY, CORE_NZ)):
    """Full-size single solve: gyroid with period in [16, 32] and threshold in [-0.4, 0.4], quad_core heat map,
    u_in in [0.01, 0.05]; tau in [0.52, 0.56] and P in [0.05, 0.20] (the generator's ranges), all from ``seed``.
    Parameters only, in the form of ``family_params`` (``design_arrays`` / ``device_design`` build the arrays)."""
    rng = np.random.default_rng(seed)
    period = float(rng.uniform(16.0, 32.0))
    thr = float(rng.uniform(-0.4, 0.4))
    u_in = float(rng.uniform(0.01, 0.05))
    tau = float(rng.uniform(0.52, 0.56))
    P = float(rng.uniform(0.05, 0.20))
    return dict(seed=seed, core=tuple(int(c) for c in core), family="gyroid", heat_name="quad_core", tau_fluid=tau,
                heat_power=P, period=period, threshold=thr, u_inlet_val=u_in)


def config2_design(seed, core=(CORE_NX, CORE_NY, CORE_NZ)):
    """``config2_params`` with the host mask and heat map."""
    p = config2_params(seed, core)
    mask, heat = design_arrays(p)
    return dict(mask=mask, heat=heat, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"],
                heat_power=p["heat_power"], period=p["period"], threshold=p["threshold"], seed=seed, family="gyroid",
                heat_name="quad_core")


FAMILIES = ("straight", "pinfin", "gyroid", "schwarz_p", "schwarz_d")
BATCH_HEATS = ("uniform", "dual_core", "quad_core", "gaussian_hotspots", "peripheral_ring")


def family_params(seed, core=None, family=None, heat_name=None):
    """The seeded parameters of ``family_design`` (every random draw, no arrays)."""
    rng = np.random.default_rng(seed)
    if core is None:
        core = (int(rng.integers(96, 149)) - 2 * BUFFER, int(rng.integers(64, 129)), int(rng.integers(16, 33)))
    family = family or FAMILIES[int(rng.integers(0, len(FAMILIES)))]
    heat_name = heat_name or BATCH_HEATS[int(rng.integers(0, len(BATCH_HEATS)))]
    tau = float(rng.uniform(0.52, 0.56))
    P = float(rng.uniform(0.05, 0.20))
    out = dict(seed=seed, core=tuple(int(c) for c in core), family=family, heat_name=heat_name, tau_fluid=tau,
               heat_power=P)
    if family in ("straight", "pinfin"):
        r = DESIGN_CONFIGS["Straight Channels" if family == "straight" else "Pin-Fin Staggered"]["ranges"]
        out["feature_size"] = float(rng.uniform(*r["feature_size"]))
        out["spacing"] = float(rng.uniform(*r["spacing"]))
        out["u_inlet_val"] = float(rng.uniform(0.02, 0.05))
    else:
        out["period"] = float(rng.uniform(16.0, 32.0))
        out["threshold"] = float(rng.uniform(-0.4, 0.4))
        out["u_inlet_val"] = float(rng.uniform(0.01, 0.025))
    return out


FAMILY_CORE_KIND = {"empty": 0, "straight": 1, "pinfin": 2, "gyroid": 3, "schwarz_p": 4, "schwarz_d": 5}


def design_arrays(params):
    """Host (mask, heat) of a parameter set from ``family_params``."""
    core, family = params["core"], params["family"]
    if family in ("straight", "pinfin"):
        mask = _frame(core_solid(FAMILY_CORE_KIND[family], params["feature_size"], params["spacing"], core), BUFFER)
    else:
        mask = tpms_mask(FAMILY_CORE_KIND[family], params["period"], params["threshold"], core)
    heat = heat_source_3d(mask.shape, params["heat_name"], core=core[:2], seed=params["seed"])
    return mask, heat


def device_design(batch, params):
    """(mask, heat) of a ``family_params`` parameter set built on the GPU by ``batch`` (a LatticeBatch).

    Same arrays as ``design_arrays(params)``: the mask bit for bit, the heat map to the last bit of the device's
    exp against the host's (the blur and the division are the same operations in the same order)."""
    core, family = params["core"], params["family"]
    shape = (core[0] + 2 * BUFFER, core[1], core[2])
    kind = FAMILY_CORE_KIND[family]
    if family in ("straight", "pinfin"):
        mask = batch.geometry_mask(shape, kind, buffer=BUFFER, base=BASE_THICKNESS, top=TOP_THICKNESS,
                                   feature_size=params["feature_size"], spacing=params["spacing"])
    else:
        mask = batch.geometry_mask(shape, kind, buffer=BUFFER, base=BASE_THICKNESS, top=TOP_THICKNESS,
                                   period=params["period"], threshold=params["threshold"], manifold=5)
    heat = batch.heat_map(heat_program(params["heat_name"], core[0], core[1], params["seed"]), shape, BUFFER)
    return mask, heat


def family_design(seed, core=None, family=None, heat_name=None):
    """A seeded design of any family on an (optionally random) core: the batched workloads.

    Lattice: nx in [96, 148] (core = nx - 20), ny in [64, 128], nz in [16, 32] unless ``core`` is given.
    TPMS u_in is capped at 0.025 as in the reference's generator (tortuous channels amplify the inlet velocity).
    """
    p = family_params(seed, core, family, heat_name)
    mask, heat = design_arrays(p)
    extra = {k: p[k] for k in ("feature_size", "spacing", "period", "threshold") if k in p}
    return dict(mask=mask, heat=heat, tau_fluid=p["tau_fluid"], u_inlet_val=p["u_inlet_val"],
                heat_power=p["heat_power"], seed=seed, family=p["family"], heat_name=p["heat_name"], **extra)
