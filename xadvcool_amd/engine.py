"""Python host layer over libxadv.so: lattice batches and the engine classes.

The reference exposes its step as Warp kernels plus one driver function
(src/fluid_engine3d.py, src/thermal_engine3d.py, src/simulation.py).  Here
the same physics sits behind a C ABI; this module gives it the shape the
reference's callers use:

* ``LatticeBatch`` — owner of one C-ABI batch (any number of lattice slots).
* ``LatticeSolver`` — one lattice (fluid, thermal or both) with the boundary
  options of the reference's kernels and test kernels, ``step(n)``,
  macroscopic accessors ``rho``, ``u``, ``p``, ``T`` and the stats behind the
  residual checks (src/simulation.py:123-137).
* ``CoupledSolver``, ``FluidEngine3D``, ``ThermalEngine3D`` — the production
  coupled solve and the single-lattice engines, constructed from a geometry
  mask, a heat-source map and an inlet velocity.

Device memory is PyTorch's (field tensors handed to the C ABI by pointer);
every kernel is the HIP library's.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .constants import BC_SOLID

STAT_NAMES = ("max_v", "sum_v", "sum_ke", "count", "sum_rho", "sum_rho_in", "sum_rho_out", "sum_ux_out", "c_in",
              "c_out", "c_outlet", "max_T", "sum_T", "count_T", "Q_out", "c_Q_out")


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("xadvcool_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


class LatticeBatch:
    """One batch of lattice slots advanced together by the fused HIP pass."""

    def __init__(self, *, slots=1, cap_shape, fluid=True, thermal=True, general=False, precision=32, block=256,
                 xcd_remap=False, graph_chunk=100, cells_per_thread=1, fluid_storage="auto", alternate_sweep=True,
                 thermal_lattice="D3Q19", slab_planes=0, thermal_prefetch=False):
        _require_gpu()
        self._lib = _lib.load()
        if graph_chunk and graph_chunk % 2:
            graph_chunk = 0
        storage = {"auto": _lib.STORAGE_AUTO, "dense": _lib.STORAGE_DENSE, "compact": _lib.STORAGE_COMPACT}
        if fluid_storage not in storage:
            raise ValueError(f"fluid_storage must be one of {sorted(storage)}")
        if thermal_lattice not in ("D3Q19", "D3Q7"):
            raise ValueError("thermal_lattice must be 'D3Q19' (the reference's) or 'D3Q7'")
        d = _lib.BatchDesc(int(precision), int(bool(fluid)), int(bool(thermal)), int(bool(general)), int(slots),
                           int(cap_shape[0]), int(cap_shape[1]), int(cap_shape[2]), int(block), int(bool(xcd_remap)),
                           int(graph_chunk), int(cells_per_thread), storage[fluid_storage], int(bool(alternate_sweep)),
                           7 if thermal_lattice == "D3Q7" else 19, int(slab_planes), int(bool(thermal_prefetch)))
        h = ctypes.c_void_p()
        torch.cuda.init()
        _lib.check(self._lib.xadv_batch_create(ctypes.byref(d), ctypes.byref(h)), "xadv_batch_create")
        self._h = h
        self.slots = int(slots)
        self.fluid, self.thermal, self.general = bool(fluid), bool(thermal), bool(general)
        self.thermal_lattice = thermal_lattice
        self.precision = int(precision)
        self.dtype = torch.float64 if precision == 64 else torch.float32
        self.np_dtype = np.float64 if precision == 64 else np.float32
        self.graph_chunk = int(graph_chunk)
        self.shapes = [None] * self.slots

    # -- lifetime -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _lib.check(self._lib.xadv_batch_destroy(h), "xadv_batch_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- instances -------------------------------------------------------
    def load(self, slot, mask, *, heat=None, tau=0.6, tau_th_fluid=0.6, tau_th_solid=0.6, heat_power=0.0, T_in=0.0,
             rho_out=1.0, u_in=0.0, periodic=(0, 0, 0), force_x=0.0, u_wall=0.0, dirichlet=None,
             frozen_velocity=False):
        m = np.ascontiguousarray(mask, dtype=np.int8)
        if m.ndim != 3:
            raise ValueError("mask must be 3-D (nx, ny, nz)")
        nx, ny, nz = m.shape
        hs = None
        if heat is not None:
            hs = np.ascontiguousarray(heat, dtype=np.float32)
            if hs.shape != m.shape:
                raise ValueError(f"heat_source shape {hs.shape} != domain shape ({nx}, {ny}, {nz})")
        dx, dT = dirichlet if dirichlet is not None else (-1, 0.0)
        d = _lib.InstanceDesc(nx, ny, nz, m.ctypes.data, hs.ctypes.data if hs is not None else None,
                              float(tau), float(tau_th_fluid), float(tau_th_solid), float(heat_power), float(T_in),
                              float(rho_out), float(u_in), (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic]),
                              float(force_x), float(u_wall), int(dx), float(dT), int(bool(frozen_velocity)))
        _lib.check(self._lib.xadv_instance_load(self._h, int(slot), ctypes.byref(d)), "xadv_instance_load")
        self.shapes[slot] = (nx, ny, nz)

    # -- designs built on the device -------------------------------------
    def geometry_mask(self, shape, kind, *, buffer=10, base=4, top=1, feature_size=3.0, spacing=3.0, period=0.0,
                      threshold=0.0, manifold=5):
        """Device int8 mask (nx, ny, nz) of a core family (``kind``: 0 empty .. 5 Schwarz D) inside the frame."""
        nx, ny, nz = (int(v) for v in shape)
        out = torch.empty((nx, ny, nz), dtype=torch.int8, device="cuda")
        g = _lib.GeometryDesc(nx, ny, nz, int(buffer), int(base), int(top), int(kind), float(feature_size),
                              float(spacing), float(period), float(threshold), int(manifold))
        torch.cuda.synchronize()
        _lib.check(self._lib.xadv_geometry_mask(self._h, ctypes.byref(g), ctypes.c_void_p(out.data_ptr())),
                   "xadv_geometry_mask")
        return out

    def heat_map(self, program, shape, x_offset):
        """Device FP32 heat field (nx, ny, nz) from a paint program (``geometry.heat_program``)."""
        from .geometry import gauss_weights
        nx, ny, nz = (int(v) for v in shape)
        prims = program["prims"]
        if len(prims) > _lib.HEAT_MAX_PRIMS:
            raise ValueError(f"a heat program holds at most {_lib.HEAT_MAX_PRIMS} primitives")
        arr = (_lib.HeatPrim * max(1, len(prims)))()
        for k, q in enumerate(prims):
            a = arr[k]
            a.kind = int(q["kind"])
            for f in ("x0", "x1", "y0", "y1", "cx", "cy", "r2"):
                setattr(a, f, int(q.get(f, 0)))
            a.value = float(q.get("value", 0.0))
            for f in ("gx", "gy", "den", "amplitude"):
                setattr(a, f, float(q.get(f, 0.0)))
        radius, w = (0, None) if program["sigma"] is None else gauss_weights(program["sigma"])
        wp = w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if w is not None else None
        h = _lib.HeatDesc(int(program["mx"]), int(program["my"]), float(program["base"]), len(prims), arr,
                          int(radius), wp, int(bool(program["normalise"])))
        out = torch.empty((nx, ny, nz), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _lib.check(self._lib.xadv_heat_map(self._h, ctypes.byref(h), nx, ny, nz, int(x_offset),
                                           ctypes.c_void_p(out.data_ptr())), "xadv_heat_map")
        return out

    def load_device(self, slot, mask, *, heat=None, tau=0.6, tau_th_fluid=0.6, tau_th_solid=0.6, heat_power=0.0,
                    T_in=0.0, rho_out=1.0, u_in=0.0, periodic=(0, 0, 0), force_x=0.0, u_wall=0.0, dirichlet=None,
                    frozen_velocity=False):
        """``load`` from a device int8 mask and an optional device FP32 heat field of the same shape."""
        if not (torch.is_tensor(mask) and mask.is_cuda and mask.dtype == torch.int8 and mask.dim() == 3):
            raise ValueError("mask must be a 3-D int8 CUDA tensor")
        mask = mask.contiguous()
        nx, ny, nz = mask.shape
        if heat is not None:
            if not (torch.is_tensor(heat) and heat.is_cuda and heat.dtype == torch.float32):
                raise ValueError("heat must be a float32 CUDA tensor")
            if tuple(heat.shape) != (nx, ny, nz):
                raise ValueError(f"heat_source shape {tuple(heat.shape)} != domain shape ({nx}, {ny}, {nz})")
            heat = heat.contiguous()
        dx, dT = dirichlet if dirichlet is not None else (-1, 0.0)
        d = _lib.InstanceDesc(nx, ny, nz, mask.data_ptr(), heat.data_ptr() if heat is not None else None,
                              float(tau), float(tau_th_fluid), float(tau_th_solid), float(heat_power), float(T_in),
                              float(rho_out), float(u_in), (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic]),
                              float(force_x), float(u_wall), int(dx), float(dT), int(bool(frozen_velocity)))
        torch.cuda.synchronize()
        _lib.check(self._lib.xadv_instance_load_device(self._h, int(slot), ctypes.byref(d)),
                   "xadv_instance_load_device")
        self.shapes[slot] = (nx, ny, nz)

    def _dev(self, a, shape):
        if a is None:
            return None
        t = torch.as_tensor(a)
        t = t.to(device="cuda", dtype=self.dtype).contiguous()
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"field shape {tuple(t.shape)} != {tuple(shape)}")
        return t

    def init(self, slot, rho=None, u=None, T=None):
        nx, ny, nz = self.shapes[slot]
        tr = self._dev(rho, (nx, ny, nz))
        tu = self._dev(u, (nx, ny, nz, 3))
        tT = self._dev(T, (nx, ny, nz))
        torch.cuda.synchronize()
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        _lib.check(self._lib.xadv_instance_init(self._h, int(slot), p(tr), p(tu), p(tT)), "xadv_instance_init")

    def set_active(self, slot, active):
        _lib.check(self._lib.xadv_set_active(self._h, int(slot), int(bool(active))), "xadv_set_active")

    def set_inlet_velocity(self, slot, u_in):
        _lib.check(self._lib.xadv_set_inlet_velocity(self._h, int(slot), float(u_in)), "xadv_set_inlet_velocity")

    def stream(self):
        """The batch's HIP stream as a torch ExternalStream (events recorded there time the kernels)."""
        ptr = ctypes.c_void_p()
        _lib.check(self._lib.xadv_get_stream(self._h, ctypes.byref(ptr)), "xadv_get_stream")
        return torch.cuda.ExternalStream(ptr.value)

    # -- stepping -------------------------------------------------------
    def step(self, n, want_fields=True):
        _lib.check(self._lib.xadv_step(self._h, int(n), int(bool(want_fields))), "xadv_step")

    def readout(self, with_stats=True):
        _lib.check(self._lib.xadv_readout(self._h, int(bool(with_stats))), "xadv_readout")

    def stats(self):
        out = np.zeros((self.slots, 16), np.float64)
        _lib.check(self._lib.xadv_stats(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "xadv_stats")
        return out

    def fields(self, slot, rho=True, u=True, T=True):
        nx, ny, nz = self.shapes[slot]
        tr = torch.empty((nx, ny, nz), dtype=self.dtype, device="cuda") if rho else None
        tu = torch.empty((nx, ny, nz, 3), dtype=self.dtype, device="cuda") if u else None
        tT = torch.empty((nx, ny, nz), dtype=self.dtype, device="cuda") if T else None
        torch.cuda.synchronize()
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        _lib.check(self._lib.xadv_fields(self._h, int(slot), p(tr), p(tu), p(tT)), "xadv_fields")
        return tr, tu, tT

    def checks_raw(self, slot, k_fluid, k_solid):
        out = np.zeros(_lib.NCHECK, np.float64)
        _lib.check(self._lib.xadv_checks(self._h, int(slot), float(k_fluid), float(k_solid),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "xadv_checks")
        return out

    def metrics(self, slot, core_x0=10, core_nx=128, core_ny=128):
        out = np.zeros(_lib.NMETRIC, np.float64)
        _lib.check(self._lib.xadv_metrics(self._h, int(slot), int(core_x0), int(core_nx), int(core_ny),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "xadv_metrics")
        return out


class LatticeSolver:
    """One lattice instance: fluid, thermal or coupled, with the reference's boundary options.

    Stepping semantics follow src/simulation.py:94-99 (coupled) and the test
    schedules of validate_analytical.py (single lattice).  ``step(n)`` advances
    n time steps; ``rho``, ``u``, ``p`` and ``T`` give the fields after the
    last step (numpy, reference layout).
    """

    def __init__(self, mask, *, fluid=True, thermal=True, heat=None, tau=0.6, tau_th_fluid=0.6, tau_th_solid=0.6,
                 heat_power=0.0, T_in=0.0, rho_out=1.0, u_in=0.0, periodic=(0, 0, 0), force_x=0.0, u_wall=0.0,
                 dirichlet=None, frozen_velocity=False, precision=32, block=256, xcd_remap=False, graph_chunk=100,
                 cells_per_thread=1, fluid_storage="auto", alternate_sweep=True, thermal_lattice="D3Q19",
                 thermal_prefetch=False):
        mask = np.asarray(mask)
        general = any(periodic) or force_x != 0.0 or u_wall != 0.0 or dirichlet is not None
        self.mask = mask
        self.shape = tuple(mask.shape)
        self.batch = LatticeBatch(slots=1, cap_shape=self.shape, fluid=fluid, thermal=thermal, general=general,
                                  precision=precision, block=block, xcd_remap=xcd_remap, graph_chunk=graph_chunk,
                                  cells_per_thread=cells_per_thread, fluid_storage=fluid_storage,
                                  alternate_sweep=alternate_sweep, thermal_lattice=thermal_lattice,
                                  thermal_prefetch=thermal_prefetch)
        self.batch.load(0, mask, heat=heat, tau=tau, tau_th_fluid=tau_th_fluid, tau_th_solid=tau_th_solid,
                        heat_power=heat_power, T_in=T_in, rho_out=rho_out, u_in=u_in, periodic=periodic,
                        force_x=force_x, u_wall=u_wall, dirichlet=dirichlet, frozen_velocity=frozen_velocity)
        self.fluid, self.thermal = bool(fluid), bool(thermal)
        self.steps = 0
        self._read = False
        self._inited = False

    def init_equilibrium(self, rho=None, u=None, T=None):
        """f = f_eq(rho, u), g = g_eq(T, u) (src/simulation.py:89-90); None = rest state at T_in."""
        self.batch.init(0, rho, u, T)
        self.steps = 0
        self._read = False
        self._inited = True
        self._init_fields = (rho, u, T)

    def set_inlet_velocity(self, u_in):
        self.batch.set_inlet_velocity(0, u_in)

    def step(self, n=1):
        if not self._inited:
            self.init_equilibrium()
        if n <= 0:
            return
        self.batch.step(n, want_fields=True)
        self.steps += n
        self._read = False

    def _readout(self, with_stats):
        if self.steps == 0:
            raise RuntimeError("no step taken yet")
        if not self._read or with_stats:
            self.batch.readout(with_stats=with_stats)
            self._read = True

    def stats(self):
        """The 16 stat slots (fluid 0-10, thermal 11-15) of the current step, as a dict."""
        self._readout(True)
        return dict(zip(STAT_NAMES, self.batch.stats()[0].tolist()))

    def fields_torch(self):
        self._readout(False)
        return self.batch.fields(0)

    @property
    def rho(self):
        return self.fields_torch()[0].cpu().numpy()

    @property
    def u(self):
        return self.fields_torch()[1].cpu().numpy()

    @property
    def p(self):
        """Pressure p = rho * c_s^2 = rho / 3 (generate_dataset.py:453)."""
        return self.rho / 3.0

    @property
    def T(self):
        return self.fields_torch()[2].cpu().numpy()

    def metrics_raw(self, core_x0=10, core_nx=128, core_ny=128):
        self._readout(False)
        return self.batch.metrics(0, core_x0, core_nx, core_ny)

    def metrics(self, *, tau_fluid, u_inlet_val, buf=10, core=(128, 128), Pr=7.0):
        """Performance metrics of the current fields (generate_dataset.py:80-179), reduced on the GPU."""
        from .metrics import device_metrics
        self._readout(False)
        return device_metrics(self.batch, 0, self.mask, tau_fluid=tau_fluid, u_inlet_val=u_inlet_val, buf=buf,
                              core=core, Pr=Pr)

    def close(self):
        self.batch.close()


def thermal_taus(tau_fluid, Pr, k_ratio):
    """tau_th = 0.5 + (tau - 0.5)/Pr in the fluid, 0.5 + k_ratio*(tau_th - 0.5) in the solid (src/simulation.py:70-71)."""
    tau_th_fluid = 0.5 + ((tau_fluid - 0.5) / Pr)
    tau_th_solid = 0.5 + k_ratio * (tau_th_fluid - 0.5)
    return tau_th_fluid, tau_th_solid


def normalized_heat(heat_source, shape):
    """Heat map as src/simulation.py:77-85 prepares it (default: the z = 0 plane; FP32, divided by its max)."""
    nx, ny, nz = shape
    if heat_source is not None:
        hm = np.array(heat_source, dtype=np.float32)
        if hm.shape != (nx, ny, nz):
            raise ValueError(f"heat_source shape {hm.shape} != domain shape ({nx}, {ny}, {nz})")
    else:
        hm = np.zeros((nx, ny, nz), dtype=np.float32)
        hm[:, :, 0] = 1.0
    if hm.max() > 0:
        hm /= hm.max()
    return hm


class CoupledSolver(LatticeSolver):
    """The production solve: D3Q19 fluid + D3Q19 thermal on one cold-plate lattice.

    Constructed like src/simulation.py:14-90 from a mask, a heat-source map
    (None: the z = 0 plane) and an inlet velocity.  ``thermal_lattice="D3Q7"``
    (keyword) swaps in the seven-velocity thermal lattice, which has no
    reference counterpart: same relaxation times in (the reference's
    diffusivities and heat input per step), 56 instead of 152 bytes per cell
    and step.
    """

    def __init__(self, mask, heat_source=None, *, tau_fluid=0.52, heat_power=0.1, t_inlet=25.0, u_inlet_val=0.03,
                 Pr=7.0, k_ratio=628.0, rho_outlet=1.0, precision=32, **kw):
        shape = np.asarray(mask).shape
        ttf, tts = thermal_taus(tau_fluid, Pr, k_ratio)
        super().__init__(mask, fluid=True, thermal=True, heat=normalized_heat(heat_source, shape), tau=tau_fluid,
                         tau_th_fluid=ttf, tau_th_solid=tts, heat_power=heat_power, T_in=t_inlet, rho_out=rho_outlet,
                         u_in=u_inlet_val, precision=precision, **kw)
        self.tau_fluid, self.tau_th_fluid, self.tau_th_solid = tau_fluid, ttf, tts
        self.t_inlet = t_inlet


class FluidEngine3D(LatticeSolver):
    """The D3Q19 fluid lattice alone (src/fluid_engine3d.py): inlet, outlet, bounce-back, optional test modes."""

    def __init__(self, mask, *, tau=0.6, u_inlet_val=0.0, rho_outlet=1.0, **kw):
        super().__init__(mask, fluid=True, thermal=False, tau=tau, u_in=u_inlet_val, rho_out=rho_outlet, **kw)


class ThermalEngine3D(LatticeSolver):
    """The D3Q19 thermal lattice alone (src/thermal_engine3d.py) with a prescribed velocity field."""

    def __init__(self, mask, *, tau_fluid=0.6, tau_solid=0.6, t_inlet=0.0, heat=None, heat_power=0.0, **kw):
        super().__init__(mask, fluid=False, thermal=True, heat=heat, tau_th_fluid=tau_fluid, tau_th_solid=tau_solid,
                         heat_power=heat_power, T_in=t_inlet, frozen_velocity=True, **kw)

    def init_equilibrium(self, rho=None, u=None, T=None):
        super().init_equilibrium(None, u, T)
