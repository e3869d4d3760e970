"""The ten closed-form cases of the reference's test suite, written once for any backend.

Each ``case_*`` restates the set-up, step count and error metric of the
matching ``run_*`` function in ``validate_analytical.py`` (cited per case) and
returns the same error figure; ``THRESHOLDS`` holds the pytest thresholds of
``validate_analytical.py:1174-1223``.

A backend provides two calls (see ``OracleBackend`` below and
``HipBackend`` in tests/hip_backend.py):

``fluid(mask, rho0, u0, tau, steps, stream=..., force_x=0, u_wall=0, u_in=0,
rho_out=1, every=None, callback=None) -> (rho, u)``
    Fluid-only run of the validate_analytical.py schedule (BGK or forced
    collide, stream, macro).  ``stream`` is "channel" (streaming_kernel),
    "periodic" (periodic_streaming_kernel), "periodic_solids"
    (periodic_streaming_with_solids_kernel) or "couette"
    (couette_streaming_kernel).  With ``every``, ``callback(step, rho, u)``
    is called after every ``every`` steps.

``thermal(mask, T0, u, tau_f, tau_s, steps, stream=..., heat=None, power=0,
T_in=0, dirichlet=None) -> T``
    Thermal-only run with a frozen velocity (thermal equilibrium, BGK +
    source, stream, optional Dirichlet plane ``(x, T)``, macro).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfc

BC_FLUID, BC_SOLID, BC_INLET, BC_OUTLET = 0, 1, 2, 3

THRESHOLDS = {
    "semi_infinite_conduction": 3.0,
    "advection_diffusion_gaussian": 3.0,
    "conjugate_slab": 3.0,
    "volumetric_heat_source": 3.0,
    "rectangular_poiseuille": 3.0,
    "taylor_green_vortex": 3.0,
    "kuwabara_pin_fin": 5.0,
    "couette_flow": 3.0,
    "grid_convergence": 1.8,  # minimum order
    "mach_sensitivity": 5.0,
}


def case_semi_infinite_conduction(be):
    """validate_analytical.py:32-113."""
    nx, ny, nz = 100, 3, 3
    mask = np.full((nx, ny, nz), BC_SOLID, np.int32)
    mask[0] = BC_INLET
    tau = 0.8
    alpha = (1.0 / 3.0) * (tau - 0.5)
    T_surface, steps = 100.0, 400
    T = be.thermal(mask, np.zeros((nx, ny, nz), np.float32), np.zeros((nx, ny, nz, 3), np.float32), tau, tau,
                   steps, stream="channel", T_in=T_surface)
    xs = np.arange(1, 50)
    T_ana = T_surface * erfc(xs / (2.0 * np.sqrt(alpha * steps)))
    return float(np.max(np.abs(T[xs, 1, 1] - T_ana) / T_surface) * 100.0)


def case_rectangular_poiseuille(be):
    """validate_analytical.py:116-252."""
    nx, ny, nz = 60, 21, 11
    mask = np.full((nx, ny, nz), BC_FLUID, np.int32)
    mask[:, 0, :] = mask[:, -1, :] = BC_SOLID
    mask[:, :, 0] = mask[:, :, -1] = BC_SOLID
    mask[0, 1:-1, 1:-1] = BC_INLET
    mask[-1, 1:-1, 1:-1] = BC_OUTLET
    tau, u_in = 0.6, 0.01
    rho, u = be.fluid(mask, np.ones((nx, ny, nz), np.float32), np.zeros((nx, ny, nz, 3), np.float32), tau, 4000,
                      stream="channel", u_in=u_in, rho_out=1.0)
    u_slice = u[nx - 10, :, :, 0]
    w, h = ny - 2, nz - 2
    a, b = w / 2.0, h / 2.0
    y = np.linspace(-w / 2 + 0.5, w / 2 - 0.5, w)
    z = np.linspace(-h / 2 + 0.5, h / 2 - 0.5, h)
    Y, Z = np.meshgrid(y, z, indexing="ij")
    shape = np.zeros_like(Y)
    for i in range(1, 60, 2):
        shape += (((-1) ** ((i - 1) // 2)) / i ** 3) * (1.0 - np.cosh(i * np.pi * Z / (2 * a)) /
                                                       np.cosh(i * np.pi * b / (2 * a))) * np.cos(i * np.pi * Y / (2 * a))
    nu = (1.0 / 3.0) * (tau - 0.5)
    x1, x2 = 10, nx - 15
    dpdx = (1.0 / 3.0) * (np.mean(rho[x2, 1:-1, 1:-1]) - np.mean(rho[x1, 1:-1, 1:-1])) / float(x2 - x1)
    u_ana = (-dpdx) * (16.0 * a ** 2 / np.pi ** 3) * shape / nu
    return float(np.max(np.abs(u_slice[1:-1, 1:-1] - u_ana) / np.max(u_slice)) * 100.0)


def _tgv_init(L, nz, U0):
    k = 2.0 * np.pi / float(L)
    X, Y = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    u = np.zeros((L, L, nz, 3), np.float32)
    rho = np.zeros((L, L, nz), np.float32)
    for zz in range(nz):
        u[:, :, zz, 0] = U0 * np.sin(k * X) * np.cos(k * Y)
        u[:, :, zz, 1] = -U0 * np.cos(k * X) * np.sin(k * Y)
        rho[:, :, zz] = 1.0 - (1.0 * U0 ** 2 / (4.0 * (1.0 / 3.0))) * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    return rho, u, k, X, Y


def case_taylor_green_vortex(be):
    """validate_analytical.py:275-374."""
    L, nz, tau, U0, steps = 32, 2, 0.55, 0.05, 1000
    nu = (tau - 0.5) / 3.0
    rho0, u0, k, _, _ = _tgv_init(L, nz, U0)
    ke = [0.5 * np.sum(u0[..., 0] ** 2 + u0[..., 1] ** 2)]
    mask = np.full((L, L, nz), BC_FLUID, np.int32)
    be.fluid(mask, rho0, u0, tau, steps, stream="periodic", every=100,
             callback=lambda s, r, u: ke.append(0.5 * np.sum(u[..., 0] ** 2 + u[..., 1] ** 2)))
    t = np.arange(0, steps + 1, 100)
    ana = ke[0] * np.exp(-4.0 * nu * k ** 2 * t)
    return float(np.max(np.abs(np.array(ke) - ana) / ke[0]) * 100.0)


def case_kuwabara_pin_fin(be):
    """validate_analytical.py:422-499."""
    L, R, nz = 40, 8.0, 3
    mask = np.full((L, L, nz), BC_FLUID, np.int32)
    xx, yy = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    cyl = (xx - L / 2.0) ** 2 + (yy - L / 2.0) ** 2 <= R ** 2
    for zz in range(nz):
        mask[cyl, zz] = BC_SOLID
    tau, F = 0.8, 1e-5
    nu = (tau - 0.5) / 3.0
    _, u = be.fluid(mask, np.ones((L, L, nz), np.float32), np.zeros((L, L, nz, 3), np.float32), tau, 4000,
                    stream="periodic_solids", force_x=F)
    U_sim = np.mean(u[:, :, 1, 0])
    phi = float(np.sum(cyl)) / float(L * L)
    K = -0.5 * np.log(phi) - 0.75 + phi - 0.25 * phi ** 2
    U_ana = (F * L ** 2 * K) / (4.0 * np.pi * nu)
    return float(abs(U_sim - U_ana) / U_ana * 100.0)


def case_couette_flow(be):
    """validate_analytical.py:540-604."""
    nx, ny, nz = 5, 5, 21
    mask = np.full((nx, ny, nz), BC_FLUID, np.int32)
    tau, uw = 0.6, 0.05
    _, u = be.fluid(mask, np.ones((nx, ny, nz), np.float32), np.zeros((nx, ny, nz, 3), np.float32), tau, 20000,
                    stream="couette", u_wall=uw)
    prof = u[2, 2, :, 0]
    ana = uw * (np.arange(nz, dtype=float) + 0.5) / float(nz)
    return float(np.max(np.abs(prof - ana) / uw) * 100.0)


def case_advection_diffusion_gaussian(be):
    """validate_analytical.py:623-708."""
    nx, ny, nz = 200, 3, 3
    tau, U0, x0, s0, steps = 0.8, 0.05, 50.0, 10.0, 1000
    alpha = (1.0 / 3.0) * (tau - 0.5)
    u = np.zeros((nx, ny, nz, 3), np.float32)
    u[..., 0] = U0
    xs = np.arange(nx, dtype=float)
    T0 = np.zeros((nx, ny, nz), np.float64)
    T0[:] = np.exp(-((xs - x0) ** 2) / (2.0 * s0))[:, None, None]
    mask = np.full((nx, ny, nz), BC_FLUID, np.int32)
    T = be.thermal(mask, T0.astype(np.float32), u, tau, tau, steps, stream="periodic")
    Tl = T[:, 1, 1].astype(np.float64)
    pdf = Tl / np.sum(Tl)
    mean = np.sum(xs * pdf)
    var = np.sum((xs - mean) ** 2 * pdf)
    e_mean = abs(mean - (x0 + U0 * steps)) / (x0 + U0 * steps) * 100.0
    e_var = abs(var - (s0 + 2.0 * alpha * steps)) / (s0 + 2.0 * alpha * steps) * 100.0
    return float(max(e_mean, e_var))


def case_conjugate_slab(be):
    """validate_analytical.py:711-832."""
    nx, ny, nz = 100, 3, 3
    xi = nx // 2
    ts, tf = 0.8, 0.6
    a_s, a_f = (ts - 0.5) / 3.0, (tf - 0.5) / 3.0
    Th, Tc = 100.0, 0.0
    mask = np.full((nx, ny, nz), BC_FLUID, np.int32)
    mask[:xi] = BC_SOLID
    mask[0] = BC_INLET
    mask[-1] = BC_INLET
    Ls, Lf = float(xi), float(nx - xi)
    Ti = (a_s * Th * Lf + a_f * Tc * Ls) / (a_s * Lf + a_f * Ls)
    ana = np.zeros(nx)
    for i in range(nx):
        ana[i] = Th + (Ti - Th) * (i / Ls) if i < xi else Ti + (Tc - Ti) * ((i - xi) / Lf)
    T0 = np.broadcast_to(ana[:, None, None], (nx, ny, nz)).astype(np.float32)
    T = be.thermal(mask, T0, np.zeros((nx, ny, nz, 3), np.float32), tf, ts, 30000, stream="channel", T_in=Tc,
                   dirichlet=(0, Th))
    sl = slice(2, nx - 2)
    return float(np.max(np.abs(T[sl, 1, 1] - ana[sl]) / (Th - Tc)) * 100.0)


def case_volumetric_heat_source(be):
    """validate_analytical.py:835-936."""
    nx, ny, nz = 52, 17, 17
    tau, Q = 0.8, 0.001
    alpha = (tau - 0.5) / 3.0
    mask = np.full((nx, ny, nz), BC_SOLID, np.int32)
    mask[0] = BC_INLET
    mask[-1] = BC_INLET
    hs = np.zeros((nx, ny, nz), np.float32)
    hs[1:-1] = 1.0
    T = be.thermal(mask, np.zeros((nx, ny, nz), np.float32), np.zeros((nx, ny, nz, 3), np.float32), tau, tau,
                   40000, stream="channel", heat=hs, power=Q, T_in=0.0)
    Tl = T[:, ny // 2, nz // 2]
    Qe = (1.0 - 0.5 / tau) * Q
    L = float(nx - 1)
    Tmax = Qe * L ** 2 / (8.0 * alpha)
    xs = np.arange(nx, dtype=float)
    ana = (Qe / (2.0 * alpha)) * xs * (L - xs)
    sl = slice(2, nx - 2)
    return float(np.max(np.abs(Tl[sl] - ana[sl]) / Tmax) * 100.0)


def case_mach_sensitivity(be):
    """validate_analytical.py:939-1039."""
    L, nz, tau, steps = 64, 2, 0.55, 500
    nu = (tau - 0.5) / 3.0
    worst = 0.0
    for U0 in [0.005, 0.01, 0.02, 0.05, 0.1]:
        rho0, u0, k, X, Y = _tgv_init(L, nz, U0)
        # the reference builds rho with c_s**2 = 1/3 through 1/sqrt(3)**2 (:977-979)
        cs = 1.0 / np.sqrt(3.0)
        for zz in range(nz):
            rho0[:, :, zz] = 1.0 - (1.0 * U0 ** 2 / (4.0 * cs ** 2)) * (np.cos(2 * k * X) + np.cos(2 * k * Y))
        mask = np.full((L, L, nz), BC_FLUID, np.int32)
        _, u = be.fluid(mask, rho0, u0, tau, steps, stream="periodic")
        dec = np.exp(-2.0 * nu * k ** 2 * steps)
        ux_a = U0 * np.sin(k * X) * np.cos(k * Y) * dec
        uy_a = -U0 * np.cos(k * X) * np.sin(k * Y) * dec
        num = np.sqrt(np.mean((u[:, :, 0, 0] - ux_a) ** 2 + (u[:, :, 0, 1] - uy_a) ** 2))
        den = np.sqrt(np.mean(ux_a ** 2 + uy_a ** 2))
        worst = max(worst, num / den * 100.0)
    return float(worst)


def case_grid_convergence(be):
    """validate_analytical.py:1042-1162; returns the mean observed order."""
    nx, nz, tau, umax = 5, 3, 0.8, 0.01
    nu = (tau - 0.5) / 3.0
    errs, dxs = [], []
    for ny in [5, 9, 13, 17]:
        a = (ny - 2) / 2.0
        F = 2.0 * nu * 1.0 * umax / a ** 2
        steps = int(20.0 * a ** 2 / nu) + 3000
        mask = np.full((nx, ny, nz), BC_FLUID, np.int32)
        mask[:, 0, :] = mask[:, -1, :] = BC_SOLID
        _, u = be.fluid(mask, np.ones((nx, ny, nz), np.float32), np.zeros((nx, ny, nz, 3), np.float32), tau,
                        steps, stream="periodic_solids", force_x=F)
        ux = u[nx - 2, :, 1, 0] + F / 2.0
        yc = (ny - 1) / 2.0
        ana = umax * (1.0 - ((np.arange(ny, dtype=float) - yc) / a) ** 2)
        sl = slice(1, ny - 1)
        num = np.sqrt(np.mean((ux[sl] - ana[sl]) ** 2))
        den = np.sqrt(np.mean(ana[sl] ** 2))
        errs.append(num / den if den >= 1e-10 else num)
        dxs.append(1.0 / (ny - 1))
    ps = [np.log(errs[i] / errs[i + 1]) / np.log(dxs[i] / dxs[i + 1]) for i in range(len(errs) - 1)]
    return float(np.mean(ps))


CASES = {
    "semi_infinite_conduction": case_semi_infinite_conduction,
    "advection_diffusion_gaussian": case_advection_diffusion_gaussian,
    "conjugate_slab": case_conjugate_slab,
    "volumetric_heat_source": case_volumetric_heat_source,
    "rectangular_poiseuille": case_rectangular_poiseuille,
    "taylor_green_vortex": case_taylor_green_vortex,
    "kuwabara_pin_fin": case_kuwabara_pin_fin,
    "couette_flow": case_couette_flow,
    "grid_convergence": case_grid_convergence,
    "mach_sensitivity": case_mach_sensitivity,
}


def passes(name, value):
    thr = THRESHOLDS[name]
    return value >= thr if name == "grid_convergence" else value <= thr


class OracleBackend:
    """The CPU oracle running the reference's test schedules (validate_analytical.py)."""

    _FS = {"channel": 0, "periodic": 1, "periodic_solids": 2, "couette": 3}
    _TS = {"channel": 0, "periodic": 1}

    def __init__(self, dtype=np.float32, mutation=0):
        from oracle import oracle as orc
        self.orc = orc
        self.dtype = dtype
        self.mutation = mutation

    def fluid(self, mask, rho0, u0, tau, steps, stream="channel", force_x=0.0, u_wall=0.0, u_in=0.0, rho_out=1.0,
              every=None, callback=None):
        st = self.orc.State(mask.shape, self.dtype)
        st.init_fluid(rho0, u0)
        kw = dict(fluid=2, fluid_stream=self._FS[stream], forced=int(force_x != 0.0), tau=tau, force_x=force_x,
                  u_wall=u_wall, u_in=u_in, rho_out=rho_out, mutation=self.mutation)
        done = 0
        while done < steps:
            n = min(every or steps, steps - done)
            self.orc.run(st, mask, n, **kw)
            done += n
            if callback is not None:
                callback(done, st.rho, st.u)
        return st.rho.copy(), st.u.copy()

    def thermal(self, mask, T0, u, tau_f, tau_s, steps, stream="channel", heat=None, power=0.0, T_in=0.0,
                dirichlet=None):
        st = self.orc.State(mask.shape, self.dtype)
        st.u[...] = u
        st.init_thermal(T0, mask)
        dx, dT = dirichlet if dirichlet is not None else (-1, 0.0)
        self.orc.run(st, mask, steps, thermal=2, thermal_stream=self._TS[stream], dirichlet_x=dx, dirichlet_T=dT,
                     tau_th_fluid=tau_f, tau_th_solid=tau_s, heat_power=power, T_in=T_in, heat=heat,
                     mutation=self.mutation)
        return st.T.copy()
