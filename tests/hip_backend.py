"""Backend of tests/analytical_cases.py on the HIP solver (xadvcool_amd)."""
import numpy as np

from xadvcool_amd.engine import LatticeSolver

_PERIODIC = {"channel": (0, 0, 0), "periodic": (1, 1, 1), "periodic_solids": (1, 1, 1), "couette": (1, 1, 0)}


class HipBackend:
    def __init__(self, precision=32):
        self.precision = precision

    def fluid(self, mask, rho0, u0, tau, steps, stream="channel", force_x=0.0, u_wall=0.0, u_in=0.0, rho_out=1.0,
              every=None, callback=None):
        s = LatticeSolver(mask, fluid=True, thermal=False, tau=tau, u_in=u_in, rho_out=rho_out,
                          periodic=_PERIODIC[stream], force_x=force_x, u_wall=u_wall, precision=self.precision)
        try:
            s.init_equilibrium(rho0, u0, None)
            done = 0
            while done < steps:
                n = min(every or steps, steps - done)
                s.step(n)
                done += n
                if callback is not None:
                    rho, u, _ = s.fields_torch()
                    callback(done, rho.cpu().numpy(), u.cpu().numpy())
            rho, u, _ = s.fields_torch()
            return rho.cpu().numpy(), u.cpu().numpy()
        finally:
            s.close()

    def thermal(self, mask, T0, u, tau_f, tau_s, steps, stream="channel", heat=None, power=0.0, T_in=0.0,
                dirichlet=None):
        s = LatticeSolver(mask, fluid=False, thermal=True, heat=heat, tau_th_fluid=tau_f, tau_th_solid=tau_s,
                          heat_power=power, T_in=T_in, periodic=(1, 1, 1) if stream == "periodic" else (0, 0, 0),
                          dirichlet=dirichlet, frozen_velocity=True, precision=self.precision)
        try:
            s.init_equilibrium(None, u, T0)
            s.step(steps)
            return s.fields_torch()[2].cpu().numpy()
        finally:
            s.close()
