"""MI355X-native coupled D3Q19 fluid + thermal lattice-Boltzmann solver.

Hot path: hand-written HIP kernels for gfx950 in ``csrc/`` behind the C ABI of
``include/xadv.h`` (``libxadv.so``), driven from Python through ctypes.
"""
from .constants import BC_FLUID, BC_INLET, BC_OUTLET, BC_SOLID, E, OPPOSITE, W  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # Engine classes need torch; import them lazily so ``import xadvcool_amd``
    # stays cheap for tools that only need constants or the build step.
    if name in ("LatticeBatch", "LatticeSolver", "CoupledSolver", "FluidEngine3D", "ThermalEngine3D"):
        from . import engine
        return getattr(engine, name)
    if name == "run_simulation":
        from .simulation import run_simulation
        return run_simulation
    raise AttributeError(name)
