"""D3Q19 lattice constants, cell codes as in src/fluid_engine3d.py:10-53.

Velocity order: rest, 6 faces, 12 edges; weights 1/3, 1/18, 1/36; OPPOSITE
pairs (1,2), (3,4), ... (17,18).  The HIP kernels carry the same table in
csrc/lattice.h.
"""
import numpy as np

BC_FLUID = 0
BC_SOLID = 1
BC_INLET = 2
BC_OUTLET = 3

E = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
    [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
], dtype=np.int32)
W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12, dtype=np.float32)
OPPOSITE = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int32)
