// D3Q19 lattice constants and per-cell flag layout shared by the kernels and
// the host code that builds the flags.
//
// Velocity order, weights and opposite table are the reference's
// (src/fluid_engine3d.py:19-53): rest, 6 faces, 12 edges.  The thermal
// lattice reuses them (src/thermal_engine3d.py:2).  The functions are
// constexpr so that fully unrolled loops fold them to immediates.
#pragma once
#include <cstdint>

namespace xadv {

enum : int { BC_FLUID = 0, BC_SOLID = 1, BC_INLET = 2, BC_OUTLET = 3 };

__host__ __device__ constexpr int ex_of(int i) {
  return (i == 1 || i == 7 || i == 9 || i == 11 || i == 13) ? 1
         : (i == 2 || i == 8 || i == 10 || i == 12 || i == 14) ? -1 : 0;
}
__host__ __device__ constexpr int ey_of(int i) {
  return (i == 3 || i == 7 || i == 10 || i == 15 || i == 17) ? 1
         : (i == 4 || i == 8 || i == 9 || i == 16 || i == 18) ? -1 : 0;
}
__host__ __device__ constexpr int ez_of(int i) {
  return (i == 5 || i == 11 || i == 14 || i == 15 || i == 18) ? 1
         : (i == 6 || i == 12 || i == 13 || i == 16 || i == 17) ? -1 : 0;
}
__host__ __device__ constexpr int opp_of(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
__host__ __device__ constexpr double w_of(int i) { return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }
// Slot of an e_x = -1 direction in the 5-wide x-high side buffer (-1 otherwise).
__host__ __device__ constexpr int xm_slot(int i) {
  return i == 2 ? 0 : i == 8 ? 1 : i == 10 ? 2 : i == 12 ? 3 : i == 14 ? 4 : -1;
}

// Per-cell flag word (uint32), built on the host from the mask.
enum : uint32_t {
  F_TYPE = 3u,              // bits 0-1: BC_* code
  F_BOUNCE0 = 2,            // bit 2+i: fluid pull of direction i bounces back
                            // (source x - e_i outside a non-periodic axis, or SOLID)
  F_FX0 = 1u << 21,         // cell on x = 0 of a non-periodic x axis
  F_FXN = 1u << 22,         // cell on x = nx-1
  F_FY0 = 1u << 23,
  F_FYN = 1u << 24,
  F_FZ0 = 1u << 25,
  F_FZN = 1u << 26,
  F_HEAT = 1u << 27,        // heat source q > 0
  F_DIR = 1u << 28,         // cell on the Dirichlet plane
  F_SNAP = 1u << 29,        // x + 1 is an OUTLET cell: publish the full post-collision set
  F_XH = 1u << 30,          // thermal x-high extrapolation cell (x = nx-1, not INLET/OUTLET)
};

}  // namespace xadv
