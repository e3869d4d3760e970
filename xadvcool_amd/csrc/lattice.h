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

// Compact storage of the fluid lattice.  SOLID cells hold no fluid
// populations (theirs never change), so the fluid planes store only the
// non-solid cells, numbered in cell order (x, y, z).  Column q = x * ny + y
// records the number of its first non-solid cell and one bit per z that is set
// for a non-solid cell (nz <= 32); the number of a non-solid (x, y, z) is
// start + popcount(bits below z).  Neighbour numbers therefore come from the
// 3 x 3 columns around a cell, without a per-cell neighbour table.
struct ColInfo {
  uint32_t start, bits;
};
constexpr int kCompactMaxNz = 32;

__host__ __device__ inline uint32_t bits_below(int z) {
  return z <= 0 ? 0u : (z >= 32 ? 0xffffffffu : ((1u << z) - 1u));
}
__host__ __device__ inline uint32_t compact_index(ColInfo c, int z) {
  return c.start + (uint32_t)__builtin_popcount(c.bits & bits_below(z));
}

// The 3 x 3 columns around column q (col[dx + 1][dy + 1] is column
// q + dx * ny + dy); indices off the column range are clamped, since the
// columns there are never used (those neighbours lie outside the lattice).
__host__ __device__ inline void around_columns(const ColInfo* cols, int q, int ny, int ncol, ColInfo (&col)[3][3]) {
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      int t = q + dx * ny + dy;
      t = t < 0 ? 0 : (t >= ncol ? ncol - 1 : t);
      col[dx + 1][dy + 1] = cols[t];
    }
}

// Number of the non-solid cell at offset (DX, DY, DZ) from non-solid cell
// number j at height z.  Meaningful only where that neighbour is inside the
// lattice and non-solid; callers discard it elsewhere.
template <int DX, int DY, int DZ>
__host__ __device__ inline uint32_t nb_number(const ColInfo (&col)[3][3], uint32_t j, int z) {
  if (DX == 0 && DY == 0) return DZ > 0 ? j + 1u : j - 1u;  // same column: next / previous non-solid cell
  return compact_index(col[DX + 1][DY + 1], z + DZ);
}

// Column table of a mask (int8 codes, C order (nx, ny, nz), nz <= 32) into
// cols[nx * ny]; returns the number of non-solid cells.
inline int64_t build_columns(const int8_t* mask, int nx, int ny, int nz, ColInfo* cols) {
  int64_t n = 0;
  const int64_t ncol = (int64_t)nx * ny;
  for (int64_t q = 0; q < ncol; ++q) {
    uint32_t bits = 0;
    for (int z = 0; z < nz; ++z)
      if (mask[q * nz + z] != BC_SOLID) bits |= 1u << z;
    cols[q].start = (uint32_t)n;
    cols[q].bits = bits;
    n += __builtin_popcount(bits);
  }
  return n;
}

// J[k] = number of the cell at c + e_k, k = 1..18.
__host__ __device__ inline void neighbor_numbers(const ColInfo (&col)[3][3], uint32_t j, int z, uint32_t* J) {
#define XADV_NB(k) J[k] = nb_number<ex_of(k), ey_of(k), ez_of(k)>(col, j, z);
  XADV_NB(1) XADV_NB(2) XADV_NB(3) XADV_NB(4) XADV_NB(5) XADV_NB(6) XADV_NB(7) XADV_NB(8) XADV_NB(9)
  XADV_NB(10) XADV_NB(11) XADV_NB(12) XADV_NB(13) XADV_NB(14) XADV_NB(15) XADV_NB(16) XADV_NB(17) XADV_NB(18)
#undef XADV_NB
}

}  // namespace xadv
