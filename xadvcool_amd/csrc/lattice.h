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
//
// The column table is padded with ny + 1 empty columns on each side, so the
// eight neighbour columns of any column are inside the table (those past the
// lattice's x ends are the padding; their numbers are never used, because the
// neighbours there lie outside the lattice and bounce back).
struct ColInfo {
  uint32_t start, bits;
};
constexpr int kCompactMaxNz = 32;

// Entries of the padded table of an nx x ny lattice.
__host__ __device__ inline int64_t padded_columns(int nx, int ny) { return (int64_t)(nx + 2) * ny + 2; }

// Padded column table of a mask (int8 codes, C order (nx, ny, nz), nz <= 32)
// into cols[padded_columns(nx, ny)]; returns the number of non-solid cells.
// Column q of the lattice is entry ny + 1 + q.
inline int64_t build_columns(const int8_t* mask, int nx, int ny, int nz, ColInfo* cols) {
  int64_t n = 0;
  const int64_t ncol = (int64_t)nx * ny, pad = ny + 1;
  for (int64_t k = 0; k < padded_columns(nx, ny); ++k) cols[k] = ColInfo{0u, 0u};
  for (int64_t q = 0; q < ncol; ++q) {
    uint32_t bits = 0;
    for (int z = 0; z < nz; ++z)
      if (mask[q * nz + z] != BC_SOLID) bits |= 1u << z;
    cols[pad + q].start = (uint32_t)n;
    cols[pad + q].bits = bits;
    n += __builtin_popcount(bits);
  }
  return n;
}

// q = c / nz for c < 2^32 / nz (so for every cell of a slot: fewer than 2^26
// cells): floor(c * m / 2^32) with m = floor((2^32 - 1) / nz) + 1.  Writing
// m = 2^32 / nz + e with 0 <= e < 1, c * m / 2^32 = c / nz + c * e / 2^32, and
// c * e / 2^32 < 1 / nz keeps the floor at c / nz.  nz = 1 is passed through.
__host__ __device__ inline uint32_t nz_magic(uint32_t nz) {
  return nz <= 1 ? 0u : (uint32_t)(0xffffffffull / nz + 1ull);
}
__host__ __device__ inline uint32_t div_nz(uint32_t c, uint32_t nz, uint32_t magic) {
  return nz <= 1 ? c : (uint32_t)(((uint64_t)c * magic) >> 32);
}

// Block order of the fused pass and the readout.  A batch of S slots has bps
// lattice blocks per slot; lattice block s * bps + lb covers cells
// [lb * cells_per_block, ...) of slot s.  launch_block(b, grid, bps, order)
// gives the lattice block that launch block b (of `grid`) works on, or a value
// >= S * bps for a padding block that does nothing:
//  * order bit 1 reverses the visiting order (grid - 1 - b);
//  * order bit 0 is the XCD-contiguous order.  The hardware hands launch block
//    b to XCD b % 8, so XCD x takes the x-th eighth of every slot (per =
//    ceil(bps / 8) blocks of it), slot after slot.  A cell's x +- 1 and
//    y +- 1 neighbours then mostly lie in blocks on the same XCD and its L2,
//    and a slot that is idle takes an equal share of work from every XCD.
//    The grid is 8 * per * S blocks.
// Both orders are bijections from the non-padding launch blocks onto the
// lattice blocks (tests/block_order_check.cpp checks it exhaustively).
__host__ __device__ inline int xcd_grid(int bps, int slots) { return 8 * ((bps + 7) >> 3) * slots; }
__host__ __device__ inline int launch_block(int b, int grid, int bps, int order) {
  if (order & 2) b = grid - 1 - b;
  if (!(order & 1)) return b;
  const int per = (bps + 7) >> 3;
  const int j = b >> 3;  // position in the XCD's sequence
  const int s = j / per;
  const int lb = (b & 7) * per + (j - s * per);
  return lb < bps ? s * bps + lb : 0x7fffffff;
}

// Column entry q + rel of a padded table (cols points at lattice column 0).
__host__ __device__ inline ColInfo load_col(const ColInfo* cols, int rel, uint32_t q) {
#if defined(__HIP_DEVICE_COMPILE__)
  // Uniform base, 32-bit per-lane offset: one global load with an SGPR base.
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const ColInfo gcol;
  gchar* base = (gchar*)(cols + rel);
  return *(gcol*)(base + q * (uint32_t)sizeof(ColInfo));
#else
  return cols[rel + (int64_t)q];
#endif
}

// Byte offsets of one non-solid cell for the compact fluid planes (RS = bytes
// per value, NFB = plane stride in bytes):
//   j4      the cell's own number x RS;
//   JB[k]   the number of the neighbour at c + e_k x RS, plus NFB for even k.
// With the planes of a direction pair addressed from the lower plane
// min(i, opp(i)), every odd-pass access of direction i is one of two such
// offsets (see lbm_kernels.h): the neighbour's slot or the cell's own.
// JB[k] is meaningful only where that neighbour is inside the lattice and
// non-solid.
template <uint32_t RS>
__host__ __device__ inline uint32_t own_offset(const ColInfo* cols, uint32_t q, int z) {
  const ColInfo own = load_col(cols, 0, q);
  return (own.start + (uint32_t)__builtin_popcount(own.bits & ((1u << z) - 1u))) * RS;
}

template <uint32_t RS>
__host__ __device__ inline void neighbour_offsets(const ColInfo* cols, uint32_t q, int z, int ny, uint32_t NFB,
                                                  uint32_t& j4, uint32_t* JB) {
  const uint32_t m0 = (1u << z) - 1u;  // z <= 31: bits below z
  const uint32_t mp = (m0 << 1) | 1u;  // bits below z + 1
  const uint32_t mm = m0 >> 1;         // bits below z - 1
  const ColInfo own = load_col(cols, 0, q);
  j4 = (own.start + (uint32_t)__builtin_popcount(own.bits & m0)) * RS;
#define XADV_NB(k)                                                                                  \
  {                                                                                                 \
    constexpr int dx = ex_of(k), dy = ey_of(k), dz = ez_of(k);                                      \
    constexpr uint32_t even = (k & 1) ? 0u : 1u;                                                    \
    if (dx == 0 && dy == 0) {                                                                       \
      JB[k] = (dz > 0 ? j4 + RS : j4 - RS) + (even ? NFB : 0u);                                     \
    } else {                                                                                        \
      const ColInfo cn = load_col(cols, dx * ny + dy, q);                                           \
      const uint32_t m = dz > 0 ? mp : (dz < 0 ? mm : m0);                                          \
      JB[k] = (cn.start + (uint32_t)__builtin_popcount(cn.bits & m)) * RS + (even ? NFB : 0u);      \
    }                                                                                               \
  }
  XADV_NB(1) XADV_NB(2) XADV_NB(3) XADV_NB(4) XADV_NB(5) XADV_NB(6) XADV_NB(7) XADV_NB(8) XADV_NB(9)
  XADV_NB(10) XADV_NB(11) XADV_NB(12) XADV_NB(13) XADV_NB(14) XADV_NB(15) XADV_NB(16) XADV_NB(17) XADV_NB(18)
#undef XADV_NB
}

}  // namespace xadv
