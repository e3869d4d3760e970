// geometry_kernels.h — cold-plate masks, heat maps and slot inputs built on
// the device.
//
// The mask kernel restates src/geometry.py:365-524 (the core carving of the
// straight-channel, staggered pin-fin and TPMS families, then the buffers,
// side walls, base, lid and inlet/outlet planes); the heat-map kernels
// restate the ten generators of src/geometry.py:69-319 as a list of painted
// primitives, a separable Gaussian blur with reflecting edges (the
// scipy.ndimage.gaussian_filter call of those generators) and a division by
// the maximum; the slot-input kernels do on the device what the host loader
// does with a host mask (flags with per-direction bounce bits, q = heat * P,
// the compact column table).  Every floating-point expression is written in
// the order the reference's numpy expressions evaluate, in FP64 where numpy
// computes in FP64, and the library is built with -ffp-contract=off, so the
// masks are the same integers and the heat maps the same FP32 values up to
// the last-bit differences of the device's exp/sin/cos against the host
// libm's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

namespace xadv {

// ---------------------------------------------------------------- masks ---

enum CoreKind : int { CORE_EMPTY = 0, CORE_STRAIGHT = 1, CORE_PINFIN = 2, CORE_GYROID = 3, CORE_SCHWARZ_P = 4,
                      CORE_SCHWARZ_D = 5 };

struct GeomParams {
  int nx, ny, nz;      // full lattice
  int buf;             // open cells before and after the core along x
  int base, top;       // solid floor and lid thickness (top >= 1)
  int kind;            // CoreKind
  int cnx;             // core x extent, nx - 2 buf
  int manifold;        // TPMS: open core cells at each x end
  // straight channels (src/geometry.py:399-421)
  int n_channels, channel_width;
  int straight_carves;  // 0 when the channels fill the whole width
  double gap;
  // pin fins (src/geometry.py:423-446)
  int pitch;
  double half_pitch, radius2;
  // TPMS (src/geometry.py:448-503)
  double omega, level;
};

// numpy's float remainder: the result takes the sign of the (positive) divisor.
__device__ inline double py_mod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0 && m < 0.0) m += b;
  return m;
}

__device__ inline bool straight_solid(const GeomParams& p, int y) {
  if (!p.straight_carves) return false;
  if ((int)p.gap > 0 && y < (int)p.gap) return true;
  double cur = p.gap;
  for (int i = 0; i < p.n_channels; ++i) {
    cur += (double)p.channel_width;
    const int y0 = (int)cur;
    int y1 = (int)(cur + p.gap);
    if (i == p.n_channels - 1) y1 = p.ny;
    if (y0 < p.ny && y >= y0 && y < y1) return true;
    cur += p.gap;
  }
  return false;
}

__device__ inline bool pinfin_solid(const GeomParams& p, int x, int y) {
  const double hp = p.half_pitch, pd = (double)p.pitch;
  const double dxa = (double)(x % p.pitch) - hp, dya = (double)(y % p.pitch) - hp;
  const double dxb = py_mod((double)x - hp, pd) - hp, dyb = py_mod((double)y - hp, pd) - hp;
  const double ra = dxa * dxa + dya * dya, rb = dxb * dxb + dyb * dyb;
  return fmin(ra, rb) < p.radius2;
}

__device__ inline bool tpms_solid(const GeomParams& p, int x, int y, int z) {
  if (x < p.manifold || x >= p.cnx - p.manifold) return false;
  const double wx = (double)x * p.omega, wy = (double)y * p.omega, wz = (double)z * p.omega;
  const double sx = sin(wx), sy = sin(wy), sz = sin(wz);
  const double cx = cos(wx), cy = cos(wy), cz = cos(wz);
  double f;
  if (p.kind == CORE_GYROID) {
    f = sx * cy + sy * cz;
    f = f + sz * cx;
  } else if (p.kind == CORE_SCHWARZ_P) {
    f = cx + cy;
    f = f + cz;
  } else {
    f = sx * sy * sz + sx * cy * cz;
    f = f + cx * sy * cz;
    f = f + cx * cy * sz;
  }
  return f < p.level;
}

__global__ __launch_bounds__(256) void k_geometry_mask(int8_t* __restrict__ mask, GeomParams p) {
  const int64_t n = (int64_t)p.nx * p.ny * p.nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int z = (int)(c % p.nz);
  const int y = (int)((c / p.nz) % p.ny);
  const int x = (int)(c / ((int64_t)p.nz * p.ny));
  int m = BC_FLUID;
  const int cx = x - p.buf;
  if (cx >= 0 && cx < p.cnx) {
    bool s = false;
    if (p.kind == CORE_STRAIGHT) s = straight_solid(p, y);
    else if (p.kind == CORE_PINFIN) s = pinfin_solid(p, cx, y);
    else if (p.kind >= CORE_GYROID) s = tpms_solid(p, cx, y, z);
    if (s) m = BC_SOLID;
  }
  if (y == 0 || y == p.ny - 1) m = BC_SOLID;
  if (z < p.base || z >= p.nz - p.top) m = BC_SOLID;
  const bool open = y >= 1 && y < p.ny - 1 && z >= p.base && z < p.nz - p.top;
  if (x == 0 && open) m = BC_INLET;
  if (x == p.nx - 1 && open) m = BC_OUTLET;
  mask[c] = (int8_t)m;
}

// ------------------------------------------------------------ heat maps ---

constexpr int kHeatMaxPrims = 16;
constexpr int kHeatMaxRadius = 63;

enum HeatPrimKind : int { HEAT_RECT = 0, HEAT_DISK = 1, HEAT_GAUSS = 2 };

struct HeatPrim {
  int kind;
  int x0, x1, y0, y1;  // RECT: x0 <= x < x1, y0 <= y < y1
  int cx, cy, r2;      // DISK: (x - cx)^2 + (y - cy)^2 < r2
  float value;         // RECT / DISK value
  double gx, gy, den, amplitude;  // GAUSS: + amplitude * exp(-((x-gx)^2 + (y-gy)^2) / den)
};

struct HeatParams {
  int mx, my;  // map size
  float base;
  int nprims;
  int radius;  // blur half-width; 0 = no blur
  int normalise;
  HeatPrim prims[kHeatMaxPrims];
  double w[kHeatMaxRadius + 1];  // w[j] weight of offset +-j
};

__global__ __launch_bounds__(256) void k_heat_paint(float* __restrict__ out, const HeatParams hp) {
  const int mx = hp.mx, my = hp.my;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= mx * my) return;
  const int x = c / my, y = c % my;
  float v = hp.base;
  for (int k = 0; k < hp.nprims; ++k) {
    const HeatPrim& q = hp.prims[k];
    if (q.kind == HEAT_RECT) {
      if (x >= q.x0 && x < q.x1 && y >= q.y0 && y < q.y1) v = q.value;
    } else if (q.kind == HEAT_DISK) {
      const int dx = x - q.cx, dy = y - q.cy;
      if (dx * dx + dy * dy < q.r2) v = q.value;
    } else {
      const double dx = (double)x - q.gx, dy = (double)y - q.gy;
      const double s = dx * dx + dy * dy;
      const double e = exp(-s / q.den);
      v = (float)((double)v + q.amplitude * e);
    }
  }
  out[c] = v;
}

// Reflecting index (scipy 'reflect': d c b a | a b c d | d c b a), |i| < n.
__device__ inline int reflect_index(int i, int n) {
  if (i < 0) return -i - 1;
  if (i >= n) return 2 * n - i - 1;
  return i;
}

// One axis of the Gaussian blur, accumulated in FP64 (centre term first,
// then the symmetric pairs from the outermost inwards), stored in FP32.
template <int AXIS>
__global__ __launch_bounds__(256) void k_heat_blur(const float* __restrict__ in, float* __restrict__ out,
                                                   const HeatParams hp) {
  const int mx = hp.mx, my = hp.my, r = hp.radius;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= mx * my) return;
  const int x = c / my, y = c % my;
  auto at = [&](int dj) -> double {
    if (AXIS == 0) return (double)in[reflect_index(x + dj, mx) * my + y];
    return (double)in[x * my + reflect_index(y + dj, my)];
  };
  double acc = at(0) * hp.w[0];
  for (int j = r; j >= 1; --j) acc += (at(-j) + at(j)) * hp.w[j];
  out[c] = (float)acc;
}

// Maximum, division by it (when positive), and placement of the map on the
// z = 0 plane of the (nx, ny, nz) heat field at x offset xo (one block).
__global__ __launch_bounds__(1024) void k_heat_place(const float* __restrict__ map, float* __restrict__ heat,
                                                     const HeatParams hp, int ny, int nz, int xo) {
  __shared__ float red[1024];
  const int mx = hp.mx, my = hp.my, n = mx * my;
  float m = -INFINITY;
  for (int c = threadIdx.x; c < n; c += blockDim.x) m = fmaxf(m, map[c]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float mx_all = red[0];
  const bool div = hp.normalise && mx_all > 0.0f;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int x = c / my, y = c % my;
    const float v = div ? map[c] / mx_all : map[c];
    heat[((int64_t)(xo + x) * ny + y) * nz] = v;
  }
}

// ------------------------------------------------ slot inputs on device ---

enum MaskErr : uint32_t { ERR_CODE = 1u, ERR_INLET = 2u, ERR_OUTLET = 4u, ERR_IO = 8u };

// Per block: OR of the mask's error bits (codes outside 0..3, INLET off
// x = 0 with the fluid on, OUTLET off x = nx-1 with either lattice on, any
// INLET/OUTLET at all).  Grid-stride; part[blockIdx.x] receives the bits.
__global__ __launch_bounds__(256) void k_mask_check(const int8_t* __restrict__ mask, int nx, int ny, int nz,
                                                    int fluid, int thermal, uint32_t* __restrict__ part) {
  __shared__ uint32_t red[4];
  const int64_t n = (int64_t)nx * ny * nz;
  uint32_t bits = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int m = mask[c];
    const int x = (int)(c / ((int64_t)ny * nz));
    if (m < 0 || m > 3) bits |= ERR_CODE;
    if (m == BC_INLET && fluid && x != 0) bits |= ERR_INLET;
    if (m == BC_OUTLET && (fluid || thermal) && x != nx - 1) bits |= ERR_OUTLET;
    if (m == BC_INLET || m == BC_OUTLET) bits |= ERR_IO;
  }
  // OR across the wave, then across the block's four waves.
  for (int off = 32; off > 0; off >>= 1) bits |= (uint32_t)__shfl_xor((int)bits, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = bits;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] | red[1] | red[2] | red[3];
}

__global__ __launch_bounds__(64) void k_or_reduce(const uint32_t* __restrict__ part, int nparts,
                                                  uint32_t* __restrict__ out) {
  uint32_t bits = 0;
  for (int k = threadIdx.x; k < nparts; k += 64) bits |= part[k];
  for (int off = 32; off > 0; off >>= 1) bits |= (uint32_t)__shfl_xor((int)bits, off, 64);
  if (threadIdx.x == 0) out[0] = bits;
}

struct FlagParams {
  int nx, ny, nz;
  int px, py, pz;  // periodic axes
  int general;
  int dirichlet_x;
  double heat_power;
};

// Cell flags and heat source of one slot from a device mask and heat map:
// the same bits, in the same order of rules, as the host loader builds.
template <class R>
__global__ __launch_bounds__(256) void k_slot_flags(const int8_t* __restrict__ mask, const float* __restrict__ heat,
                                                    FlagParams p, uint32_t* __restrict__ flags, R* __restrict__ q) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int z = (int)(c % nz), y = (int)((c / nz) % ny), x = (int)(c / ((int64_t)nz * ny));
  auto at = [&](int xx, int yy, int zz) { return ((int64_t)xx * ny + yy) * nz + zz; };
  const int m = mask[c];
  uint32_t w = (uint32_t)m;
#pragma unroll
  for (int i = 1; i < 19; ++i) {
    int sx = x - ex_of(i), sy = y - ey_of(i), sz = z - ez_of(i);
    bool oob = false;
    if (sx < 0 || sx >= nx) { if (p.px) sx = (sx + nx) % nx; else oob = true; }
    if (sy < 0 || sy >= ny) { if (p.py) sy = (sy + ny) % ny; else oob = true; }
    if (sz < 0 || sz >= nz) { if (p.pz) sz = (sz + nz) % nz; else oob = true; }
    const bool bb = oob || mask[at(sx, sy, sz)] == BC_SOLID;
    if (bb) w |= 1u << (F_BOUNCE0 + i);
  }
  if (!p.px && x == 0) w |= F_FX0;
  if (!p.px && x == nx - 1) w |= F_FXN;
  if (!p.py && y == 0) w |= F_FY0;
  if (!p.py && y == ny - 1) w |= F_FYN;
  if (!p.pz && z == 0) w |= F_FZ0;
  if (!p.pz && z == nz - 1) w |= F_FZN;
  if (heat && (double)heat[c] * p.heat_power > 0.0) w |= F_HEAT;
  if (p.general && p.dirichlet_x >= 0 && x == p.dirichlet_x) w |= F_DIR;
  if (x + 1 < nx && mask[at(x + 1, y, z)] == BC_OUTLET) w |= F_SNAP;
  if (!p.px && x == nx - 1 && m != BC_INLET && m != BC_OUTLET) w |= F_XH;
  flags[c] = w;
  q[c] = heat ? (R)heat[c] * (R)p.heat_power : (R)0;
}

// Column bits of the compact fluid table (one thread per lattice column).
__global__ __launch_bounds__(256) void k_col_bits(const int8_t* __restrict__ mask, int64_t ncol, int nz,
                                                  ColInfo* __restrict__ cols_lattice) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncol) return;
  uint32_t bits = 0;
  for (int z = 0; z < nz; ++z)
    if (mask[q * nz + z] != BC_SOLID) bits |= 1u << z;
  cols_lattice[q].bits = bits;
}

// Exclusive prefix sum of the columns' popcounts into .start, zeroed padding
// entries, and the slot's non-solid cell count (one block of 1024 threads:
// each thread scans a contiguous segment, the block scans the segment sums).
__global__ __launch_bounds__(1024) void k_col_scan(ColInfo* __restrict__ cols, int64_t ncol, int64_t pad,
                                                   int64_t padded, uint64_t* __restrict__ nfluid) {
  __shared__ uint32_t sums[1024];
  const int t = threadIdx.x, T = blockDim.x;
  for (int64_t k = t; k < pad; k += T) cols[k] = ColInfo{0u, 0u};
  for (int64_t k = pad + ncol + t; k < padded; k += T) cols[k] = ColInfo{0u, 0u};
  ColInfo* lat = cols + pad;
  const int64_t seg = (ncol + T - 1) / T;
  const int64_t a = t * seg < ncol ? t * seg : ncol, e = a + seg < ncol ? a + seg : ncol;
  uint32_t s = 0;
  for (int64_t q = a; q < e; ++q) s += (uint32_t)__popc(lat[q].bits);
  sums[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan over the segment sums.
  for (int off = 1; off < T; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = t > 0 ? sums[t - 1] : 0u;
  for (int64_t q = a; q < e; ++q) {
    lat[q].start = run;
    run += (uint32_t)__popc(lat[q].bits);
  }
  if (t == T - 1) nfluid[0] = (uint64_t)sums[T - 1];
}

}  // namespace xadv
