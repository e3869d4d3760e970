// MI355X (gfx950) kernels of the warm-start surrogate: a 3-D convolutional
// encoder-decoder trained and run with BF16 MFMA and FP32 accumulation.
//
// Tensors are channels-last (N, X, Y, Z, C) with C a multiple of 16 and the
// spatial dims of each level given explicitly.  Convolutions are 3x3x3, stride
// 1, zero padding 1, as implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward   out[v][co] = b[co] + sum_k W[co][k] * in[v + tap(k)][ci(k)],
//             k = tap * Cin + ci, K padded to a multiple of 32 with zero weights;
//   data grad = forward with W'[ci][tap' * Cout + co] = W[co][(26 - tap') * Cin + ci];
//   weight grad dW[co][k] = sum_v g[v][co] * in[v + tap(k)][ci(k)], split over
//             voxel chunks (partials reduced in a fixed order), db[co] = sum_v g.
// MFMA operand layout (checked on the device): lane l holds A[l%16][8*(l/16)+j],
// B[8*(l/16)+j][l%16] (j = 0..7) and D[4*(l/16)+j][l%16] (j = 0..3).  The forward
// kernels pass the weights as A and the input as B, so D is (channel x voxel) and
// a lane ends with four consecutive channels of one voxel: one 16-byte store.
// Activations use SiLU, x * sigmoid(x) with HIP's expf.  Adam follows Kingma &
// Ba (2015) with bias correction; master weights and moments are FP32.
#include <hip/hip_runtime.h>

#include <initializer_list>
#include <mutex>
#include <type_traits>

#include <cmath>
#include <cstdint>

#include "../../include/xadv_nn.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(__bf16 h) { return (float)h; }
__device__ __forceinline__ __bf16 f2bf(float f) { return (__bf16)f; }
__device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }
__device__ __forceinline__ float dsilu(float x) {
  const float s = 1.0f / (1.0f + expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

struct Geo {
  int X, Y, Z;  // spatial dims of this level
  int N;        // batch
  __host__ __device__ __forceinline__ int64_t V() const { return (int64_t)N * X * Y * Z; }
};

// Neighbour voxel of (n,x,y,z) for tap t (t = (dx+1)*9 + (dy+1)*3 + (dz+1)),
// or -1 outside the lattice (zero padding).
__device__ __forceinline__ int64_t tap_voxel(const Geo& g, int n, int x, int y, int z, int t) {
  const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < 0 || xx >= g.X || yy < 0 || yy >= g.Y || zz < 0 || zz >= g.Z) return -1;
  return (((int64_t)n * g.X + xx) * g.Y + yy) * g.Z + zz;
}

__device__ __forceinline__ void decompose(const Geo& g, int64_t v, int& n, int& x, int& y, int& z) {
  z = (int)(v % g.Z);
  int64_t r = v / g.Z;
  y = (int)(r % g.Y);
  r /= g.Y;
  x = (int)(r % g.X);
  n = (int)(r / g.X);
}

// --------------------------------------------------------------- forward ---
// One block = 4 waves; each wave computes 16 voxels x NT*16 output channels.
template <int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(const __bf16* __restrict__ in, const __bf16* __restrict__ W,
                                                  const float* __restrict__ bias, float* __restrict__ out, Geo g,
                                                  int Cin, int Kp) {
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t V = g.V();
  const int64_t v0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  if (v0 >= V) return;
  const int Cout = NT * 16;
  // A-operand voxel of this lane
  const int64_t va = v0 + (lane & 15);
  int n = 0, x = 0, y = 0, z = 0;
  const bool va_ok = va < V;
  if (va_ok) decompose(g, va, n, x, y, z);
  const int kg = (lane >> 4) * 8;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0, 0, 0, 0};
  const int K = 27 * Cin;
  for (int kb = 0; kb < Kp; kb += 32) {
    const int k0 = kb + kg;
    bf16x8 a;
    bool have = false;
    if (va_ok && k0 < K) {
      const int tap = k0 / Cin, ci = k0 - tap * Cin;
      const int64_t nb = tap_voxel(g, n, x, y, z, tap);
      if (nb >= 0) {
        const u16x8 raw = *reinterpret_cast<const u16x8*>(in + nb * Cin + ci);
        a = __builtin_bit_cast(bf16x8, raw);
        have = true;
      }
    }
    if (!have) a = __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0});
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = t * 16 + (lane & 15);
      const u16x8 rb = *reinterpret_cast<const u16x8*>(W + (int64_t)co * Kp + k0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, rb), a, acc[t], 0, 0, 0);
    }
  }
  if (!va_ok) return;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c0 = t * 16 + (lane >> 4) * 4;  // this lane's four channels of voxel va
    f32x4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = acc[t][j] + (bias ? bias[c0 + j] : 0.0f);
    *reinterpret_cast<f32x4*>(out + va * Cout + c0) = o;
  }
}

// The same convolution for Cout = 16 and Cin in {16, 32} (the full-resolution
// layers and their data gradients): a wave keeps all of its B fragments (the
// whole weight matrix, KS x 8 values per lane) in registers and sweeps
// `tiles` 16-voxel tiles, so per MFMA only the A operand is loaded.  Same
// MFMA sequence per tile as k_conv_fwd<1>, hence the same results.
template <int CIN>
__global__ __launch_bounds__(256) void k_conv_fwd16(const __bf16* __restrict__ in, const __bf16* __restrict__ W,
                                                    const float* __restrict__ bias, float* __restrict__ out, Geo g,
                                                    int Kp, int tiles) {
  constexpr int K = 27 * CIN, KS = (K + 31) / 32;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = lane & 15, kg = (lane >> 4) * 8;
  const int64_t V = g.V();
  const int64_t t0 = ((int64_t)blockIdx.x * 4 + wave) * tiles;
  if (t0 * 16 >= V) return;
  bf16x8 B[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
    B[ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(W + (int64_t)m * Kp + ks * 32 + kg));
  f32x4 b;  // this lane's four output channels 4 (l / 16) + j
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = bias ? bias[kg / 2 + j] : 0.0f;
  for (int t = 0; t < tiles; ++t) {
    const int64_t v0 = (t0 + t) * 16;
    if (v0 >= V) break;
    const int64_t va = v0 + m;
    int n = 0, x = 0, y = 0, z = 0;
    const bool va_ok = va < V;
    if (va_ok) decompose(g, va, n, x, y, z);
    floatx4 acc = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k0 = ks * 32 + kg;
      u16x8 raw = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
      if (va_ok && k0 < K) {
        const int tap = k0 / CIN, ci = k0 - tap * CIN;
        const int64_t nb = tap_voxel(g, n, x, y, z, tap);
        if (nb >= 0) raw = *reinterpret_cast<const u16x8*>(in + nb * CIN + ci);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[ks], __builtin_bit_cast(bf16x8, raw), acc, 0, 0, 0);
    }
    if (va_ok) *reinterpret_cast<f32x4*>(out + va * 16 + kg / 2) = acc + b;
  }
}

// Loads a block's input halo into registers: 3 x HY x HZ voxels of CH * 8 BF16 channels around (n, x,
// y0 .., z = 0 ..), voxel hv = (ix * HY + iy) * HZ + iz from (x + ix - 1, y0 + iy - 1, iz - 1), zero outside
// the lattice; chunk e = threadIdx.x + 256 i is 16 bytes (part e % CH) of voxel e / CH.  The coordinates
// of a thread's chunks are stepped from one to the next (256 / CH voxels apart) instead of divided out, and
// every load is issued (a position outside the lattice reads the block's own voxel (x, y0, 0) and is
// replaced by zeros), as a 32-bit offset from that voxel's address.
template <int CH, int HY, int HZ, int PER>
__device__ __forceinline__ void load_halo(const __bf16* __restrict__ in, u16x8 (&tmp)[PER], const Geo& g, int n,
                                          int x, int y0) {
  constexpr int NCH = 3 * HY * HZ * CH, DV = 256 / CH;
  constexpr int DZ = DV % HZ, DY = (DV / HZ) % HY, DX = DV / (HY * HZ);
  static_assert(256 % CH == 0, "chunks per voxel");
  const int part = threadIdx.x % CH, hv = threadIdx.x / CH;
  int ix = hv / (HY * HZ), iy = (hv / HZ) % HY, iz = hv % HZ;
  const int sz = CH * 8, sy = g.Z * sz, sx = g.Y * sy;  // element strides of z, y and x
  const __bf16* blk = in + (((int64_t)n * g.X + x) * g.Y + y0) * g.Z * sz;  // voxel (n, x, y0, 0)
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const bool ok = ((int)threadIdx.x + i * 256 < NCH) & ((unsigned)(x + ix - 1) < (unsigned)g.X) &
                    ((unsigned)(y0 + iy - 1) < (unsigned)g.Y) & ((unsigned)(iz - 1) < (unsigned)g.Z);
    const int off = (ok ? (ix - 1) * sx + (iy - 1) * sy + (iz - 1) * sz : 0) + part * 8;
    const u16x8 v = *reinterpret_cast<const u16x8*>(blk + off);
    tmp[i] = ok ? v : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    iz += DZ;
    iy += DY;
    ix += DX;
    if (iz >= HZ) iz -= HZ, ++iy;
    if (iy >= HY) iy -= HY, ++ix;
  }
}

// CIN -> 16 convolution for lattices with Z = 32 and Y a multiple of YB (the
// full-resolution layers): a block owns the YB * 32 voxels of one (n, x) and
// YB consecutive y columns, all z.  It stages their input halo (x - 1 .. x + 1,
// YB + 2 y columns, z = -1 .. 32; zeros outside the lattice) in LDS once, and
// its waves then read every tap's A operand from LDS instead of gathering it
// from global memory 27 times.  Each wave computes YB / 2 tiles of 16 voxels
// (half a z column each) with the whole weight matrix held as MFMA B
// fragments, exactly the MFMA sequence of k_conv_fwd16<CIN> per tile, so the
// outputs are the same bits.  Halo: 3 * (YB + 2) * 34 voxels of CIN BF16
// channels (CIN 16, YB 16: 58,752 bytes; CIN 32, YB 8: 65,280 bytes).
template <int CIN, int YB, bool PERSIST>
__global__ __launch_bounds__(256) void k_conv_fwd16_lds(const __bf16* __restrict__ in, const __bf16* __restrict__ W,
                                                        const float* __restrict__ bias, float* __restrict__ out,
                                                        __bf16* __restrict__ act, Geo g, int Kp) {
  constexpr int K = 27 * CIN, KS = (K + 31) / 32;
  constexpr int HY = YB + 2, HZ = 34, HV = 3 * HY * HZ;
  constexpr int CH = CIN / 8;                 // 16-byte chunks per voxel
  constexpr int kChunks = HV * CH;
  constexpr int kPer = (kChunks + 255) / 256;  // per thread
  constexpr int TPW = YB / 2;                  // tiles per wave
  static_assert(HV * CIN * 2 <= 65536, "halo must fit the static LDS limit");
  static_assert(YB % 2 == 0 && YB * 2 == 4 * TPW, "a block is 4 waves of whole tiles");
  __shared__ __attribute__((aligned(16))) __bf16 halo[HV * CIN];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = lane & 15, kg = (lane >> 4) * 8;
  const int ybl = g.Y / YB;  // work items per (n, x)
  const int nitems = g.N * g.X * ybl;
  // Work item `it`: one (n, x) and YB y columns from y0.  PERSIST: the grid is persistent and a block's
  // next item's halo is loaded into registers while it computes the current one (more registers).
  u16x8 tmp[kPer];
  auto fetch = [&](int it) {
    const int nx = it / ybl, y0 = (it - nx * ybl) * YB;
    const int n = nx / g.X, x = nx - n * g.X;
    load_halo<CH, HY, HZ>(in, tmp, g, n, x, y0);
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = threadIdx.x + i * 256;
      if (e < kChunks) *reinterpret_cast<u16x8*>(halo + e * 8) = tmp[i];
    }
  };
  if ((int)blockIdx.x < nitems) fetch(blockIdx.x);
  if (!PERSIST) stage();  // before the weights are loaded: the staging registers are free again
  bf16x8 B[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
    B[ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(W + (int64_t)m * Kp + ks * 32 + kg));
  f32x4 b;  // this lane's four output channels 4 (l / 16) + j
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = bias ? bias[kg / 2 + j] : 0.0f;
  for (int it = blockIdx.x; it < nitems; it = PERSIST ? it + (int)gridDim.x : nitems) {  // else one item per block
  if (PERSIST) stage();
  __syncthreads();
  if (PERSIST && it + (int)gridDim.x < nitems) fetch(it + gridDim.x);
  const int64_t vblock = (int64_t)it * (YB * 32);
  // The wave's tiles are computed G at a time with independent accumulators, and the A operands of
  // the next K-step are read from LDS before this step's MFMAs, so neither the LDS latency nor the
  // accumulator dependency of one tile's MFMA chain stalls the wave.  Each tile still gets the MFMA
  // sequence above (K-steps in order), so the outputs are the same bits.
  constexpr int G = TPW < 4 ? TPW : 4;
  static_assert(TPW % G == 0 && TPW % 2 == 0, "tile groups");
  // LDS offset (elements) of this lane's A operand for K-step ks in the wave's first tile; the other tiles
  // add a constant (tile t: column t / 2, z from (t % 2) * 16).  Lanes past K in the last step read zeros.
  int lo[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k0 = ks * 32 + kg;
    const int tap = k0 / CIN, ci = k0 - tap * CIN;
    const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;  // offsets + 1
    lo[ks] = ((dx * HY + wave * (TPW / 2) + dy) * HZ + m + dz) * CIN + ci;
  }
  auto a_op = [&](int ks, int t) {
    if (ks == KS - 1 && (KS - 1) * 32 + kg >= K) return u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    return *reinterpret_cast<const u16x8*>(halo + lo[ks] + ((t >> 1) * HZ + (t & 1) * 16) * CIN);
  };
#pragma unroll
  for (int t0 = 0; t0 < TPW; t0 += G) {
    floatx4 acc[G];
    u16x8 raw[2][G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      acc[q] = floatx4{0, 0, 0, 0};
      raw[0][q] = a_op(0, t0 + q);
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < KS) {
#pragma unroll
        for (int q = 0; q < G; ++q) raw[cur ^ 1][q] = a_op(ks + 1, t0 + q);
      }
#pragma unroll
      for (int q = 0; q < G; ++q)
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[ks], __builtin_bit_cast(bf16x8, raw[cur][q]), acc[q],
                                                         0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int64_t o = (vblock + (wave * TPW + t0 + q) * 16 + m) * 16 + kg / 2;
      const f32x4 pv = acc[q] + b;
      *reinterpret_cast<f32x4*>(out + o) = pv;
      if (act) {
        u16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = __builtin_bit_cast(unsigned short, f2bf(silu(pv[j])));
        *reinterpret_cast<u16x4*>(act + o) = a;
      }
    }
  }
  __syncthreads();
  }
}

// CIN -> NT * 16 convolution for lattices with Z = ZL and Y a multiple of YB:
// the input halo of a block (one (n, x), YB y columns, all ZL z; 3 x (YB + 2)
// x (ZL + 2) voxels, zeros outside the lattice) is staged in LDS once, as in
// k_conv_fwd16_lds.  Each wave owns YB * ZL / 64 tiles of 16 consecutive
// voxels (with ZL = 8 a tile spans two z columns) and walks the K steps in
// order, loading each step's B fragments once for all of its tiles (blocks of
// four tiles: the waves share the tiles and split the output channels): every
// output element gets the same MFMA sequence as in k_conv_fwd<NT>, so the
// outputs are the same bits.
template <int CIN, int NT, int ZL, int YB, bool PERSIST>
__global__ __launch_bounds__(256) void k_conv_fwd_lds(const __bf16* __restrict__ in, const __bf16* __restrict__ W,
                                                      const float* __restrict__ bias, float* __restrict__ out,
                                                      __bf16* __restrict__ act, Geo g, int Kp) {
  constexpr int K = 27 * CIN, KS = (K + 31) / 32, COUT = NT * 16;
  constexpr int HY = YB + 2, HZ = ZL + 2, HV = 3 * HY * HZ;
  constexpr int CH = CIN / 8;
  constexpr int kChunks = HV * CH;
  constexpr int kPer = (kChunks + 255) / 256;
  constexpr int TPW = YB * ZL / 64;        // tiles per wave
  static_assert(HV * CIN * 2 <= 65536, "halo must fit the static LDS limit");
  static_assert((ZL == 8 || ZL % 16 == 0) && TPW * 64 == YB * ZL, "a block is 4 waves of whole tiles");
  __shared__ __attribute__((aligned(16))) __bf16 halo[HV * CIN];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = lane & 15, kg = (lane >> 4) * 8;
  const int ybl = g.Y / YB;
  const int nitems = g.N * g.X * ybl;
  // Work items and PERSIST as in k_conv_fwd16_lds.
  u16x8 tmp[kPer];
  auto fetch = [&](int it) {
    const int nx = it / ybl, y0 = (it - nx * ybl) * YB;
    const int n = nx / g.X, x = nx - n * g.X;
    load_halo<CH, HY, HZ>(in, tmp, g, n, x, y0);
  };
  if ((int)blockIdx.x < nitems) fetch(blockIdx.x);
  for (int it = blockIdx.x; it < nitems; it = PERSIST ? it + (int)gridDim.x : nitems) {  // else one item per block
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int e = threadIdx.x + i * 256;
    if (e < kChunks) *reinterpret_cast<u16x8*>(halo + e * 8) = tmp[i];
  }
  __syncthreads();
  if (PERSIST && it + (int)gridDim.x < nitems) fetch(it + gridDim.x);
  // The wave's share: TW tiles x GW 16-channel output groups.  With 4 tiles per block (one per wave) the
  // waves take all tiles and split the output groups instead, so that each B fragment (a load from global
  // memory) serves every tile of the block; every output element still gets the same MFMA sequence.
  constexpr int TB = YB * ZL / 16;
  constexpr bool kShare = TB == 4 && NT >= 2;
  constexpr int GW = kShare ? (NT >= 4 ? NT / 4 : 1) : NT;
  constexpr int TW = kShare ? (NT >= 4 ? 4 : 2) : TPW;
  static_assert(TW * GW * 4 == TB * NT, "the waves cover the block");
  auto tile_of = [&](int i) { return kShare ? (NT >= 4 ? i : (wave >> 1) * 2 + i) : wave * TPW + i; };
  auto group_of = [&](int j) { return kShare ? (NT >= 4 ? wave * GW + j : (wave & 1)) : j; };
  floatx4 acc[TW][GW];
#pragma unroll
  for (int tl = 0; tl < TW; ++tl)
#pragma unroll
    for (int t = 0; t < GW; ++t) acc[tl][t] = floatx4{0, 0, 0, 0};
  // Operands of K-step ks: the B fragments (global memory) and each tile's A operand (LDS); the next step's
  // are loaded during this step's MFMAs.  Same MFMA sequence per output element.
  auto load_ops = [&](int ks, bf16x8 (&Bt)[GW], u16x8 (&raw)[TW]) {
    const int k0 = ks * 32 + kg;
#pragma unroll
    for (int t = 0; t < GW; ++t)
      Bt[t] = __builtin_bit_cast(bf16x8,
                                 *reinterpret_cast<const u16x8*>(W + (int64_t)(group_of(t) * 16 + m) * Kp + k0));
    int hoff = -1;  // LDS offset of this lane's tap relative to its voxel, or -1 (zero A)
    if (k0 < K) {
      const int tap = k0 / CIN, ci = k0 - tap * CIN;
      const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;  // offsets + 1
      hoff = ((dx * HY + dy) * HZ + dz) * CIN + ci;
    }
#pragma unroll
    for (int tl = 0; tl < TW; ++tl) {
      const int off = tile_of(tl) * 16 + m;  // this lane's voxel in the block
      const int yy = off / ZL, z = off - yy * ZL;
      raw[tl] = hoff >= 0 ? *reinterpret_cast<const u16x8*>(halo + hoff + (yy * HZ + z) * CIN)
                          : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
  };
  auto mfmas = [&](const bf16x8 (&Bt)[GW], const u16x8 (&raw)[TW]) {
#pragma unroll
    for (int tl = 0; tl < TW; ++tl)
#pragma unroll
      for (int t = 0; t < GW; ++t)
        acc[tl][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Bt[t], __builtin_bit_cast(bf16x8, raw[tl]),
                                                            acc[tl][t], 0, 0, 0);
  };
  bf16x8 B0[GW], B1[GW];
  u16x8 R0[TW], R1[TW];
  load_ops(0, B0, R0);
  for (int ks = 0; ks < KS; ks += 2) {
    if (ks + 1 < KS) load_ops(ks + 1, B1, R1);
    mfmas(B0, R0);
    if (ks + 1 < KS) {
      if (ks + 2 < KS) load_ops(ks + 2, B0, R0);
      mfmas(B1, R1);
    }
  }
  const int64_t vblock = (int64_t)it * (YB * ZL);
#pragma unroll
  for (int t = 0; t < GW; ++t) {
    const int c0 = group_of(t) * 16 + kg / 2;  // this lane's four output channels
    f32x4 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = bias ? bias[c0 + j] : 0.0f;
#pragma unroll
    for (int tl = 0; tl < TW; ++tl) {
      const int64_t o = (vblock + tile_of(tl) * 16 + m) * COUT + c0;
      const f32x4 pv = acc[tl][t] + b;
      *reinterpret_cast<f32x4*>(out + o) = pv;
      if (act) {
        u16x4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = __builtin_bit_cast(unsigned short, f2bf(silu(pv[j])));
        *reinterpret_cast<u16x4*>(act + o) = a;
      }
    }
  }
  __syncthreads();
  }
}

// Split-K gather convolution for the coarse levels (few voxels, 64 or 128 input channels): there a block per
// 64 voxels leaves most of the GPU idle while each of its waves walks 27 Cin / 32 K steps, one dependent
// round of loads after another.  Here a block computes MT tiles of 16 voxels x NT * 16 output channels
// (channel group blockIdx.y); its four waves take every fourth K step (wave w: steps w, w + 4, ...), with
// the operands of the next two of its steps in flight during the current step's MFMAs, and the four partial
// sums are added in wave order through LDS.  The MT tiles share each B fragment (weights: with one tile per
// block the weights are fetched again for every 16 voxels).  Operands as in k_conv_fwd; every output element
// gets the same MFMA sequence and sum whatever MT.
template <int NT, int MT>
__global__ __launch_bounds__(256) void k_conv_fwd_sk(const __bf16* __restrict__ in, const __bf16* __restrict__ W,
                                                     const float* __restrict__ bias, float* __restrict__ out, Geo g,
                                                     int Cin, int Cout, int Kp) {
  constexpr int D = 3;  // K steps of this wave in registers: the current one and the next two
  __shared__ floatx4 red[3][MT][NT][64];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
  const int m = lane & 15, kg = (lane >> 4) * 8;
  const int64_t V = g.V();
  const int64_t v0 = (int64_t)blockIdx.x * (MT * 16);
  const int co0 = (int)blockIdx.y * NT * 16;
  int n[MT], x[MT], y[MT], z[MT];
  bool ok[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int64_t va = v0 + t * 16 + m;
    ok[t] = va < V;
    n[t] = x[t] = y[t] = z[t] = 0;
    if (ok[t]) decompose(g, va, n[t], x[t], y[t], z[t]);
  }
  const int K = 27 * Cin, KS = Kp / 32;
  auto load_a = [&](int ks, int t) {
    const int k0 = ks * 32 + kg;
    u16x8 raw = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    if (ok[t] && k0 < K) {
      const int tap = k0 / Cin, ci = k0 - tap * Cin;
      const int64_t nb = tap_voxel(g, n[t], x[t], y[t], z[t], tap);
      if (nb >= 0) raw = *reinterpret_cast<const u16x8*>(in + nb * Cin + ci);
    }
    return raw;
  };
  const __bf16* wrow = W + (int64_t)(co0 + m) * Kp + kg;
  floatx4 acc[MT][NT];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[t][u] = floatx4{0, 0, 0, 0};
  const int cnt = KS > wave ? (KS - wave + 3) / 4 : 0;  // this wave's K steps: wave + 4 i, i < cnt
  u16x8 a[D][MT], b[D][NT];
  auto load = [&](int i, u16x8 (&as)[MT], u16x8 (&bs)[NT]) {
    const int ks = wave + 4 * i;
#pragma unroll
    for (int t = 0; t < MT; ++t) as[t] = load_a(ks, t);
#pragma unroll
    for (int u = 0; u < NT; ++u) bs[u] = *reinterpret_cast<const u16x8*>(wrow + (int64_t)u * 16 * Kp + ks * 32);
  };
#pragma unroll
  for (int d = 0; d < D - 1; ++d)
    if (d < cnt) load(d, a[d], b[d]);
  for (int i = 0; i < cnt; i += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      if (i + s < cnt) {
        if (i + s + D - 1 < cnt) load(i + s + D - 1, a[(s + D - 1) % D], b[(s + D - 1) % D]);
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
          for (int u = 0; u < NT; ++u)
            acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, b[s][u]),
                                                                __builtin_bit_cast(bf16x8, a[s][t]), acc[t][u], 0, 0, 0);
      }
    }
  }
  if (wave > 0) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int u = 0; u < NT; ++u) red[wave - 1][t][u][lane] = acc[t][u];
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      floatx4 sm = acc[t][u];
      sm += red[0][t][u][lane];
      sm += red[1][t][u][lane];
      sm += red[2][t][u][lane];
      const int c0 = co0 + u * 16 + kg / 2;  // this lane's four channels of voxel v
      const int64_t v = v0 + t * 16 + m;
#pragma unroll
      for (int j = 0; j < 4; ++j) sm[j] += bias ? bias[c0 + j] : 0.0f;
      if (v < V) *reinterpret_cast<f32x4*>(out + v * Cout + c0) = sm;
    }
}

// W'[ci][tap' * Cout + co] = W[co][(26 - tap') * Cin + ci], K' padded with zeros.
__global__ void k_wflip(const __bf16* __restrict__ W, __bf16* __restrict__ Wt, int Cin, int Cout, int Kp,
                        int Kpt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)Cin * Kpt) return;
  const int ci = (int)(i / Kpt), k = (int)(i - (int64_t)ci * Kpt);
  __bf16 v = f2bf(0.0f);
  if (k < 27 * Cout) {
    const int tp = k / Cout, co = k - tp * Cout;
    v = W[(int64_t)co * Kp + (26 - tp) * Cin + ci];
  }
  Wt[i] = v;
}

// ----------------------------------------------------------- weight grad ---
// Block: one tap, one 16-wide Cout tile, all Cin tiles, a chunk of voxels.
// Per 32-voxel step the block stages g (32 x 16) and the shifted input
// (32 x Cin) in LDS transposed (channel-major), then each wave runs MFMAs for
// its share of the Cin tiles.  Partials: part[chunk][co][tap*Cin + ci].
template <int CIT>
__global__ __launch_bounds__(256) void k_conv_wgrad(const __bf16* __restrict__ in, const __bf16* __restrict__ gout,
                                                    float* __restrict__ part, Geo g, int Cout, int vchunk,
                                                    int Kp) {
  constexpr int Cin = CIT * 16;
  __shared__ __attribute__((aligned(16))) __bf16 sg[16][40];       // [co][v], padded rows
  __shared__ __attribute__((aligned(16))) __bf16 si[Cin][40];      // [ci][v]
  const int tap = blockIdx.x;
  const int cot = blockIdx.y;
  const int chunk = blockIdx.z;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t V = g.V();
  const int64_t vb = (int64_t)chunk * vchunk;
  const int64_t ve = vb + vchunk < V ? vb + vchunk : V;
  floatx4 acc[(CIT + 3) / 4];
#pragma unroll
  for (int t = 0; t < (CIT + 3) / 4; ++t) acc[t] = floatx4{0, 0, 0, 0};
  for (int64_t v0 = vb; v0 < ve; v0 += 32) {
    // stage: 32 voxels x (16 g channels + Cin input channels)
    for (int e = threadIdx.x; e < 32 * (16 + Cin); e += 256) {
      const int vi = e / (16 + Cin), c = e - vi * (16 + Cin);
      const int64_t v = v0 + vi;
      __bf16 val = f2bf(0.0f);
      if (v < ve) {
        if (c < 16) {
          val = gout[v * Cout + cot * 16 + c];
        } else {
          int n, x, y, z;
          decompose(g, v, n, x, y, z);
          const int64_t nb = tap_voxel(g, n, x, y, z, tap);
          if (nb >= 0) val = in[nb * Cin + (c - 16)];
        }
      }
      if (c < 16)
        sg[c][vi] = val;
      else
        si[c - 16][vi] = val;
    }
    __syncthreads();
    const int kg = (lane >> 4) * 8;
    const u16x8 ra = *reinterpret_cast<const u16x8*>(&sg[lane & 15][kg]);
    const bf16x8 a = __builtin_bit_cast(bf16x8, ra);
#pragma unroll
    for (int t = 0; t < (CIT + 3) / 4; ++t) {
      const int cit = wave + 4 * t;
      if (cit < CIT) {
        const u16x8 rb = *reinterpret_cast<const u16x8*>(&si[cit * 16 + (lane & 15)][kg]);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, rb), acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D[m = co][n = ci]: lane holds co = 4*(l/16)+j, ci = l%16
#pragma unroll
  for (int t = 0; t < (CIT + 3) / 4; ++t) {
    const int cit = wave + 4 * t;
    if (cit >= CIT) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = cot * 16 + (lane >> 4) * 4 + j;
      const int ci = cit * 16 + (lane & 15);
      part[((int64_t)chunk * Cout + co) * Kp + tap * Cin + ci] = acc[t][j];
    }
  }
}

__global__ void k_wgrad_reduce(const float* __restrict__ part, float* __restrict__ dW, int nchunk, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * n + i];
  dW[i] = s;
}

// ------------------------------------------- weight grad, z-column tiles ---
// For Z in {4, 8, 16, 32}: a block owns tiles of TX x TY (x, y) columns over the
// full z range and the 9 taps that share one dx (grid.y = dx + 1; grid.z splits
// wide layers' output channels), and loops over tiles (grid.x blocks),
// accumulating in registers.  Per tile it stages in LDS the upstream gradient
// of the tile's columns (A operand, co x voxel) and the input of the TX x
// (TY + 2) source columns at x + dx (B operand, voxel x ci), with the next
// tile's loads in flight during the current tile's MFMAs.  A K-step is 32
// consecutive voxels of the tile in (column, z) order; lane group l / 16 takes
// 8 of them.  Partials: part[dx + 1][block][co][p * Cin + ci], p = (dy + 1) * 3
// + dz + 1, summed over the blocks in a fixed order by k_wgrad_tile_reduce.
constexpr int kWTX = 2, kWTY = 4;

// Weight gradient with both operands kept in LDS in their natural channels-last
// layout and read transposed by ds_read_b64_tr_b16 (gfx950).  Within each
// 16-lane group that instruction returns, to lane i, element i % 4 of the 8-byte
// rows addressed by lanes i / 4, i / 4 + 4, i / 4 + 8 and i / 4 + 12 (checked
// on the device).  With lane r of the group addressing voxel k0 + r / 4 at
// channel 4 (r % 4), lane i receives channel i of voxels k0 .. k0 + 3: two such
// reads give the 8 consecutive voxels of channel i that an MFMA operand lane
// holds (A: rows = output channels, B: columns = input channels, K = voxels).
// Staging is then a plain 16-byte copy per 8 channels of a voxel, the input is
// stored once (the tap's z shift is an address offset into its column, which
// carries zero rows at z = -1 and z = ZZ), and rows of 32 or more channels are
// padded by 16 channels so that the four rows of a read fall in distinct banks.
// Wave w takes taps w and w + 4 of the block's nine over every K-step; the
// ninth tap's K-steps go to wave ks % 4 when there are a multiple of four of
// them (their four sums folded in wave order at the end), else all to wave 0.
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

__device__ __forceinline__ s16x4 lds_read_tr16(const __bf16* p) {
  return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)(p));
}

template <int CIT, int COT, int ZZ>
__global__ __launch_bounds__(256) void k_conv_wgrad_tr(const __bf16* __restrict__ in, const __bf16* __restrict__ gout,
                                                       float* __restrict__ part, Geo g, int ntx, int nty, int CoutT) {
  constexpr int Cin = CIT * 16, Cout = COT * 16;
  constexpr int OC = kWTX * kWTY;         // output columns per tile
  constexpr int SY = kWTY + 2;            // source columns per x row (y halo)
  constexpr int SC = kWTX * SY;           // source columns per tile
  constexpr int ZH = ZZ + 2;              // rows per source column: z = -1 .. ZZ
  constexpr int KS = OC * ZZ / 32;        // K-steps per tile
  constexpr int GR = Cout + (Cout >= 32 ? 16 : 0), XR = Cin + (Cin >= 32 ? 16 : 0);  // row lengths (elements)
  static_assert(ZZ % 4 == 0 && (OC * ZZ) % 32 == 0, "tile shape");
  constexpr bool kBal = KS % 4 == 0;
  const int co0 = (int)blockIdx.z * Cout;
  __shared__ __attribute__((aligned(16))) __bf16 sg[OC * ZZ * GR];
  __shared__ __attribute__((aligned(16))) __bf16 sx[SC * ZH * XR];
  __shared__ __attribute__((aligned(16))) floatx4 red[kBal ? 4 * COT * CIT * 64 : 1];
  const int dx = (int)blockIdx.y - 1;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
  const int64_t ntiles = (int64_t)g.N * ntx * nty;
  floatx4 acc[3][COT][CIT];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int u = 0; u < COT; ++u)
#pragma unroll
      for (int w = 0; w < CIT; ++w) acc[a][u][w] = floatx4{0, 0, 0, 0};

  // Staging: 16-byte chunks (8 channels of one voxel), loaded to registers (all in flight together, the
  // next tile's during this tile's MFMAs), then stored to LDS.
  constexpr int NEG = OC * ZZ * (Cout / 8), NEX = SC * ZH * (Cin / 8);
  constexpr int NG = (NEG + 255) / 256, NX = (NEX + 255) / 256;
  u16x8 rg[NG], rx[NX];
  auto fetch = [&](int64_t tile) {
    const int n = (int)(tile / ((int64_t)ntx * nty));
    const int r = (int)(tile - (int64_t)n * ntx * nty);
    const int x0 = (r / nty) * kWTX, y0 = (r % nty) * kWTY;
#pragma unroll
    for (int q = 0; q < NG; ++q) {  // upstream gradient of the tile's columns
      const int e = threadIdx.x + q * 256;
      const int cg = e % (Cout / 8), v = e / (Cout / 8);
      const int z = v % ZZ, oc = v / ZZ;
      const int x = x0 + oc / kWTY, y = y0 + oc % kWTY;
      rg[q] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
      if (e < NEG && x < g.X && y < g.Y)
        rg[q] = *reinterpret_cast<const u16x8*>(gout + ((((int64_t)n * g.X + x) * g.Y + y) * ZZ + z) * CoutT + co0 +
                                                cg * 8);
    }
#pragma unroll
    for (int q = 0; q < NX; ++q) {  // input of the source columns at x + dx, with the zero z rows
      const int e = threadIdx.x + q * 256;
      const int cg = e % (Cin / 8), v = e / (Cin / 8);
      const int zh = v % ZH, sc = v / ZH;
      const int x = x0 + sc / SY + dx, y = y0 + sc % SY - 1, z = zh - 1;
      rx[q] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
      if (e < NEX && x >= 0 && x < g.X && y >= 0 && y < g.Y && z >= 0 && z < ZZ)
        rx[q] = *reinterpret_cast<const u16x8*>(in + ((((int64_t)n * g.X + x) * g.Y + y) * ZZ + z) * Cin + cg * 8);
    }
  };
  const int a = lane >> 4, r4 = lane & 15;
  const int rrow = r4 >> 2, rcol = (r4 & 3) * 4;  // this lane's row and channel offset in a transposed read
  // One half (4 voxels) of a K-step's operands: the upstream gradient A[u] and the input B[w] of tap p.
  auto put = [](bf16x8& d, int h, s16x4 v) {
    u16x8 t = __builtin_bit_cast(u16x8, d);
#pragma unroll
    for (int j = 0; j < 4; ++j) t[h * 4 + j] = (unsigned short)v[j];
    d = __builtin_bit_cast(bf16x8, t);
  };
  auto voxel = [&](int ks, int h, int& oc, int& z) {  // the voxel this lane addresses, (column, z) order
    const int L = ks * 32 + a * 8 + h * 4 + rrow;
    oc = L / ZZ;
    z = L % ZZ;
  };
  auto load_a = [&](int ks, bf16x8 (&A)[COT]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int oc, z;
      voxel(ks, h, oc, z);
#pragma unroll
      for (int u = 0; u < COT; ++u) put(A[u], h, lds_read_tr16(&sg[(oc * ZZ + z) * GR + u * 16 + rcol]));
    }
  };
  auto load_b = [&](int ks, int p, bf16x8 (&B)[CIT]) {
    const int dy = p / 3 - 1, dzi = p % 3;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int oc, z;
      voxel(ks, h, oc, z);
      const int row = ((oc / kWTY) * SY + oc % kWTY + dy + 1) * ZH + z + dzi;
#pragma unroll
      for (int w = 0; w < CIT; ++w) put(B[w], h, lds_read_tr16(&sx[row * XR + w * 16 + rcol]));
    }
  };
  auto mfma = [&](floatx4 (&c)[COT][CIT], const bf16x8 (&A)[COT], const bf16x8 (&B)[CIT]) {
#pragma unroll
    for (int w = 0; w < CIT; ++w)
#pragma unroll
      for (int u = 0; u < COT; ++u) c[u][w] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[u], B[w], c[u][w], 0, 0, 0);
  };
  if (blockIdx.x < ntiles) fetch(blockIdx.x);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int e = threadIdx.x + q * 256;
      if (e < NEG) *reinterpret_cast<u16x8*>(&sg[(e / (Cout / 8)) * GR + (e % (Cout / 8)) * 8]) = rg[q];
    }
#pragma unroll
    for (int q = 0; q < NX; ++q) {
      const int e = threadIdx.x + q * 256;
      if (e < NEX) *reinterpret_cast<u16x8*>(&sx[(e / (Cin / 8)) * XR + (e % (Cin / 8)) * 8]) = rx[q];
    }
    if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);  // in flight during this tile's MFMAs
    __syncthreads();
    // Taps wave and wave + 4 over every K-step, the next step's operands read during this step's MFMAs.
    {
      bf16x8 A[2][COT], B0[2][CIT], B1[2][CIT];
      load_a(0, A[0]);
      load_b(0, wave, B0[0]);
      load_b(0, wave + 4, B1[0]);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int cur = ks & 1;
        if (ks + 1 < KS) {
          load_a(ks + 1, A[cur ^ 1]);
          load_b(ks + 1, wave, B0[cur ^ 1]);
          load_b(ks + 1, wave + 4, B1[cur ^ 1]);
        }
        mfma(acc[0], A[cur], B0[cur]);
        mfma(acc[1], A[cur], B1[cur]);
      }
    }
    // The ninth tap: K-steps dealt out over the waves (balanced), or all of them on wave 0.
    if (kBal || wave == 0) {
      constexpr int NK = kBal ? KS / 4 : KS;
      bf16x8 A[2][COT], B[2][CIT];
      const int k0 = kBal ? wave : 0, kstep = kBal ? 4 : 1;
      load_a(k0, A[0]);
      load_b(k0, 8, B[0]);
#pragma unroll
      for (int j = 0; j < NK; ++j) {
        const int cur = j & 1;
        if (j + 1 < NK) {
          load_a(k0 + (j + 1) * kstep, A[cur ^ 1]);
          load_b(k0 + (j + 1) * kstep, 8, B[cur ^ 1]);
        }
        mfma(acc[2], A[cur], B[cur]);
      }
    }
    __syncthreads();
  }
  const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  if constexpr (kBal) {  // the ninth tap's four per-wave sums, folded in wave order
#pragma unroll
    for (int u = 0; u < COT; ++u)
#pragma unroll
      for (int w = 0; w < CIT; ++w) red[((wave * COT + u) * CIT + w) * 64 + lane] = acc[2][u][w];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int u = 0; u < COT; ++u)
#pragma unroll
        for (int w = 0; w < CIT; ++w) {
          floatx4 s = acc[2][u][w];
#pragma unroll
          for (int wv = 1; wv < 4; ++wv) s += red[((wv * COT + u) * CIT + w) * 64 + lane];
          acc[2][u][w] = s;
        }
    }
  }
#pragma unroll
  for (int pi = 0; pi < 3; ++pi) {
    const int p = wave + 4 * pi;
    if (p >= 9) continue;
#pragma unroll
    for (int u = 0; u < COT; ++u)
#pragma unroll
      for (int w = 0; w < CIT; ++w)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int co = u * 16 + (lane >> 4) * 4 + j, ci = w * 16 + (lane & 15);
          part[(blk * CoutT + co0 + co) * (9 * Cin) + p * Cin + ci] = acc[pi][u][w][j];
        }
  }
}

// dW[co][tap * Cin + ci] = sum over the blocks of the tap's dx group (fixed
// order) for tap * Cin + ci < 27 Cin; zero in the K padding.
__global__ void k_wgrad_tile_reduce(const float* __restrict__ part, float* __restrict__ dW, int nblk, int Cout,
                                    int Cin, int Kp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)Cout * Kp) return;
  const int co = (int)(i / Kp), k = (int)(i - (int64_t)co * Kp);
  float s = 0.0f;
  if (k < 27 * Cin) {
    const int grp = k / (9 * Cin), loc = k - grp * 9 * Cin;
    const int64_t stride = (int64_t)Cout * (9 * Cin);
    const float* p = part + ((int64_t)grp * nblk * Cout + co) * (9 * Cin) + loc;
    // block partials summed in block order; 16 loads in flight ahead of the adds
    int b = 0;
    for (; b + 16 <= nblk; b += 16) {
      float w[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) w[q] = p[(int64_t)(b + q) * stride];
#pragma unroll
      for (int q = 0; q < 16; ++q) s += w[q];
    }
    for (; b < nblk; ++b) s += p[(int64_t)b * stride];
  }
  dW[i] = s;
}

// db[co] = sum_v g[v][co], two passes with fixed order: per-block partials
// over contiguous voxel ranges, each thread summing 8 channels loaded as one
// 16-byte vector (thread t: channel group t % (C/8), voxels t / (C/8) + k R),
// then one thread per channel sums the blocks.
__global__ __launch_bounds__(256) void k_bias_part(const __bf16* __restrict__ g, float* __restrict__ part,
                                                   int64_t V, int C, int64_t vper) {
  __shared__ float red[256][9];  // 8 channel sums per thread, padded row
  const int64_t v0 = (int64_t)blockIdx.x * vper;
  const int64_t v1 = v0 + vper < V ? v0 + vper : V;
  const int G = C / 8;    // channel groups (G divides 256, checked by the host)
  const int R = 256 / G;  // voxels per sweep
  const int cg = threadIdx.x % G, r = threadIdx.x / G;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t v = v0 + r; v < v1; v += R) {
    const u16x8 raw = *reinterpret_cast<const u16x8*>(g + v * C + cg * 8);
    const bf16x8 h = __builtin_bit_cast(bf16x8, raw);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += bf2f(h[j]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) red[threadIdx.x][j] = s[j];
  __syncthreads();
  if (threadIdx.x < C) {
    const int c = threadIdx.x, cgc = c / 8, j = c % 8;
    float a = 0.0f;
    for (int k = 0; k < R; ++k) a += red[k * G + cgc][j];
    part[(int64_t)blockIdx.x * C + c] = a;
  }
}

__global__ void k_bias_final(const float* __restrict__ part, float* __restrict__ db, int nb, int C) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float a = 0.0f;
  int b = 0;  // block partials summed in block order; 16 loads in flight ahead of the adds
  for (; b + 16 <= nb; b += 16) {
    float w[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) w[q] = part[(int64_t)(b + q) * C + c];
#pragma unroll
    for (int q = 0; q < 16; ++q) a += w[q];
  }
  for (; b < nb; ++b) a += part[(int64_t)b * C + c];
  db[c] = a;
}

// ------------------------------------------------------------ elementwise ---



__global__ void k_silu_fwd(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(silu(x[i]));
}
// gx = gy * silu'(x); written as bf16 (operand of the next MFMA) and optionally fp32.
__global__ void k_silu_bwd(const float* __restrict__ x, const float* __restrict__ gy, __bf16* __restrict__ gxb,
                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gxb[i] = f2bf(gy[i] * dsilu(x[i]));
}
// gx = gy * silu'(x) in FP32 (for a gradient that still goes through a FiLM layer before the next MFMA).
__global__ void k_silu_bwd_f32(const float* __restrict__ x, const float* __restrict__ gy, float* __restrict__ gx,
                               int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gx[i] = gy[i] * dsilu(x[i]);
}
__global__ void k_to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}
__global__ void k_add_f32(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ y,
                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] + b[i];
}
// y_bf16 = bf16(a_bf16 + b_bf16) computed in fp32 (additive skip of the decoder)
__global__ void k_add_bf16(const __bf16* __restrict__ a, const __bf16* __restrict__ b, __bf16* __restrict__ y,
                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// 2x2x2 average pooling, fine (N,X,Y,Z,C) -> coarse (N,X/2,Y/2,Z/2,C).
__global__ void k_pool_fwd(const __bf16* __restrict__ in, __bf16* __restrict__ out, Geo gc, int C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nco = gc.V() * C;
  if (i >= nco) return;
  const int c = (int)(i % C);
  int n, x, y, z;
  decompose(gc, i / C, n, x, y, z);
  const int FX = gc.X * 2, FY = gc.Y * 2, FZ = gc.Z * 2;
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) {
    const int xx = 2 * x + (d >> 2), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d & 1);
    s += bf2f(in[((((int64_t)n * FX + xx) * FY + yy) * FZ + zz) * C + c]);
  }
  out[i] = f2bf(s * 0.125f);
}
// grad of pooling: g_fine = g_coarse / 8 (fp32 in, fp32 out)
__global__ void k_pool_bwd(const float* __restrict__ gc_, float* __restrict__ gf, Geo gf_geo, int C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gf_geo.V() * C) return;
  const int c = (int)(i % C);
  int n, x, y, z;
  decompose(gf_geo, i / C, n, x, y, z);
  const int CX = gf_geo.X / 2, CY = gf_geo.Y / 2, CZ = gf_geo.Z / 2;
  gf[i] = 0.125f * gc_[((((int64_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2) * C + c];
}
// nearest x2 upsampling coarse -> fine (bf16)
__global__ void k_up_fwd(const __bf16* __restrict__ in, __bf16* __restrict__ out, Geo gf_geo, int C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gf_geo.V() * C) return;
  const int c = (int)(i % C);
  int n, x, y, z;
  decompose(gf_geo, i / C, n, x, y, z);
  const int CX = gf_geo.X / 2, CY = gf_geo.Y / 2, CZ = gf_geo.Z / 2;
  out[i] = in[((((int64_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2) * C + c];
}
// ------------------------------------------------- vectorised elementwise ---
// The same arithmetic as the scalar kernels above, element for element (so the
// same bits), with 16-byte accesses and 32-bit index math: one thread per 4
// fp32 or 8 bf16 values, or per voxel and 8 (bf16) / 4 (fp32) channels.  The
// host uses them when the element count fits 32-bit indices and the pointers
// are 16-byte aligned; the scalar kernels remain for everything else.

__device__ __forceinline__ void dec32(uint32_t v, int X, int Y, int Z, int& n, int& x, int& y, int& z) {
  z = (int)(v % (uint32_t)Z);
  uint32_t r = v / (uint32_t)Z;
  y = (int)(r % (uint32_t)Y);
  r /= (uint32_t)Y;
  x = (int)(r % (uint32_t)X);
  n = (int)(r / (uint32_t)X);
}

// n4 = n / 4 full groups; the tail (n % 4 elements) is done by the first threads.
__global__ void k_silu_fwd4(const float* __restrict__ x, __bf16* __restrict__ y, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n4 = n >> 2;
  if (i < n4) {
    const f32x4 v = reinterpret_cast<const f32x4*>(x)[i];
    u16x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = __builtin_bit_cast(unsigned short, f2bf(silu(v[k])));
    reinterpret_cast<u16x4*>(y)[i] = o;
  }
  const uint32_t t = 4 * n4 + i;
  if (i < (n & 3)) y[t] = f2bf(silu(x[t]));
}
__global__ void k_silu_bwd4(const float* __restrict__ x, const float* __restrict__ gy, __bf16* __restrict__ gxb,
                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n4 = n >> 2;
  if (i < n4) {
    const f32x4 v = reinterpret_cast<const f32x4*>(x)[i];
    const f32x4 g = reinterpret_cast<const f32x4*>(gy)[i];
    u16x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = __builtin_bit_cast(unsigned short, f2bf(g[k] * dsilu(v[k])));
    reinterpret_cast<u16x4*>(gxb)[i] = o;
  }
  const uint32_t t = 4 * n4 + i;
  if (i < (n & 3)) gxb[t] = f2bf(gy[t] * dsilu(x[t]));
}
__global__ void k_add_f32_4(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ y,
                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n4 = n >> 2;
  if (i < n4) reinterpret_cast<f32x4*>(y)[i] = reinterpret_cast<const f32x4*>(a)[i] + reinterpret_cast<const f32x4*>(b)[i];
  const uint32_t t = 4 * n4 + i;
  if (i < (n & 3)) y[t] = a[t] + b[t];
}
__global__ void k_add_bf16_8(const __bf16* __restrict__ a, const __bf16* __restrict__ b, __bf16* __restrict__ y,
                             uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n8 = n >> 3;
  if (i < n8) {
    const bf16x8 va = __builtin_bit_cast(bf16x8, reinterpret_cast<const u16x8*>(a)[i]);
    const bf16x8 vb = __builtin_bit_cast(bf16x8, reinterpret_cast<const u16x8*>(b)[i]);
    u16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = __builtin_bit_cast(unsigned short, f2bf(bf2f(va[k]) + bf2f(vb[k])));
    reinterpret_cast<u16x8*>(y)[i] = o;
  }
  const uint32_t t = 8 * n8 + i;
  if (i < (n & 7)) y[t] = f2bf(bf2f(a[t]) + bf2f(b[t]));
}
// nearest x2 upsampling, one thread per fine voxel and 8 channels (C % 8 == 0)
__global__ void k_up_fwd8(const __bf16* __restrict__ in, __bf16* __restrict__ out, Geo gf, int C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C8 = (uint32_t)C >> 3;
  if (i >= (uint32_t)gf.V() * C8) return;
  const uint32_t v = i / C8, c8 = i - v * C8;
  int n, x, y, z;
  dec32(v, gf.X, gf.Y, gf.Z, n, x, y, z);
  const int CX = gf.X / 2, CY = gf.Y / 2, CZ = gf.Z / 2;
  const uint32_t cv = (((uint32_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2;
  reinterpret_cast<u16x8*>(out)[i] = reinterpret_cast<const u16x8*>(in)[cv * C8 + c8];
}
// grad of upsampling, one thread per coarse voxel and 4 channels (C % 4 == 0): children summed in the
// scalar kernel's order
__global__ void k_up_bwd4(const float* __restrict__ gf, float* __restrict__ gc_, Geo gc, int C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C4 = (uint32_t)C >> 2;
  if (i >= (uint32_t)gc.V() * C4) return;
  const uint32_t v = i / C4, c4 = i - v * C4;
  int n, x, y, z;
  dec32(v, gc.X, gc.Y, gc.Z, n, x, y, z);
  const int FX = gc.X * 2, FY = gc.Y * 2, FZ = gc.Z * 2;
  f32x4 s = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int d = 0; d < 8; ++d) {
    const int xx = 2 * x + (d >> 2), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d & 1);
    s += reinterpret_cast<const f32x4*>(gf)[((((uint32_t)n * FX + xx) * FY + yy) * FZ + zz) * C4 + c4];
  }
  reinterpret_cast<f32x4*>(gc_)[i] = s;
}
// 2x2x2 average pooling, one thread per coarse voxel and 8 channels (C % 8 == 0)
__global__ void k_pool_fwd8(const __bf16* __restrict__ in, __bf16* __restrict__ out, Geo gc, int C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C8 = (uint32_t)C >> 3;
  if (i >= (uint32_t)gc.V() * C8) return;
  const uint32_t v = i / C8, c8 = i - v * C8;
  int n, x, y, z;
  dec32(v, gc.X, gc.Y, gc.Z, n, x, y, z);
  const int FX = gc.X * 2, FY = gc.Y * 2, FZ = gc.Z * 2;
  float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int d = 0; d < 8; ++d) {
    const int xx = 2 * x + (d >> 2), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d & 1);
    const bf16x8 w = __builtin_bit_cast(
        bf16x8, reinterpret_cast<const u16x8*>(in)[((((uint32_t)n * FX + xx) * FY + yy) * FZ + zz) * C8 + c8]);
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += bf2f(w[k]);
  }
  u16x8 o;
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = __builtin_bit_cast(unsigned short, f2bf(s[k] * 0.125f));
  reinterpret_cast<u16x8*>(out)[i] = o;
}
// grad of pooling, one thread per fine voxel and 4 channels (C % 4 == 0)
__global__ void k_pool_bwd4(const float* __restrict__ gc_, float* __restrict__ gf, Geo gfg, int C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C4 = (uint32_t)C >> 2;
  if (i >= (uint32_t)gfg.V() * C4) return;
  const uint32_t v = i / C4, c4 = i - v * C4;
  int n, x, y, z;
  dec32(v, gfg.X, gfg.Y, gfg.Z, n, x, y, z);
  const int CX = gfg.X / 2, CY = gfg.Y / 2, CZ = gfg.Z / 2;
  const f32x4 g = reinterpret_cast<const f32x4*>(gc_)[((((uint32_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2) * C4 +
                                                      c4];
  reinterpret_cast<f32x4*>(gf)[i] = 0.125f * g;
}

// acc + the pooling gradient (a skip connection's gradient plus the one through the pooling), in one
// pass: the bits of k_add_f32 of acc and k_pool_bwd's output, without the fine temporary.
__global__ void k_pool_bwd_add(const float* __restrict__ gc_, const float* __restrict__ acc, float* __restrict__ out,
                               Geo gf_geo, int C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gf_geo.V() * C) return;
  const int c = (int)(i % C);
  int n, x, y, z;
  decompose(gf_geo, i / C, n, x, y, z);
  const int CX = gf_geo.X / 2, CY = gf_geo.Y / 2, CZ = gf_geo.Z / 2;
  out[i] = acc[i] + 0.125f * gc_[((((int64_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2) * C + c];
}
__global__ void k_pool_bwd_add4(const float* __restrict__ gc_, const float* __restrict__ acc, float* __restrict__ out,
                                Geo gfg, int C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C4 = (uint32_t)C >> 2;
  if (i >= (uint32_t)gfg.V() * C4) return;
  const uint32_t v = i / C4, c4 = i - v * C4;
  int n, x, y, z;
  dec32(v, gfg.X, gfg.Y, gfg.Z, n, x, y, z);
  const int CX = gfg.X / 2, CY = gfg.Y / 2, CZ = gfg.Z / 2;
  const f32x4 g = reinterpret_cast<const f32x4*>(gc_)[((((uint32_t)n * CX + x / 2) * CY + y / 2) * CZ + z / 2) * C4 +
                                                      c4];
  reinterpret_cast<f32x4*>(out)[i] = reinterpret_cast<const f32x4*>(acc)[i] + 0.125f * g;
}

// grad of upsampling: g_coarse = sum of the 8 fine children (fp32)
__global__ void k_up_bwd(const float* __restrict__ gf, float* __restrict__ gc_, Geo gc, int C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gc.V() * C) return;
  const int c = (int)(i % C);
  int n, x, y, z;
  decompose(gc, i / C, n, x, y, z);
  const int FX = gc.X * 2, FY = gc.Y * 2, FZ = gc.Z * 2;
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) {
    const int xx = 2 * x + (d >> 2), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d & 1);
    s += gf[((((int64_t)n * FX + xx) * FY + yy) * FZ + zz) * C + c];
  }
  gc_[i] = s;
}

// Weighted MSE over the first `nout` of C channels:
// loss partials (one per block) and grad = 2 w (pred - target) / norm.
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      const float* __restrict__ wmask, float* __restrict__ grad, float* __restrict__ part, int64_t V,
                      int C, int nout, float inv_norm) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < V * C) {
    const int64_t v = i / C;
    const int c = (int)(i - v * C);
    float gr = 0.0f;
    if (c < nout) {
      const float w = wmask[v * nout + c];
      const float d = pred[i] - target[v * nout + c];
      l = w * d * d * inv_norm;
      gr = 2.0f * w * d * inv_norm;
    }
    grad[i] = gr;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// k_mse with the gradient written in BF16 (the operand of the head's backward, so no FP32 copy is written and
// converted) and the normaliser taken from device memory: inv_norm = 1 / max(*wsum, 1).
__global__ void k_mse_bf16(const float* __restrict__ pred, const float* __restrict__ target,
                           const float* __restrict__ wmask, const float* __restrict__ wsum, __bf16* __restrict__ grad,
                           float* __restrict__ part, int64_t V, int C, int nout) {
  __shared__ float red[256];
  const float inv_norm = 1.0f / fmaxf(*wsum, 1.0f);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < V * C) {
    const int64_t v = i / C;
    const int c = (int)(i - v * C);
    float gr = 0.0f;
    if (c < nout) {
      const float w = wmask[v * nout + c];
      const float d = pred[i] - target[v * nout + c];
      l = w * d * d * inv_norm;
      gr = 2.0f * w * d * inv_norm;
    }
    grad[i] = f2bf(gr);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// k_mse_bf16 for C = 16 with one voxel per thread: the prediction row in four 16-byte loads, the gradient
// row in two 16-byte stores.  The gradient has the same bits; the loss partials sum a voxel's channels first.
// Partials nblk(V) .. npart - 1 (the scalar kernel's count, which the caller allocates) are set to zero.
__global__ void k_mse16_bf16(const float* __restrict__ pred, const float* __restrict__ target,
                             const float* __restrict__ wmask, const float* __restrict__ wsum,
                             __bf16* __restrict__ grad, float* __restrict__ part, int64_t V, int nout,
                             int64_t npart) {
  __shared__ float red[256];
  const float inv_norm = 1.0f / fmaxf(*wsum, 1.0f);
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t k = gridDim.x + v; k < npart; k += (int64_t)gridDim.x * blockDim.x) part[k] = 0.0f;
  float l = 0.0f;
  if (v < V) {
    float pr[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f32x4 a = reinterpret_cast<const f32x4*>(pred + v * 16)[q];
#pragma unroll
      for (int k = 0; k < 4; ++k) pr[q * 4 + k] = a[k];
    }
    u16x8 o[2];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      float gr = 0.0f;
      if (c < nout) {
        const float w = wmask[v * nout + c];
        const float d = pr[c] - target[v * nout + c];
        l += w * d * d * inv_norm;
        gr = 2.0f * w * d * inv_norm;
      }
      o[c / 8][c % 8] = __builtin_bit_cast(unsigned short, f2bf(gr));
    }
    reinterpret_cast<u16x8*>(grad + v * 16)[0] = o[0];
    reinterpret_cast<u16x8*>(grad + v * 16)[1] = o[1];
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// Blocks of a persistent launch over `items` work items: per_cu blocks on every CU of the current device
// (the number its LDS allows), at most one per item.
unsigned persistent_grid(int64_t items, int per_cu) {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cus[dev] <= 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  const int64_t g = (int64_t)cus[dev] * per_cu;
  return (unsigned)(items < g ? items : g);
}

// The vectorised elementwise kernels index elements with 32 bits and access 16 bytes at a time.
inline bool vec_ok(int64_t nelem, std::initializer_list<const void*> ptrs) {
  if (nelem <= 0 || nelem >= (int64_t(1) << 31)) return false;
  for (const void* p : ptrs)
    if (reinterpret_cast<uintptr_t>(p) % 16) return false;
  return true;
}

}  // namespace

// ================================================================= C ABI ===

extern "C" {

}  // extern "C"
namespace {
// The convolution entry points' body; act (optional) receives bf16(SiLU(out)) from the LDS-halo kernels'
// epilogue, or from a separate pass after the other kernels.
int conv_fwd(const void* in, const void* W, const float* bias, float* out, __bf16* act, int N, int X, int Y, int Z,
             int Cin, int Cout, int Kp, hipStream_t s);
}  // namespace
extern "C" {
int xnn_conv_fwd(const void* in, const void* W, const float* bias, float* out, int N, int X, int Y, int Z, int Cin,
                 int Cout, int Kp, void* stream) {
  return conv_fwd(in, W, bias, out, nullptr, N, X, Y, Z, Cin, Cout, Kp, (hipStream_t)stream);
}
int xnn_conv_fwd_silu(const void* in, const void* W, const float* bias, float* out, void* act, int N, int X, int Y,
                      int Z, int Cin, int Cout, int Kp, void* stream) {
  if (!act) return 1;
  return conv_fwd(in, W, bias, out, (__bf16*)act, N, X, Y, Z, Cin, Cout, Kp, (hipStream_t)stream);
}
}  // extern "C"
namespace {
int conv_fwd_kernels(const void* in, const void* W, const float* bias, float* out, __bf16* act, bool& fused, int N,
                     int X, int Y, int Z, int Cin, int Cout, int Kp, hipStream_t s) {
  fused = true;
  if (Cin % 16 || Cout % 16 || Cout > 128 || Kp % 32 || Kp < 27 * Cin) return 1;
  Geo g{X, Y, Z, N};
  const int64_t V = g.V();
  const unsigned grid = (unsigned)((V + 63) / 64);
  const __bf16* i = (const __bf16*)in;
  const __bf16* w = (const __bf16*)W;
  if (Cout == 16 && Cin == 16 && Kp == (27 * Cin + 31) / 32 * 32 && Z == 32 && Y % 16 == 0) {
    // one block per (n, x) and 16 y columns: V / 512 blocks
    hipLaunchKernelGGL((k_conv_fwd16_lds<16, 16, true>), dim3(persistent_grid(V / 512, 2)), dim3(256), 0, s, i, w,
                       bias, out, act, g, Kp);
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  if (Cout == 16 && Cin == 32 && Kp == (27 * Cin + 31) / 32 * 32 && Z == 32 && Y % 8 == 0) {
    // one block per (n, x) and 8 y columns: V / 256 blocks
    hipLaunchKernelGGL((k_conv_fwd16_lds<32, 8, false>), dim3((unsigned)(V / 256)), dim3(256), 0, s, i, w, bias,
                       out, act, g, Kp);
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  if (Kp == (27 * Cin + 31) / 32 * 32 && Y % 16 == 0) {
    // LDS-halo variants of k_conv_fwd<2> and k_conv_fwd16<32> for the half- and full-resolution layers
    if (Cout == 32 && Cin == 16 && Z == 32) {
      hipLaunchKernelGGL((k_conv_fwd_lds<16, 2, 32, 16, false>), dim3((unsigned)(V / 512)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Cout == 32 && Cin == 16 && Z == 16) {
      hipLaunchKernelGGL((k_conv_fwd_lds<16, 2, 16, 16, false>), dim3((unsigned)(V / 256)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Cout == 32 && Cin == 32 && Z == 16) {
      hipLaunchKernelGGL((k_conv_fwd_lds<32, 2, 16, 16, false>), dim3((unsigned)(V / 256)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Cout == 16 && Cin == 32 && Z == 16) {
      hipLaunchKernelGGL((k_conv_fwd_lds<32, 1, 16, 16, false>), dim3((unsigned)(V / 256)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Cout == 64 && Cin == 32 && Z == 16) {
      hipLaunchKernelGGL((k_conv_fwd_lds<32, 4, 16, 16, false>), dim3((unsigned)(V / 256)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Cout == 32 && Cin == 64 && Z == 16) {  // 4 y columns: the 64-channel halo fits the LDS limit
      hipLaunchKernelGGL((k_conv_fwd_lds<64, 2, 16, 4, false>), dim3((unsigned)(V / 64)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Z == 8 && Cin == 32 && Cout == 64) {
      hipLaunchKernelGGL((k_conv_fwd_lds<32, 4, 8, 16, false>), dim3((unsigned)(V / 128)), dim3(256), 0, s, i, w, bias,
                         out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (Z == 8 && Cin == 64 && (Cout == 32 || Cout == 64 || Cout == 128)) {  // 8 y columns per block
      const dim3 grid((unsigned)(V / 64));
      if (Cout == 32)
        hipLaunchKernelGGL((k_conv_fwd_lds<64, 2, 8, 8, false>), grid, dim3(256), 0, s, i, w, bias, out, act, g, Kp);
      else if (Cout == 64)
        hipLaunchKernelGGL((k_conv_fwd_lds<64, 4, 8, 8, false>), grid, dim3(256), 0, s, i, w, bias, out, act, g, Kp);
      else
        hipLaunchKernelGGL((k_conv_fwd_lds<64, 8, 8, 8, false>), grid, dim3(256), 0, s, i, w, bias, out, act, g, Kp);
      return hipGetLastError() == hipSuccess ? 0 : 2;
    }
  }
  fused = false;  // the kernels below have no SiLU epilogue
  if (Cin >= 64 && (Z <= 4 || Cin == 128)) {
    // the coarsest level (Z <= 4) and 128 input channels: split-K over the waves of a block
    // tiles per block: as many as keep at least 512 blocks (two per CU) in the launch
    const int nt = Cout % 64 == 0 ? 4 : Cout % 32 == 0 ? 2 : 1;
    const int64_t groups = Cout / (16 * nt);
    const int mt = (V + 63) / 64 * groups >= 512 ? 4 : (V + 31) / 32 * groups >= 512 ? 2 : 1;
    const dim3 grid((unsigned)((V + 16 * mt - 1) / (16 * mt)), (unsigned)groups);
    switch (nt * 10 + mt) {
      case 44: hipLaunchKernelGGL((k_conv_fwd_sk<4, 4>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 42: hipLaunchKernelGGL((k_conv_fwd_sk<4, 2>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 41: hipLaunchKernelGGL((k_conv_fwd_sk<4, 1>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 24: hipLaunchKernelGGL((k_conv_fwd_sk<2, 4>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 22: hipLaunchKernelGGL((k_conv_fwd_sk<2, 2>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 21: hipLaunchKernelGGL((k_conv_fwd_sk<2, 1>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 14: hipLaunchKernelGGL((k_conv_fwd_sk<1, 4>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      case 12: hipLaunchKernelGGL((k_conv_fwd_sk<1, 2>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
      default: hipLaunchKernelGGL((k_conv_fwd_sk<1, 1>), grid, dim3(256), 0, s, i, w, bias, out, g, Cin, Cout, Kp); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  if (Cout == 16 && (Cin == 16 || Cin == 32) && Kp == (27 * Cin + 31) / 32 * 32) {
    const int tiles = 8;
    const unsigned gw = (unsigned)(((V + 15) / 16 + 4 * tiles - 1) / (4 * tiles));
    if (Cin == 16)
      hipLaunchKernelGGL(k_conv_fwd16<16>, dim3(gw), dim3(256), 0, s, i, w, bias, out, g, Kp, tiles);
    else
      hipLaunchKernelGGL(k_conv_fwd16<32>, dim3(gw), dim3(256), 0, s, i, w, bias, out, g, Kp, tiles);
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  switch (Cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv_fwd<1>, dim3(grid), dim3(256), 0, s, i, w, bias, out, g, Cin, Kp); break;
    case 2: hipLaunchKernelGGL(k_conv_fwd<2>, dim3(grid), dim3(256), 0, s, i, w, bias, out, g, Cin, Kp); break;
    case 4: hipLaunchKernelGGL(k_conv_fwd<4>, dim3(grid), dim3(256), 0, s, i, w, bias, out, g, Cin, Kp); break;
    case 8: hipLaunchKernelGGL(k_conv_fwd<8>, dim3(grid), dim3(256), 0, s, i, w, bias, out, g, Cin, Kp); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int conv_fwd(const void* in, const void* W, const float* bias, float* out, __bf16* act, int N, int X, int Y, int Z,
             int Cin, int Cout, int Kp, hipStream_t s) {
  bool fused = false;
  const int rc = conv_fwd_kernels(in, W, bias, out, act, fused, N, X, Y, Z, Cin, Cout, Kp, s);
  if (rc != 0 || !act || fused) return rc;
  return xnn_silu_fwd(out, act, (int64_t)N * X * Y * Z * Cout, s);
}
}  // namespace
extern "C" {

int xnn_wflip(const void* W, void* Wt, int Cin, int Cout, int Kp, int Kpt, void* stream) {
  if (Kpt % 32 || Kpt < 27 * Cout) return 1;
  hipLaunchKernelGGL(k_wflip, dim3(nblk((int64_t)Cin * Kpt)), dim3(256), 0, (hipStream_t)stream,
                     (const __bf16*)W, (__bf16*)Wt, Cin, Cout, Kp, Kpt);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"

namespace {

// Tiled weight gradient for a layer shape: output tiles per block (cotb) and blocks along z (split); false if the
// shape takes the chunked kernel.  Accumulators per wave: 3 taps x cotb x Cin/16 tiles.
bool wgrad_tile_plan(int Z, int Cin, int Cout, int& cotb, int& split) {
  if (Cin % 16 || Cout % 16) return false;
  const int cit = Cin / 16, cot = Cout / 16;
  cotb = cot;
  split = 1;
  if (Z == 32 && cot == 1 && (cit == 1 || cit == 2)) return true;
  if (Z == 16 && cot == 2 && (cit == 1 || cit == 2 || cit == 4)) return true;
  if (Z == 8 && cot == 4 && (cit == 2 || cit == 4)) return true;
  if (Z == 8 && cot == 4 && cit == 8) return cotb = 1, split = 4, true;
  if (Z == 4 && cot == 8 && cit == 4) return cotb = 2, split = 4, true;
  if (Z == 4 && cot == 8 && cit == 8) return cotb = 1, split = 8, true;
  return false;
}

// Blocks per dx group of the tiled weight gradient (0: the shape takes the chunked kernel).
int wgrad_tile_blocks(int N, int X, int Y, int Z, int Cin, int Cout) {
  int cotb, split;
  if (!wgrad_tile_plan(Z, Cin, Cout, cotb, split)) return 0;
  const int64_t ntiles = (int64_t)N * ((X + kWTX - 1) / kWTX) * ((Y + kWTY - 1) / kWTY);
  int64_t nb = ntiles / 4;
  nb = nb < 1 ? 1 : (nb > 256 ? 256 : nb);
  return (int)nb;
}

template <int CIT, int COT, int ZZ>
void launch_wgrad_tr(const __bf16* i, const __bf16* go, float* part, Geo g, int nb, int split, int CoutT,
                     hipStream_t s) {
  const int ntx = (g.X + kWTX - 1) / kWTX, nty = (g.Y + kWTY - 1) / kWTY;
  hipLaunchKernelGGL((k_conv_wgrad_tr<CIT, COT, ZZ>), dim3(nb, 3, split), dim3(256), 0, s, i, go, part, g, ntx, nty,
                     CoutT);
}


// ------------------------------------------------------------------ FiLM ---
// Feature-wise affine modulation of a layer's pre-activations by per-sample
// vectors computed from the design's scalars (Perez et al. 2018, FiLM):
//   y[n][v][c] = p[n][v][c] * (1 + gam[n][c]) + bet[n][c],
// gam[n][c] = film[n * L + off + c], bet[n][c] = film[n * L + off + C + c].
__global__ void k_film_fwd(const float* __restrict__ p, const float* __restrict__ film, float* __restrict__ y,
                           int64_t Vn, int C, int off, int L, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const int n = (int)(i / ((int64_t)C * Vn));
  const float* f = film + (int64_t)n * L + off;
  y[i] = p[i] * (1.0f + f[c]) + f[C + c];
}

// Backward: gx = gy * (1 + gam), and per-block partial sums of gy * p (-> dgam) and of gy (-> dbet) over a
// contiguous voxel range of one sample (grid.y = sample); thread t sums channel t % C over voxels
// t / C + k (256 / C) (C divides 256), then the block folds its rows in a fixed order.
__global__ __launch_bounds__(256) void k_film_bwd_part(const float* __restrict__ p, const float* __restrict__ gy,
                                                       const float* __restrict__ film, float* __restrict__ gx,
                                                       float* __restrict__ part, int64_t Vn, int C, int off, int L,
                                                       int64_t vper) {
  __shared__ float red[2][256];
  const int n = blockIdx.y, b = blockIdx.x, nbn = gridDim.x;
  const int64_t v0 = (int64_t)b * vper;
  const int64_t v1 = v0 + vper < Vn ? v0 + vper : Vn;
  const int c = threadIdx.x % C, r = threadIdx.x / C, R = 256 / C;
  const float g1 = 1.0f + film[(int64_t)n * L + off + c];
  float sgp = 0.0f, sg = 0.0f;
  for (int64_t v = v0 + r; v < v1; v += R) {
    const int64_t i = ((int64_t)n * Vn + v) * C + c;
    const float g = gy[i];
    sgp += g * p[i];
    sg += g;
    gx[i] = g * g1;
  }
  red[0][threadIdx.x] = sgp;
  red[1][threadIdx.x] = sg;
  __syncthreads();
  if (threadIdx.x < C) {
    float a = 0.0f, e = 0.0f;
    for (int k = 0; k < R; ++k) {
      a += red[0][k * C + c];
      e += red[1][k * C + c];
    }
    float* q = part + ((int64_t)n * nbn + b) * (2 * C);
    q[c] = a;
    q[C + c] = e;
  }
}

// FiLM then SiLU in one pass: y = silu(p (1 + gam) + bet) in BF16 (the same operations, in the same order, as
// k_film_fwd followed by k_silu_fwd, so the same bits).
__global__ void k_film_silu_fwd(const float* __restrict__ p, const float* __restrict__ film, __bf16* __restrict__ y,
                                int64_t Vn, int C, int off, int L, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const int n = (int)(i / ((int64_t)C * Vn));
  const float* f = film + (int64_t)n * L + off;
  y[i] = f2bf(silu(p[i] * (1.0f + f[c]) + f[C + c]));
}

// k_film_silu_fwd with 4 values per thread (16-byte loads, 8-byte stores), C a power of two >= 4 and
// Vn * C / 4 < 2^31: sample blockIdx.y, so the index math is 32-bit with no division.  Same bits.
__global__ void k_film_silu_fwd4(const float* __restrict__ p, const float* __restrict__ film, __bf16* __restrict__ y,
                                 uint32_t n4, int C, int off, int L) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const int n = blockIdx.y;
  const int64_t e = (int64_t)n * n4 + i;  // 4-value group in the whole tensor
  const int c = (int)((i * 4) & (uint32_t)(C - 1));
  const float* f = film + (int64_t)n * L + off;
  const f32x4 v = reinterpret_cast<const f32x4*>(p)[e];
  u16x4 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = __builtin_bit_cast(unsigned short, f2bf(silu(v[k] * (1.0f + f[c + k]) + f[C + c + k])));
  reinterpret_cast<u16x4*>(y)[e] = o;
}

// Backward of k_film_silu_fwd from the gradient gy w.r.t. its output: the modulated pre-activation is formed
// again, gm = gy * silu'(pm), gx = gm * (1 + gam) in BF16, and the per-block partials of gm * p and gm as in
// k_film_bwd_part (the same bits as k_silu_bwd_f32, k_film_bwd_part and the BF16 cast in sequence).
__global__ __launch_bounds__(256) void k_film_silu_bwd_part(const float* __restrict__ p,
                                                            const float* __restrict__ gy,
                                                            const float* __restrict__ film,
                                                            __bf16* __restrict__ gxb, float* __restrict__ part,
                                                            int64_t Vn, int C, int off, int L, int64_t vper) {
  __shared__ float red[2][256];
  const int n = blockIdx.y, b = blockIdx.x, nbn = gridDim.x;
  const int64_t v0 = (int64_t)b * vper;
  const int64_t v1 = v0 + vper < Vn ? v0 + vper : Vn;
  const int c = threadIdx.x % C, r = threadIdx.x / C, R = 256 / C;
  const float* f = film + (int64_t)n * L + off;
  const float g1 = 1.0f + f[c], bt = f[C + c];
  float sgp = 0.0f, sg = 0.0f;
  for (int64_t v = v0 + r; v < v1; v += R) {
    const int64_t i = ((int64_t)n * Vn + v) * C + c;
    const float pv = p[i];
    const float g = gy[i] * dsilu(pv * g1 + bt);
    sgp += g * pv;
    sg += g;
    gxb[i] = f2bf(g * g1);
  }
  red[0][threadIdx.x] = sgp;
  red[1][threadIdx.x] = sg;
  __syncthreads();
  if (threadIdx.x < C) {
    float a = 0.0f, e = 0.0f;
    for (int k = 0; k < R; ++k) {
      a += red[0][k * C + c];
      e += red[1][k * C + c];
    }
    float* q = part + ((int64_t)n * nbn + b) * (2 * C);
    q[c] = a;
    q[C + c] = e;
  }
}

// dfilm[n][off + j] = sum over the sample's blocks (in block order) of part[n][b][j], j < 2C.
__global__ void k_film_bwd_final(const float* __restrict__ part, float* __restrict__ dfilm, int N, int nbn, int C,
                                 int off, int L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * 2 * C) return;
  const int n = t / (2 * C), j = t % (2 * C);
  float a = 0.0f;
  for (int b = 0; b < nbn; ++b) a += part[((int64_t)n * nbn + b) * (2 * C) + j];
  dfilm[(int64_t)n * L + off + j] = a;
}

// ----------------------------------------------------------------- dense ---
// The small fully connected layers that map the design's scalars to the FiLM
// vectors (FP32): pre[n][o] = sum_i W[o][i] x[n][i] (i in order) + b[o];
// post = silu(pre) with act, else pre.
__global__ void k_dense_fwd(const float* __restrict__ x, const float* __restrict__ W, const float* __restrict__ b,
                            float* __restrict__ pre, float* __restrict__ post, int N, int I, int O, int act) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * O) return;
  const int n = t / O, o = t % O;
  float a = 0.0f;
  for (int i = 0; i < I; ++i) a += W[(int64_t)o * I + i] * x[(int64_t)n * I + i];
  a += b[o];
  pre[t] = a;
  if (post) post[t] = act ? silu(a) : a;
}

// dW[o][i] = sum_n g[n][o] x[n][i], db[o] = sum_n g[n][o], with g = gy * silu'(pre) (act) or gy.
__global__ void k_dense_bwd_w(const float* __restrict__ x, const float* __restrict__ gy,
                              const float* __restrict__ pre, float* __restrict__ dW, float* __restrict__ db, int N,
                              int I, int O, int act) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= O * (I + 1)) return;
  const int o = t / (I + 1), i = t % (I + 1);
  float a = 0.0f;
  for (int n = 0; n < N; ++n) {
    float g = gy[(int64_t)n * O + o];
    if (act) g *= dsilu(pre[(int64_t)n * O + o]);
    a += g * (i < I ? x[(int64_t)n * I + i] : 1.0f);
  }
  if (i < I)
    dW[(int64_t)o * I + i] = a;
  else
    db[o] = a;
}

// The sum of a over the 64 lanes of a wave, every lane holding the result: a butterfly of pairwise adds
// (the same pairs in every call, so the same bits).
__device__ __forceinline__ float wave_sum(float a) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m, 64);
  return a;
}

// gx[n][i] = sum_o W[o][i] g[n][o]: one wave per (n, i); lane l sums o = l, l + 64, ... in order, then
// wave_sum.  (A thread walking all O outputs alone waits on O dependent loads in turn.)
__global__ __launch_bounds__(256) void k_dense_bwd_x(const float* __restrict__ W, const float* __restrict__ gy,
                                                     const float* __restrict__ pre, float* __restrict__ gx, int N,
                                                     int I, int O, int act) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= N * I) return;  // the whole wave
  const int n = w / I, i = w % I;
  float a = 0.0f;
  for (int o = lane; o < O; o += 64) {
    float g = gy[(int64_t)n * O + o];
    if (act) g *= dsilu(pre[(int64_t)n * O + o]);
    a += W[(int64_t)o * I + i] * g;
  }
  a = wave_sum(a);
  if (lane == 0) gx[w] = a;
}

// ------------------------------------------------- activation backward ---
// One activation site's backward in one pass, with the per-channel sums its parameters need:
//   pm = p (1 + gam) + bet (FILM) or p,  gm = gy * silu'(pm),  gx = bf16(gm (1 + gam)) or bf16(gm)
// (gx: the gradient w.r.t. the convolution output, operand of the next MFMAs; elementwise the bits of
// k_film_silu_bwd_part / k_silu_bwd), and per (sample, block) partial sums over the block's voxels of
// gm * p and gm (FILM; -> dgam, dbet) and of gx as BF16 (-> the convolution's bias gradient).
// Thread t: channels 4 (t % C4) .. 4 (t % C4) + 3 as 16-byte vectors, voxel rows t / C4 + k R of the block's
// range (R = 256 / C4, C4 = C / 4 a power of two); the block then folds its rows pairwise in LDS (a fixed
// tree).  part[(n * nbn + b) * Q * C + q * C + c]: q = 0 sum gm p, 1 sum gm, 2 sum gx (FILM); q = 0 sum gx.
template <bool FILM>
__global__ __launch_bounds__(256) void k_act_bwd_part(const float* __restrict__ p, const float* __restrict__ gy,
                                                      const float* __restrict__ film, __bf16* __restrict__ gxb,
                                                      float* __restrict__ part, int Vn, int C, int off, int L,
                                                      int vper) {
  constexpr int Q = FILM ? 3 : 1;
  __shared__ f32x4 red[Q][256];
  const int n = blockIdx.y, b = blockIdx.x, nbn = gridDim.x;
  const int C4 = C >> 2, R = 256 / C4;
  const int cq = threadIdx.x % C4, r = threadIdx.x / C4;
  f32x4 g1 = f32x4{1.0f, 1.0f, 1.0f, 1.0f}, bt = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (FILM) {
    const float* f = film + (int64_t)n * L + off;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      g1[k] = 1.0f + f[4 * cq + k];
      bt[k] = f[C + 4 * cq + k];
    }
  }
  f32x4 sgp = f32x4{0.0f, 0.0f, 0.0f, 0.0f}, sg = sgp, sb = sgp;
  const int v0 = b * vper;
  const int v1 = v0 + vper < Vn ? v0 + vper : Vn;
  const int64_t base = (int64_t)n * Vn * C4 + cq;
  const f32x4* p4 = reinterpret_cast<const f32x4*>(p) + base;
  const f32x4* g4 = reinterpret_cast<const f32x4*>(gy) + base;
  u16x4* o4 = reinterpret_cast<u16x4*>(gxb) + base;
#pragma unroll 4
  for (int v = v0 + r; v < v1; v += R) {
    const f32x4 pv = p4[(int64_t)v * C4];
    const f32x4 gv = g4[(int64_t)v * C4];
    u16x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float g;
      __bf16 h;
      if constexpr (FILM) {
        g = gv[k] * dsilu(pv[k] * g1[k] + bt[k]);
        sgp[k] += g * pv[k];
        sg[k] += g;
        h = f2bf(g * g1[k]);
      } else {
        g = gv[k] * dsilu(pv[k]);
        h = f2bf(g);
      }
      sb[k] += bf2f(h);
      o[k] = __builtin_bit_cast(unsigned short, h);
    }
    o4[(int64_t)v * C4] = o;
  }
  if constexpr (FILM) {
    red[0][threadIdx.x] = sgp;
    red[1][threadIdx.x] = sg;
  }
  red[Q - 1][threadIdx.x] = sb;
  __syncthreads();
  for (int s = R / 2; s > 0; s >>= 1) {
    if (r < s) {
#pragma unroll
      for (int q = 0; q < Q; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s * C4];
    }
    __syncthreads();
  }
  if (r == 0) {
    float* o = part + ((int64_t)n * nbn + b) * (Q * C);
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) o[q * C + 4 * cq + k] = red[q][cq][k];
  }
}

// out[gi * ostride + ooff + j] = sum over the rows q of group gi (rows gi * rpg .. gi * rpg + rpg - 1) of
// part[q * rowlen + col0 + j], j < ncols: one wave per (gi, j), lane l summing rows l, l + 64, ... in order,
// then wave_sum.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ part, float* __restrict__ out, int ngroups,
                                                int rpg, int rowlen, int col0, int ncols, int ostride, int ooff) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= ngroups * ncols) return;  // the whole wave
  const int gi = w / ncols, j = w % ncols;
  const float* src = part + (int64_t)gi * rpg * rowlen + col0 + j;
  float a = 0.0f;
#pragma unroll 4
  for (int q = lane; q < rpg; q += 64) a += src[(int64_t)q * rowlen];
  a = wave_sum(a);
  if (lane == 0) out[(int64_t)gi * ostride + ooff + j] = a;
}

// Voxel rows per block of k_act_bwd_part: 16 per thread.
inline int act_vper(int C) { return 16 * (256 / (C / 4)); }
inline int act_blocks(int64_t Vn, int C) { return (int)((Vn + act_vper(C) - 1) / act_vper(C)); }

int film_blocks(int64_t Vn) {
  int64_t nb = (Vn + 4095) / 4096;
  return (int)(nb < 1 ? 1 : (nb > 256 ? 256 : nb));
}

int64_t bias_blocks(int64_t V) {
  int64_t nb = (V + 4095) / 4096;
  return nb < 1 ? 1 : (nb > 1024 ? 1024 : nb);
}

// Device scratch of the allocation-free bias gradient: the per-block partials of k_bias_part, at most
// 1024 blocks x 256 channels (bias_blocks caps the blocks at 1024 and the kernels take C <= 256), 1 MiB per
// GPU for the life of the process.  Every call on a GPU shares it: calls that overlap on one GPU race.
constexpr int64_t kScratchFloats = 1024 * 256;
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// The device address of ``scratch`` on the current device, looked up once per device (each GPU holds its own
// copy of the array); the host-side name of the array is not a device address.
int scratch_on_current_device(float** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 2;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return 2;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int xnn_conv_wgrad(const void* in, const void* gout, float* part, float* dW, int N, int X, int Y, int Z, int Cin,
                   int Cout, int Kp, int vchunk, void* stream) {
  if (Cin % 16 || Cout % 16 || Cin > 128 || vchunk % 32 || vchunk <= 0 || Kp % 32 || Kp < 27 * Cin) return 1;
  Geo g{X, Y, Z, N};
  hipStream_t s = (hipStream_t)stream;
  const __bf16* i = (const __bf16*)in;
  const __bf16* go = (const __bf16*)gout;
  const int64_t n = (int64_t)Cout * Kp;
  const int nb = wgrad_tile_blocks(N, X, Y, Z, Cin, Cout);
  if (nb > 0) {
    int cotb = 0, split = 0;
    wgrad_tile_plan(Z, Cin, Cout, cotb, split);
    const int key = Z * 100 + (Cin / 16) * 10 + cotb;
    switch (key) {
      case 3211: launch_wgrad_tr<1, 1, 32>(i, go, part, g, nb, split, Cout, s); break;
      case 3221: launch_wgrad_tr<2, 1, 32>(i, go, part, g, nb, split, Cout, s); break;
      case 1612: launch_wgrad_tr<1, 2, 16>(i, go, part, g, nb, split, Cout, s); break;
      case 1622: launch_wgrad_tr<2, 2, 16>(i, go, part, g, nb, split, Cout, s); break;
      case 1642: launch_wgrad_tr<4, 2, 16>(i, go, part, g, nb, split, Cout, s); break;
      case 824: launch_wgrad_tr<2, 4, 8>(i, go, part, g, nb, split, Cout, s); break;
      case 844: launch_wgrad_tr<4, 4, 8>(i, go, part, g, nb, split, Cout, s); break;
      case 881: launch_wgrad_tr<8, 1, 8>(i, go, part, g, nb, split, Cout, s); break;
      case 442: launch_wgrad_tr<4, 2, 4>(i, go, part, g, nb, split, Cout, s); break;
      case 481: launch_wgrad_tr<8, 1, 4>(i, go, part, g, nb, split, Cout, s); break;
      default: return 1;
    }
    hipLaunchKernelGGL(k_wgrad_tile_reduce, dim3(nblk(n)), dim3(256), 0, s, part, dW, nb, Cout, Cin, Kp);
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  const int64_t V = g.V();
  const int nchunk = (int)((V + vchunk - 1) / vchunk);
  dim3 grid(27, Cout / 16, nchunk);
  // padded K entries of the partials stay zero: the caller allocates part zeroed
  switch (Cin / 16) {
    case 1: hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(256), 0, s, i, go, part, g, Cout, vchunk, Kp); break;
    case 2: hipLaunchKernelGGL(k_conv_wgrad<2>, grid, dim3(256), 0, s, i, go, part, g, Cout, vchunk, Kp); break;
    case 4: hipLaunchKernelGGL(k_conv_wgrad<4>, grid, dim3(256), 0, s, i, go, part, g, Cout, vchunk, Kp); break;
    case 8: hipLaunchKernelGGL(k_conv_wgrad<8>, grid, dim3(256), 0, s, i, go, part, g, Cout, vchunk, Kp); break;
    default: return 1;
  }
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblk(n)), dim3(256), 0, s, part, dW, nchunk, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int64_t xnn_wgrad_part_floats(int N, int X, int Y, int Z, int Cin, int Cout, int Kp, int vchunk) {
  const int nb = wgrad_tile_blocks(N, X, Y, Z, Cin, Cout);
  if (nb > 0) return (int64_t)3 * nb * Cout * 9 * Cin;
  const int64_t V = (int64_t)N * X * Y * Z;
  return (V + vchunk - 1) / vchunk * Cout * Kp;
}

int xnn_bias_grad(const void* g, float* db, float* part, int64_t V, int C, void* stream) {
  if (C < 8 || C > 256 || C % 8 || 256 % (C / 8)) return 1;
  const int64_t nb = bias_blocks(V);
  const int64_t vper = (V + nb - 1) / nb;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bias_part, dim3((unsigned)nb), dim3(256), 0, s, (const __bf16*)g, part, V, C, vper);
  hipLaunchKernelGGL(k_bias_final, dim3(1), dim3(256), 0, s, part, db, (int)nb, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int64_t xnn_bias_part_floats(int64_t V, int C) { return bias_blocks(V) * C; }
int xnn_bias_grad_noalloc(const void* g, float* db, int64_t V, int C, void* stream) {
  if (C < 8 || C > 256 || C % 8 || 256 % (C / 8)) return 1;
  const int64_t nb = bias_blocks(V);
  if (nb * C > kScratchFloats) return 3;
  float* part = nullptr;
  if (scratch_on_current_device(&part) != 0) return 2;
  const int64_t vper = (V + nb - 1) / nb;
  hipStream_t s = (hipStream_t)stream;
  // k_bias_part writes all nb * C partials before k_bias_final reads them: nothing to clear between calls
  hipLaunchKernelGGL(k_bias_part, dim3((unsigned)nb), dim3(256), 0, s, (const __bf16*)g, part, V, C, vper);
  hipLaunchKernelGGL(k_bias_final, dim3(1), dim3(256), 0, s, part, db, (int)nb, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int xnn_silu_fwd(const float* x, void* y, int64_t n, void* stream) {
  if (vec_ok(n, {x, y}))
    hipLaunchKernelGGL(k_silu_fwd4, dim3(nblk((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, x, (__bf16*)y,
                       (uint32_t)n);
  else
    hipLaunchKernelGGL(k_silu_fwd, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, x, (__bf16*)y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_silu_bwd(const float* x, const float* gy, void* gxb, int64_t n, void* stream) {
  if (vec_ok(n, {x, gy, gxb}))
    hipLaunchKernelGGL(k_silu_bwd4, dim3(nblk((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, x, gy,
                       (__bf16*)gxb, (uint32_t)n);
  else
    hipLaunchKernelGGL(k_silu_bwd, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, x, gy, (__bf16*)gxb, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_to_bf16(const float* x, void* y, int64_t n, void* stream) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, x, (__bf16*)y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_add_f32(const float* a, const float* b, float* y, int64_t n, void* stream) {
  if (vec_ok(n, {a, b, y}))
    hipLaunchKernelGGL(k_add_f32_4, dim3(nblk((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, a, b, y,
                       (uint32_t)n);
  else
    hipLaunchKernelGGL(k_add_f32, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, a, b, y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_add_bf16(const void* a, const void* b, void* y, int64_t n, void* stream) {
  if (vec_ok(n, {a, b, y}))
    hipLaunchKernelGGL(k_add_bf16_8, dim3(nblk((n + 7) / 8)), dim3(256), 0, (hipStream_t)stream, (const __bf16*)a,
                       (const __bf16*)b, (__bf16*)y, (uint32_t)n);
  else
    hipLaunchKernelGGL(k_add_bf16, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16*)a,
                       (const __bf16*)b, (__bf16*)y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_pool_fwd(const void* in, void* out, int N, int CX, int CY, int CZ, int C, void* stream) {
  Geo gc{CX, CY, CZ, N};
  if (C % 8 == 0 && vec_ok(8 * gc.V() * C, {in, out}))  // the fine input has 8x the coarse elements
    hipLaunchKernelGGL(k_pool_fwd8, dim3(nblk(gc.V() * (C / 8))), dim3(256), 0, (hipStream_t)stream,
                       (const __bf16*)in, (__bf16*)out, gc, C);
  else
    hipLaunchKernelGGL(k_pool_fwd, dim3(nblk(gc.V() * C)), dim3(256), 0, (hipStream_t)stream, (const __bf16*)in,
                       (__bf16*)out, gc, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_pool_bwd(const float* gcoarse, float* gfine, int N, int FX, int FY, int FZ, int C, void* stream) {
  Geo gf{FX, FY, FZ, N};
  if (C % 4 == 0 && FX % 2 == 0 && FY % 2 == 0 && FZ % 2 == 0 && vec_ok(gf.V() * C, {gcoarse, gfine}))
    hipLaunchKernelGGL(k_pool_bwd4, dim3(nblk(gf.V() * (C / 4))), dim3(256), 0, (hipStream_t)stream, gcoarse,
                       gfine, gf, C);
  else
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(gf.V() * C)), dim3(256), 0, (hipStream_t)stream, gcoarse, gfine, gf,
                       C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_pool_bwd_add(const float* gcoarse, const float* acc, float* out, int N, int FX, int FY, int FZ, int C,
                     void* stream) {
  Geo gf{FX, FY, FZ, N};
  if (C % 4 == 0 && FX % 2 == 0 && FY % 2 == 0 && FZ % 2 == 0 && vec_ok(gf.V() * C, {gcoarse, acc, out}))
    hipLaunchKernelGGL(k_pool_bwd_add4, dim3(nblk(gf.V() * (C / 4))), dim3(256), 0, (hipStream_t)stream, gcoarse,
                       acc, out, gf, C);
  else
    hipLaunchKernelGGL(k_pool_bwd_add, dim3(nblk(gf.V() * C)), dim3(256), 0, (hipStream_t)stream, gcoarse, acc,
                       out, gf, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_up_fwd(const void* in, void* out, int N, int FX, int FY, int FZ, int C, void* stream) {
  Geo gf{FX, FY, FZ, N};
  if (C % 8 == 0 && FX % 2 == 0 && FY % 2 == 0 && FZ % 2 == 0 && vec_ok(gf.V() * C, {in, out}))
    hipLaunchKernelGGL(k_up_fwd8, dim3(nblk(gf.V() * (C / 8))), dim3(256), 0, (hipStream_t)stream,
                       (const __bf16*)in, (__bf16*)out, gf, C);
  else
    hipLaunchKernelGGL(k_up_fwd, dim3(nblk(gf.V() * C)), dim3(256), 0, (hipStream_t)stream, (const __bf16*)in,
                       (__bf16*)out, gf, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_up_bwd(const float* gfine, float* gcoarse, int N, int CX, int CY, int CZ, int C, void* stream) {
  Geo gc{CX, CY, CZ, N};
  if (C % 4 == 0 && vec_ok(8 * gc.V() * C, {gfine, gcoarse}))  // the fine gradient has 8x the elements
    hipLaunchKernelGGL(k_up_bwd4, dim3(nblk(gc.V() * (C / 4))), dim3(256), 0, (hipStream_t)stream, gfine,
                       gcoarse, gc, C);
  else
    hipLaunchKernelGGL(k_up_bwd, dim3(nblk(gc.V() * C)), dim3(256), 0, (hipStream_t)stream, gfine, gcoarse, gc, C);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
}  // extern "C"
namespace {
// The one-voxel-per-thread MSE: 16 channels and 16-byte aligned rows.
bool mse16(const float* pred, const void* grad, int C) {
  return C == 16 && reinterpret_cast<uintptr_t>(pred) % 16 == 0 && reinterpret_cast<uintptr_t>(grad) % 16 == 0;
}
}  // namespace
extern "C" {
int xnn_mse(const float* pred, const float* target, const float* wmask, float* grad, float* part, int64_t V, int C,
            int nout, float inv_norm, void* stream) {
  hipLaunchKernelGGL(k_mse, dim3(nblk(V * C)), dim3(256), 0, (hipStream_t)stream, pred, target, wmask, grad, part,
                     V, C, nout, inv_norm);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_mse_bf16(const float* pred, const float* target, const float* wmask, const float* wsum, void* grad,
                 float* part, int64_t V, int C, int nout, void* stream) {
  if (V < 1 || C < 1 || nout < 1 || nout > C) return 1;
  if (mse16(pred, grad, C)) {
    hipLaunchKernelGGL(k_mse16_bf16, dim3(nblk(V)), dim3(256), 0, (hipStream_t)stream, pred, target, wmask, wsum,
                       (__bf16*)grad, part, V, nout, (int64_t)nblk(V * C));
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  hipLaunchKernelGGL(k_mse_bf16, dim3(nblk(V * C)), dim3(256), 0, (hipStream_t)stream, pred, target, wmask, wsum,
                     (__bf16*)grad, part, V, C, nout);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_mse_blocks(int64_t V, int C) { return (int)nblk(V * C); }
int xnn_adam(float* w, const float* g, float* m, float* v, void* wb, int64_t n, float lr, float b1, float b2,
             float eps, int step, float wd, void* stream) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, (__bf16*)wb, n, lr, b1,
                     b2, eps, c1, c2, wd);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int xnn_silu_bwd_f32(const float* x, const float* gy, float* gx, int64_t n, void* stream) {
  hipLaunchKernelGGL(k_silu_bwd_f32, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, x, gy, gx, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_film_fwd(const float* p, const float* film, float* y, int N, int64_t Vn, int C, int off, int L,
                 void* stream) {
  if (N < 1 || Vn < 1 || C < 1 || off < 0 || off + 2 * C > L) return 1;
  const int64_t total = (int64_t)N * Vn * C;
  hipLaunchKernelGGL(k_film_fwd, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, p, film, y, Vn, C, off, L,
                     total);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int64_t xnn_film_part_floats(int N, int64_t Vn, int C) { return (int64_t)N * film_blocks(Vn) * 2 * C; }
int xnn_film_silu_fwd(const float* p, const float* film, void* y, int N, int64_t Vn, int C, int off, int L,
                      void* stream) {
  if (N < 1 || Vn < 1 || C < 1 || off < 0 || off + 2 * C > L) return 1;
  const int64_t total = (int64_t)N * Vn * C;
  const int64_t n4 = Vn * C / 4;
  if (C >= 4 && (C & (C - 1)) == 0 && n4 < ((int64_t)1 << 31) && N <= 65535 &&
      reinterpret_cast<uintptr_t>(p) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 8 == 0) {
    hipLaunchKernelGGL(k_film_silu_fwd4, dim3(nblk(n4), N), dim3(256), 0, (hipStream_t)stream, p, film, (__bf16*)y,
                       (uint32_t)n4, C, off, L);
    return hipGetLastError() == hipSuccess ? 0 : 2;
  }
  hipLaunchKernelGGL(k_film_silu_fwd, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, p, film, (__bf16*)y,
                     Vn, C, off, L, total);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_film_silu_bwd(const float* p, const float* gy, const float* film, void* gxb, float* part, float* dfilm,
                      int N, int64_t Vn, int C, int off, int L, void* stream) {
  if (N < 1 || N > 65535 || Vn < 1 || C < 1 || C > 256 || 256 % C || off < 0 || off + 2 * C > L) return 1;
  const int nbn = film_blocks(Vn);
  const int64_t vper = (Vn + nbn - 1) / nbn;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_film_silu_bwd_part, dim3(nbn, N), dim3(256), 0, s, p, gy, film, (__bf16*)gxb, part, Vn, C,
                     off, L, vper);
  hipLaunchKernelGGL(k_film_bwd_final, dim3(nblk((int64_t)N * 2 * C)), dim3(256), 0, s, part, dfilm, N, nbn, C, off,
                     L);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_film_bwd(const float* p, const float* gy, const float* film, float* gx, float* part, float* dfilm, int N,
                 int64_t Vn, int C, int off, int L, void* stream) {
  if (N < 1 || N > 65535 || Vn < 1 || C < 1 || C > 256 || 256 % C || off < 0 || off + 2 * C > L) return 1;
  const int nbn = film_blocks(Vn);
  const int64_t vper = (Vn + nbn - 1) / nbn;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_film_bwd_part, dim3(nbn, N), dim3(256), 0, s, p, gy, film, gx, part, Vn, C, off, L, vper);
  hipLaunchKernelGGL(k_film_bwd_final, dim3(nblk((int64_t)N * 2 * C)), dim3(256), 0, s, part, dfilm, N, nbn, C, off,
                     L);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int64_t xnn_act_part_floats(int N, int64_t Vn, int C, int film) {
  if (N < 1 || Vn < 1 || C < 4 || C % 4 || C > 256) return 0;
  return (int64_t)N * act_blocks(Vn, C) * (film ? 3 : 1) * C;
}
int xnn_act_bwd(const float* p, const float* gy, const float* film, void* gxb, float* part, float* dfilm, float* db,
                int N, int64_t Vn, int C, int off, int L, void* stream) {
  const int C4 = C / 4;
  if (N < 1 || N > 65535 || Vn < 1 || Vn >= (int64_t(1) << 31) || C < 4 || C % 4 || C > 256 || (C4 & (C4 - 1)))
    return 1;
  if (film && (dfilm == nullptr || off < 0 || off + 2 * C > L)) return 1;
  if (reinterpret_cast<uintptr_t>(p) % 16 || reinterpret_cast<uintptr_t>(gy) % 16 ||
      reinterpret_cast<uintptr_t>(gxb) % 8)
    return 1;
  const int nbn = act_blocks(Vn, C), vper = act_vper(C), Q = film ? 3 : 1;
  if ((int64_t)N * nbn >= (int64_t(1) << 31) / (Q * C)) return 1;
  hipStream_t s = (hipStream_t)stream;
  if (film)
    hipLaunchKernelGGL(k_act_bwd_part<true>, dim3(nbn, N), dim3(256), 0, s, p, gy, film, (__bf16*)gxb, part,
                       (int)Vn, C, off, L, vper);
  else
    hipLaunchKernelGGL(k_act_bwd_part<false>, dim3(nbn, N), dim3(256), 0, s, p, gy, film, (__bf16*)gxb, part,
                       (int)Vn, C, off, L, vper);
  if (film)
    hipLaunchKernelGGL(k_colsum, dim3(nblk((int64_t)N * 2 * C, 4)), dim3(256), 0, s, part, dfilm, N, nbn, Q * C, 0,
                       2 * C, L, off);
  if (db)
    hipLaunchKernelGGL(k_colsum, dim3(nblk(C, 4)), dim3(256), 0, s, part, db, 1, N * nbn, Q * C, (Q - 1) * C, C, 0,
                       0);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_dense_fwd(const float* x, const float* W, const float* b, float* pre, float* post, int N, int I, int O,
                  int act, void* stream) {
  if (N < 1 || I < 1 || O < 1) return 1;
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((int64_t)N * O)), dim3(256), 0, (hipStream_t)stream, x, W, b, pre, post,
                     N, I, O, act);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
int xnn_dense_bwd(const float* x, const float* W, const float* pre, const float* gy, float* dW, float* db, float* gx,
                  int N, int I, int O, int act, void* stream) {
  if (N < 1 || I < 1 || O < 1) return 1;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((int64_t)O * (I + 1))), dim3(256), 0, s, x, gy, pre, dW, db, N, I, O,
                     act);
  if (gx)
    hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk((int64_t)N * I, 4)), dim3(256), 0, s, W, gy, pre, gx, N, I, O, act);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
