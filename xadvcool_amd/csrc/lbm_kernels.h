// MI355X (gfx950) lattice-Boltzmann kernels: coupled D3Q19 fluid + D3Q19
// thermal step in one AA-pattern pass.
//
// What each time step computes is the reference's six-launch step
// (src/simulation.py:94-99):
//   fluid:   f* = BGK(f, rho, u)              src/fluid_engine3d.py:141-176
//            f  = stream(f*) with inlet/outlet/bounce-back   :181-272
//            rho, u = moments(f)                              :85-113
//   thermal: g* = BGK(g, T, u_new) + source   src/thermal_engine3d.py:91-145
//            g  = stream(g*) with Dirichlet inlet, zero-gradient outlet,
//                 adiabatic walls                             :147-211
//            T  = sum(g)                                      :214-224
// Every + and * is the reference's, in its order (built with
// -ffp-contract=off), so a pass reproduces the six-kernel schedule bit for bit
// in FP32, and the same source built for double is the FP64 reference.
//
// How it is laid out for MI355X:
//  * One array of 19 SoA planes per lattice (AA pattern, in place).  Pass n
//    with n odd gathers post-collision values from the neighbours' opposite
//    slots and scatters its own post-collision values into the neighbours'
//    slots; with n even it reads and writes only its own slots (swapped).
//    Halfway bounce-back is a redirect of that read or write to the cell's
//    own slot (one precomputed bit per direction).
//  * The reference's boundary rules that read values no AA slot holds go
//    through small side planes: INLET cells keep their own post-collision
//    set (fin), cells left of an OUTLET publish their full post-collision set
//    (fsnap / gsnap, double-buffered by step parity), and thermal cells on
//    x = nx-1 keep their own e_x = -1 values (gxh).
//  * Thermal runs one step behind the fluid inside the same pass: pass n
//    streams g*^{n-2}, forms T^{n-1} and collides it with u^n, which the fluid
//    half of the same thread has just formed.  Its AA parity is the opposite
//    of the fluid's.  Pass 1 takes the initial g and T as given ("raw").
//  * Solid cells never run the fluid half (their populations are frozen, as
//    in the reference) but do run the thermal half (conjugate heat transfer).
//  * A batch is an array of slots; each block belongs to one slot, so lattices
//    of different sizes advance in one launch.  Blocks can be remapped so
//    that consecutive lattice blocks share an XCD (and its L2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "lattice.h"

namespace xadv {

template <class R>
struct Slot {
  int nx, ny, nz, P;  // P = ny * nz (cells per x-plane)
  int64_t ncell, npad;
  int active;
  int raw;  // next odd pass takes the initial thermal state as is
  int px, py, pz, dir_x;
  R omega, rho_out, u_in, force_x, u_wall;
  R omega_tf, omega_ts, sf_tf, sf_ts;  // thermal 1/tau and (1 - 0.5*omega): fluid, solid
  R T_in, dir_T;
  R* f;
  R* g;
  const uint32_t* flags;
  const R* q;
  R* fin;    // [19][P]   INLET own post-collision set
  R* fsnap;  // [2][19][P] fluid post-collision sets of the cells left of OUTLET
  R* gsnap;  // [2][19][P] thermal ditto
  R* gxh;    // [5][P]    thermal e_x=-1 post-collision values on x = nx-1
  R* rho;
  R* u;  // (ncell, 3)
  R* T;
};

__device__ __forceinline__ int map_block(int G, int xcd) {
  int b = blockIdx.x;
  if (xcd) {
    const int per = (G + 7) >> 3;
    b = (b & 7) * per + (b >> 3);
  }
  return b;
}

template <class R, int i>
__device__ __forceinline__ R feq_i(R rho, R ux, R uy, R uz, R usq) {
  // src/fluid_engine3d.py:74-78 (same expression for the thermal lattice with T)
  const R eu = (R)ex_of(i) * ux + (R)ey_of(i) * uy + (R)ez_of(i) * uz;
  return (R)w_of(i) * rho * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
}

template <class R>
__device__ __forceinline__ R wall_mom(int i, R u_wall) {
  // validate_analytical.py:531-532
  return (R)6.0 * (R)w_of(i) * (R)ex_of(i) * u_wall;
}

template <int i>
__device__ __forceinline__ int64_t nb_off(int P, int nz) {
  return (int64_t)ex_of(i) * P + (int64_t)ey_of(i) * nz + ez_of(i);
}

// Index of c + s*e_i with periodic wrap on the periodic axes (general mode).
// On a non-periodic axis the result may leave the lattice; callers only use it
// where the precomputed bits say the neighbour exists.
template <int i>
__device__ __forceinline__ int64_t nb_wrap(int x, int y, int z, int s, int nx, int ny, int nz) {
  int sx = x + s * ex_of(i), sy = y + s * ey_of(i), sz = z + s * ez_of(i);
  sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
  sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
  sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
  return ((int64_t)sx * ny + sy) * nz + sz;
}

template <int i>
__device__ __forceinline__ bool th_src_xlow(uint32_t fl) { return ex_of(i) == 1 && (fl & F_FX0); }
template <int i>
__device__ __forceinline__ bool th_src_xhigh(uint32_t fl) { return ex_of(i) == -1 && (fl & F_FXN); }
template <int i>
__device__ __forceinline__ bool th_src_yz(uint32_t fl) {
  return (ey_of(i) == 1 && (fl & F_FY0)) || (ey_of(i) == -1 && (fl & F_FYN)) ||
         (ez_of(i) == 1 && (fl & F_FZ0)) || (ez_of(i) == -1 && (fl & F_FZN));
}
// Target c + e_i outside a non-periodic axis.
template <int i>
__device__ __forceinline__ bool th_dst_oob(uint32_t fl) {
  return (ex_of(i) == 1 && (fl & F_FXN)) || (ex_of(i) == -1 && (fl & F_FX0)) ||
         (ey_of(i) == 1 && (fl & F_FYN)) || (ey_of(i) == -1 && (fl & F_FY0)) ||
         (ez_of(i) == 1 && (fl & F_FZN)) || (ez_of(i) == -1 && (fl & F_FZ0));
}

// ------------------------------------------------------------------ fluid ---

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ R fluid_load(const R* __restrict__ A, int64_t N, int64_t c, uint32_t fl, int P, int nz,
                                        int x, int y, int zc, int nx, int ny, R u_wall) {
  if (NODD == 0 || i == 0) return A[(int64_t)i * N + c];
  const bool bb = (fl >> (F_BOUNCE0 + i)) & 1u;
  int64_t src;
  if (GEN)
    src = nb_wrap<i>(x, y, zc, -1, nx, ny, nz);
  else
    src = c - nb_off<i>(P, nz);
  const int64_t a = bb ? ((int64_t)i * N + c) : ((int64_t)opp_of(i) * N + src);
  R v = A[a];
  if (GEN && ez_of(i) == -1) {
    if (bb && (fl & F_FZN)) v = v + wall_mom<R>(i, u_wall);
  }
  return v;
}

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ void fluid_store(R* __restrict__ A, int64_t N, int64_t c, uint32_t fl, int P, int nz, int x,
                                            int y, int zc, int nx, int ny, R u_wall, R v) {
  if (NODD == 0 || i == 0) {
    A[(int64_t)opp_of(i) * N + c] = v;
    return;
  }
  const bool rb = (fl >> (F_BOUNCE0 + opp_of(i))) & 1u;  // target c + e_i is outside or SOLID
  int64_t dst;
  if (GEN)
    dst = nb_wrap<i>(x, y, zc, +1, nx, ny, nz);
  else
    dst = c + nb_off<i>(P, nz);
  if (GEN && ez_of(i) == 1) {
    if (rb && (fl & F_FZN)) v = v + wall_mom<R>(opp_of(i), u_wall);
  }
  const int64_t a = rb ? ((int64_t)opp_of(i) * N + c) : ((int64_t)i * N + dst);
  A[a] = v;
}

// ---------------------------------------------------------------- thermal ---

// Streamed thermal populations of cell c for the pass of parity NODD (the
// thermal half uses the opposite AA parity: TODD = !NODD).
template <class R, int GEN, int NODD>
__device__ __forceinline__ void thermal_gather(const Slot<R>& S, int64_t c, uint32_t fl, int type, int x, int y,
                                               int zc, R* g) {
  constexpr int TODD = 1 - NODD;
  const R* __restrict__ B = S.g;
  const int64_t N = S.npad;
  const int P = S.P, nz = S.nz;
  if (type == BC_INLET) {  // thermal_engine3d.py:160-165
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = (R)w_of(i) * S.T_in;
  } else if (type == BC_OUTLET) {  // :170-177, values of x-1 published into gsnap
    const int64_t p = c - (int64_t)(S.nx - 1) * P;
    const R* __restrict__ sn = S.gsnap + (int64_t)NODD * 19 * P;
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = sn[(int64_t)i * P + p];
  } else {
    const int64_t pxh = c - (int64_t)(S.nx - 1) * P;
#define XADV_TG(i)                                                                     \
  {                                                                                    \
    R v;                                                                               \
    if (TODD == 0 || i == 0) {                                                         \
      v = B[(int64_t)i * N + c];                                                       \
    } else {                                                                           \
      const bool oob = th_src_xlow<i>(fl) || th_src_xhigh<i>(fl) || th_src_yz<i>(fl);   \
      int64_t src;                                                                     \
      if (GEN)                                                                         \
        src = nb_wrap<i>(x, y, zc, -1, S.nx, S.ny, nz);                                \
      else                                                                             \
        src = c - nb_off<i>(P, nz);                                                    \
      const int64_t a = oob ? ((int64_t)i * N + c) : ((int64_t)opp_of(i) * N + src);   \
      v = B[a];                                                                        \
    }                                                                                  \
    if (th_src_xlow<i>(fl)) v = (R)w_of(i) * S.T_in;                                   \
    if (xm_slot(i) >= 0) {                                                             \
      if (th_src_xhigh<i>(fl)) v = S.gxh[(int64_t)(xm_slot(i) < 0 ? 0 : xm_slot(i)) * P + pxh]; \
    }                                                                                  \
    g[i] = v;                                                                          \
  }
    XADV_TG(0) XADV_TG(1) XADV_TG(2) XADV_TG(3) XADV_TG(4) XADV_TG(5) XADV_TG(6) XADV_TG(7) XADV_TG(8)
    XADV_TG(9) XADV_TG(10) XADV_TG(11) XADV_TG(12) XADV_TG(13) XADV_TG(14) XADV_TG(15) XADV_TG(16)
    XADV_TG(17) XADV_TG(18)
#undef XADV_TG
  }
  if (GEN && (fl & F_DIR)) {  // validate_analytical.py:21-30, applied after streaming
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = (R)w_of(i) * S.dir_T;
  }
}

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ void thermal_store(R* __restrict__ B, int64_t N, int64_t c, uint32_t fl, int P, int nz,
                                              int x, int y, int zc, int nx, int ny, R v) {
  constexpr int TODD = 1 - NODD;
  if (TODD == 0 || i == 0) {
    B[(int64_t)opp_of(i) * N + c] = v;
    return;
  }
  const bool oob = th_dst_oob<i>(fl);
  int64_t dst;
  if (GEN)
    dst = nb_wrap<i>(x, y, zc, +1, nx, ny, nz);
  else
    dst = c + nb_off<i>(P, nz);
  const int64_t a = oob ? ((int64_t)opp_of(i) * N + c) : ((int64_t)i * N + dst);
  B[a] = v;
}

template <class R>
__device__ __forceinline__ void coords(int64_t c, int P, int nz, int& x, int& y, int& z) {
  x = (int)(c / P);
  const int r = (int)(c - (int64_t)x * P);
  y = r / nz;
  z = r - y * nz;
}

// ------------------------------------------------------------- the pass ----

template <class R, int FL, int TH, int GEN, int NODD>
__global__ __launch_bounds__(256) void k_pass(Slot<R>* __restrict__ slots, int bps, int G, int xcd,
                                              int want_fields) {
  const int b = map_block(G, xcd);
  if (b >= G) return;
  const int si = b / bps;
  const int lb = b - si * bps;
  Slot<R>& S = slots[si];
  if (!S.active) return;
  // The first thermal half after an init reads the initial state (odd pass);
  // the following even pass clears the flag (no even pass reads it).
  int raw = 0;
  if (TH) {
    if (NODD)
      raw = S.raw;
    else if (lb == 0 && threadIdx.x == 0 && S.raw)
      S.raw = 0;
  }
  const int64_t c = (int64_t)lb * blockDim.x + threadIdx.x;
  if (c >= S.ncell) return;

  const uint32_t fl = S.flags[c];
  const int type = (int)(fl & F_TYPE);
  const int P = S.P, nz = S.nz, nx = S.nx, ny = S.ny;
  const int64_t N = S.npad;
  int x = 0, y = 0, zc = 0;
  if (GEN) coords<R>(c, P, nz, x, y, zc);

  R ux = 0, uy = 0, uz = 0;
  if (FL) {
    R* __restrict__ A = S.f;
    if (type != BC_SOLID) {
      R f[19];
      if (type == BC_INLET) {
        // Velocity inlet from the cell's own post-collision set
        // (src/fluid_engine3d.py:197-244); x = 0 so the plane index is c.
        const R* __restrict__ fi = S.fin;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = fi[(int64_t)i * P + c];
        const R uin = S.u_in;
        const R rho_new = (f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18] +
                           (R)2.0 * (f[2] + f[8] + f[10] + f[12] + f[14])) /
                          ((R)1.0 - uin);
        const R c23 = (R)(2.0 / 3.0), c16 = (R)(1.0 / 6.0);
        f[1] = f[2] + c23 * rho_new * uin;
        f[7] = f[10] + c16 * rho_new * uin;
        f[9] = f[8] + c16 * rho_new * uin;
        f[11] = f[14] + c16 * rho_new * uin;
        f[13] = f[12] + c16 * rho_new * uin;
      } else if (type == BC_OUTLET) {
        // Pressure outlet: all 19 post-collision values of x-1, rescaled
        // (src/fluid_engine3d.py:247-259).
        const int64_t p = c - (int64_t)(nx - 1) * P;
        const R* __restrict__ sn = S.fsnap + (int64_t)(1 - NODD) * 19 * P;
        R rs = 0;
#pragma unroll
        for (int i = 0; i < 19; ++i) {
          f[i] = sn[(int64_t)i * P + p];
          rs += f[i];
        }
        if (rs > (R)1e-6) {
          const R sc = S.rho_out / rs;
#pragma unroll
          for (int i = 0; i < 19; ++i) f[i] *= sc;
        }
      } else {
        const R uw = GEN ? S.u_wall : (R)0;
#define XADV_FL(i) f[i] = fluid_load<R, GEN, NODD, i>(A, N, c, fl, P, nz, x, y, zc, nx, ny, uw);
        XADV_FL(0) XADV_FL(1) XADV_FL(2) XADV_FL(3) XADV_FL(4) XADV_FL(5) XADV_FL(6) XADV_FL(7) XADV_FL(8)
        XADV_FL(9) XADV_FL(10) XADV_FL(11) XADV_FL(12) XADV_FL(13) XADV_FL(14) XADV_FL(15) XADV_FL(16)
        XADV_FL(17) XADV_FL(18)
#undef XADV_FL
      }
      // Moments (src/fluid_engine3d.py:92-113)
      R lr = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        lr += f[i];
        mx += f[i] * (R)ex_of(i);
        my += f[i] * (R)ey_of(i);
        mz += f[i] * (R)ez_of(i);
      }
      if (lr > (R)1e-10) {
        ux = mx / lr;
        uy = my / lr;
        uz = mz / lr;
      }
      if (want_fields) {
        S.rho[c] = lr;
        S.u[3 * c] = ux;
        S.u[3 * c + 1] = uy;
        S.u[3 * c + 2] = uz;
      }
      // BGK with the equilibrium in registers (src/fluid_engine3d.py:157-173)
      const R om = S.omega;
      const R usq = ux * ux + uy * uy + uz * uz;
#define XADV_FC(i) f[i] = f[i] - om * (f[i] - feq_i<R, i>(lr, ux, uy, uz, usq));
      XADV_FC(0) XADV_FC(1) XADV_FC(2) XADV_FC(3) XADV_FC(4) XADV_FC(5) XADV_FC(6) XADV_FC(7) XADV_FC(8)
      XADV_FC(9) XADV_FC(10) XADV_FC(11) XADV_FC(12) XADV_FC(13) XADV_FC(14) XADV_FC(15) XADV_FC(16)
      XADV_FC(17) XADV_FC(18)
#undef XADV_FC
      if (GEN) {  // validate_analytical.py:388-390
        const R F = S.force_x;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = f[i] + (R)w_of(i) * (R)3.0 * (R)ex_of(i) * F;
      }
      {
        const R uw = GEN ? S.u_wall : (R)0;
#define XADV_FS(i) fluid_store<R, GEN, NODD, i>(A, N, c, fl, P, nz, x, y, zc, nx, ny, uw, f[i]);
        XADV_FS(0) XADV_FS(1) XADV_FS(2) XADV_FS(3) XADV_FS(4) XADV_FS(5) XADV_FS(6) XADV_FS(7) XADV_FS(8)
        XADV_FS(9) XADV_FS(10) XADV_FS(11) XADV_FS(12) XADV_FS(13) XADV_FS(14) XADV_FS(15) XADV_FS(16)
        XADV_FS(17) XADV_FS(18)
#undef XADV_FS
      }
      if (type == BC_INLET) {
        R* __restrict__ fi = S.fin;
#pragma unroll
        for (int i = 0; i < 19; ++i) fi[(int64_t)i * P + c] = f[i];
      }
      if (fl & F_SNAP) {
        const int64_t p = c - (int64_t)(nx - 2) * P;
        R* __restrict__ sn = S.fsnap + (int64_t)NODD * 19 * P;
#pragma unroll
        for (int i = 0; i < 19; ++i) sn[(int64_t)i * P + p] = f[i];
      }
    } else if (want_fields) {
      // Solid cells keep their initial populations, stored in the even layout
      // (slot opp(i)); the reference's moment kernel still runs on them.
      R lr = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        const R v = A[(int64_t)opp_of(i) * N + c];
        lr += v;
        mx += v * (R)ex_of(i);
        my += v * (R)ey_of(i);
        mz += v * (R)ez_of(i);
      }
      R sx = 0, sy = 0, sz = 0;
      if (lr > (R)1e-10) {
        sx = mx / lr;
        sy = my / lr;
        sz = mz / lr;
      }
      S.rho[c] = lr;
      S.u[3 * c] = sx;
      S.u[3 * c + 1] = sy;
      S.u[3 * c + 2] = sz;
    }
  } else if (TH) {
    // Thermal lattice alone: collide with the velocity field given at init.
    ux = S.u[3 * c];
    uy = S.u[3 * c + 1];
    uz = S.u[3 * c + 2];
  }

  if (TH) {
    R g[19];
    R Tl;
    if (raw) {
      // First thermal half: the initial populations and T as given
      // (src/simulation.py:73,89), natural layout.
      const R* __restrict__ B = S.g;
#pragma unroll
      for (int i = 0; i < 19; ++i) g[i] = B[(int64_t)i * N + c];
      Tl = S.T[c];
    } else {
      thermal_gather<R, GEN, NODD>(S, c, fl, type, x, y, zc, g);
      Tl = 0;
#pragma unroll
      for (int i = 0; i < 19; ++i) Tl += g[i];  // thermal_engine3d.py:220-224
    }
    const bool solid = type == BC_SOLID;
    if (solid) ux = uy = uz = 0;  // thermal_engine3d.py:123-125
    const R om = solid ? S.omega_ts : S.omega_tf;
    const R sf = solid ? S.sf_ts : S.sf_tf;
    const R q = (fl & F_HEAT) ? S.q[c] : (R)0;
    const R usq = ux * ux + uy * uy + uz * uz;
    // thermal_engine3d.py:128-145
#define XADV_TC(i)                                                  \
  {                                                                 \
    R v = g[i] - om * (g[i] - feq_i<R, i>(Tl, ux, uy, uz, usq));    \
    if (q > (R)0.0) v += (R)w_of(i) * sf * q;                       \
    g[i] = v;                                                       \
  }
    XADV_TC(0) XADV_TC(1) XADV_TC(2) XADV_TC(3) XADV_TC(4) XADV_TC(5) XADV_TC(6) XADV_TC(7) XADV_TC(8)
    XADV_TC(9) XADV_TC(10) XADV_TC(11) XADV_TC(12) XADV_TC(13) XADV_TC(14) XADV_TC(15) XADV_TC(16)
    XADV_TC(17) XADV_TC(18)
#undef XADV_TC
    R* __restrict__ B = S.g;
#define XADV_TS(i) thermal_store<R, GEN, NODD, i>(B, N, c, fl, P, nz, x, y, zc, nx, ny, g[i]);
    XADV_TS(0) XADV_TS(1) XADV_TS(2) XADV_TS(3) XADV_TS(4) XADV_TS(5) XADV_TS(6) XADV_TS(7) XADV_TS(8)
    XADV_TS(9) XADV_TS(10) XADV_TS(11) XADV_TS(12) XADV_TS(13) XADV_TS(14) XADV_TS(15) XADV_TS(16)
    XADV_TS(17) XADV_TS(18)
#undef XADV_TS
    if (fl & F_XH) {
      const int64_t p = c - (int64_t)(nx - 1) * P;
      R* __restrict__ xh = S.gxh;
      xh[0 * (int64_t)P + p] = g[2];
      xh[1 * (int64_t)P + p] = g[8];
      xh[2 * (int64_t)P + p] = g[10];
      xh[3 * (int64_t)P + p] = g[12];
      xh[4 * (int64_t)P + p] = g[14];
    }
    if (fl & F_SNAP) {
      const int64_t p = c - (int64_t)(nx - 2) * P;
      R* __restrict__ sn = S.gsnap + (int64_t)(1 - NODD) * 19 * P;
#pragma unroll
      for (int i = 0; i < 19; ++i) sn[(int64_t)i * P + p] = g[i];
    }
  }
}

// ------------------------------------------------------------------- init ---

// Populations from macroscopic fields, stored as the post-collision state of
// step 1 in the even layout (what an even pass would have written).  For the
// fluid, BGK(f_eq(rho,u), rho, u) returns f_eq exactly, so only the body force
// term changes the stored values.  The thermal lattice keeps g_eq in the
// natural layout for the raw first pass.
template <class R, int FL, int TH, int GEN>
__global__ __launch_bounds__(256) void k_init(Slot<R>* __restrict__ slots, int slot, int frozen_u) {
  Slot<R>& S = slots[slot];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncell) return;
  const uint32_t fl = S.flags[c];
  const int type = (int)(fl & F_TYPE);
  const int64_t N = S.npad;
  const int P = S.P;
  const R ux = S.u[3 * c], uy = S.u[3 * c + 1], uz = S.u[3 * c + 2];
  if (FL) {
    const R lr = S.rho[c];
    const R usq = ux * ux + uy * uy + uz * uz;
    R f[19];
#define XADV_FE(i) f[i] = feq_i<R, i>(lr, ux, uy, uz, usq);
    XADV_FE(0) XADV_FE(1) XADV_FE(2) XADV_FE(3) XADV_FE(4) XADV_FE(5) XADV_FE(6) XADV_FE(7) XADV_FE(8)
    XADV_FE(9) XADV_FE(10) XADV_FE(11) XADV_FE(12) XADV_FE(13) XADV_FE(14) XADV_FE(15) XADV_FE(16)
    XADV_FE(17) XADV_FE(18)
#undef XADV_FE
    if (type != BC_SOLID) {
      const R om = S.omega;
#pragma unroll
      for (int i = 0; i < 19; ++i) f[i] = f[i] - om * (f[i] - f[i]);
      if (GEN) {
        const R F = S.force_x;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = f[i] + (R)w_of(i) * (R)3.0 * (R)ex_of(i) * F;
      }
    }
    R* __restrict__ A = S.f;
#pragma unroll
    for (int i = 0; i < 19; ++i) A[(int64_t)opp_of(i) * N + c] = f[i];
    if (type == BC_INLET) {
#pragma unroll
      for (int i = 0; i < 19; ++i) S.fin[(int64_t)i * P + c] = f[i];
    }
    if (fl & F_SNAP) {
      const int64_t p = c - (int64_t)(S.nx - 2) * P;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        S.fsnap[(int64_t)i * P + p] = f[i];
        S.fsnap[(int64_t)(19 + i) * P + p] = f[i];
      }
    }
  }
  if (TH) {
    // src/thermal_engine3d.py:24-47 (u = 0 in solids)
    R vx = ux, vy = uy, vz = uz;
    if (type == BC_SOLID) vx = vy = vz = 0;
    const R usq = vx * vx + vy * vy + vz * vz;
    const R Tl = S.T[c];
    R* __restrict__ B = S.g;
#define XADV_GE(i) B[(int64_t)i * N + c] = feq_i<R, i>(Tl, vx, vy, vz, usq);
    XADV_GE(0) XADV_GE(1) XADV_GE(2) XADV_GE(3) XADV_GE(4) XADV_GE(5) XADV_GE(6) XADV_GE(7) XADV_GE(8)
    XADV_GE(9) XADV_GE(10) XADV_GE(11) XADV_GE(12) XADV_GE(13) XADV_GE(14) XADV_GE(15) XADV_GE(16)
    XADV_GE(17) XADV_GE(18)
#undef XADV_GE
  }
  (void)frozen_u;
}

template <class R>
__global__ void k_fill(R* __restrict__ p, int64_t n, R v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// -------------------------------------------------------------- readout ----

// Wave (64 lanes) then block reduction of NV doubles; op[k] = 0 sum, 1 max, 2 min.
template <int NV>
__device__ __forceinline__ void block_reduce(double* v, const int* op, double* out) {
  __shared__ double red[4][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_xor(v[k], o, 64);
      v[k] = op[k] == 0 ? v[k] + w : (op[k] == 1 ? fmax(v[k], w) : fmin(v[k], w));
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wave][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    const int nw = blockDim.x >> 6;
    double a = red[0][k];
    for (int w = 1; w < nw; ++w) a = op[k] == 0 ? a + red[w][k] : (op[k] == 1 ? fmax(a, red[w][k]) : fmin(a, red[w][k]));
    out[k] = a;
  }
}

__constant__ int kStatOp[16] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0};

// T of the current step (thermal streaming + moments) and the 16 stat slots
// of src/fluid_engine3d.py:283-312 and src/thermal_engine3d.py:234-243.
template <class R, int FL, int TH, int GEN, int NODD>
__global__ __launch_bounds__(256) void k_readout(Slot<R>* __restrict__ slots, int bps, int G, int xcd,
                                                 int with_stats, double* __restrict__ partials) {
  const int b = map_block(G, xcd);
  if (b >= G) return;
  const int si = b / bps;
  const int lb = b - si * bps;
  Slot<R>& S = slots[si];
  if (!S.active) return;
  const int64_t c = (int64_t)lb * blockDim.x + threadIdx.x;
  const bool valid = c < S.ncell;
  double v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = 0.0;
  if (valid) {
    const uint32_t fl = S.flags[c];
    const int type = (int)(fl & F_TYPE);
    int x, y, zc;
    coords<R>(c, S.P, S.nz, x, y, zc);
    R Tl = 0;
    if (TH) {
      R g[19];
      thermal_gather<R, GEN, NODD>(S, c, fl, type, x, y, zc, g);
#pragma unroll
      for (int i = 0; i < 19; ++i) Tl += g[i];
      S.T[c] = Tl;
    }
    if (with_stats) {
      const float ux = (float)S.u[3 * c], uy = (float)S.u[3 * c + 1], uz = (float)S.u[3 * c + 2];
      const float vsq = ux * ux + uy * uy + uz * uz;
      const float vmag = sqrtf(vsq);
      const double r = (double)S.rho[c];
      v[2] = vsq;
      if (type != BC_SOLID) {
        v[0] = vmag;
        v[1] = vmag;
        v[3] = 1.0;
        v[4] = r;
      }
      if (x == 1) {
        v[5] = r;
        v[8] = 1.0;
      }
      if (x == S.nx - 2) {
        v[6] = r;
        v[9] = 1.0;
      }
      if (type == BC_OUTLET) {
        v[7] = ux;
        v[10] = 1.0;
      }
      if (TH) {
        const float t = (float)Tl;
        v[11] = t > 0.0f ? t : 0.0f;
        v[12] = t;
        v[13] = 1.0;
        if (type == BC_OUTLET) {
          v[14] = (double)(ux * t);
          v[15] = 1.0;
        }
      }
    }
  }
  if (with_stats) {
    int op[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) op[k] = kStatOp[k];
    block_reduce<16>(v, op, partials + (int64_t)b * 16);
  }
}

// One block per slot: fixed-order reduction of the per-block partials.
template <class R>
__global__ void k_stats_final(const Slot<R>* __restrict__ slots, int bps, const double* __restrict__ partials,
                              double* __restrict__ out) {
  const int si = blockIdx.x;
  const int k = threadIdx.x;
  if (k >= 16) return;
  double a = 0.0;
  if (slots[si].active) {
    for (int j = 0; j < bps; ++j) {
      const double w = partials[((int64_t)si * bps + j) * 16 + k];
      a = (k == 0 || k == 11) ? fmax(a, w) : a + w;
    }
  }
  out[si * 16 + k] = a;
}

// --------------------------------------------------------------- metrics ---

constexpr int kNMet = 24;
// Index meaning: XADV_M_* in include/xadv.h.  0 sum, 1 max, 2 min.
__constant__ int kMetOp[kNMet] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0};

// Reductions behind compute_macroscopic_metrics (generate_dataset.py:80-179)
// and the pressure_drop column (:556), over one slot's fields.
template <class R>
__global__ __launch_bounds__(256) void k_metrics(Slot<R>* __restrict__ slots, int slot, int x0, int cnx, int cny,
                                                 double* __restrict__ partials) {
  const Slot<R>& S = slots[slot];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kNMet];
  int op[kNMet];
#pragma unroll
  for (int k = 0; k < kNMet; ++k) {
    op[k] = kMetOp[k];
    v[k] = op[k] == 1 ? -INFINITY : (op[k] == 2 ? INFINITY : 0.0);
  }
  if (c < S.ncell) {
    int x, y, z;
    coords<R>(c, S.P, S.nz, x, y, z);
    const int type = (int)(S.flags[c] & F_TYPE);
    const R r = S.rho[c];
    const float ux = (float)S.u[3 * c], uy = (float)S.u[3 * c + 1], uz = (float)S.u[3 * c + 2];
    const float t = (float)S.T[c];
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    if (x == nx - 1 && type == BC_OUTLET) {
      v[0] = ux;
      v[13] = (double)(ux * t);
    }
    if (x == 1) {
      v[5] = r;
      v[7] = 1.0;
      if (type != BC_SOLID) {
        v[1] = r;
        v[2] = 1.0;
      }
    }
    if (x == nx - 2) {
      v[6] = r;
      if (type != BC_SOLID) {
        v[3] = r;
        v[4] = 1.0;
      }
    }
    if (z == 0 && x >= x0 && x < x0 + cnx && y < cny) {
      v[8] = t;
      v[9] = t;
      v[10] = t;
      v[11] = (double)t * (double)t;
      v[12] = 1.0;
    }
    if (type != BC_SOLID && x >= x0 && x < nx - x0 && y >= 1 && y < ny - 1 && z >= 4 && z < nz - 2) {
      v[14] = t;
      v[15] = 1.0;
    }
    v[16] = sqrtf(ux * ux + uy * uy + uz * uz);
    v[17] = t;
    v[18] = t;
    v[19] = 1.0;
    const bool bad = !(isfinite((double)r) && isfinite(ux) && isfinite(uy) && isfinite(uz) && isfinite(t));
    v[20] = bad ? 1.0 : 0.0;
  }
  block_reduce<kNMet>(v, op, partials + (int64_t)blockIdx.x * kNMet);
}

__global__ void k_metrics_final(int nblocks, const double* __restrict__ partials, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= kNMet) return;
  const int o = kMetOp[k];
  double a = o == 1 ? -INFINITY : (o == 2 ? INFINITY : 0.0);
  for (int j = 0; j < nblocks; ++j) {
    const double w = partials[(int64_t)j * kNMet + k];
    a = o == 0 ? a + w : (o == 1 ? fmax(a, w) : fmin(a, w));
  }
  out[k] = a;
}

}  // namespace xadv
