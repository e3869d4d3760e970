// CPU oracle for the coupled D3Q19 fluid + D3Q19 thermal lattice-Boltzmann step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
// solver in xadvcool_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it; the product path never does.
//
// It restates, kernel by kernel and in the reference's operation order, the
// Warp kernels of the reference:
//   src/fluid_engine3d.py   (D3Q19 constants, equilibrium, macroscopic, BGK,
//                            fused collision, streaming with velocity inlet /
//                            pressure outlet / halfway bounce-back, stats)
//   src/thermal_engine3d.py (thermal equilibrium, BGK + source, fused
//                            collision, streaming with Dirichlet inlet /
//                            zero-gradient outlet / adiabatic walls, T macro,
//                            thermal stats)
//   validate_analytical.py  (the six test-only kernels)
// and the six-launch step order of src/simulation.py:94-99.
//
// It also holds the D3Q7 thermal variant, which has no reference kernel (the
// reference's thermal lattice is D3Q19): the same collide / stream / moments
// on the seven velocities 0..6 of E_CONST with weights 1/4 and 1/8, pinned by
// the closed forms of tests/test_oracle_d3q7.py.
//
// Layout is the reference's: populations (19, nx, ny, nz) C-order, fields
// (nx, ny, nz), velocity (nx, ny, nz, 3), mask int32 (nx, ny, nz).
// Every kernel reads one buffer and writes another (no in-place update), as
// the reference does.  Build with -ffp-contract=off so each + and * rounds
// the way the source reads; template on float (the reference's FP32) and
// double (the FP64 error reference).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

// Planted-mistake selector for the mutation tests (see orc_cfg.mutation).
static int g_mutation = 0;

namespace {

enum { BC_FLUID = 0, BC_SOLID = 1, BC_INLET = 2, BC_OUTLET = 3 };  // fluid_engine3d.py:11-14

// fluid_engine3d.py:19-39 (E_CONST), :42-47 (W_CONST), :51-53 (OPPOSITE_CONST)
const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

template <class R> struct Wt {
  R w[19];
  Wt() {
    // The reference builds W_CONST from Python doubles rounded to the lattice
    // precision: 1/3, 1/18 (x6), 1/36 (x12).
    w[0] = (R)(1.0 / 3.0);
    for (int i = 1; i < 7; ++i) w[i] = (R)(1.0 / 18.0);
    for (int i = 7; i < 19; ++i) w[i] = (R)(1.0 / 36.0);
  }
};

struct Dims {
  int nx, ny, nz;
  int64_t n() const { return (int64_t)nx * ny * nz; }
  int64_t c(int x, int y, int z) const { return ((int64_t)x * ny + y) * nz + z; }
};

// ---------------------------------------------------------------- fluid ----

// fluid_engine3d.py:58-80 compute_equilibrium_kernel
template <class R>
void equilibrium(const Dims& d, const R* rho, const R* u, R* feq) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R lr = rho[c];
    R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
    R usq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < 19; ++i) {
      R eu = (R)EX[i] * ux + (R)EY[i] * uy + (R)EZ[i] * uz;
      feq[i * N + c] = W.w[i] * lr * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
    }
  }
}

// fluid_engine3d.py:85-113 compute_macroscopic_kernel
template <class R>
void macroscopic(const Dims& d, const R* f, R* rho, R* u) {
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R lr = 0, mx = 0, my = 0, mz = 0;
    for (int i = 0; i < 19; ++i) {
      R v = f[i * N + c];
      lr += v;
      mx += v * (R)EX[i];
      my += v * (R)EY[i];
      mz += v * (R)EZ[i];
    }
    rho[c] = lr;
    if (lr > (R)1e-10) {
      u[3 * c] = mx / lr;
      u[3 * c + 1] = my / lr;
      u[3 * c + 2] = mz / lr;
    } else {
      u[3 * c] = u[3 * c + 1] = u[3 * c + 2] = 0;
    }
  }
}

// fluid_engine3d.py:118-136 bgk_collision_kernel (f_eq from an array)
template <class R>
void bgk_collision(const Dims& d, const R* fo, const R* feq, R* fpc, const int* mask, R tau) {
  const int64_t N = d.n();
  const R omega = (R)1.0 / tau;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    int m = mask[c];
    if (m == BC_FLUID || m == BC_INLET || m == BC_OUTLET) {
      for (int i = 0; i < 19; ++i) fpc[i * N + c] = fo[i * N + c] - omega * (fo[i * N + c] - feq[i * N + c]);
    } else {
      for (int i = 0; i < 19; ++i) fpc[i * N + c] = fo[i * N + c];
    }
  }
}

// fluid_engine3d.py:141-176 fused_equilibrium_collision_kernel
template <class R>
void fused_collision(const Dims& d, const R* fo, const R* rho, const R* u, R* fpc, const int* mask, R tau) {
  static const Wt<R> W;
  const int64_t N = d.n();
  const R omega = (R)1.0 / tau;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    int m = mask[c];
    if (m == BC_FLUID || m == BC_INLET || m == BC_OUTLET) {
      R lr = rho[c];
      R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
      R usq = ux * ux + uy * uy + uz * uz;
      for (int i = 0; i < 19; ++i) {
        R eu = (R)EX[i] * ux + (R)EY[i] * uy + (R)EZ[i] * uz;
        R feq = W.w[i] * lr * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
        fpc[i * N + c] = fo[i * N + c] - omega * (fo[i * N + c] - feq);
      }
    } else {
      for (int i = 0; i < 19; ++i) fpc[i * N + c] = fo[i * N + c];
    }
  }
}

// validate_analytical.py:377-395 forced_bgk_collision_kernel
template <class R>
void forced_collision(const Dims& d, const R* fo, const R* feq, R* fn, const int* mask, R tau, R force_x) {
  static const Wt<R> W;
  const int64_t N = d.n();
  const R omega = (R)1.0 / tau;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    if (mask[c] != BC_SOLID) {
      for (int i = 0; i < 19; ++i) {
        R src = W.w[i] * (R)3.0 * (R)EX[i] * force_x;
        fn[i * N + c] = fo[i * N + c] - omega * (fo[i * N + c] - feq[i * N + c]) + src;
      }
    } else {
      for (int i = 0; i < 19; ++i) fn[i * N + c] = fo[i * N + c];
    }
  }
}

// fluid_engine3d.py:181-272 streaming_kernel
template <class R>
void streaming(const Dims& d, const R* fpc, R* fs, const int* mask, R ux_in, R rho_out) {
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z) {
        const int64_t c = d.c(x, y, z);
        const int m = mask[c];
        if (m == BC_SOLID) continue;  // :193-194
        if (m == BC_INLET) {          // :197-244
          R f[19];
          for (int i = 0; i < 19; ++i) f[i] = fpc[i * N + c];
          R rho_new = (f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18] +
                       (R)2.0 * (f[2] + f[8] + f[10] + f[12] + f[14])) /
                      ((R)1.0 - ux_in);
          const R c23 = (R)(2.0 / 3.0), c16 = (R)(1.0 / 6.0);
          R o[19];
          for (int i = 0; i < 19; ++i) o[i] = f[i];
          o[1] = f[2] + c23 * rho_new * ux_in;
          o[7] = f[10] + c16 * rho_new * ux_in;
          o[9] = f[8] + c16 * rho_new * ux_in;
          o[11] = f[14] + c16 * rho_new * ux_in;
          o[13] = f[12] + c16 * rho_new * ux_in;
          for (int i = 0; i < 19; ++i) fs[i * N + c] = o[i];
          continue;
        }
        if (m == BC_OUTLET) {  // :247-259
          int sx = x - 1;
          if (sx < 0) sx = 0;
          const int64_t s = d.c(sx, y, z);
          R rs = 0;
          for (int i = 0; i < 19; ++i) {
            R v = fpc[i * N + s];
            fs[i * N + c] = v;
            rs += v;
          }
          if (rs > (R)1e-6) {
            R scale = rho_out / rs;
            for (int i = 0; i < 19; ++i) fs[i * N + c] *= scale;
          }
          continue;
        }
        for (int i = 0; i < 19; ++i) {  // :262-272
          int sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
          if (sx < 0 || sx >= d.nx || sy < 0 || sy >= d.ny || sz < 0 || sz >= d.nz) {
            fs[i * N + c] = fpc[OPP[i] * N + c];
            continue;
          }
          if (mask[d.c(sx, sy, sz)] == BC_SOLID)
            fs[i * N + c] = fpc[OPP[i] * N + c];
          else
            fs[i * N + c] = fpc[i * N + d.c(sx, sy, sz)];
        }
      }
}

// validate_analytical.py:255-273 periodic_streaming_kernel
template <class R>
void periodic_streaming(const Dims& d, const R* fpc, R* fs) {
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z)
        for (int i = 0; i < 19; ++i) {
          int sx = (x - EX[i] + d.nx) % d.nx, sy = (y - EY[i] + d.ny) % d.ny, sz = (z - EZ[i] + d.nz) % d.nz;
          fs[i * N + d.c(x, y, z)] = fpc[i * N + d.c(sx, sy, sz)];
        }
}

// validate_analytical.py:397-420 periodic_streaming_with_solids_kernel
template <class R>
void periodic_streaming_solids(const Dims& d, const R* fpc, R* fs, const int* mask) {
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z) {
        const int64_t c = d.c(x, y, z);
        if (mask[c] == BC_SOLID) continue;
        for (int i = 0; i < 19; ++i) {
          int sx = (x - EX[i] + d.nx) % d.nx, sy = (y - EY[i] + d.ny) % d.ny, sz = (z - EZ[i] + d.nz) % d.nz;
          int64_t s = d.c(sx, sy, sz);
          if (mask[s] == BC_SOLID)
            fs[i * N + c] = fpc[(g_mutation == 2 ? i : OPP[i]) * N + c];
          else
            fs[i * N + c] = fpc[i * N + s];
        }
      }
}

// validate_analytical.py:502-538 couette_streaming_kernel
template <class R>
void couette_streaming(const Dims& d, const R* fpc, R* fs, const int* mask, R u_wall_x) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z) {
        const int64_t c = d.c(x, y, z);
        if (mask[c] == BC_SOLID) continue;
        for (int i = 0; i < 19; ++i) {
          int sx = (x - EX[i] + d.nx) % d.nx, sy = (y - EY[i] + d.ny) % d.ny, sz = z - EZ[i];
          if (sz < 0) {
            fs[i * N + c] = fpc[OPP[i] * N + c];
          } else if (sz >= d.nz) {
            R mom = (R)6.0 * W.w[i] * (R)EX[i] * u_wall_x;
            if (g_mutation == 3) mom = -mom;
            fs[i * N + c] = fpc[OPP[i] * N + c] + mom;
          } else {
            int64_t s = d.c(sx, sy, sz);
            if (mask[s] == BC_SOLID)
              fs[i * N + c] = fpc[OPP[i] * N + c];
            else
              fs[i * N + c] = fpc[i * N + s];
          }
        }
      }
}

// -------------------------------------------------------------- thermal ----

// thermal_engine3d.py:7-47 compute_thermal_equilibrium_kernel (rho unused)
template <class R>
void thermal_equilibrium(const Dims& d, const R* u, const R* T, R* geq, const int* mask) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R lT = T[c];
    R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
    if (mask[c] == BC_SOLID) ux = uy = uz = 0;
    R usq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < 19; ++i) {
      R eu = (R)EX[i] * ux + (R)EY[i] * uy + (R)EZ[i] * uz;
      geq[i * N + c] = W.w[i] * lT * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
    }
  }
}

// thermal_engine3d.py:49-86 thermal_bgk_collision_kernel
template <class R>
void thermal_bgk(const Dims& d, const R* go, const R* geq, R* gn, const int* mask, R tau_f, R tau_s,
                 const R* hs, R power) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R tau = (mask[c] == BC_SOLID) ? tau_s : tau_f;
    R omega = (R)1.0 / tau;
    R q = hs[c] * power;
    for (int i = 0; i < 19; ++i) {
      R v = go[i * N + c] - omega * (go[i * N + c] - geq[i * N + c]);
      if (q > (R)0.0) v += (g_mutation == 1) ? W.w[i] * q : W.w[i] * ((R)1.0 - (R)0.5 * omega) * q;
      gn[i * N + c] = v;
    }
  }
}

// thermal_engine3d.py:91-145 fused_thermal_equilibrium_collision_kernel
template <class R>
void fused_thermal_collision(const Dims& d, const R* go, const R* u, const R* T, R* gn, const int* mask, R tau_f,
                             R tau_s, const R* hs, R power) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    const int m = mask[c];
    R tau = (m == BC_SOLID) ? tau_s : tau_f;
    R omega = (R)1.0 / tau;
    R q = hs[c] * power;
    R lT = T[c];
    R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
    if (m == BC_SOLID) ux = uy = uz = 0;
    R usq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < 19; ++i) {
      R eu = (R)EX[i] * ux + (R)EY[i] * uy + (R)EZ[i] * uz;
      R geq = W.w[i] * lT * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
      R v = go[i * N + c] - omega * (go[i * N + c] - geq);
      if (q > (R)0.0) v += W.w[i] * ((R)1.0 - (R)0.5 * omega) * q;
      gn[i * N + c] = v;
    }
  }
}

// thermal_engine3d.py:147-211 thermal_streaming_kernel
template <class R>
void thermal_streaming(const Dims& d, const R* gpc, R* gs, const int* mask, R T_in) {
  static const Wt<R> W;
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z) {
        const int64_t c = d.c(x, y, z);
        const int m = mask[c];
        if (m == BC_INLET) {  // :160-165
          for (int i = 0; i < 19; ++i) gs[i * N + c] = W.w[i] * T_in;
          continue;
        }
        if (m == BC_OUTLET) {  // :170-177
          int sx = x - 1;
          if (sx < 0) sx = 0;
          const int64_t s = d.c(sx, y, z);
          for (int i = 0; i < 19; ++i) gs[i * N + c] = gpc[i * N + s];
          continue;
        }
        for (int i = 0; i < 19; ++i) {  // :180-211
          int sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
          if (sx < 0 || sx >= d.nx || sy < 0 || sy >= d.ny || sz < 0 || sz >= d.nz) {
            if (sx < 0)
              gs[i * N + c] = W.w[i] * T_in;
            else if (sx >= d.nx)
              gs[i * N + c] = gpc[i * N + c];
            else
              gs[i * N + c] = gpc[OPP[i] * N + c];
          } else {
            gs[i * N + c] = gpc[i * N + d.c(sx, sy, sz)];
          }
        }
      }
}

// validate_analytical.py:607-621 periodic_thermal_streaming_kernel
template <class R>
void periodic_thermal_streaming(const Dims& d, const R* gpc, R* gs) { periodic_streaming<R>(d, gpc, gs); }

// validate_analytical.py:21-30 dirichlet_thermal_bc_kernel
template <class R>
void dirichlet_plane(const Dims& d, R* g, R T_val, int xp) {
  static const Wt<R> W;
  const int64_t N = d.n();
  for (int y = 0; y < d.ny; ++y)
    for (int z = 0; z < d.nz; ++z)
      for (int i = 0; i < 19; ++i) g[i * N + d.c(xp, y, z)] = W.w[i] * T_val;
}

// thermal_engine3d.py:214-224 compute_thermal_macroscopic_kernel
template <class R>
void thermal_macroscopic(const Dims& d, const R* g, R* T) {
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R s = 0;
    for (int i = 0; i < 19; ++i) s += g[i * N + c];
    T[c] = s;
  }
}

// ------------------------------------------------------ D3Q7 thermal variant ----
// Not a reference kernel: the reference's thermal lattice is D3Q19.  The same
// advection-diffusion problem on the rest and six face velocities (directions
// 0..6 of E_CONST, with their OPPOSITE_CONST pairs), weights 1/4 and 1/8
// (c_s^2 = 1/4), the linear equilibrium g_eq_i = w_i T (1 + 4 e_i.u) (the
// quadratic terms of the D3Q19 equilibrium are not isotropic on seven
// velocities), and the streaming rules of thermal_engine3d.py:147-211 for those
// seven directions.  The relaxation times passed in are the reference's (D3Q19,
// diffusivity (tau - 1/2) / 3): the D3Q7 step relaxes with
// tau7 = 1/2 + 4/3 (tau - 1/2), the same diffusivity, and injects the source as
// w_i (1 - 1/(2 tau)) q, the heat the D3Q19 step adds per step (the factor
// depends on tau, so with (1 - 1/(2 tau7)) the two lattices would heat
// differently).  Planted mistakes: 1 drops (1 - omega/2) from the
// source, 2 bounces back the same direction at walls, 4 takes 1 + 3 e.u (the
// D3Q19 factor), 5 takes the weights 1/3 and 1/9.

template <class R> struct Wt7 {
  R w[7];
  Wt7() {
    const bool m5 = g_mutation == 5;
    w[0] = m5 ? (R)(1.0 / 3.0) : (R)0.25;
    for (int i = 1; i < 7; ++i) w[i] = m5 ? (R)(1.0 / 9.0) : (R)0.125;
  }
};

template <class R>
inline R geq7(const Wt7<R>& W, int i, R T, R ux, R uy, R uz) {
  const R eu = (R)EX[i] * ux + (R)EY[i] * uy + (R)EZ[i] * uz;
  return W.w[i] * T * ((R)1.0 + (g_mutation == 4 ? (R)3.0 : (R)4.0) * eu);
}

template <class R>
void thermal_equilibrium7(const Dims& d, const R* u, const R* T, R* geq, const int* mask) {
  const Wt7<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
    if (mask[c] == BC_SOLID) ux = uy = uz = 0;
    for (int i = 0; i < 7; ++i) geq[i * N + c] = geq7(W, i, T[c], ux, uy, uz);
  }
}

template <class R>
void fused_thermal_collision7(const Dims& d, const R* go, const R* u, const R* T, R* gn, const int* mask, R tau_f,
                              R tau_s, const R* hs, R power) {
  const Wt7<R> W;
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    const int m = mask[c];
    R tau = (m == BC_SOLID) ? tau_s : tau_f;
    R omega = (R)1.0 / ((R)0.5 + (tau - (R)0.5) * (R)(4.0 / 3.0));
    R sf = (R)1.0 - (R)0.5 * ((R)1.0 / tau);
    R q = hs[c] * power;
    R ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
    if (m == BC_SOLID) ux = uy = uz = 0;
    for (int i = 0; i < 7; ++i) {
      R v = go[i * N + c] - omega * (go[i * N + c] - geq7(W, i, T[c], ux, uy, uz));
      if (q > (R)0.0) v += (g_mutation == 1) ? W.w[i] * q : W.w[i] * sf * q;
      gn[i * N + c] = v;
    }
  }
}

template <class R>
void thermal_streaming7(const Dims& d, const R* gpc, R* gs, const int* mask, R T_in) {
  const Wt7<R> W;
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z) {
        const int64_t c = d.c(x, y, z);
        const int m = mask[c];
        if (m == BC_INLET) {
          for (int i = 0; i < 7; ++i) gs[i * N + c] = W.w[i] * T_in;
          continue;
        }
        if (m == BC_OUTLET) {
          const int64_t s = d.c(x > 0 ? x - 1 : 0, y, z);
          for (int i = 0; i < 7; ++i) gs[i * N + c] = gpc[i * N + s];
          continue;
        }
        for (int i = 0; i < 7; ++i) {
          int sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
          if (sx < 0 || sx >= d.nx || sy < 0 || sy >= d.ny || sz < 0 || sz >= d.nz) {
            if (sx < 0)
              gs[i * N + c] = W.w[i] * T_in;
            else if (sx >= d.nx)
              gs[i * N + c] = gpc[i * N + c];
            else
              gs[i * N + c] = gpc[(g_mutation == 2 ? i : OPP[i]) * N + c];
          } else {
            gs[i * N + c] = gpc[i * N + d.c(sx, sy, sz)];
          }
        }
      }
}

template <class R>
void periodic_thermal_streaming7(const Dims& d, const R* gpc, R* gs) {
  const int64_t N = d.n();
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < d.nx; ++x)
    for (int y = 0; y < d.ny; ++y)
      for (int z = 0; z < d.nz; ++z)
        for (int i = 0; i < 7; ++i) {
          const int sx = (x - EX[i] + d.nx) % d.nx, sy = (y - EY[i] + d.ny) % d.ny, sz = (z - EZ[i] + d.nz) % d.nz;
          gs[i * N + d.c(x, y, z)] = gpc[i * N + d.c(sx, sy, sz)];
        }
}

template <class R>
void thermal_macroscopic7(const Dims& d, const R* g, R* T) {
  const int64_t N = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < N; ++c) {
    R s = 0;
    for (int i = 0; i < 7; ++i) s += g[i * N + c];
    T[c] = s;
  }
}

}  // namespace

// ------------------------------------------------------------------ C ABI ----

extern "C" {

// Step configuration.  Fluid and thermal schedules follow the reference's
// callers: production = src/simulation.py:94-99; the test schedules follow the
// graphs and loops of validate_analytical.py (cited per field).
typedef struct {
  int nx, ny, nz;
  // fluid: 0 off; 1 = simulation.py schedule (fused collide, stream, macro);
  //        2 = validate_analytical.py schedule (bgk or forced collide with an
  //            f_eq array, stream, macro, equilibrium).
  int fluid;
  // fluid streaming: 0 streaming_kernel (channel), 1 periodic_streaming_kernel,
  //                  2 periodic_streaming_with_solids_kernel, 3 couette_streaming_kernel
  int fluid_stream;
  int forced;  // 1: forced_bgk_collision_kernel (fluid == 2 only)
  // thermal: 0 off; 1 = simulation.py schedule (fused thermal collide with the
  //          velocity just computed, stream, macro); 2 = validate_analytical.py
  //          schedule (thermal equilibrium, thermal BGK, stream, [dirichlet], macro)
  //          with a frozen velocity field; 3 = schedule 1 on the D3Q7 variant
  //          (populations in planes 0..6 of g); 4 = the D3Q7 collide, stream
  //          (thermal_stream 1: periodic) and macro with the velocity field as given.
  int thermal;
  int thermal_stream;  // 0 thermal_streaming_kernel, 1 periodic_thermal_streaming_kernel
  int dirichlet_x;     // -1: none; else dirichlet_thermal_bc_kernel plane after streaming
  double tau, force_x, u_wall, u_in, rho_out;
  double tau_th_fluid, tau_th_solid, heat_power, T_in, dirichlet_T;
  // Planted mistakes, for tests that show the closed-form cases catch them:
  // 1 drops the (1 - omega/2) factor of the thermal source, 2 bounces back
  // the same direction instead of the opposite one, 3 flips the sign of the
  // moving-wall momentum; D3Q7: 4 takes 1 + 3 e.u in the equilibrium, 5 the
  // weights 1/3 and 1/9.  0 = the reference.
  int mutation;
} orc_cfg;
}  // extern "C"

template <class R>
static int run_steps(const orc_cfg* cfg, int nsteps, const int* mask, const R* hs, R* f_old, R* f_new, R* f_eq,
                     R* rho, R* u, R* g_old, R* g_new, R* g_eq, R* T) {
  Dims d{cfg->nx, cfg->ny, cfg->nz};
  g_mutation = cfg->mutation;
  const R tau = (R)cfg->tau, ttf = (R)cfg->tau_th_fluid, tts = (R)cfg->tau_th_solid;
  const R P = (R)cfg->heat_power, Tin = (R)cfg->T_in;
  for (int s = 0; s < nsteps; ++s) {
    if (cfg->fluid == 1) {
      // simulation.py:94-96
      fused_collision<R>(d, f_old, rho, u, f_new, mask, tau);
      streaming<R>(d, f_new, f_old, mask, (R)cfg->u_in, (R)cfg->rho_out);
      macroscopic<R>(d, f_old, rho, u);
    } else if (cfg->fluid == 2) {
      // validate_analytical.py graphs, e.g. :157-161, :337-341, :468-472, :570-574
      if (cfg->forced)
        forced_collision<R>(d, f_old, f_eq, f_new, mask, tau, (R)cfg->force_x);
      else
        bgk_collision<R>(d, f_old, f_eq, f_new, mask, tau);
      switch (cfg->fluid_stream) {
        case 0: streaming<R>(d, f_new, f_old, mask, (R)cfg->u_in, (R)cfg->rho_out); break;
        case 1: periodic_streaming<R>(d, f_new, f_old); break;
        case 2: periodic_streaming_solids<R>(d, f_new, f_old, mask); break;
        case 3: couette_streaming<R>(d, f_new, f_old, mask, (R)cfg->u_wall); break;
        default: return -1;
      }
      macroscopic<R>(d, f_old, rho, u);
      equilibrium<R>(d, rho, u, f_eq);
    }
    if (cfg->thermal == 1) {
      // simulation.py:97-99
      fused_thermal_collision<R>(d, g_old, u, T, g_new, mask, ttf, tts, hs, P);
      thermal_streaming<R>(d, g_new, g_old, mask, Tin);
      thermal_macroscopic<R>(d, g_old, T);
    } else if (cfg->thermal == 2) {
      // validate_analytical.py:86-90, :680-684, :776-788, :886-892
      thermal_equilibrium<R>(d, u, T, g_eq, mask);
      thermal_bgk<R>(d, g_old, g_eq, g_new, mask, ttf, tts, hs, P);
      if (cfg->thermal_stream == 1)
        periodic_thermal_streaming<R>(d, g_new, g_old);
      else
        thermal_streaming<R>(d, g_new, g_old, mask, Tin);
      if (cfg->dirichlet_x >= 0) dirichlet_plane<R>(d, g_old, (R)cfg->dirichlet_T, cfg->dirichlet_x);
      thermal_macroscopic<R>(d, g_old, T);
    } else if (cfg->thermal == 3 || cfg->thermal == 4) {
      fused_thermal_collision7<R>(d, g_old, u, T, g_new, mask, ttf, tts, hs, P);
      if (cfg->thermal == 4 && cfg->thermal_stream == 1)
        periodic_thermal_streaming7<R>(d, g_new, g_old);
      else
        thermal_streaming7<R>(d, g_new, g_old, mask, Tin);
      thermal_macroscopic7<R>(d, g_old, T);
    }
  }
  return 0;
}

extern "C" {
int orc_run_f32(const orc_cfg* cfg, int nsteps, const int* mask, const float* hs, float* f_old, float* f_new,
                float* f_eq, float* rho, float* u, float* g_old, float* g_new, float* g_eq, float* T) {
  return run_steps<float>(cfg, nsteps, mask, hs, f_old, f_new, f_eq, rho, u, g_old, g_new, g_eq, T);
}
int orc_run_f64(const orc_cfg* cfg, int nsteps, const int* mask, const double* hs, double* f_old, double* f_new,
                double* f_eq, double* rho, double* u, double* g_old, double* g_new, double* g_eq, double* T) {
  return run_steps<double>(cfg, nsteps, mask, hs, f_old, f_new, f_eq, rho, u, g_old, g_new, g_eq, T);
}

// Single kernels, for kernel-level tests.
void orc_equilibrium_f32(int nx, int ny, int nz, const float* rho, const float* u, float* feq) {
  equilibrium<float>(Dims{nx, ny, nz}, rho, u, feq);
}
void orc_equilibrium_f64(int nx, int ny, int nz, const double* rho, const double* u, double* feq) {
  equilibrium<double>(Dims{nx, ny, nz}, rho, u, feq);
}
void orc_thermal_equilibrium_f32(int nx, int ny, int nz, const float* u, const float* T, float* geq,
                                 const int* mask) {
  thermal_equilibrium<float>(Dims{nx, ny, nz}, u, T, geq, mask);
}
void orc_thermal_equilibrium_f64(int nx, int ny, int nz, const double* u, const double* T, double* geq,
                                 const int* mask) {
  thermal_equilibrium<double>(Dims{nx, ny, nz}, u, T, geq, mask);
}
void orc_thermal_equilibrium7_f32(int nx, int ny, int nz, const float* u, const float* T, float* geq,
                                  const int* mask) {
  thermal_equilibrium7<float>(Dims{nx, ny, nz}, u, T, geq, mask);
}
void orc_thermal_equilibrium7_f64(int nx, int ny, int nz, const double* u, const double* T, double* geq,
                                  const int* mask) {
  thermal_equilibrium7<double>(Dims{nx, ny, nz}, u, T, geq, mask);
}
void orc_macroscopic_f32(int nx, int ny, int nz, const float* f, float* rho, float* u) {
  macroscopic<float>(Dims{nx, ny, nz}, f, rho, u);
}
void orc_thermal_macroscopic_f32(int nx, int ny, int nz, const float* g, float* T) {
  thermal_macroscopic<float>(Dims{nx, ny, nz}, g, T);
}

// fluid_engine3d.py:277-312 compute_stats_kernel and thermal_engine3d.py:226-243
// compute_thermal_stats_kernel.  The reference accumulates with unordered FP32
// atomics; here the per-cell terms are formed in FP32 exactly as there and
// summed in order in FP64.  out[0..10] = fluid slots, out[11..15] = thermal.
void orc_stats_f32(int nx, int ny, int nz, const float* rho, const float* u, const float* T, const int* mask,
                   double* out) {
  Dims d{nx, ny, nz};
  for (int k = 0; k < 16; ++k) out[k] = 0.0;
  double vmax = 0.0, tmax = -INFINITY;
  int first_t = 1;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        int64_t c = d.c(x, y, z);
        int m = mask[c];
        float ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
        float vsq = ux * ux + uy * uy + uz * uz;
        float vmag = std::sqrt(vsq);
        float r = rho[c];
        out[2] += vsq;
        if (m != BC_SOLID) {
          if (vmag > vmax) vmax = vmag;
          out[1] += vmag;
          out[3] += 1.0;
          out[4] += r;
        }
        if (x == 1) { out[5] += r; out[8] += 1.0; }
        if (x == nx - 2) { out[6] += r; out[9] += 1.0; }
        if (m == BC_OUTLET) { out[7] += ux; out[10] += 1.0; }
        float t = T[c];
        if (first_t || t > tmax) { tmax = t; first_t = 0; }
        out[12] += t;
        out[13] += 1.0;
        if (m == BC_OUTLET) { out[14] += (double)(ux * t); out[15] += 1.0; }
      }
  out[0] = vmax;
  // Warp's atomic_max starts from the zeroed slot (simulation.py:124-125).
  out[11] = tmax > 0.0 ? tmax : 0.0;
}

}  // extern "C"

#include <omp.h>
extern "C" {
// Thread count for the OpenMP loops above (1 = the scalar schedule).
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }
}
