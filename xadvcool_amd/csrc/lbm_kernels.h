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
//    of different sizes advance in one launch.  The block order is linear or
//    XCD-contiguous (each XCD takes one eighth of every slot), forward or
//    reversed on even passes (lattice.h, launch_block).
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
  uint32_t nf;        // plane stride of f: npad (dense) or the padded non-solid count (compact)
  uint32_t nz_magic;  // c / nz = div_nz(c, nz, nz_magic) (compact storage)
  R* f;
  R* g;
  const ColInfo* cols;  // compact fluid storage: padded column table at lattice column 0, else null
  const uint32_t* flags;
  const R* q;
  R* fin;    // [19][P]   INLET own post-collision set
  R* fsnap;  // [2][19][P] fluid post-collision sets of the cells left of OUTLET
  R* gsnap;  // [2][19][P] thermal ditto
  R* gxh;    // [5][P]    thermal e_x=-1 post-collision values on x = nx-1
  R* rho;
  R* u;  // (ncell, 3)
  R* T;
  R* g7;  // D3Q7 thermal variant: [2][7][npad], the two buffers a step reads and writes (null otherwise)
};

// Lattice block of this launch block (block order bits: lattice.h,
// launch_block); values >= G are padding blocks.
__device__ __forceinline__ int map_block(int bps, int order) {
  return launch_block((int)blockIdx.x, (int)gridDim.x, bps, order);
}

// e_i . u with the zero components left out.  The reference forms
// ex*ux + ey*uy + ez*uz; a 0*u term only adds a signed zero, which cannot
// change a non-zero sum and disappears in 1 + 3*eu, so the result is the
// reference's (the parity tests check the bits).  Non-zero terms keep the
// reference's x, y, z order; -1 factors are free negation modifiers.
template <class R, int i>
__device__ __forceinline__ R edot(R ux, R uy, R uz) {
  constexpr int a = ex_of(i), b = ey_of(i), c = ez_of(i);
  const R tx = a > 0 ? ux : -ux, ty = b > 0 ? uy : -uy, tz = c > 0 ? uz : -uz;
  if (a != 0 && b != 0) return tx + ty;
  if (a != 0 && c != 0) return tx + tz;
  if (b != 0 && c != 0) return ty + tz;
  if (a != 0) return tx;
  if (b != 0) return ty;
  if (c != 0) return tz;
  return (R)0;
}

template <class R, int i>
__device__ __forceinline__ R feq_i(R rho, R ux, R uy, R uz, R usq) {
  // src/fluid_engine3d.py:74-78 (same expression for the thermal lattice with T)
  const R eu = edot<R, i>(ux, uy, uz);
  return (R)w_of(i) * rho * ((R)1.0 + (R)3.0 * eu + (R)4.5 * (eu * eu) - (R)1.5 * usq);
}

template <class R>
__device__ __forceinline__ R wall_mom(int i, R u_wall) {
  // validate_analytical.py:531-532
  return (R)6.0 * (R)w_of(i) * (R)ex_of(i) * u_wall;
}

// Loads and stores through the global address space with a wave-uniform
// base (SGPRs) and a 32-bit per-lane byte offset, so the compiler emits
// global_load/store with saddr instead of FLAT with 64-bit address math.
// Every lattice array of a slot stays below 4 GiB (checked on the host).
#define XADV_GLOBAL __attribute__((address_space(1)))
template <class R>
__device__ __forceinline__ R ldb(const void* base, uint32_t byte_off) {
  // C-style cast: an address-space cast from the generic pointer.
  const XADV_GLOBAL char* g = (const XADV_GLOBAL char*)base;
  return *(const XADV_GLOBAL R*)(g + byte_off);
}
template <class R>
__device__ __forceinline__ void stb(void* base, uint32_t byte_off, R v) {
  XADV_GLOBAL char* g = (XADV_GLOBAL char*)base;
  *(XADV_GLOBAL R*)(g + byte_off) = v;
}
template <class R>
__device__ __forceinline__ R ld(const R* base, uint32_t idx) {
  return ldb<R>(base, idx * (uint32_t)sizeof(R));
}
template <class R>
__device__ __forceinline__ void st(R* base, uint32_t idx, R v) {
  stb<R>(base, idx * (uint32_t)sizeof(R), v);
}
// All-ones if bit `b` of w is set, else 0 (one v_bfe_i32).
__device__ __forceinline__ uint32_t bitmask(uint32_t w, int b) { return (uint32_t)(((int32_t)(w << (31 - b))) >> 31); }
// All-ones if any bit of m is set in w, else 0.
__device__ __forceinline__ uint32_t anymask(uint32_t w, uint32_t m) { return (w & m) ? 0xffffffffu : 0u; }

// Element c + ka (lanes with sel == 0) or c + kb (lanes with sel == ~0) of A,
// where ka and kb are wave-uniform and KB_GT says kb > ka.  The base is the
// lower of the two, so the per-lane byte offset is non-negative and below the
// array's size (< 4 GiB).
template <class R, bool KB_GT>
__device__ __forceinline__ R ld2(const R* A, int64_t ka, int64_t kb, uint32_t c, uint32_t sel) {
  constexpr uint32_t rs = sizeof(R);
  const char* a = reinterpret_cast<const char*>(A);
  if (KB_GT) return ldb<R>(a + ka * (int64_t)rs, c * rs + (sel & (uint32_t)((kb - ka) * rs)));
  return ldb<R>(a + kb * (int64_t)rs, c * rs + (~sel & (uint32_t)((ka - kb) * rs)));
}
template <class R, bool KB_GT>
__device__ __forceinline__ void st2(R* A, int64_t ka, int64_t kb, uint32_t c, uint32_t sel, R v) {
  constexpr uint32_t rs = sizeof(R);
  char* a = reinterpret_cast<char*>(A);
  if (KB_GT)
    stb<R>(a + ka * (int64_t)rs, c * rs + (sel & (uint32_t)((kb - ka) * rs)), v);
  else
    stb<R>(a + kb * (int64_t)rs, c * rs + (~sel & (uint32_t)((ka - kb) * rs)), v);
}

template <int i>
__device__ __forceinline__ int nb_off(int P, int nz) {
  return ex_of(i) * P + ey_of(i) * nz + ez_of(i);
}

// Index of c + s*e_i with periodic wrap on the periodic axes (general mode).
// On a non-periodic axis the result may leave the lattice; callers only use it
// where the precomputed bits say the neighbour exists.
template <int i>
__device__ __forceinline__ uint32_t nb_wrap(int x, int y, int z, int s, int nx, int ny, int nz) {
  int sx = x + s * ex_of(i), sy = y + s * ey_of(i), sz = z + s * ez_of(i);
  sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
  sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
  sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
  return (uint32_t)((sx * ny + sy) * nz + sz);
}

// Face bits under which the source x - e_i of direction i lies outside a
// non-periodic axis (the face flags are only set on non-periodic axes).
__host__ __device__ constexpr uint32_t src_face_mask(int i) {
  return (ex_of(i) == 1 ? F_FX0 : 0u) | (ex_of(i) == -1 ? F_FXN : 0u) | (ey_of(i) == 1 ? F_FY0 : 0u) |
         (ey_of(i) == -1 ? F_FYN : 0u) | (ez_of(i) == 1 ? F_FZ0 : 0u) | (ez_of(i) == -1 ? F_FZN : 0u);
}

// Per-cell context of one thread's cell.
struct Cell {
  uint32_t c;   // cell index within the slot
  uint32_t fl;  // flag word
  int type;
  int x, y, z;  // coordinates (general mode only)
  uint32_t j4;  // compact fluid storage: the cell's number among non-solid cells, in bytes
  uint32_t jp;  // j4 + one fluid plane in bytes (odd passes)
  bool valid;
};

__device__ __forceinline__ void coords(uint32_t c, int P, int nz, int& x, int& y, int& z) {
  x = (int)(c / (uint32_t)P);
  const int r = (int)(c - (uint32_t)x * (uint32_t)P);
  y = r / nz;
  z = r - y * nz;
}

// ------------------------------------------------------------------ fluid ---

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ R fluid_load(const R* __restrict__ A, uint32_t N, const Cell& k, int P, int nz, int nx,
                                        int ny, R u_wall) {
  constexpr uint32_t rs = sizeof(R);
  if (NODD == 0 || i == 0) return ldb<R>(reinterpret_cast<const char*>(A) + (size_t)i * N * rs, k.c * rs);
  if (!GEN) {
    // Pull from the neighbour's opposite slot, or (bounce-back) the own slot i.
    const int64_t pull = (int64_t)opp_of(i) * N - nb_off<i>(P, nz);
    return ld2<R, (i > opp_of(i))>(A, pull, (int64_t)i * N, k.c, bitmask(k.fl, F_BOUNCE0 + i));
  }
  const bool bb = (k.fl >> (F_BOUNCE0 + i)) & 1u;
  const uint32_t src = nb_wrap<i>(k.x, k.y, k.z, -1, nx, ny, nz);
  const uint32_t a = bb ? ((uint32_t)i * N + k.c) : ((uint32_t)opp_of(i) * N + src);
  R v = ld(A, a);
  if (ez_of(i) == -1) {
    if (bb && (k.fl & F_FZN)) v = v + wall_mom<R>(i, u_wall);
  }
  return v;
}

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ void fluid_store(R* __restrict__ A, uint32_t N, const Cell& k, int P, int nz, int nx,
                                            int ny, R u_wall, R v) {
  constexpr uint32_t rs = sizeof(R);
  if (NODD == 0 || i == 0) {
    stb<R>(reinterpret_cast<char*>(A) + (size_t)opp_of(i) * N * rs, k.c * rs, v);
    return;
  }
  if (!GEN) {
    // Push into the neighbour's slot i, or (target outside / SOLID) into the
    // own opposite slot for the next pass's bounce-back.
    const int64_t push = (int64_t)i * N + nb_off<i>(P, nz);
    st2<R, (opp_of(i) > i)>(A, push, (int64_t)opp_of(i) * N, k.c, bitmask(k.fl, F_BOUNCE0 + opp_of(i)), v);
    return;
  }
  const bool rb = (k.fl >> (F_BOUNCE0 + opp_of(i))) & 1u;  // target c + e_i is outside or SOLID
  const uint32_t dst = nb_wrap<i>(k.x, k.y, k.z, +1, nx, ny, nz);
  if (ez_of(i) == 1) {
    if (rb && (k.fl & F_FZN)) v = v + wall_mom<R>(opp_of(i), u_wall);
  }
  const uint32_t a = rb ? ((uint32_t)opp_of(i) * N + k.c) : ((uint32_t)i * N + dst);
  st(A, a, v);
}

// ------------------------------------------------- compact fluid storage ---
//
// Same AA slots as above, addressed by non-solid cell number: slot (i, cell)
// lives at element i * nf + number(cell).  No non-solid cell ever reads or
// writes a slot of a SOLID cell (those accesses are the bounce-back redirects
// to the cell's own slots), so dropping the SOLID cells changes no value.

// The cell's own byte offset and, for odd passes, the offsets of its 18
// neighbours (JB[k]: neighbour c + e_k, plus one plane for even k).
template <class R, int NODD>
__device__ __forceinline__ void compact_offsets(const Slot<R>& S, Cell& k, uint32_t* JB) {
  const uint32_t nz = (uint32_t)S.nz;
  const uint32_t q = div_nz(k.c, nz, S.nz_magic);
  const int z = (int)(k.c - q * nz);
  if (!NODD) {
    k.j4 = own_offset<sizeof(R)>(S.cols, q, z);
    return;
  }
  const uint32_t NFB = S.nf * (uint32_t)sizeof(R);
  neighbour_offsets<sizeof(R)>(S.cols, q, z, S.ny, NFB, k.j4, JB);
  k.jp = k.j4 + NFB;
}

// Odd passes address the two planes of a direction pair from the lower one,
// lo = min(i, opp(i)): slot (i, own) is j4 (i odd) or jp (i even) from plane
// lo, slot (opp(i), neighbour c - e_i) is JB[opp(i)].  One bit-extract and
// one bitfield insert per access.
template <class R, int NODD, int i>
__device__ __forceinline__ R fluid_load_c(const R* __restrict__ A, uint32_t NF, const Cell& k, const uint32_t* JB) {
  constexpr uint32_t rs = sizeof(R);
  const char* Ab = reinterpret_cast<const char*>(A);
  if (NODD == 0 || i == 0) return ldb<R>(Ab + (size_t)i * NF * rs, k.j4);
  // Pull from the neighbour's opposite slot, or (bounce-back) the own slot i.
  constexpr int lo = i < opp_of(i) ? i : opp_of(i);
  const uint32_t own = (i & 1) ? k.j4 : k.jp;
  const uint32_t bb = bitmask(k.fl, F_BOUNCE0 + i);
  return ldb<R>(Ab + (size_t)lo * NF * rs, (bb & own) | (~bb & JB[opp_of(i)]));
}

template <class R, int NODD, int i>
__device__ __forceinline__ void fluid_store_c(R* __restrict__ A, uint32_t NF, const Cell& k, const uint32_t* JB, R v) {
  constexpr uint32_t rs = sizeof(R);
  char* Ab = reinterpret_cast<char*>(A);
  if (NODD == 0 || i == 0) {
    stb<R>(Ab + (size_t)opp_of(i) * NF * rs, k.j4, v);
    return;
  }
  // Push into the neighbour's slot i, or (target outside / SOLID) the own slot opp(i).
  constexpr int lo = i < opp_of(i) ? i : opp_of(i);
  const uint32_t own = (i & 1) ? k.jp : k.j4;
  const uint32_t bb = bitmask(k.fl, F_BOUNCE0 + opp_of(i));
  stb<R>(Ab + (size_t)lo * NF * rs, (bb & own) | (~bb & JB[i]), v);
}

// Moments of a SOLID cell's frozen populations (src/fluid_engine3d.py:92-113,
// the full sums over i = 0..18 as the reference forms them).
template <class R>
__device__ __forceinline__ void solid_moments(const R* f, R& lr, R& sx, R& sy, R& sz) {
  R mx = 0, my = 0, mz = 0;
  lr = 0;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    lr += f[i];
    mx += f[i] * (R)ex_of(i);
    my += f[i] * (R)ey_of(i);
    mz += f[i] * (R)ez_of(i);
  }
  sx = sy = sz = 0;
  if (lr > (R)1e-10) {
    sx = mx / lr;
    sy = my / lr;
    sz = mz / lr;
  }
}

// Streamed fluid populations of a non-solid cell for the pass of parity NODD.
template <class R, int GEN, int NODD, int CMP>
__device__ __forceinline__ void fluid_gather(const Slot<R>& S, const Cell& k, const uint32_t* J, R* f) {
  const int P = S.P, nz = S.nz;
  const uint32_t N = (uint32_t)S.npad;
  if (k.type == BC_INLET) {
    // Velocity inlet from the cell's own post-collision set
    // (src/fluid_engine3d.py:197-244); x = 0 so the plane index is c.
#pragma unroll
    for (int i = 0; i < 19; ++i) f[i] = ld(S.fin, (uint32_t)(i * P) + k.c);
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
  } else if (k.type == BC_OUTLET) {
    // Pressure outlet: all 19 post-collision values of x-1, rescaled
    // (src/fluid_engine3d.py:247-259).
    const uint32_t p = k.c - (uint32_t)((S.nx - 1) * P);
    const R* __restrict__ sn = S.fsnap + (1 - NODD) * 19 * P;
    R rs = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      f[i] = ld(sn, (uint32_t)(i * P) + p);
      rs += f[i];
    }
    if (rs > (R)1e-6) {
      const R sc = S.rho_out / rs;
#pragma unroll
      for (int i = 0; i < 19; ++i) f[i] *= sc;
    }
  } else if (CMP) {
    const R* __restrict__ A = S.f;
    const uint32_t NF = S.nf;
#define XADV_FL(i) f[i] = fluid_load_c<R, NODD, i>(A, NF, k, J);
    XADV_FL(0) XADV_FL(1) XADV_FL(2) XADV_FL(3) XADV_FL(4) XADV_FL(5) XADV_FL(6) XADV_FL(7) XADV_FL(8)
    XADV_FL(9) XADV_FL(10) XADV_FL(11) XADV_FL(12) XADV_FL(13) XADV_FL(14) XADV_FL(15) XADV_FL(16)
    XADV_FL(17) XADV_FL(18)
#undef XADV_FL
  } else {
    const R* __restrict__ A = S.f;
    const R uw = GEN ? S.u_wall : (R)0;
    const int nx = S.nx, ny = S.ny;
#define XADV_FL(i) f[i] = fluid_load<R, GEN, NODD, i>(A, N, k, P, nz, nx, ny, uw);
    XADV_FL(0) XADV_FL(1) XADV_FL(2) XADV_FL(3) XADV_FL(4) XADV_FL(5) XADV_FL(6) XADV_FL(7) XADV_FL(8)
    XADV_FL(9) XADV_FL(10) XADV_FL(11) XADV_FL(12) XADV_FL(13) XADV_FL(14) XADV_FL(15) XADV_FL(16)
    XADV_FL(17) XADV_FL(18)
#undef XADV_FL
  }
}

// Moments (src/fluid_engine3d.py:92-113), then BGK with the equilibrium in
// registers (:157-173) and, in general mode, the body force
// (validate_analytical.py:388-390).
template <class R, int GEN>
__device__ __forceinline__ void fluid_collide(const Slot<R>& S, R* f, R& lr, R& ux, R& uy, R& uz) {
  // The reference sums f_i * e_i over i = 0..18; the zero terms only add
  // signed zeros, so the sums keep its order over the non-zero terms.
  lr = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
       f[14] + f[15] + f[16] + f[17] + f[18];
  const R mx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
  const R my = f[3] - f[4] + f[7] - f[8] - f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
  const R mz = f[5] - f[6] + f[11] - f[12] - f[13] + f[14] + f[15] - f[16] - f[17] + f[18];
  ux = uy = uz = 0;
  if (lr > (R)1e-10) {
    ux = mx / lr;
    uy = my / lr;
    uz = mz / lr;
  }
  const R om = S.omega;
  const R usq = ux * ux + uy * uy + uz * uz;
#define XADV_FC(i) f[i] = f[i] - om * (f[i] - feq_i<R, i>(lr, ux, uy, uz, usq));
  XADV_FC(0) XADV_FC(1) XADV_FC(2) XADV_FC(3) XADV_FC(4) XADV_FC(5) XADV_FC(6) XADV_FC(7) XADV_FC(8)
  XADV_FC(9) XADV_FC(10) XADV_FC(11) XADV_FC(12) XADV_FC(13) XADV_FC(14) XADV_FC(15) XADV_FC(16)
  XADV_FC(17) XADV_FC(18)
#undef XADV_FC
  if (GEN) {
    const R F = S.force_x;
#pragma unroll
    for (int i = 0; i < 19; ++i) f[i] = f[i] + (R)w_of(i) * (R)3.0 * (R)ex_of(i) * F;
  }
}

template <class R, int GEN, int NODD, int CMP>
__device__ __forceinline__ void fluid_scatter(const Slot<R>& S, const Cell& k, const uint32_t* J, const R* f) {
  R* __restrict__ A = S.f;
  const uint32_t N = (uint32_t)S.npad;
  const int P = S.P, nz = S.nz, nx = S.nx, ny = S.ny;
  const R uw = GEN ? S.u_wall : (R)0;
  if (CMP) {
    const uint32_t NF = S.nf;
#define XADV_FS(i) fluid_store_c<R, NODD, i>(A, NF, k, J, f[i]);
    XADV_FS(0) XADV_FS(1) XADV_FS(2) XADV_FS(3) XADV_FS(4) XADV_FS(5) XADV_FS(6) XADV_FS(7) XADV_FS(8)
    XADV_FS(9) XADV_FS(10) XADV_FS(11) XADV_FS(12) XADV_FS(13) XADV_FS(14) XADV_FS(15) XADV_FS(16)
    XADV_FS(17) XADV_FS(18)
#undef XADV_FS
  } else {
#define XADV_FS(i) fluid_store<R, GEN, NODD, i>(A, N, k, P, nz, nx, ny, uw, f[i]);
    XADV_FS(0) XADV_FS(1) XADV_FS(2) XADV_FS(3) XADV_FS(4) XADV_FS(5) XADV_FS(6) XADV_FS(7) XADV_FS(8)
    XADV_FS(9) XADV_FS(10) XADV_FS(11) XADV_FS(12) XADV_FS(13) XADV_FS(14) XADV_FS(15) XADV_FS(16)
    XADV_FS(17) XADV_FS(18)
#undef XADV_FS
  }
  if (k.type == BC_INLET) {
#pragma unroll
    for (int i = 0; i < 19; ++i) st(S.fin, (uint32_t)(i * P) + k.c, f[i]);
  }
  if (k.fl & F_SNAP) {
    const uint32_t p = k.c - (uint32_t)((nx - 2) * P);
    R* __restrict__ sn = S.fsnap + NODD * 19 * P;
#pragma unroll
    for (int i = 0; i < 19; ++i) st(sn, (uint32_t)(i * P) + p, f[i]);
  }
}

// ---------------------------------------------------------------- thermal ---

// All-ones words for the cell's face flags, in the order x = 0, x = nx-1,
// y = 0, y = ny-1, z = 0, z = nz-1 (one bit-extract each).
struct Faces {
  uint32_t m[6];
};
__device__ __forceinline__ Faces faces_of(uint32_t fl) {
  Faces f;
#pragma unroll
  for (int a = 0; a < 6; ++a) f.m[a] = bitmask(fl, 21 + a);  // F_FX0 .. F_FZN are bits 21..26
  return f;
}
static_assert(F_FX0 == 1u << 21 && F_FXN == 1u << 22 && F_FY0 == 1u << 23 && F_FYN == 1u << 24 &&
                  F_FZ0 == 1u << 25 && F_FZN == 1u << 26,
              "face bits");
// All-ones where the source c - e_i of direction i lies outside a
// non-periodic axis (the same test as src_face_mask(i), from the face words).
template <int i>
__device__ __forceinline__ uint32_t src_outside(const Faces& f) {
  uint32_t r = 0;
  if (ex_of(i) == 1) r |= f.m[0];
  if (ex_of(i) == -1) r |= f.m[1];
  if (ey_of(i) == 1) r |= f.m[2];
  if (ey_of(i) == -1) r |= f.m[3];
  if (ez_of(i) == 1) r |= f.m[4];
  if (ez_of(i) == -1) r |= f.m[5];
  return r;
}

// Streamed thermal populations of cell k for the pass of parity NODD (the
// thermal half uses the opposite AA parity: TODD = !NODD).
template <class R, int GEN, int NODD>
__device__ __forceinline__ void thermal_gather(const Slot<R>& S, const Cell& k, R* g) {
  constexpr int TODD = 1 - NODD;
  const R* __restrict__ B = S.g;
  const uint32_t N = (uint32_t)S.npad;
  const int P = S.P, nz = S.nz;
  if (k.type == BC_INLET) {  // thermal_engine3d.py:160-165
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = (R)w_of(i) * S.T_in;
  } else if (k.type == BC_OUTLET) {  // :170-177, values of x-1 published into gsnap
    const uint32_t p = k.c - (uint32_t)((S.nx - 1) * P);
    const R* __restrict__ sn = S.gsnap + NODD * 19 * P;
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = ld(sn, (uint32_t)(i * P) + p);
  } else {
    const uint32_t fl = k.fl;
    constexpr uint32_t rs = sizeof(R);
    const Faces fs = faces_of(fl);
#define XADV_TG(i)                                                                                          \
  {                                                                                                         \
    if (TODD == 0 || i == 0) {                                                                              \
      g[i] = ldb<R>(reinterpret_cast<const char*>(B) + (size_t)i * N * rs, k.c * rs);                        \
    } else if (!GEN) {                                                                                      \
      const int64_t pull = (int64_t)opp_of(i) * N - nb_off<i>(P, nz);                                       \
      g[i] = ld2<R, (i > opp_of(i))>(B, pull, (int64_t)i * N, k.c, src_outside<i>(fs));                    \
    } else {                                                                                                \
      const bool oob = (fl & src_face_mask(i)) != 0;                                                        \
      const uint32_t src = nb_wrap<i>(k.x, k.y, k.z, -1, S.nx, S.ny, nz);                                   \
      g[i] = ld(B, oob ? ((uint32_t)i * N + k.c) : ((uint32_t)opp_of(i) * N + src));                        \
    }                                                                                                       \
  }
    XADV_TG(0) XADV_TG(1) XADV_TG(2) XADV_TG(3) XADV_TG(4) XADV_TG(5) XADV_TG(6) XADV_TG(7) XADV_TG(8)
    XADV_TG(9) XADV_TG(10) XADV_TG(11) XADV_TG(12) XADV_TG(13) XADV_TG(14) XADV_TG(15) XADV_TG(16)
    XADV_TG(17) XADV_TG(18)
#undef XADV_TG
    if (fl & (F_FX0 | F_FXN)) {
      // x = 0: e_x = +1 sources lie before the inlet (w_i T_in); x = nx-1:
      // e_x = -1 sources lie past the outlet (the cell's own values, kept in
      // gxh).  These rules take precedence over the y/z walls
      // (thermal_engine3d.py:193-207).
      if (fl & F_FX0) {
        g[1] = (R)w_of(1) * S.T_in;
        g[7] = (R)w_of(7) * S.T_in;
        g[9] = (R)w_of(9) * S.T_in;
        g[11] = (R)w_of(11) * S.T_in;
        g[13] = (R)w_of(13) * S.T_in;
      }
      if (fl & F_FXN) {
        const uint32_t pxh = k.c - (uint32_t)((S.nx - 1) * P);
        g[2] = ld(S.gxh, 0u * P + pxh);
        g[8] = ld(S.gxh, 1u * P + pxh);
        g[10] = ld(S.gxh, 2u * P + pxh);
        g[12] = ld(S.gxh, 3u * P + pxh);
        g[14] = ld(S.gxh, 4u * P + pxh);
      }
    }
  }
  if (GEN && (k.fl & F_DIR)) {  // validate_analytical.py:21-30, applied after streaming
#pragma unroll
    for (int i = 0; i < 19; ++i) g[i] = (R)w_of(i) * S.dir_T;
  }
}

template <class R, int GEN, int NODD, int i>
__device__ __forceinline__ void thermal_store(R* __restrict__ B, uint32_t N, const Cell& k, const Faces& fs, int P,
                                              int nz, int nx, int ny, R v) {
  constexpr int TODD = 1 - NODD;
  constexpr uint32_t rs = sizeof(R);
  if (TODD == 0 || i == 0) {
    stb<R>(reinterpret_cast<char*>(B) + (size_t)opp_of(i) * N * rs, k.c * rs, v);
    return;
  }
  // Target c + e_i outside a non-periodic axis = source of opp(i) outside.
  if (!GEN) {
    const int64_t push = (int64_t)i * N + nb_off<i>(P, nz);
    st2<R, (opp_of(i) > i)>(B, push, (int64_t)opp_of(i) * N, k.c, src_outside<opp_of(i)>(fs), v);
    return;
  }
  const bool oob = (k.fl & src_face_mask(opp_of(i))) != 0;
  const uint32_t dst = nb_wrap<i>(k.x, k.y, k.z, +1, nx, ny, nz);
  st(B, oob ? ((uint32_t)opp_of(i) * N + k.c) : ((uint32_t)i * N + dst), v);
}

// BGK with the source term (thermal_engine3d.py:107-145).
template <class R>
__device__ __forceinline__ void thermal_collide(const Slot<R>& S, const Cell& k, R* g, R Tl, R ux, R uy, R uz) {
  const bool solid = k.type == BC_SOLID;
  if (solid) ux = uy = uz = 0;  // thermal_engine3d.py:123-125
  const R om = solid ? S.omega_ts : S.omega_tf;
  const R sf = solid ? S.sf_ts : S.sf_tf;
  const R usq = ux * ux + uy * uy + uz * uz;
#define XADV_TC(i) g[i] = g[i] - om * (g[i] - feq_i<R, i>(Tl, ux, uy, uz, usq));
  XADV_TC(0) XADV_TC(1) XADV_TC(2) XADV_TC(3) XADV_TC(4) XADV_TC(5) XADV_TC(6) XADV_TC(7) XADV_TC(8)
  XADV_TC(9) XADV_TC(10) XADV_TC(11) XADV_TC(12) XADV_TC(13) XADV_TC(14) XADV_TC(15) XADV_TC(16)
  XADV_TC(17) XADV_TC(18)
#undef XADV_TC
  // The source term, added to each population after its BGK update
  // (thermal_engine3d.py:141-143); cells without heat skip the whole block.
  if (k.fl & F_HEAT) {
    const R q = ld(S.q, k.c);
    if (q > (R)0.0) {
#pragma unroll
      for (int i = 0; i < 19; ++i) g[i] += (R)w_of(i) * sf * q;
    }
  }
}

template <class R, int GEN, int NODD>
__device__ __forceinline__ void thermal_scatter(const Slot<R>& S, const Cell& k, const R* g) {
  R* __restrict__ B = S.g;
  const uint32_t N = (uint32_t)S.npad;
  const int P = S.P, nz = S.nz, nx = S.nx, ny = S.ny;
  const Faces fs = faces_of(k.fl);
#define XADV_TS(i) thermal_store<R, GEN, NODD, i>(B, N, k, fs, P, nz, nx, ny, g[i]);
  XADV_TS(0) XADV_TS(1) XADV_TS(2) XADV_TS(3) XADV_TS(4) XADV_TS(5) XADV_TS(6) XADV_TS(7) XADV_TS(8)
  XADV_TS(9) XADV_TS(10) XADV_TS(11) XADV_TS(12) XADV_TS(13) XADV_TS(14) XADV_TS(15) XADV_TS(16)
  XADV_TS(17) XADV_TS(18)
#undef XADV_TS
  if (k.fl & F_XH) {
    const uint32_t p = k.c - (uint32_t)((nx - 1) * P);
    st(S.gxh, 0u * P + p, g[2]);
    st(S.gxh, 1u * P + p, g[8]);
    st(S.gxh, 2u * P + p, g[10]);
    st(S.gxh, 3u * P + p, g[12]);
    st(S.gxh, 4u * P + p, g[14]);
  }
  if (k.fl & F_SNAP) {
    const uint32_t p = k.c - (uint32_t)((nx - 2) * P);
    R* __restrict__ sn = S.gsnap + (1 - NODD) * 19 * P;
#pragma unroll
    for (int i = 0; i < 19; ++i) st(sn, (uint32_t)(i * P) + p, g[i]);
  }
}

// ---------------------------------------------------------- D3Q7 thermal ----
// D3Q7 variant of the thermal lattice.  Not a reference kernel: the
// reference's thermal lattice is D3Q19.  It solves the same advection-diffusion
// problem on directions 0..6 of the D3Q19 set (rest and the six faces, with
// their opposite pairs), weights 1/4 and 1/8 (c_s^2 = 1/4), the linear
// equilibrium g_eq_i = w_i T (1 + 4 e_i.u) (the quadratic terms are not
// isotropic on seven velocities).  A slot's relaxation times are the
// reference's: omega_tf / omega_ts hold 1/tau7 with tau7 = 1/2 + 4/3 (tau - 1/2)
// (the same diffusivity) and sf_tf / sf_ts the reference step's source factor
// 1 - 1/(2 tau), so the heat injected per step is the reference's.  Streaming
// follows thermal_engine3d.py:147-211 for these directions.  Two buffers of 7
// planes per slot (g7, g7 + 7 npad): odd steps read the first and write the
// second, even steps the reverse.  The fused pass (thermal lattice 7) runs a
// cell's D3Q7 step after the fluid half of the same step, with the velocity that
// half has just computed (u = 0 in solids).  The first odd step after an init
// takes the initial g and T as given (raw), as the D3Q19 thermal half does; the
// next even step clears the flag.
template <class R>
__device__ __forceinline__ R w7(int i) {
  return i == 0 ? (R)0.25 : (R)0.125;
}

// Streamed D3Q7 populations of cell c from the post-collision buffer src.
template <class R>
__device__ __forceinline__ void gather7(const Slot<R>& S, const R* __restrict__ src, uint32_t c, uint32_t fl,
                                        R* g) {
  const int type = (int)(fl & F_TYPE);
  const uint32_t N = (uint32_t)S.npad;
  if (type == BC_INLET) {  // thermal_engine3d.py:160-165
#pragma unroll
    for (int i = 0; i < 7; ++i) g[i] = w7<R>(i) * S.T_in;
    return;
  }
  if (type == BC_OUTLET) {  // :170-177 (x - 1, or the cell itself on x = 0)
    const uint32_t s = (fl & F_FX0) ? c : c - (uint32_t)S.P;
#pragma unroll
    for (int i = 0; i < 7; ++i) g[i] = ld(src, i * N + s);
    return;
  }
  // :180-211: the source c - e_i; off the lattice: x < 0 -> w_i T_in, x >= nx -> own value, y/z -> opposite.
  // Every load's index is selected first (the neighbour, or the cell's own slot), so no load leaves the slot.
  const uint32_t P = (uint32_t)S.P, nz = (uint32_t)S.nz;
  g[0] = ld(src, c);
  const R v1 = ld(src, N + ((fl & F_FX0) ? c : c - P));
  g[1] = (fl & F_FX0) ? w7<R>(1) * S.T_in : v1;
  g[2] = ld(src, 2 * N + ((fl & F_FXN) ? c : c + P));
  g[3] = ld(src, (fl & F_FY0) ? 4 * N + c : 3 * N + c - nz);
  g[4] = ld(src, (fl & F_FYN) ? 3 * N + c : 4 * N + c + nz);
  g[5] = ld(src, (fl & F_FZ0) ? 6 * N + c : 5 * N + c - 1);
  g[6] = ld(src, (fl & F_FZN) ? 5 * N + c : 6 * N + c + 1);
}

// The streamed populations of a thread's cells from the post-collision buffer
// src (or, on the raw step, the initial g and T as given) and T = sum of g in
// order from g_0.
template <class R, int CPT>
__device__ __forceinline__ void thermal7_load(const Slot<R>& S, const R* __restrict__ src, const Cell* cell, int raw,
                                              R (*g)[7], R* T) {
  const uint32_t N = (uint32_t)S.npad;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    if (!cell[k].valid) continue;
    const uint32_t c = cell[k].c;
    if (raw) {
#pragma unroll
      for (int i = 0; i < 7; ++i) g[k][i] = ld(src, i * N + c);
      T[k] = ld(S.T, c);
    } else {
      gather7<R>(S, src, c, cell[k].fl, g[k]);
      T[k] = 0;
#pragma unroll
      for (int i = 0; i < 7; ++i) T[k] += g[k][i];
    }
  }
}

// BGK relaxation towards g_eq(T, u) at the relaxation time of the cell's kind,
// plus w_i (1 - 1/(2 tau)) q on heated cells, stored into dst.
template <class R>
__device__ __forceinline__ void thermal7_collide(const Slot<R>& S, R* __restrict__ dst, const Cell& k, const R* g, R T,
                                                 R ux, R uy, R uz) {
  const uint32_t N = (uint32_t)S.npad;
  const R om = k.type == BC_SOLID ? S.omega_ts : S.omega_tf;
  const R sf = k.type == BC_SOLID ? S.sf_ts : S.sf_tf;
  const R q = (k.fl & F_HEAT) ? ld(S.q, k.c) : (R)0;
  const R eu[7] = {(R)0, ux, -ux, uy, -uy, uz, -uz};
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const R geq = w7<R>(i) * T * ((R)1.0 + (R)4.0 * eu[i]);
    R v = g[i] - om * (g[i] - geq);
    if (q > (R)0.0) v += w7<R>(i) * sf * q;
    st(dst, i * N + k.c, v);
  }
}

// ------------------------------------------------------------- the pass ----

// Thermal gather for a thread's cells: post-collision g of the previous step
// through the streaming rules, or (first thermal half after an init) the
// initial populations and T as given (src/simulation.py:73,89).
template <class R, int GEN, int NODD, int CPT>
__device__ __forceinline__ void thermal_load(const Slot<R>& S, const Cell* cell, int raw, R (*g)[19], R* Tl) {
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    if (!cell[k].valid) continue;
    if (raw) {
      const uint32_t N = (uint32_t)S.npad;
#pragma unroll
      for (int i = 0; i < 19; ++i) g[k][i] = ld(S.g, (uint32_t)i * N + cell[k].c);
      Tl[k] = ld(S.T, cell[k].c);
    } else {
      thermal_gather<R, GEN, NODD>(S, cell[k], g[k]);
    }
  }
}

// One time step for CPT cells per thread.  The phases (fluid gather, fluid
// collide + scatter, thermal gather, thermal collide + scatter) run for all
// of a thread's cells before the next phase, so each thread keeps
// CPT x 19 loads in flight.  With PRE the D3Q19 thermal gather is issued
// right after the fluid gather (38 loads in flight, more registers; the two
// lattices are separate arrays and each slot read here is written only by
// this cell, so the order of the fluid stores and thermal loads does not
// matter).
// Cell k of a thread is block_base + k*blockDim + threadIdx, so every load
// instruction of a wavefront covers 64 consecutive cells.  A launch covers the
// x-planes [x0, x1) of every slot (cells x0*P .. x1*P - 1; the whole lattice
// for x0 = 0, x1 >= nx); clear_raw says whether this even launch may clear the
// raw flag (only the last launch of an even pass split over x-slabs does, so
// every launch of the odd pass before it has read the flag).  TH = 1: the D3Q19
// thermal lattice one step behind the fluid (AA in place); TH = 7: the D3Q7
// variant of the same step (two buffers), its 7 loads per cell issued right
// after the fluid gather so they overlap the fluid lattice's.
template <class R, int FL, int TH, int GEN, int NODD, int CPT, int CMP, int PRE = 0>
__global__ __launch_bounds__(256) void k_pass(Slot<R>* __restrict__ slots, int bps, int G, int order,
                                              int want_fields, int x0, int x1, int clear_raw) {
  static_assert(!(CMP && (GEN || !FL)), "compact fluid storage is built for the channel kernels with the fluid on");
  static_assert(TH == 0 || TH == 1 || (TH == 7 && FL && !GEN), "thermal lattice: none, D3Q19, or D3Q7 with the fluid");
  static_assert(!PRE || (FL && TH == 1), "the thermal prefetch is for the fused D3Q19 pass");
  const int b = map_block(bps, order);
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
    else if (clear_raw && lb == 0 && threadIdx.x == 0 && S.raw)
      S.raw = 0;
  }
  const uint32_t ncell = min((uint32_t)S.ncell, (uint32_t)min(x1, S.nx) * (uint32_t)S.P);
  const uint32_t base = (uint32_t)x0 * (uint32_t)S.P + (uint32_t)lb * blockDim.x * CPT + threadIdx.x;
  if (base >= ncell) return;

  Cell cell[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    Cell& q = cell[k];
    q.c = base + (uint32_t)k * blockDim.x;
    q.valid = q.c < ncell;
    q.fl = q.valid ? ld(S.flags, q.c) : (uint32_t)BC_SOLID;
    q.type = (int)(q.fl & F_TYPE);
    q.x = q.y = q.z = 0;
    q.j4 = q.jp = 0;
    if (GEN && q.valid) coords(q.c, S.P, S.nz, q.x, q.y, q.z);
  }

  R ux[CPT], uy[CPT], uz[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) ux[k] = uy[k] = uz[k] = 0;
  R g[CPT][19];
  R Tl[CPT];
  R g7[TH == 7 ? CPT : 1][7];
  R T7[TH == 7 ? CPT : 1];
  const uint32_t N7 = TH == 7 ? (uint32_t)S.npad : 0u;
  const R* __restrict__ src7 = TH == 7 ? (NODD ? S.g7 : S.g7 + 7 * (size_t)N7) : nullptr;
  R* __restrict__ dst7 = TH == 7 ? (NODD ? S.g7 + 7 * (size_t)N7 : S.g7) : nullptr;

  if (FL) {
    R f[CPT][19];
    uint32_t J[CPT][CMP && NODD ? 19 : 1];
#pragma unroll
    for (int k = 0; k < CPT; ++k)
      if (cell[k].valid && cell[k].type != BC_SOLID) {
        if (CMP) compact_offsets<R, NODD>(S, cell[k], J[k]);
        fluid_gather<R, GEN, NODD, CMP>(S, cell[k], J[k], f[k]);
      }
    if (PRE) thermal_load<R, GEN, NODD, CPT>(S, cell, raw, g, Tl);
    if constexpr (TH == 7) thermal7_load<R, CPT>(S, src7, cell, raw, g7, T7);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const Cell& q = cell[k];
      if (!q.valid) continue;
      if (q.type != BC_SOLID) {
        R lr;
        fluid_collide<R, GEN>(S, f[k], lr, ux[k], uy[k], uz[k]);
        if (want_fields) {
          st(S.rho, q.c, lr);
          st(S.u, 3 * q.c, ux[k]);
          st(S.u, 3 * q.c + 1, uy[k]);
          st(S.u, 3 * q.c + 2, uz[k]);
        }
        fluid_scatter<R, GEN, NODD, CMP>(S, q, J[k], f[k]);
      } else if (want_fields && !CMP) {
        // Solid cells keep their initial populations, stored in the even
        // layout (slot opp(i)); the reference's moment kernel still runs on
        // them.  (With compact storage the init writes these moments once.)
        const uint32_t N = (uint32_t)S.npad;
        R fs[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) fs[i] = ld(S.f, (uint32_t)opp_of(i) * N + q.c);
        R lr, sx, sy, sz;
        solid_moments<R>(fs, lr, sx, sy, sz);
        st(S.rho, q.c, lr);
        st(S.u, 3 * q.c, sx);
        st(S.u, 3 * q.c + 1, sy);
        st(S.u, 3 * q.c + 2, sz);
      }
    }
  } else if (TH == 1) {
    // Thermal lattice alone: collide with the velocity field given at init.
#pragma unroll
    for (int k = 0; k < CPT; ++k)
      if (cell[k].valid) {
        ux[k] = ld(S.u, 3 * cell[k].c);
        uy[k] = ld(S.u, 3 * cell[k].c + 1);
        uz[k] = ld(S.u, 3 * cell[k].c + 2);
      }
  }

  if constexpr (TH == 7) {
#pragma unroll
    for (int k = 0; k < CPT; ++k)
      if (cell[k].valid) thermal7_collide<R>(S, dst7, cell[k], g7[k], T7[k], ux[k], uy[k], uz[k]);
  }

  if (TH == 1) {
    if (!PRE) thermal_load<R, GEN, NODD, CPT>(S, cell, raw, g, Tl);
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      if (!cell[k].valid) continue;
      if (!raw) {  // thermal_engine3d.py:220-224, in order from g_0
        Tl[k] = g[k][0];
#pragma unroll
        for (int i = 1; i < 19; ++i) Tl[k] += g[k][i];
      }
      thermal_collide<R>(S, cell[k], g[k], Tl[k], ux[k], uy[k], uz[k]);
      thermal_scatter<R, GEN, NODD>(S, cell[k], g[k]);
    }
  }
}

// ------------------------------------------------------------------- init ---

// Populations from macroscopic fields, stored as the post-collision state of
// step 1 in the even layout (what an even pass would have written).  For the
// fluid, BGK(f_eq(rho,u), rho, u) returns f_eq exactly, so only the body force
// term changes the stored values.  The thermal lattice keeps g_eq in the
// natural layout for the raw first pass.
template <class R, int FL, int TH, int GEN, int CMP>
__global__ __launch_bounds__(256) void k_init(Slot<R>* __restrict__ slots, int slot, int frozen_u) {
  Slot<R>& S = slots[slot];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncell) return;
  const uint32_t fl = S.flags[c];
  const int type = (int)(fl & F_TYPE);
  const int64_t N = S.npad;
  const int P = S.P;
  const R ux = S.u[3 * c], uy = S.u[3 * c + 1], uz = S.u[3 * c + 2];
  if (FL && CMP) {
    // Compact storage: non-solid cells get f_eq in their own slots; SOLID
    // cells store nothing and their (frozen) moments go to the fields now.
    const R lr = S.rho[c];
    const R usq = ux * ux + uy * uy + uz * uz;
    R f[19];
#define XADV_FE(i) f[i] = feq_i<R, i>(lr, ux, uy, uz, usq);
    XADV_FE(0) XADV_FE(1) XADV_FE(2) XADV_FE(3) XADV_FE(4) XADV_FE(5) XADV_FE(6) XADV_FE(7) XADV_FE(8)
    XADV_FE(9) XADV_FE(10) XADV_FE(11) XADV_FE(12) XADV_FE(13) XADV_FE(14) XADV_FE(15) XADV_FE(16)
    XADV_FE(17) XADV_FE(18)
#undef XADV_FE
    if (type == BC_SOLID) {
      R sr, sx, sy, sz;
      solid_moments<R>(f, sr, sx, sy, sz);
      S.rho[c] = sr;
      S.u[3 * c] = sx;
      S.u[3 * c + 1] = sy;
      S.u[3 * c + 2] = sz;
    } else {
      const R om = S.omega;
#pragma unroll
      for (int i = 0; i < 19; ++i) f[i] = f[i] - om * (f[i] - f[i]);
      const uint32_t nz = (uint32_t)S.nz;
      const uint32_t q = div_nz((uint32_t)c, nz, S.nz_magic);
      const uint32_t j4 = own_offset<sizeof(R)>(S.cols, q, (int)((uint32_t)c - q * nz));
      char* __restrict__ A = reinterpret_cast<char*>(S.f);
      const size_t NFB = (size_t)S.nf * sizeof(R);
#pragma unroll
      for (int i = 0; i < 19; ++i) *reinterpret_cast<R*>(A + (size_t)opp_of(i) * NFB + j4) = f[i];
      if (type == BC_INLET) {
#pragma unroll
        for (int i = 0; i < 19; ++i) S.fin[(int64_t)i * P + c] = f[i];
      }
    }
    if (fl & F_SNAP) {
      // Also for a SOLID cell left of an OUTLET: the outlet copies its frozen
      // populations every step, and passes never rewrite them.
      const int64_t p = c - (int64_t)(S.nx - 2) * P;
#pragma unroll
      for (int i = 0; i < 19; ++i) {
        S.fsnap[(int64_t)i * P + p] = f[i];
        S.fsnap[(int64_t)(19 + i) * P + p] = f[i];
      }
    }
  } else if (FL) {
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

// ---------------------------------------------------------- D3Q7 init ----

// Initial D3Q7 populations g_eq(T, u) (u = 0 in solids) into the buffer the first odd step reads.
template <class R>
__global__ __launch_bounds__(256) void k_init7(Slot<R>* __restrict__ slots, int slot) {
  Slot<R>& S = slots[slot];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncell) return;
  const int type = (int)(S.flags[c] & F_TYPE);
  R ux = S.u[3 * c], uy = S.u[3 * c + 1], uz = S.u[3 * c + 2];
  if (type == BC_SOLID) ux = uy = uz = 0;
  const R T = S.T[c];
  const R eu[7] = {(R)0, ux, -ux, uy, -uy, uz, -uz};
#pragma unroll
  for (int i = 0; i < 7; ++i) S.g7[(int64_t)i * S.npad + c] = w7<R>(i) * T * ((R)1.0 + (R)4.0 * eu[i]);
}

// T of the current step: streaming of the buffer the last step wrote (the second after an odd step) + moments.
template <class R>
__global__ __launch_bounds__(256) void k_readout7(Slot<R>* __restrict__ slots, int bps, int G, int order,
                                                  int last_odd) {
  const int b = map_block(bps, order);
  if (b >= G) return;
  const int si = b / bps;
  const int lb = b - si * bps;
  Slot<R>& S = slots[si];
  if (!S.active) return;
  const uint32_t c = (uint32_t)lb * blockDim.x + threadIdx.x;
  if (c >= (uint32_t)S.ncell) return;
  const R* src = last_odd ? S.g7 + 7 * (size_t)S.npad : S.g7;
  R g[7];
  gather7<R>(S, src, c, ld(S.flags, c), g);
  R T = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i) T += g[i];
  st(S.T, c, T);
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
// Each thread takes kReadCells cells (c = block base + j * blockDim + thread,
// so every load instruction covers consecutive cells), folds their stat
// values in j order, and the block reduces once: one partial row per
// kReadCells * blockDim cells, which amortises the FP64 block reduction and
// shortens the per-slot fold of k_stats_final by the same factor.
constexpr int kReadCells = 4;

template <class R, int FL, int TH, int GEN, int NODD>
__global__ __launch_bounds__(256) void k_readout(Slot<R>* __restrict__ slots, int bps, int G, int order,
                                                 int with_stats, double* __restrict__ partials, int t_field) {
  const int b = map_block(bps, order);
  if (b >= G) return;
  const int si = b / bps;
  const int lb = b - si * bps;
  Slot<R>& S = slots[si];
  if (!S.active) return;
  const uint32_t base = (uint32_t)lb * blockDim.x * kReadCells + threadIdx.x;
  int op[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) op[k] = kStatOp[k];
  double v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = 0.0;  // every max slot holds values >= 0
  R Tl[kReadCells];
  Cell q[kReadCells];
#pragma unroll
  for (int j = 0; j < kReadCells; ++j) {
    q[j].c = base + (uint32_t)j * blockDim.x;
    q[j].valid = q[j].c < (uint32_t)S.ncell;
    q[j].fl = q[j].valid ? ld(S.flags, q[j].c) : (uint32_t)BC_SOLID;
    q[j].type = (int)(q[j].fl & F_TYPE);
    q[j].x = q[j].y = q[j].z = 0;
    q[j].j4 = q[j].jp = 0;
    if (q[j].valid) coords(q[j].c, S.P, S.nz, q[j].x, q[j].y, q[j].z);
    Tl[j] = 0;
  }
  if (TH) {  // all gathers first: kReadCells x 19 loads in flight
    R g[kReadCells][19];
#pragma unroll
    for (int j = 0; j < kReadCells; ++j)
      if (q[j].valid) thermal_gather<R, GEN, NODD>(S, q[j], g[j]);
#pragma unroll
    for (int j = 0; j < kReadCells; ++j)
      if (q[j].valid) {
#pragma unroll
        for (int i = 0; i < 19; ++i) Tl[j] += g[j][i];
        st(S.T, q[j].c, Tl[j]);
      }
  } else if (t_field) {  // T formed by another kernel (the D3Q7 readout)
#pragma unroll
    for (int j = 0; j < kReadCells; ++j)
      if (q[j].valid) Tl[j] = S.T[q[j].c];
  }
  if (!with_stats) return;
#pragma unroll
  for (int j = 0; j < kReadCells; ++j) {
    if (!q[j].valid) continue;
    const uint32_t c = q[j].c;
    const int type = q[j].type;
    const int x = q[j].x;
    double w[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = 0.0;
    const float ux = (float)S.u[3 * c], uy = (float)S.u[3 * c + 1], uz = (float)S.u[3 * c + 2];
    const float vsq = ux * ux + uy * uy + uz * uz;
    const float vmag = sqrtf(vsq);
    const double r = (double)S.rho[c];
    w[2] = vsq;
    if (type != BC_SOLID) {
      w[0] = vmag;
      w[1] = vmag;
      w[3] = 1.0;
      w[4] = r;
    }
    if (x == 1) {
      w[5] = r;
      w[8] = 1.0;
    }
    if (x == S.nx - 2) {
      w[6] = r;
      w[9] = 1.0;
    }
    if (type == BC_OUTLET) {
      w[7] = ux;
      w[10] = 1.0;
    }
    if (TH || t_field) {
      const float t = (float)Tl[j];
      w[11] = t > 0.0f ? t : 0.0f;
      w[12] = t;
      w[13] = 1.0;
      if (type == BC_OUTLET) {
        w[14] = (double)(ux * t);
        w[15] = 1.0;
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = op[k] == 0 ? v[k] + w[k] : fmax(v[k], w[k]);
  }
  block_reduce<16>(v, op, partials + (int64_t)b * 16);
}

// Fixed-order fold of nrows rows of NV doubles (row-major) by one block of
// kFoldThreads threads: the block stages kFoldRows rows at a time in LDS with
// coalesced loads, then thread k < NV folds them row after row in index order
// (op 0 sum, 1 max, 2 min), so each result is the same sequence of operations
// as one thread walking the rows in global memory, without its latency-bound
// chain of dependent loads.  Each thread issues all of its loads of a chunk
// before it stores any (the partials were written by blocks on every XCD, so
// each load goes out to memory), and the next chunk's loads are in flight
// while the current one is folded.
constexpr int kFoldThreads = 256;
constexpr int kFoldRows = 256;

template <int NV>
__device__ double ordered_fold(const double* __restrict__ rows, int nrows, int op, double a, double* lds) {
  constexpr int U = kFoldRows * NV / kFoldThreads;  // values per thread and chunk
  static_assert(kFoldRows * NV % kFoldThreads == 0, "a chunk must split evenly over the block");
  const int t = threadIdx.x;
  double nxt[U];
  auto fetch = [&](int j0) {
    const int n = min(kFoldRows, nrows - j0) * NV;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = t + u * kFoldThreads;
      nxt[u] = i < n ? rows[(int64_t)j0 * NV + i] : 0.0;
    }
  };
  if (nrows > 0) fetch(0);
  for (int j0 = 0; j0 < nrows; j0 += kFoldRows) {
    const int nr = min(kFoldRows, nrows - j0);
    __syncthreads();  // the previous chunk has been folded
#pragma unroll
    for (int u = 0; u < U; ++u) lds[t + u * kFoldThreads] = nxt[u];
    __syncthreads();
    if (j0 + kFoldRows < nrows) fetch(j0 + kFoldRows);
    if (t < NV) {
      // Eight LDS reads in flight ahead of eight folds, in row order.
      int j = 0;
      for (; j + 8 <= nr; j += 8) {
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = lds[(j + q) * NV + t];
#pragma unroll
        for (int q = 0; q < 8; ++q) a = op == 0 ? a + w[q] : (op == 1 ? fmax(a, w[q]) : fmin(a, w[q]));
      }
      for (; j < nr; ++j) {
        const double w = lds[j * NV + t];
        a = op == 0 ? a + w : (op == 1 ? fmax(a, w) : fmin(a, w));
      }
    }
  }
  return a;
}

// One block per slot: fixed-order reduction of the per-block partials.
template <class R>
__global__ __launch_bounds__(kFoldThreads) void k_stats_final(const Slot<R>* __restrict__ slots, int bps,
                                                              const double* __restrict__ partials,
                                                              double* __restrict__ out) {
  __shared__ double lds[kFoldRows * 16];
  const int si = blockIdx.x;
  const int k = threadIdx.x;
  if (!slots[si].active) {  // uniform per block
    if (k < 16) out[si * 16 + k] = 0.0;
    return;
  }
  const double a = ordered_fold<16>(partials + (int64_t)si * bps * 16, bps, (k == 0 || k == 11) ? 1 : 0, 0.0, lds);
  if (k < 16) out[si * 16 + k] = a;
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
  const uint32_t c = (uint32_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kNMet];
  int op[kNMet];
#pragma unroll
  for (int k = 0; k < kNMet; ++k) {
    op[k] = kMetOp[k];
    v[k] = op[k] == 1 ? -INFINITY : (op[k] == 2 ? INFINITY : 0.0);
  }
  if (c < (uint32_t)S.ncell) {
    int x, y, z;
    coords(c, S.P, S.nz, x, y, z);
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

// Reductions behind the physics acceptance checks of analyze_dataset.py:184-325
// (mass balance, Mach, density deviation, fluid temperature bound, velocity
// divergence, energy balance terms).  Index meaning: XADV_C_* in include/xadv.h.
constexpr int kNChk = 16;
__constant__ int kChkOp[kNChk] = {0, 0, 1, 1, 2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};

template <class R>
__global__ __launch_bounds__(256) void k_checks(Slot<R>* __restrict__ slots, int slot, double k_fluid,
                                                double k_solid, double* __restrict__ partials) {
  const Slot<R>& S = slots[slot];
  const uint32_t c = (uint32_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[kNChk];
  int op[kNChk];
#pragma unroll
  for (int k = 0; k < kNChk; ++k) {
    op[k] = kChkOp[k];
    v[k] = op[k] == 1 ? -INFINITY : (op[k] == 2 ? INFINITY : 0.0);
  }
  if (c < (uint32_t)S.ncell) {
    int x, y, z;
    coords(c, S.P, S.nz, x, y, z);
    const int nx = S.nx, ny = S.ny, nz = S.nz, P = S.P;
    const int type = (int)(S.flags[c] & F_TYPE);
    const bool fluid = type != BC_SOLID;
    const double r = (double)S.rho[c];
    const float ux = (float)S.u[3 * c], uy = (float)S.u[3 * c + 1], uz = (float)S.u[3 * c + 2];
    const double t = (double)S.T[c];
    if (x == 1 && fluid) v[0] = r * ux;          // mass flow one cell inside the inlet
    if (x == nx - 2 && fluid) v[1] = r * ux;     // and the outlet
    if (fluid) {
      v[2] = sqrtf(ux * ux + uy * uy + uz * uz);
      v[3] = fabs(r - 1.0);
      v[4] = t;
    }
    if (x == 1) v[5] = r;
    if (x == nx - 2) v[6] = r;
    if (fluid && x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2) {
      const double d = ((double)S.u[3 * (c + P)] - (double)S.u[3 * (c - P)]) / 2.0 +
                       ((double)S.u[3 * (c + nz) + 1] - (double)S.u[3 * (c - nz) + 1]) / 2.0 +
                       ((double)S.u[3 * (c + 1) + 2] - (double)S.u[3 * (c - 1) + 2]) / 2.0;
      v[7] = fabs(d);
      v[8] = 1.0;
      v[9] = fabs(d);
    }
    if (x == 0 && type == BC_INLET) v[10] = ux;
    if (x == nx - 1 && type == BC_OUTLET) v[11] = (double)ux * t;
    if (x == 1) {
      const double kk = type == BC_SOLID ? k_solid : k_fluid;
      v[12] = -kk * (t - (double)S.T[c - P]);
    }
    if (x == nx - 2) {
      const double kk = type == BC_SOLID ? k_solid : k_fluid;
      v[13] = -kk * ((double)S.T[c + P] - t);
    }
    if (fluid) v[14] = 1.0;
  }
  block_reduce<kNChk>(v, op, partials + (int64_t)blockIdx.x * kNChk);
}

__global__ __launch_bounds__(kFoldThreads) void k_checks_final(int nblocks, const double* __restrict__ partials,
                                                               double* __restrict__ out) {
  __shared__ double lds[kFoldRows * kNChk];
  const int k = threadIdx.x;
  const int o = k < kNChk ? kChkOp[k] : 0;
  const double a0 = o == 1 ? -INFINITY : (o == 2 ? INFINITY : 0.0);
  const double a = ordered_fold<kNChk>(partials, nblocks, o, a0, lds);
  if (k < kNChk) out[k] = a;
}

__global__ __launch_bounds__(kFoldThreads) void k_metrics_final(int nblocks, const double* __restrict__ partials,
                                                                double* __restrict__ out) {
  __shared__ double lds[kFoldRows * kNMet];
  const int k = threadIdx.x;
  const int o = k < kNMet ? kMetOp[k] : 0;
  const double a0 = o == 1 ? -INFINITY : (o == 2 ? INFINITY : 0.0);
  const double a = ordered_fold<kNMet>(partials, nblocks, o, a0, lds);
  if (k < kNMet) out[k] = a;
}

}  // namespace xadv
