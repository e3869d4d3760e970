// Host side of libxadv.so: the C ABI declared in include/xadv.h.
//
// A batch owns one HIP stream (default priority), the slot table in device
// memory, and per-slot arenas sized for the slot capacity.  Steps are fused
// AA passes (lbm_kernels.h); runs of graph_chunk steps that start on an odd
// pass are captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/xadv.h"
#include "lattice.h"
#include "lbm_kernels.h"
#include "geometry_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(XADV_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

int cleanup_early(xadv_batch* b, int rc);

}  // namespace

struct xadv_batch {
  xadv_batch_desc d{};
  hipStream_t stream = nullptr;
  int64_t cap_cells = 0, cap_npad = 0;
  int cap_P = 0;
  int bps = 0, G = 0, G_launch = 0;        // pass grid (block * cpt cells per block)
  int bps_r = 0, G_r = 0, G_r_launch = 0;  // one cell per thread (D3Q7 readout)
  int bps_s = 0, G_s = 0, G_s_launch = 0;  // k_readout: kReadCells cells per thread, one stats row per block
  size_t rsize = 4;
  void* slots_dev = nullptr;
  void* slots_host = nullptr;  // pinned copy of the slot table
  size_t slot_bytes = 0;
  int compact = 0;  // fluid lattice stored for non-solid cells only
  int q7 = 0;       // thermal lattice is the D3Q7 variant (fused into the fluid pass; d.thermal is 0)
  int64_t cols_per_slot = 0;  // padded column-table entries per slot
  // arenas (bytes)
  char *f = nullptr, *g = nullptr, *flags = nullptr, *q = nullptr, *fin = nullptr, *fsnap = nullptr, *gsnap = nullptr,
       *gxh = nullptr, *rho = nullptr, *u = nullptr, *T = nullptr, *cols = nullptr, *g7 = nullptr;
  double* partials = nullptr;
  double* stats_dev = nullptr;
  double* met_partials = nullptr;
  double* met_dev = nullptr;
  double* host_small = nullptr;  // pinned, for stats/metrics read-back
  char* gen = nullptr;           // device scratch of the design builders (grown on demand)
  size_t gen_bytes = 0;
  int64_t npass = 0;             // passes since the batch's last init of all slots
  bool fields_fresh = false;
  std::vector<int> loaded, inited;
  std::vector<int> frozen;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
};

namespace {

using xadv::Slot;

template <class R>
Slot<R>* host_slots(xadv_batch* b) { return reinterpret_cast<Slot<R>*>(b->slots_host); }

int push_slots(xadv_batch* b) {
  HIPCHK(hipMemcpyAsync(b->slots_dev, b->slots_host, b->slot_bytes * b->d.slots, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

template <class R>
int push_slot(xadv_batch* b, int s) {
  Slot<R>* hs = host_slots<R>(b);
  HIPCHK(hipMemcpyAsync(reinterpret_cast<Slot<R>*>(b->slots_dev) + s, hs + s, sizeof(Slot<R>), hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

// Push one field of a slot.  Passes change `raw` in the device copy, so
// after init only single fields are pushed, never the whole (stale) struct.
template <class R>
int push_field(xadv_batch* b, int s, size_t off, size_t bytes) {
  const char* src = reinterpret_cast<const char*>(host_slots<R>(b) + s) + off;
  char* dst = reinterpret_cast<char*>(reinterpret_cast<Slot<R>*>(b->slots_dev) + s) + off;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

// ----------------------------------------------------------- dispatch -----

// The x-planes a pass launch covers (all of them, or one slab of a pass split over x-slabs) and its grid.
struct PassRange {
  int x0 = 0, x1 = 1 << 30;  // planes [x0, x1) of every slot
  int clear_raw = 1;         // an even launch may clear the raw flag
  int bps = 0, G = 0, G_launch = 0;  // 0: the batch's whole-lattice grid
  bool slab = false;
};

template <class R, int FL, int TH, int GEN, int CPT, int CMP, int PRE = 0>
void launch_pass_c(xadv_batch* b, int nodd, int want_fields, const PassRange& r) {
  auto* sl = reinterpret_cast<Slot<R>*>(b->slots_dev);
  const int bps = r.bps ? r.bps : b->bps, G = r.G ? r.G : b->G;
  dim3 grid(r.G_launch ? r.G_launch : b->G_launch), block(b->d.block);
  const int order = (b->d.xcd_remap ? 1 : 0) | ((b->d.alternate_sweep && !nodd && !r.slab) ? 2 : 0);
  if (nodd)
    hipLaunchKernelGGL((xadv::k_pass<R, FL, TH, GEN, 1, CPT, CMP, PRE>), grid, block, 0, b->stream, sl, bps, G,
                       order, want_fields, r.x0, r.x1, r.clear_raw);
  else
    hipLaunchKernelGGL((xadv::k_pass<R, FL, TH, GEN, 0, CPT, CMP, PRE>), grid, block, 0, b->stream, sl, bps, G,
                       order, want_fields, r.x0, r.x1, r.clear_raw);
}

// Compact storage exists for the channel kernels with the fluid lattice on; the thermal prefetch for the
// fused D3Q19 pass on compact storage.
template <class R, int FL, int TH, int GEN, int CPT>
void launch_pass_s(xadv_batch* b, int nodd, int want_fields, const PassRange& r) {
  if constexpr (FL && !GEN) {
    if constexpr (TH == 1) {
      if (b->compact && b->d.thermal_prefetch)
        return launch_pass_c<R, FL, TH, GEN, CPT, 1, 1>(b, nodd, want_fields, r);
    }
    if (b->compact) return launch_pass_c<R, FL, TH, GEN, CPT, 1>(b, nodd, want_fields, r);
  }
  launch_pass_c<R, FL, TH, GEN, CPT, 0>(b, nodd, want_fields, r);
}

template <class R, int FL, int TH, int GEN>
void launch_pass_t(xadv_batch* b, int nodd, int want_fields, const PassRange& r) {
  if (b->d.cells_per_thread == 2)
    launch_pass_s<R, FL, TH, GEN, 2>(b, nodd, want_fields, r);
  else
    launch_pass_s<R, FL, TH, GEN, 1>(b, nodd, want_fields, r);
}

template <class R, int FL, int TH, int GEN>
void launch_readout_t(xadv_batch* b, int nodd, int with_stats) {
  auto* sl = reinterpret_cast<Slot<R>*>(b->slots_dev);
  dim3 grid(b->G_s_launch), block(b->d.block);
  if (nodd)
    hipLaunchKernelGGL((xadv::k_readout<R, FL, TH, GEN, 1>), grid, block, 0, b->stream, sl, b->bps_s, b->G_s,
                       b->d.xcd_remap, with_stats, b->partials, b->q7);
  else
    hipLaunchKernelGGL((xadv::k_readout<R, FL, TH, GEN, 0>), grid, block, 0, b->stream, sl, b->bps_s, b->G_s,
                       b->d.xcd_remap, with_stats, b->partials, b->q7);
}

template <class R, int FL, int TH, int GEN>
void launch_init_t(xadv_batch* b, int slot, int64_t ncell, int frozen) {
  auto* sl = reinterpret_cast<Slot<R>*>(b->slots_dev);
  const int bs = 256;
  dim3 grid((unsigned)((ncell + bs - 1) / bs)), block(bs);
  if constexpr (FL && !GEN) {
    if (b->compact) {
      hipLaunchKernelGGL((xadv::k_init<R, FL, TH, GEN, 1>), grid, block, 0, b->stream, sl, slot, frozen);
      return;
    }
  }
  hipLaunchKernelGGL((xadv::k_init<R, FL, TH, GEN, 0>), grid, block, 0, b->stream, sl, slot, frozen);
}

#define XADV_DISPATCH(FN, R, ...)                                        \
  do {                                                                   \
    const int fl_ = b->d.fluid, th_ = b->d.thermal, ge_ = b->d.general;  \
    if (fl_ && th_ && ge_) FN<R, 1, 1, 1>(__VA_ARGS__);                  \
    else if (fl_ && th_) FN<R, 1, 1, 0>(__VA_ARGS__);                    \
    else if (fl_ && ge_) FN<R, 1, 0, 1>(__VA_ARGS__);                    \
    else if (fl_) FN<R, 1, 0, 0>(__VA_ARGS__);                           \
    else if (ge_) FN<R, 0, 1, 1>(__VA_ARGS__);                           \
    else FN<R, 0, 1, 0>(__VA_ARGS__);                                    \
  } while (0)

void launch_pass(xadv_batch* b, int nodd, int want_fields, const PassRange& r = PassRange{}) {
  if (b->q7) {  // fluid + D3Q7 thermal in one pass (d.thermal is 0 for the other kernels)
    if (b->d.precision == 64)
      launch_pass_t<double, 1, 7, 0>(b, nodd, want_fields, r);
    else
      launch_pass_t<float, 1, 7, 0>(b, nodd, want_fields, r);
    return;
  }
  if (b->d.precision == 64)
    XADV_DISPATCH(launch_pass_t, double, b, nodd, want_fields, r);
  else
    XADV_DISPATCH(launch_pass_t, float, b, nodd, want_fields, r);
}

// Two passes (odd, then even) split over x-slabs of slab_planes planes and run as a wavefront: A_k is the
// odd pass on slab k, B_k the even pass on slab k, launched A_0, A_1, B_0, A_2, B_1, ..., B_{K-1} (or the
// mirror order).  Pass n touches the neighbour slots of x +- 1 only, so B_k needs A_{k-1}, A_k and A_{k+1},
// all launched before it, and writes nothing that a later A launch reads (the odd pass reads only the
// slots of its own cells and their x - 1 / x + 1 planes, which B has not touched yet or no longer needs).
// Every cell computes what it computes in the whole-lattice pass, so the bits are the same; the even pass
// of a slab runs while the odd pass's lines of that slab are still in the Infinity Cache.
void launch_pass_pair_slabs(xadv_batch* b, int want_fields, bool descending) {
  const int S = b->d.slab_planes;
  const int K = (b->d.cap_nx + S - 1) / S;
  const int cpb = b->d.block * b->d.cells_per_thread;
  PassRange r;
  r.slab = true;
  r.bps = (int)(((int64_t)b->cap_P * S + cpb - 1) / cpb);
  r.G = r.bps * b->d.slots;
  r.G_launch = b->d.xcd_remap ? xadv::xcd_grid(r.bps, b->d.slots) : r.G;
  for (int t = 0; t <= K; ++t) {
    if (t < K) {
      const int k = descending ? K - 1 - t : t;
      r.x0 = k * S;
      r.x1 = (k + 1) * S;
      r.clear_raw = 0;
      launch_pass(b, 1, 0, r);
    }
    if (t >= 1) {
      const int k = descending ? K - t : t - 1;
      r.x0 = k * S;
      r.x1 = (k + 1) * S;
      r.clear_raw = t == K;  // the last launch of the pair: every odd launch has read the raw flag
      launch_pass(b, 0, want_fields, r);
    }
  }
}

void launch_readout(xadv_batch* b, int nodd, int with_stats) {
  if (b->d.precision == 64)
    XADV_DISPATCH(launch_readout_t, double, b, nodd, with_stats);
  else
    XADV_DISPATCH(launch_readout_t, float, b, nodd, with_stats);
}

// ------------------------------------------------------------- flags ------

// Per-cell flag words for one instance (see lattice.h).  Also validates the
// mask against what the AA side planes support.
int build_flags(const xadv_instance_desc* d, int fluid, int thermal, int general, std::vector<uint32_t>& out) {
  using namespace xadv;
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const int px = general ? d->periodic[0] : 0, py = general ? d->periodic[1] : 0, pz = general ? d->periodic[2] : 0;
  const int64_t n = (int64_t)nx * ny * nz;
  auto at = [&](int x, int y, int z) { return ((int64_t)x * ny + y) * nz + z; };
  out.assign(n, 0u);
  bool any_io = false;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int m = d->mask[at(x, y, z)];
        if (m < 0 || m > 3) return fail(XADV_E_ARG, "mask holds a code other than 0..3");
        if (m == BC_INLET && fluid && x != 0)
          return fail(XADV_E_ARG, "INLET cells must lie on x = 0 when the fluid lattice runs");
        if (m == BC_OUTLET && (fluid || thermal) && x != nx - 1)
          return fail(XADV_E_ARG, "OUTLET cells must lie on x = nx-1");
        if (m == BC_INLET || m == BC_OUTLET) any_io = true;
      }
  if (any_io && px) return fail(XADV_E_ARG, "INLET/OUTLET cells need a non-periodic x axis");
  if ((fluid || thermal) && nx < 3) return fail(XADV_E_ARG, "nx must be at least 3");
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t c = at(x, y, z);
        const int m = d->mask[c];
        uint32_t w = (uint32_t)m;
        for (int i = 1; i < 19; ++i) {
          int sx = x - ex_of(i), sy = y - ey_of(i), sz = z - ez_of(i);
          bool oob = false;
          if (sx < 0 || sx >= nx) { if (px) sx = (sx + nx) % nx; else oob = true; }
          if (sy < 0 || sy >= ny) { if (py) sy = (sy + ny) % ny; else oob = true; }
          if (sz < 0 || sz >= nz) { if (pz) sz = (sz + nz) % nz; else oob = true; }
          const bool bb = oob || d->mask[at(sx, sy, sz)] == BC_SOLID;
          if (bb) w |= 1u << (F_BOUNCE0 + i);
        }
        if (!px && x == 0) w |= F_FX0;
        if (!px && x == nx - 1) w |= F_FXN;
        if (!py && y == 0) w |= F_FY0;
        if (!py && y == ny - 1) w |= F_FYN;
        if (!pz && z == 0) w |= F_FZ0;
        if (!pz && z == nz - 1) w |= F_FZN;
        if (d->heat && (double)d->heat[c] * d->heat_power > 0.0) w |= F_HEAT;
        if (general && d->dirichlet_x >= 0 && x == d->dirichlet_x) w |= F_DIR;
        if (x + 1 < nx && d->mask[at(x + 1, y, z)] == BC_OUTLET) w |= F_SNAP;
        if (!px && x == nx - 1 && m != BC_INLET && m != BC_OUTLET) w |= F_XH;
        out[c] = w;
      }
  return XADV_OK;
}

// Device scratch of at least `bytes` for the design builders (kept, grown on demand).
int ensure_gen(xadv_batch* b, size_t bytes) {
  if (b->gen_bytes >= bytes) return XADV_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->gen) {
    HIPCHK(hipFree(b->gen));
    b->gen = nullptr;
    b->gen_bytes = 0;
  }
  HIPCHK(hipMalloc((void**)&b->gen, bytes));
  b->gen_bytes = bytes;
  return XADV_OK;
}

constexpr int kCheckBlocks = 512;                 // blocks of the mask check
constexpr size_t kGenHead = 4096;                 // check partials, error word, cell count
const char* compact_nz_message() {
  return "this batch stores the fluid lattice compactly, which needs nz <= 32; create the batch with "
         "fluid_storage = XADV_STORAGE_DENSE (or cap_nz > 32) for taller lattices";
}

// Slot inputs from device-resident mask and heat map: checks first (nothing
// of the slot is touched when they fail), then flags, q and the column table
// written straight into the slot's arenas.  Returns the non-solid cell count
// through *nfluid.
template <class R>
int build_inputs_device(xadv_batch* b, int s, const xadv_instance_desc* d, int64_t* nfluid) {
  using namespace xadv;
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const int64_t n = (int64_t)nx * ny * nz;
  const int px = b->d.general ? d->periodic[0] : 0;
  int rc = ensure_gen(b, kGenHead);
  if (rc) return rc;
  uint32_t* part = reinterpret_cast<uint32_t*>(b->gen);
  uint32_t* err = part + kCheckBlocks;
  uint64_t* cnt = reinterpret_cast<uint64_t*>(b->gen + 2048 + 64);
  const int nb = (int)std::min<int64_t>(kCheckBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(k_mask_check, dim3(nb), dim3(256), 0, b->stream, d->mask, nx, ny, nz, b->d.fluid,
                     b->d.thermal, part);
  hipLaunchKernelGGL(k_or_reduce, dim3(1), dim3(64), 0, b->stream, part, nb, err);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->host_small, err, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  uint32_t bits;
  memcpy(&bits, b->host_small, sizeof(bits));
  if (bits & ERR_CODE) return fail(XADV_E_ARG, "mask holds a code other than 0..3");
  if (bits & ERR_INLET) return fail(XADV_E_ARG, "INLET cells must lie on x = 0 when the fluid lattice runs");
  if (bits & ERR_OUTLET) return fail(XADV_E_ARG, "OUTLET cells must lie on x = nx-1");
  if ((bits & ERR_IO) && px) return fail(XADV_E_ARG, "INLET/OUTLET cells need a non-periodic x axis");
  if ((b->d.fluid || b->d.thermal) && nx < 3) return fail(XADV_E_ARG, "nx must be at least 3");
  if (b->compact && nz > kCompactMaxNz) return fail(XADV_E_ARG, compact_nz_message());

  FlagParams fp{};
  fp.nx = nx;
  fp.ny = ny;
  fp.nz = nz;
  fp.px = px;
  fp.py = b->d.general ? d->periodic[1] : 0;
  fp.pz = b->d.general ? d->periodic[2] : 0;
  fp.general = b->d.general;
  fp.dirichlet_x = d->dirichlet_x;
  fp.heat_power = d->heat_power;
  uint32_t* flags = reinterpret_cast<uint32_t*>(b->flags + (size_t)s * b->cap_cells * 4);
  R* q = reinterpret_cast<R*>(b->q + (size_t)s * b->cap_cells * sizeof(R));
  hipLaunchKernelGGL((k_slot_flags<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, d->mask,
                     d->heat, fp, flags, q);
  *nfluid = 0;
  if (b->compact) {
    ColInfo* arena = reinterpret_cast<ColInfo*>(b->cols) + (size_t)s * b->cols_per_slot;
    const int64_t ncol = (int64_t)nx * ny, pad = ny + 1;
    hipLaunchKernelGGL(k_col_bits, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, b->stream, d->mask, ncol,
                       nz, arena + pad);
    hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(1024), 0, b->stream, arena, ncol, pad, padded_columns(nx, ny),
                       cnt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->host_small, cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->compact) {
    uint64_t v;
    memcpy(&v, b->host_small, sizeof(v));
    *nfluid = (int64_t)v;
  }
  return XADV_OK;
}

template <class R>
int load_t(xadv_batch* b, int s, const xadv_instance_desc* d, bool device_inputs) {
  const int64_t n = (int64_t)d->nx * d->ny * d->nz;
  std::vector<uint32_t> fl;
  std::vector<xadv::ColInfo> cols;
  std::vector<R> q;
  int64_t nfluid = 0;
  if (device_inputs) {
    int rc = build_inputs_device<R>(b, s, d, &nfluid);
    if (rc) return rc;
  } else {
    int rc = build_flags(d, b->d.fluid, b->d.thermal, b->d.general, fl);
    if (rc) return rc;
    // Compact fluid storage: per (x, y) column, the number of its first
    // non-solid cell and a bit per z for its non-solid cells.
    if (b->compact) {
      if (d->nz > xadv::kCompactMaxNz) return fail(XADV_E_ARG, compact_nz_message());
      cols.resize((size_t)xadv::padded_columns(d->nx, d->ny));
      nfluid = xadv::build_columns(d->mask, d->nx, d->ny, d->nz, cols.data());
    }
    q.assign(n, (R)0);
    if (d->heat) {
      for (int64_t c = 0; c < n; ++c) q[c] = (R)d->heat[c] * (R)d->heat_power;  // thermal_engine3d.py:116-117
    }
  }
  Slot<R>& S = host_slots<R>(b)[s];
  const int64_t cap = b->cap_cells, capn = b->cap_npad, capP = b->cap_P;
  const size_t rs = sizeof(R);
  S.nx = d->nx;
  S.ny = d->ny;
  S.nz = d->nz;
  S.P = d->ny * d->nz;
  S.ncell = n;
  S.npad = round_up(n, 64);
  S.active = 0;
  S.raw = 0;
  S.px = b->d.general ? d->periodic[0] : 0;
  S.py = b->d.general ? d->periodic[1] : 0;
  S.pz = b->d.general ? d->periodic[2] : 0;
  S.dir_x = b->d.general ? d->dirichlet_x : -1;
  // Scalars rounded to the lattice precision and derived there, as the
  // reference's kernels do with their float arguments.
  const R tau = (R)d->tau, ttf = (R)d->tau_th_fluid, tts = (R)d->tau_th_solid;
  S.omega = (R)1.0 / tau;
  S.rho_out = (R)d->rho_out;
  S.u_in = (R)d->u_in;
  S.force_x = b->d.general ? (R)d->force_x : (R)0;
  S.u_wall = b->d.general ? (R)d->u_wall : (R)0;
  S.omega_tf = (R)1.0 / ttf;
  S.omega_ts = (R)1.0 / tts;
  S.sf_tf = (R)1.0 - (R)0.5 * S.omega_tf;
  S.sf_ts = (R)1.0 - (R)0.5 * S.omega_ts;
  if (b->q7) {
    // D3Q7 thermal variant: 1/tau7 with tau7 = 1/2 + 4/3 (tau - 1/2), the diffusivity of the reference's tau;
    // sf_tf / sf_ts keep the reference step's source factor 1 - 1/(2 tau).
    S.omega_tf = (R)1.0 / ((R)0.5 + (ttf - (R)0.5) * (R)(4.0 / 3.0));
    S.omega_ts = (R)1.0 / ((R)0.5 + (tts - (R)0.5) * (R)(4.0 / 3.0));
  }
  S.T_in = (R)d->T_in;
  S.dir_T = (R)d->dirichlet_T;
  S.nf = (uint32_t)(b->compact ? round_up(nfluid, 64) : S.npad);
  S.nz_magic = xadv::nz_magic((uint32_t)d->nz);
  S.f = reinterpret_cast<R*>(b->f + (size_t)s * 19 * capn * rs);
  S.g = reinterpret_cast<R*>(b->g + (size_t)s * 19 * capn * rs);
  // The slot's padded column table; S.cols points at lattice column 0 (entry ny + 1).
  xadv::ColInfo* col_arena =
      b->compact ? reinterpret_cast<xadv::ColInfo*>(b->cols) + (size_t)s * b->cols_per_slot : nullptr;
  S.cols = b->compact ? col_arena + d->ny + 1 : nullptr;
  S.flags = reinterpret_cast<const uint32_t*>(b->flags + (size_t)s * cap * 4);
  S.q = reinterpret_cast<const R*>(b->q + (size_t)s * cap * rs);
  S.fin = reinterpret_cast<R*>(b->fin + (size_t)s * 19 * capP * rs);
  S.fsnap = reinterpret_cast<R*>(b->fsnap + (size_t)s * 2 * 19 * capP * rs);
  S.gsnap = reinterpret_cast<R*>(b->gsnap + (size_t)s * 2 * 19 * capP * rs);
  S.gxh = reinterpret_cast<R*>(b->gxh + (size_t)s * 5 * capP * rs);
  S.rho = reinterpret_cast<R*>(b->rho + (size_t)s * cap * rs);
  S.u = reinterpret_cast<R*>(b->u + (size_t)s * 3 * cap * rs);
  S.T = reinterpret_cast<R*>(b->T + (size_t)s * cap * rs);
  S.g7 = b->q7 ? reinterpret_cast<R*>(b->g7 + (size_t)s * 2 * 7 * capn * rs) : nullptr;
  if (!device_inputs) {
    HIPCHK(hipMemcpyAsync((void*)S.flags, fl.data(), n * 4, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync((void*)S.q, q.data(), n * rs, hipMemcpyHostToDevice, b->stream));
    if (b->compact)
      HIPCHK(hipMemcpyAsync(col_arena, cols.data(), cols.size() * sizeof(xadv::ColInfo), hipMemcpyHostToDevice,
                            b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  b->loaded[s] = 1;
  b->inited[s] = 0;
  b->frozen[s] = d->frozen_velocity;
  return push_slot<R>(b, s);
}

template <class R>
int init_t(xadv_batch* b, int s, const void* rho, const void* u, const void* T) {
  Slot<R>& S = host_slots<R>(b)[s];
  const int64_t n = S.ncell;
  const size_t rs = sizeof(R);
  const int bs = 256;
  const unsigned gb = (unsigned)((n + bs - 1) / bs), gb3 = (unsigned)((3 * n + bs - 1) / bs);
  if (rho)
    HIPCHK(hipMemcpyAsync(S.rho, rho, n * rs, hipMemcpyDeviceToDevice, b->stream));
  else
    hipLaunchKernelGGL((xadv::k_fill<R>), dim3(gb), dim3(bs), 0, b->stream, S.rho, n, (R)1);
  if (u)
    HIPCHK(hipMemcpyAsync(S.u, u, 3 * n * rs, hipMemcpyDeviceToDevice, b->stream));
  else
    hipLaunchKernelGGL((xadv::k_fill<R>), dim3(gb3), dim3(bs), 0, b->stream, S.u, 3 * n, (R)0);
  if (T)
    HIPCHK(hipMemcpyAsync(S.T, T, n * rs, hipMemcpyDeviceToDevice, b->stream));
  else
    hipLaunchKernelGGL((xadv::k_fill<R>), dim3(gb), dim3(bs), 0, b->stream, S.T, n, S.T_in);
  HIPCHK(hipGetLastError());
  XADV_DISPATCH(launch_init_t, R, b, s, n, b->frozen[s]);
  HIPCHK(hipGetLastError());
  if (b->q7) {
    hipLaunchKernelGGL((xadv::k_init7<R>), dim3(gb), dim3(bs), 0, b->stream, reinterpret_cast<Slot<R>*>(b->slots_dev),
                       s);
    HIPCHK(hipGetLastError());
  }
  S.active = 1;
  S.raw = (b->d.thermal || b->q7) ? 1 : 0;
  b->inited[s] = 1;
  return push_slot<R>(b, s);
}

}  // namespace

namespace {
int cleanup_early(xadv_batch* b, int rc) {
  delete b;
  return rc;
}

// Frees everything a batch holds, including partially created batches. Every
// release is attempted; the first HIP error is returned and the last-error
// text is left untouched, so a failed create keeps its own message.
hipError_t release_batch(xadv_batch* b) {
  if (!b) return hipSuccess;
  hipError_t first = hipSuccess;
  auto keep = [&](hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  };
  if (b->stream) keep(hipStreamSynchronize(b->stream));
  for (auto& kv : b->graphs) keep(hipGraphExecDestroy(kv.second));
  char** ps[] = {&b->f, &b->g, &b->flags, &b->q, &b->fin, &b->fsnap, &b->gsnap, &b->gxh, &b->rho, &b->u, &b->T,
                 &b->cols, &b->g7};
  for (char** p : ps)
    if (*p) keep(hipFree(*p));
  if (b->partials) keep(hipFree(b->partials));
  if (b->stats_dev) keep(hipFree(b->stats_dev));
  if (b->met_partials) keep(hipFree(b->met_partials));
  if (b->met_dev) keep(hipFree(b->met_dev));
  if (b->slots_dev) keep(hipFree(b->slots_dev));
  if (b->slots_host) keep(hipHostFree(b->slots_host));
  if (b->host_small) keep(hipHostFree(b->host_small));
  if (b->gen) keep(hipFree(b->gen));
  if (b->stream) keep(hipStreamDestroy(b->stream));
  delete b;
  return first;
}
}  // namespace

// ================================================================ C ABI ====

extern "C" {

const char* xadv_last_error(void) { return g_err.c_str(); }
const char* xadv_version(void) { return "xadvcool-amd 0.1 (gfx950)"; }

int xadv_get_stream(xadv_batch* b, void** stream) {
  if (!b || !stream) return fail(XADV_E_ARG, "null argument");
  *stream = (void*)b->stream;
  return XADV_OK;
}

int xadv_batch_create(const xadv_batch_desc* desc, xadv_batch** out) {
  if (!desc || !out) return fail(XADV_E_ARG, "null argument");
  *out = nullptr;
  xadv_batch_desc d = *desc;
  if (d.precision != 32 && d.precision != 64) return fail(XADV_E_ARG, "precision must be 32 or 64");
  if (!d.fluid && !d.thermal) return fail(XADV_E_ARG, "enable the fluid or the thermal lattice");
  if (d.slots < 1) return fail(XADV_E_ARG, "slots must be >= 1");
  if (d.cap_nx < 1 || d.cap_ny < 1 || d.cap_nz < 1) return fail(XADV_E_ARG, "capacity dims must be >= 1");
  if (d.block == 0) d.block = 256;
  if (d.block != 64 && d.block != 128 && d.block != 256) return fail(XADV_E_ARG, "block must be 64, 128 or 256");
  if (d.graph_chunk < 0 || (d.graph_chunk % 2)) return fail(XADV_E_ARG, "graph_chunk must be even and >= 0");
  if (d.cells_per_thread == 0) d.cells_per_thread = 2;
  if (d.cells_per_thread != 1 && d.cells_per_thread != 2) return fail(XADV_E_ARG, "cells_per_thread must be 1 or 2");
  if (d.xcd_remap != 0 && d.xcd_remap != 1) return fail(XADV_E_ARG, "xcd_remap must be 0 or 1");
  if (d.alternate_sweep != 0 && d.alternate_sweep != 1) return fail(XADV_E_ARG, "alternate_sweep must be 0 or 1");
  if (d.fluid_storage < XADV_STORAGE_AUTO || d.fluid_storage > XADV_STORAGE_COMPACT)
    return fail(XADV_E_ARG, "fluid_storage must be XADV_STORAGE_AUTO, _DENSE or _COMPACT");
  if (d.thermal_q != 0 && d.thermal_q != 19 && d.thermal_q != 7)
    return fail(XADV_E_ARG, "thermal_q must be 19 (or 0) or 7");
  const bool q7 = d.thermal && d.thermal_q == 7;
  if (q7 && (!d.fluid || d.general))
    return fail(XADV_E_ARG, "the D3Q7 thermal lattice runs coupled to the fluid lattice, channel boundaries only");
  if (q7) d.thermal = 0;  // init and readout run their fluid-only kernels and the D3Q7 ones beside them
  if (d.slab_planes < 0) return fail(XADV_E_ARG, "slab_planes must be >= 0");
  if (d.slab_planes > 0 && (q7 || d.general))
    return fail(XADV_E_ARG, "x-slab pairs of passes need the D3Q19 lattices and general = 0 (no periodic x axis)");
  const bool can_compact = d.fluid && !d.general;
  if (d.fluid_storage == XADV_STORAGE_COMPACT && !can_compact)
    return fail(XADV_E_ARG, "compact fluid storage needs the fluid lattice on and general = 0");
  const bool compact = can_compact && (d.fluid_storage == XADV_STORAGE_COMPACT ||
                                       (d.fluid_storage == XADV_STORAGE_AUTO && d.cap_nz <= xadv::kCompactMaxNz));
  if (d.thermal_prefetch != 0 && d.thermal_prefetch != 1) return fail(XADV_E_ARG, "thermal_prefetch must be 0 or 1");
  if (d.thermal_prefetch && !(compact && d.thermal))
    return fail(XADV_E_ARG, "thermal_prefetch is for the fused D3Q19 pass with compact fluid storage");
  auto* b = new xadv_batch();
  b->d = d;
  b->q7 = q7 ? 1 : 0;
  b->compact = compact;
  b->rsize = d.precision == 64 ? 8 : 4;
  b->cap_cells = (int64_t)d.cap_nx * d.cap_ny * d.cap_nz;
  b->cap_npad = round_up(b->cap_cells, 64);
  b->cap_P = d.cap_ny * d.cap_nz;
  // A slot's lattice has nx * ny <= cap_cells columns and ny <= cap_P, so its
  // padded table, (nx + 2) * ny + 2 entries, fits in cap_cells + 2 * cap_P + 2.
  b->cols_per_slot = b->cap_cells + 2 * (int64_t)b->cap_P + 2;
  // Every lattice array of a slot is addressed with 32-bit byte offsets.
  if ((double)19 * round_up(b->cap_cells, 64) * b->rsize >= 4294967296.0)
    return cleanup_early(b, fail(XADV_E_ARG, "slot capacity too large for 32-bit addressing (max ~56M cells FP32)"));
  const int cpb = d.block * d.cells_per_thread;
  b->bps = (int)((b->cap_cells + cpb - 1) / cpb);
  b->G = b->bps * d.slots;
  b->G_launch = d.xcd_remap ? xadv::xcd_grid(b->bps, d.slots) : b->G;
  b->bps_r = (int)((b->cap_cells + d.block - 1) / d.block);
  b->G_r = b->bps_r * d.slots;
  b->G_r_launch = d.xcd_remap ? xadv::xcd_grid(b->bps_r, d.slots) : b->G_r;
  b->bps_s = (int)((b->cap_cells + (int64_t)d.block * xadv::kReadCells - 1) / ((int64_t)d.block * xadv::kReadCells));
  b->G_s = b->bps_s * d.slots;
  b->G_s_launch = d.xcd_remap ? xadv::xcd_grid(b->bps_s, d.slots) : b->G_s;
  b->slot_bytes = d.precision == 64 ? sizeof(Slot<double>) : sizeof(Slot<float>);
  b->loaded.assign(d.slots, 0);
  b->inited.assign(d.slots, 0);
  b->frozen.assign(d.slots, 0);
  const size_t rs = b->rsize, S = d.slots;
  const size_t capn = b->cap_npad, cap = b->cap_cells, capP = b->cap_P;
  auto cleanup = [&](int rc) {
    (void)release_batch(b);
    return rc;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
    return cleanup(fail(XADV_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  struct A {
    char** p;
    size_t bytes;
  } allocs[] = {
      {&b->f, d.fluid ? S * 19 * capn * rs : 0},
      {&b->g, d.thermal ? S * 19 * capn * rs : 0},
      {&b->flags, S * cap * 4},
      {&b->q, S * cap * rs},
      {&b->fin, S * 19 * capP * rs},
      {&b->fsnap, S * 2 * 19 * capP * rs},
      {&b->gsnap, S * 2 * 19 * capP * rs},
      {&b->gxh, S * 5 * capP * rs},
      {&b->rho, S * cap * rs},
      {&b->u, S * 3 * cap * rs},
      {&b->T, S * cap * rs},
      {&b->cols, b->compact ? S * (size_t)b->cols_per_slot * sizeof(xadv::ColInfo) : 0},
      {&b->g7, q7 ? S * 2 * 7 * capn * rs : 0},
  };
  for (auto& a : allocs) {
    if (!a.bytes) continue;
    if ((e = hipMalloc((void**)a.p, a.bytes)) != hipSuccess)
      return cleanup(fail(XADV_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)));
    if ((e = hipMemsetAsync(*a.p, 0, a.bytes, b->stream)) != hipSuccess)
      return cleanup(fail(XADV_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e)));
  }
  const int nmb = (int)((b->cap_cells + 255) / 256);
  if ((e = hipMalloc((void**)&b->partials, (size_t)b->G_r * 16 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&b->stats_dev, S * 16 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&b->met_partials, (size_t)nmb * xadv::kNMet * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&b->met_dev, xadv::kNMet * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&b->slots_dev, b->slot_bytes * S)) != hipSuccess ||
      (e = hipHostMalloc(&b->slots_host, b->slot_bytes * S, hipHostMallocDefault)) != hipSuccess ||
      (e = hipHostMalloc((void**)&b->host_small, S * 16 * sizeof(double) + 64 * sizeof(double),
                         hipHostMallocDefault)) != hipSuccess)
    return cleanup(fail(XADV_E_HIP, std::string("allocation: ") + hipGetErrorString(e)));
  memset(b->slots_host, 0, b->slot_bytes * S);
  if ((e = hipMemsetAsync(b->stats_dev, 0, S * 16 * sizeof(double), b->stream)) != hipSuccess)
    return cleanup(fail(XADV_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e)));
  int rc = push_slots(b);
  if (rc) return cleanup(rc);
  *out = b;
  return XADV_OK;
}

int xadv_batch_destroy(xadv_batch* b) {
  hipError_t e = release_batch(b);
  if (e != hipSuccess) return fail(XADV_E_HIP, std::string("xadv_batch_destroy: ") + hipGetErrorString(e));
  return XADV_OK;
}

}  // extern "C"

namespace {
int instance_load(xadv_batch* b, int slot, const xadv_instance_desc* d, bool device_inputs) {
  if (!b || !d || !d->mask) return fail(XADV_E_ARG, "null argument");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(XADV_E_ARG, "dims must be >= 1");
  if ((int64_t)d->nx * d->ny * d->nz > b->cap_cells || (int64_t)d->ny * d->nz > b->cap_P)
    return fail(XADV_E_ARG, "lattice larger than the slot capacity");
  if (!(d->tau > 0.5) || !(d->tau_th_fluid > 0.5) || !(d->tau_th_solid > 0.5))
    return fail(XADV_E_ARG, "relaxation times must exceed 0.5");
  if (d->frozen_velocity && b->d.fluid) return fail(XADV_E_ARG, "frozen_velocity needs a thermal-only batch");
  return b->d.precision == 64 ? load_t<double>(b, slot, d, device_inputs) : load_t<float>(b, slot, d, device_inputs);
}
}  // namespace

extern "C" {

int xadv_instance_load(xadv_batch* b, int slot, const xadv_instance_desc* d) {
  return instance_load(b, slot, d, false);
}

int xadv_instance_load_device(xadv_batch* b, int slot, const xadv_instance_desc* d) {
  return instance_load(b, slot, d, true);
}

int xadv_geometry_mask(xadv_batch* b, const xadv_geometry_desc* g, int8_t* mask) {
  using namespace xadv;
  if (!b || !g || !mask) return fail(XADV_E_ARG, "null argument");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(XADV_E_ARG, "dims must be >= 1");
  if (g->buffer < 0 || 2 * (int64_t)g->buffer >= g->nx) return fail(XADV_E_ARG, "need 0 <= buffer and 2 buffer < nx");
  if (g->base < 0 || g->top < 1) return fail(XADV_E_ARG, "need base >= 0 and top >= 1");
  if (g->kind < XADV_CORE_EMPTY || g->kind > XADV_CORE_SCHWARZ_D) return fail(XADV_E_ARG, "unknown core kind");
  GeomParams p{};
  p.nx = g->nx;
  p.ny = g->ny;
  p.nz = g->nz;
  p.buf = g->buffer;
  p.base = g->base;
  p.top = g->top;
  p.kind = g->kind;
  p.cnx = g->nx - 2 * g->buffer;
  p.manifold = g->manifold;
  // Parameter mapping of src/geometry.py:394-398 (raw_width = feature_size,
  // raw_count = spacing), evaluated as Python does (truncating int()).
  const double raw_width = g->feature_size, raw_count = g->spacing;
  if (g->kind == XADV_CORE_STRAIGHT || g->kind == XADV_CORE_PINFIN) {
    if (!(raw_width > 0.0) || !(raw_count > 0.0) || raw_width > 1e6 || raw_count > 1e6)
      return fail(XADV_E_ARG, "feature_size and spacing must be positive");
  }
  if (g->kind == XADV_CORE_STRAIGHT) {
    const int cw = std::max(2, (int)(raw_width * 2.0));
    const int target = std::max(4, (int)(raw_count * 4.0));
    const int max_n = (int)((double)(g->ny - 1) / (double)(cw + 1));
    const int nch = std::max(2, std::min(target, max_n));
    p.channel_width = cw;
    p.n_channels = nch;
    p.straight_carves = (int64_t)nch * cw < g->ny;
    p.gap = (double)(g->ny - nch * cw) / (double)(nch + 1);
  } else if (g->kind == XADV_CORE_PINFIN) {
    const int d_pin = std::max(2, (int)(raw_width * 2.0));
    p.pitch = std::max(d_pin + 2, (int)(raw_count * 3.0));
    p.half_pitch = (double)p.pitch / 2.0;
    const double radius = (double)d_pin / 2.0;
    p.radius2 = radius * radius;
  } else if (g->kind >= XADV_CORE_GYROID) {
    if (g->manifold < 0) return fail(XADV_E_ARG, "manifold must be >= 0");
    if (g->period > 0.0) {
      p.omega = 2.0 * M_PI / g->period;
      p.level = g->threshold;
    } else {
      const double cell = 20.0 + ((raw_count - 1.0) / 9.0) * 80.0;
      if (!(cell > 0.0)) return fail(XADV_E_ARG, "TPMS cell size must be positive");
      p.level = 0.8 - ((raw_width - 1.0) / 9.0) * 1.6;
      p.omega = 2.0 * M_PI / cell;
    }
  }
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_geometry_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, mask, p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

int xadv_heat_map(xadv_batch* b, const xadv_heat_desc* h, int nx, int ny, int nz, int x_offset, float* heat) {
  using namespace xadv;
  if (!b || !h || !heat) return fail(XADV_E_ARG, "null argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(XADV_E_ARG, "dims must be >= 1");
  if (h->mx < 1 || h->my < 1 || x_offset < 0 || x_offset + h->mx > nx || h->my > ny)
    return fail(XADV_E_ARG, "the map must fit the lattice's z = 0 plane at x_offset");
  if (h->nprims < 0 || h->nprims > XADV_HEAT_MAX_PRIMS || (h->nprims && !h->prims))
    return fail(XADV_E_ARG, "nprims must be 0..XADV_HEAT_MAX_PRIMS with a prims array");
  if (h->radius < 0 || h->radius > XADV_HEAT_MAX_RADIUS || (h->radius && !h->weights))
    return fail(XADV_E_ARG, "radius must be 0..XADV_HEAT_MAX_RADIUS with a weights array");
  if (h->radius >= h->mx || h->radius >= h->my) return fail(XADV_E_ARG, "blur radius must be below the map size");
  HeatParams hp{};
  hp.mx = h->mx;
  hp.my = h->my;
  hp.base = h->base;
  hp.nprims = h->nprims;
  hp.radius = h->radius;
  hp.normalise = h->normalise ? 1 : 0;
  for (int k = 0; k < h->nprims; ++k) {
    const xadv_heat_prim& a = h->prims[k];
    if (a.kind < XADV_HEAT_RECT || a.kind > XADV_HEAT_GAUSS) return fail(XADV_E_ARG, "unknown heat primitive");
    if (a.kind == XADV_HEAT_GAUSS && !(a.den > 0.0)) return fail(XADV_E_ARG, "Gaussian primitive needs den > 0");
    HeatPrim& o = hp.prims[k];
    o.kind = a.kind;
    o.x0 = a.x0;
    o.x1 = a.x1;
    o.y0 = a.y0;
    o.y1 = a.y1;
    o.cx = a.cx;
    o.cy = a.cy;
    o.r2 = a.r2;
    o.value = a.value;
    o.gx = a.gx;
    o.gy = a.gy;
    o.den = a.den;
    o.amplitude = a.amplitude;
  }
  for (int j = 0; j <= h->radius && h->radius; ++j) hp.w[j] = h->weights[j];
  const int64_t m = (int64_t)h->mx * h->my, n = (int64_t)nx * ny * nz;
  int rc = ensure_gen(b, kGenHead + 2 * (size_t)m * sizeof(float));
  if (rc) return rc;
  float* A = reinterpret_cast<float*>(b->gen + kGenHead);
  float* B = A + m;
  const unsigned gm = (unsigned)((m + 255) / 256);
  hipLaunchKernelGGL((k_fill<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, heat, n, 0.0f);
  hipLaunchKernelGGL(k_heat_paint, dim3(gm), dim3(256), 0, b->stream, A, hp);
  const float* src = A;
  if (hp.radius) {
    hipLaunchKernelGGL(k_heat_blur<0>, dim3(gm), dim3(256), 0, b->stream, A, B, hp);
    hipLaunchKernelGGL(k_heat_blur<1>, dim3(gm), dim3(256), 0, b->stream, B, A, hp);
  }
  hipLaunchKernelGGL(k_heat_place, dim3(1), dim3(1024), 0, b->stream, src, heat, hp, ny, nz, x_offset);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

int xadv_instance_init(xadv_batch* b, int slot, const void* rho, const void* u, const void* T) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (!b->loaded[slot]) return fail(XADV_E_STATE, "slot not loaded");
  if (b->d.slots == 1) b->npass = 0;
  if (b->npass % 2) return fail(XADV_E_STATE, "init needs an even step count of the batch; step once more first");
  return b->d.precision == 64 ? init_t<double>(b, slot, rho, u, T) : init_t<float>(b, slot, rho, u, T);
}

int xadv_set_active(xadv_batch* b, int slot, int active) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (active && !b->inited[slot]) return fail(XADV_E_STATE, "slot not initialised");
  if (b->d.precision == 64) {
    host_slots<double>(b)[slot].active = active ? 1 : 0;
    return push_field<double>(b, slot, offsetof(Slot<double>, active), sizeof(int));
  }
  host_slots<float>(b)[slot].active = active ? 1 : 0;
  return push_field<float>(b, slot, offsetof(Slot<float>, active), sizeof(int));
}

int xadv_set_inlet_velocity(xadv_batch* b, int slot, double u_in) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (b->d.precision == 64) {
    host_slots<double>(b)[slot].u_in = (double)u_in;
    return push_field<double>(b, slot, offsetof(Slot<double>, u_in), sizeof(double));
  }
  host_slots<float>(b)[slot].u_in = (float)u_in;
  return push_field<float>(b, slot, offsetof(Slot<float>, u_in), sizeof(float));
}

int xadv_step(xadv_batch* b, int nsteps, int want_fields) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (nsteps < 0) return fail(XADV_E_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return XADV_OK;
  const int chunk = b->d.graph_chunk;
  // x-slab pairs (slab_planes > 0) need pairs of passes that start on an odd pass; other runs take whole passes.
  const bool slabs = b->d.slab_planes > 0 && b->npass % 2 == 0 && nsteps % 2 == 0;
  if (chunk && nsteps % chunk == 0 && b->npass % 2 == 0) {
    const std::pair<int, int> key(chunk, want_fields ? 1 : 0);
    const std::pair<int, int> key0(chunk, 0);
    for (int k = 0; k < nsteps / chunk; ++k) {
      const bool last = k == nsteps / chunk - 1;
      const auto kk = (last && want_fields) ? key : key0;
      auto it = b->graphs.find(kk);
      if (it == b->graphs.end()) {
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        if (b->d.slab_planes > 0) {  // chunk is even: chunk / 2 pairs, the sweep direction alternating
          for (int j = 0; j < chunk / 2; ++j)
            launch_pass_pair_slabs(b, (kk.second && j == chunk / 2 - 1) ? 1 : 0, (j & 1) != 0);
        } else {
          for (int j = 0; j < chunk; ++j) launch_pass(b, (j + 1) & 1, (kk.second && j == chunk - 1) ? 1 : 0);
        }
        hipError_t le = hipGetLastError();
        HIPCHK(hipStreamEndCapture(b->stream, &gr));
        if (le != hipSuccess) return fail(XADV_E_HIP, std::string("launch in capture: ") + hipGetErrorString(le));
        hipGraphExec_t ex;
        HIPCHK(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(gr));
        it = b->graphs.emplace(kk, ex).first;
      }
      HIPCHK(hipGraphLaunch(it->second, b->stream));
      b->npass += chunk;
    }
  } else if (slabs) {
    for (int k = 0; k < nsteps / 2; ++k) {
      launch_pass_pair_slabs(b, (want_fields && k == nsteps / 2 - 1) ? 1 : 0, ((b->npass / 2) & 1) != 0);
      b->npass += 2;
    }
    HIPCHK(hipGetLastError());
  } else {
    for (int k = 0; k < nsteps; ++k) {
      b->npass += 1;
      launch_pass(b, (int)(b->npass & 1), (want_fields && k == nsteps - 1) ? 1 : 0);
    }
    HIPCHK(hipGetLastError());
  }
  b->fields_fresh = want_fields != 0;
  return XADV_OK;
}

int xadv_readout(xadv_batch* b, int with_stats) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (b->npass < 1) return fail(XADV_E_STATE, "readout needs at least one step");
  if (!b->fields_fresh) return fail(XADV_E_STATE, "readout needs the last xadv_step with want_fields = 1");
  if (b->q7) {
    const int last_odd = (int)(b->npass & 1);
    if (b->d.precision == 64)
      hipLaunchKernelGGL((xadv::k_readout7<double>), dim3(b->G_r_launch), dim3(b->d.block), 0, b->stream,
                         reinterpret_cast<Slot<double>*>(b->slots_dev), b->bps_r, b->G_r, b->d.xcd_remap, last_odd);
    else
      hipLaunchKernelGGL((xadv::k_readout7<float>), dim3(b->G_r_launch), dim3(b->d.block), 0, b->stream,
                         reinterpret_cast<Slot<float>*>(b->slots_dev), b->bps_r, b->G_r, b->d.xcd_remap, last_odd);
  }
  launch_readout(b, (int)((b->npass + 1) & 1), with_stats);
  HIPCHK(hipGetLastError());
  if (with_stats) {
    if (b->d.precision == 64)
      hipLaunchKernelGGL((xadv::k_stats_final<double>), dim3(b->d.slots), dim3(xadv::kFoldThreads), 0, b->stream,
                         reinterpret_cast<const Slot<double>*>(b->slots_dev), b->bps_s, b->partials, b->stats_dev);
    else
      hipLaunchKernelGGL((xadv::k_stats_final<float>), dim3(b->d.slots), dim3(xadv::kFoldThreads), 0, b->stream,
                         reinterpret_cast<const Slot<float>*>(b->slots_dev), b->bps_s, b->partials, b->stats_dev);
    HIPCHK(hipGetLastError());
  }
  return XADV_OK;
}

int xadv_stats(xadv_batch* b, double* out) {
  if (!b || !out) return fail(XADV_E_ARG, "null argument");
  const size_t bytes = (size_t)b->d.slots * 16 * sizeof(double);
  HIPCHK(hipMemcpyAsync(b->host_small, b->stats_dev, bytes, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  memcpy(out, b->host_small, bytes);
  return XADV_OK;
}

int xadv_fields(xadv_batch* b, int slot, void* rho, void* u, void* T) {
  if (!b) return fail(XADV_E_ARG, "null batch");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (!b->inited[slot]) return fail(XADV_E_STATE, "slot not initialised");
  const size_t rs = b->rsize;
  int64_t n;
  void *dr, *du, *dT;
  if (b->d.precision == 64) {
    auto& S = host_slots<double>(b)[slot];
    n = S.ncell; dr = S.rho; du = S.u; dT = S.T;
  } else {
    auto& S = host_slots<float>(b)[slot];
    n = S.ncell; dr = S.rho; du = S.u; dT = S.T;
  }
  if (rho) HIPCHK(hipMemcpyAsync(rho, dr, n * rs, hipMemcpyDeviceToDevice, b->stream));
  if (u) HIPCHK(hipMemcpyAsync(u, du, 3 * n * rs, hipMemcpyDeviceToDevice, b->stream));
  if (T) HIPCHK(hipMemcpyAsync(T, dT, n * rs, hipMemcpyDeviceToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return XADV_OK;
}

int xadv_metrics(xadv_batch* b, int slot, int core_x0, int core_nx, int core_ny, double* out) {
  if (!b || !out) return fail(XADV_E_ARG, "null argument");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (!b->inited[slot]) return fail(XADV_E_STATE, "slot not initialised");
  int64_t n = b->d.precision == 64 ? host_slots<double>(b)[slot].ncell : host_slots<float>(b)[slot].ncell;
  const int nb = (int)((n + 255) / 256);
  if (b->d.precision == 64)
    hipLaunchKernelGGL((xadv::k_metrics<double>), dim3(nb), dim3(256), 0, b->stream,
                       reinterpret_cast<Slot<double>*>(b->slots_dev), slot, core_x0, core_nx, core_ny, b->met_partials);
  else
    hipLaunchKernelGGL((xadv::k_metrics<float>), dim3(nb), dim3(256), 0, b->stream,
                       reinterpret_cast<Slot<float>*>(b->slots_dev), slot, core_x0, core_nx, core_ny, b->met_partials);
  hipLaunchKernelGGL(xadv::k_metrics_final, dim3(1), dim3(xadv::kFoldThreads), 0, b->stream, nb, b->met_partials,
                     b->met_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->host_small, b->met_dev, xadv::kNMet * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  memcpy(out, b->host_small, XADV_NMETRIC * sizeof(double));
  return XADV_OK;
}

int xadv_checks(xadv_batch* b, int slot, double k_fluid, double k_solid, double* out) {
  if (!b || !out) return fail(XADV_E_ARG, "null argument");
  if (slot < 0 || slot >= b->d.slots) return fail(XADV_E_ARG, "slot out of range");
  if (!b->inited[slot]) return fail(XADV_E_STATE, "slot not initialised");
  int64_t n = b->d.precision == 64 ? host_slots<double>(b)[slot].ncell : host_slots<float>(b)[slot].ncell;
  const int nb = (int)((n + 255) / 256);
  static_assert(xadv::kNChk <= xadv::kNMet, "checks reuse the metrics partial buffer");
  if (b->d.precision == 64)
    hipLaunchKernelGGL((xadv::k_checks<double>), dim3(nb), dim3(256), 0, b->stream,
                       reinterpret_cast<Slot<double>*>(b->slots_dev), slot, k_fluid, k_solid, b->met_partials);
  else
    hipLaunchKernelGGL((xadv::k_checks<float>), dim3(nb), dim3(256), 0, b->stream,
                       reinterpret_cast<Slot<float>*>(b->slots_dev), slot, k_fluid, k_solid, b->met_partials);
  hipLaunchKernelGGL(xadv::k_checks_final, dim3(1), dim3(xadv::kFoldThreads), 0, b->stream, nb, b->met_partials,
                     b->met_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->host_small, b->met_dev, xadv::kNChk * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  memcpy(out, b->host_small, XADV_NCHECK * sizeof(double));
  return XADV_OK;
}

}  // extern "C"
