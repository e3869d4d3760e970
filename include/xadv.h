/* xadv.h — C ABI of the MI355X lattice-Boltzmann cold-plate solver (libxadv.so).
 *
 * The reference (tgautam03/xAdvCool) has no C ABI: its hot path is a set of
 * Warp kernels (src/fluid_engine3d.py, src/thermal_engine3d.py) launched by
 * the Python driver src/simulation.py.  Each entry point below names the
 * reference code it replaces.  Conventions:
 *   - every function returns 0 on success, non-zero on error (XADV_E_*), and
 *     never throws; xadv_last_error() gives the message (thread-local);
 *   - host pointers are marked "host", device pointers "device"; device
 *     buffers passed in are owned by the caller, buffers inside a batch by the
 *     batch;
 *   - all work runs on the batch's own HIP stream (created with default
 *     priority); calls that hand data back to the caller synchronise it;
 *   - one batch per host thread; calls on one batch must not overlap.
 * Lattice values are FP32 (precision 32) or FP64 (precision 64); the
 * "lattice precision" below is that type.
 */
#ifndef XADV_H
#define XADV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XADV_OK 0
#define XADV_E_ARG 1     /* invalid argument or unsupported mask layout */
#define XADV_E_HIP 2     /* a HIP runtime call failed */
#define XADV_E_STATE 3   /* call out of order (e.g. step before init) */

/* Cell codes of the mask: src/fluid_engine3d.py:11-14 (BC_FLUID/SOLID/INLET/OUTLET). */
#define XADV_FLUID 0
#define XADV_SOLID 1
#define XADV_INLET 2
#define XADV_OUTLET 3

typedef struct xadv_batch xadv_batch;

/* A batch of lattice slots advanced together.  A single solve is a batch of
 * one slot sized to its lattice.  Replaces the per-solve allocations of
 * src/simulation.py:60-75 and the captured six-launch graph of
 * src/simulation.py:93-100. */
typedef struct {
  int precision;        /* 32 or 64 */
  int fluid;            /* 1: advance the D3Q19 fluid lattice */
  int thermal;          /* 1: advance the D3Q19 thermal lattice */
  int general;          /* 1: build the kernels that also take periodic axes, a moving
                           top wall, a body force and a Dirichlet plane (test modes of
                           validate_analytical.py); 0: channel boundaries only */
  int slots;            /* number of lattice slots */
  int cap_nx, cap_ny, cap_nz;  /* capacity of each slot (largest lattice it can hold) */
  int block;            /* threads per block: 64, 128 or 256; 0 = default (256) */
  int xcd_remap;        /* 1: XCD-contiguous block order (each of the 8 XCDs takes one eighth of
                           every slot, so neighbour lines stay in its L2); 0: linear order */
  int graph_chunk;      /* steps captured per hipGraph (even, >= 2); 0 = no graphs */
  int cells_per_thread; /* lattice cells per thread in the fused pass: 1 or 2; 0 = default (2) */
  int fluid_storage;    /* fluid populations: XADV_STORAGE_AUTO (compact where it can be used),
                           XADV_STORAGE_DENSE (every cell) or XADV_STORAGE_COMPACT (non-solid
                           cells only: fluid on, general 0, every lattice nz <= 32).  Results are
                           the same bits either way; compact moves fewer bytes per step. */
  int alternate_sweep;  /* 1: even passes visit the lattice blocks in reverse order, so a pass
                           starts on the lines the previous pass touched last (still in the
                           Infinity Cache when the batch's state is larger than it).  Same bits. */
  int thermal_q;        /* thermal lattice: 19 (or 0), the reference's D3Q19; 7, the D3Q7 variant
                           (no reference counterpart): seven velocities, weights 1/4 and 1/8,
                           g_eq = w T (1 + 4 e.u), relaxed at the diffusivity of the instance's
                           tau_th and heated with the reference step's factor 1 - 1/(2 tau_th);
                           needs the fluid lattice and general = 0.  A step is one fused pass:
                           the fluid lattice, then the D3Q7 step with the velocity just computed. */
  int slab_planes;      /* 0: each pass is one launch over whole lattices.  n > 0: runs of an even
                           number of steps starting on an odd pass go as pairs of passes split into
                           x-slabs of n planes, launched as a wavefront (odd pass on slab k, then the
                           even pass on slab k - 1), so a slab's even pass finds the lines its odd pass
                           touched still in the Infinity Cache; the sweep direction alternates between
                           pairs.  Same bits.  Needs thermal_q 19 and general = 0. */
  int thermal_prefetch; /* 1: the fused D3Q19 pass issues its thermal loads right after the fluid
                           gather (38 loads per cell in flight, more registers).  Measured on
                           148x128x32 designs: 1.6 % less time per pass at three lattices per
                           launch (partly resident in the Infinity Cache), 3.6 % more at 64
                           (HBM).  Same bits.  Needs the D3Q19 thermal lattice and compact fluid
                           storage. */
} xadv_batch_desc;

#define XADV_STORAGE_AUTO 0
#define XADV_STORAGE_DENSE 1
#define XADV_STORAGE_COMPACT 2

/* One lattice instance.  Mirrors the arguments of src/simulation.py:14-29
 * after its derivations (:70-71 thermal taus, :77-86 heat map normalised,
 * :61 rho_outlet), plus the boundary options of validate_analytical.py's
 * test kernels. */
typedef struct {
  int nx, ny, nz;
  const int8_t* mask;   /* host, (nx,ny,nz) C-order, XADV_* codes */
  const float* heat;    /* host, (nx,ny,nz) C-order, normalised heat map, or NULL (none) */
  double tau;           /* fluid relaxation time */
  double tau_th_fluid, tau_th_solid;
  double heat_power;    /* q = heat * heat_power (thermal_engine3d.py:116-117) */
  double T_in;          /* inlet temperature */
  double rho_out;       /* outlet density (fluid_engine3d.py:186) */
  double u_in;          /* inlet velocity (device-resident, see xadv_set_inlet_velocity) */
  /* general mode only */
  int periodic[3];      /* periodic x, y, z (validate_analytical.py:255-273, :397-420) */
  double force_x;       /* body force (validate_analytical.py:377-395) */
  double u_wall;        /* top-z moving wall velocity (validate_analytical.py:502-538) */
  int dirichlet_x;      /* -1 or the x-plane held at dirichlet_T (validate_analytical.py:21-30) */
  double dirichlet_T;
  int frozen_velocity;  /* thermal-only: collide with the velocity given at init */
} xadv_instance_desc;

/* Create / destroy a batch. */
int xadv_batch_create(const xadv_batch_desc* desc, xadv_batch** out);
int xadv_batch_destroy(xadv_batch* b);

/* Upload one instance into a slot (mask, heat map and parameters).  Checks
 * the mask: with the fluid on, INLET cells only on x=0 and OUTLET cells only
 * on x=nx-1; with the thermal on, OUTLET cells only on x=nx-1; no INLET or
 * OUTLET on a periodic x axis.  The slot is inactive until xadv_instance_init.
 * Replaces the host->device copies of src/simulation.py:62-86. */
int xadv_instance_load(xadv_batch* b, int slot, const xadv_instance_desc* d);

/* Initial populations from macroscopic fields: f = f_eq(rho, u), g = g_eq(T, u)
 * with u = 0 in solids.  rho (nx,ny,nz), u (nx,ny,nz,3), T (nx,ny,nz) are
 * device pointers in lattice precision, or NULL for rho = 1, u = 0, T = T_in.
 * Activates the slot and restarts its step count.
 * Replaces compute_equilibrium_kernel and compute_thermal_equilibrium_kernel
 * (src/fluid_engine3d.py:58-80, src/thermal_engine3d.py:7-47) as launched at
 * src/simulation.py:89-90. */
int xadv_instance_init(xadv_batch* b, int slot, const void* rho, const void* u, const void* T);

/* Mark a slot active (1) or retired (0); retired slots are skipped by every kernel. */
int xadv_set_active(xadv_batch* b, int slot, int active);

/* Inlet velocity of a slot; device-resident, so captured graphs stay valid.
 * Replaces u_inlet_array.assign (src/simulation.py:146). */
int xadv_set_inlet_velocity(xadv_batch* b, int slot, double u_in);

/* Advance every active slot by nsteps time steps (one fused AA pass per step).
 * With want_fields != 0 the last pass also stores rho and u.
 * Replaces the graph replay loop of src/simulation.py:120-122 (the six
 * kernels of :94-99: collide, stream, macroscopic for both lattices). */
int xadv_step(xadv_batch* b, int nsteps, int want_fields);

/* After xadv_step(..., want_fields=1): form T (thermal streaming + macroscopic
 * of the current step) and, if with_stats, the 16 reduction slots per active
 * slot (11 fluid + 5 thermal, same meaning as the reference).
 * Replaces compute_stats_kernel and compute_thermal_stats_kernel
 * (src/fluid_engine3d.py:277-312, src/thermal_engine3d.py:226-243). */
int xadv_readout(xadv_batch* b, int with_stats);

/* Copy the 16 stats of every slot to host out[slots*16] (doubles; slots of
 * retired instances are zero).  Synchronises.  Replaces stats_warp.numpy()
 * (src/simulation.py:128-129). */
int xadv_stats(xadv_batch* b, double* out);

/* Copy one slot's macroscopic fields into caller device buffers (any may be
 * NULL): rho (nx,ny,nz), u (nx,ny,nz,3), T (nx,ny,nz), lattice precision.
 * Synchronises.  Replaces u.numpy(), rho_warp.numpy(), T.numpy()
 * (src/simulation.py:161). */
int xadv_fields(xadv_batch* b, int slot, void* rho, void* u, void* T);

/* Post-solve reductions for one slot, from the fields of the last readout.
 * out[XADV_NMETRIC] doubles, see XADV_M_* (host, synchronises).  core_x0 is
 * the inlet buffer length and core_nx, core_ny the heated core size.
 * Replaces the numpy reductions of compute_macroscopic_metrics
 * (generate_dataset.py:80-179) and the pressure_drop column (:556). */
#define XADV_NMETRIC 24
enum {
  XADV_M_QVOL = 0,        /* sum u_x over OUTLET cells on x=nx-1 */
  XADV_M_RHO_IN_SUM,      /* sum rho over non-solid cells on x=1 */
  XADV_M_RHO_IN_N,
  XADV_M_RHO_OUT_SUM,     /* sum rho over non-solid cells on x=nx-2 */
  XADV_M_RHO_OUT_N,
  XADV_M_RHO_P1_SUM,      /* sum rho over the whole plane x=1 */
  XADV_M_RHO_P2_SUM,      /* sum rho over the whole plane x=nx-2 */
  XADV_M_PLANE_N,         /* ny*nz */
  XADV_M_TS_MAX,          /* heated surface T[x0:x0+cnx, :cny, 0]: max, min, sum, sum of squares, count */
  XADV_M_TS_MIN,
  XADV_M_TS_SUM,
  XADV_M_TS_SUMSQ,
  XADV_M_TS_N,
  XADV_M_QOUT_FLUX,       /* sum u_x*T over OUTLET cells */
  XADV_M_TF_SUM,          /* sum T over non-solid cells of [x0:-x0, 1:-1, 4:-2] */
  XADV_M_TF_N,
  XADV_M_UMAX,            /* max |u| over all cells */
  XADV_M_TMAX,            /* max T over all cells */
  XADV_M_TSUM,            /* sum T over all cells */
  XADV_M_NCELL,
  XADV_M_NAN              /* count of cells with a non-finite rho, u or T */
};
int xadv_metrics(xadv_batch* b, int slot, int core_x0, int core_nx, int core_ny, double* out);

/* Reductions behind the physics acceptance checks of analyze_dataset.py:184-325,
 * over one slot's fields of the last readout; out[XADV_NCHECK] doubles, see
 * XADV_C_*.  k_fluid / k_solid are the thermal conductivities (tau_th - 0.5)/3.
 * Replaces the numpy checks of validate_sample (analyze_dataset.py:184-325). */
#define XADV_NCHECK 16
enum {
  XADV_C_MDOT_IN = 0,     /* sum rho*u_x over non-solid cells on x=1 */
  XADV_C_MDOT_OUT,        /* ... on x=nx-2 */
  XADV_C_VMAX_FLUID,      /* max |u| over non-solid cells */
  XADV_C_RHODEV_FLUID,    /* max |rho - 1| over non-solid cells */
  XADV_C_TMIN_FLUID,      /* min T over non-solid cells */
  XADV_C_RHO_P1_SUM,      /* sum rho over the plane x=1 */
  XADV_C_RHO_P2_SUM,      /* sum rho over the plane x=nx-2 */
  XADV_C_DIV_SUM,         /* sum |div u| (central differences) over interior non-solid cells */
  XADV_C_DIV_N,
  XADV_C_DIV_MAX,
  XADV_C_UX_INLET,        /* sum u_x over INLET cells */
  XADV_C_UXT_OUTLET,      /* sum u_x*T over OUTLET cells */
  XADV_C_QDIFF_IN,        /* sum -k (T[1] - T[0]) over the plane x=1 */
  XADV_C_QDIFF_OUT,       /* sum -k (T[nx-1] - T[nx-2]) over the plane x=nx-2 */
  XADV_C_NFLUID
};
int xadv_checks(xadv_batch* b, int slot, double k_fluid, double k_solid, double* out);

/* The batch's HIP stream (hipStream_t), for callers that time or order work
 * against it (e.g. hipEventRecord around xadv_step).  No reference
 * counterpart: Warp's default stream was implicit (src/simulation.py:93-121). */
int xadv_get_stream(xadv_batch* b, void** stream);

/* ---- Designs built on the device ---------------------------------------
 * Masks and heat maps generated on the GPU from design parameters, and a
 * slot loaded from device-resident inputs, so that a design goes from its
 * parameters to a loaded slot without a host pass over its cells. */

/* Core families of src/geometry.py:365-503 (design_idx). */
#define XADV_CORE_EMPTY 0
#define XADV_CORE_STRAIGHT 1
#define XADV_CORE_PINFIN 2
#define XADV_CORE_GYROID 3
#define XADV_CORE_SCHWARZ_P 4
#define XADV_CORE_SCHWARZ_D 5

typedef struct {
  int nx, ny, nz;       /* full lattice; the core spans x in [buffer, nx - buffer), all y and z */
  int buffer;           /* open cells before and after the core along x (>= 0, 2 buffer < nx) */
  int base, top;        /* solid floor and lid thickness (base >= 0, top >= 1) */
  int kind;             /* XADV_CORE_* */
  double feature_size;  /* straight / pin-fin feature size (and the TPMS level when period <= 0) */
  double spacing;       /* straight / pin-fin spacing (and the TPMS cell size when period <= 0) */
  double period;        /* TPMS: > 0 gives omega = 2 pi / period and solid where f < threshold */
  double threshold;
  int manifold;         /* TPMS: open core cells at each x end (the reference's 5) */
} xadv_geometry_desc;

/* Write the (nx, ny, nz) int8 mask of a design into a caller device buffer:
 * the core of src/geometry.py:365-503 (mapped from feature_size/spacing as
 * there, or an explicit TPMS period and threshold), then the frame of
 * src/geometry.py:506-524 (buffers, side walls y = 0 and ny-1, base, lid,
 * INLET on x = 0 and OUTLET on x = nx-1 over the open y, z cells).  Runs on
 * the batch's stream and synchronises it. */
int xadv_geometry_mask(xadv_batch* b, const xadv_geometry_desc* g, int8_t* mask);

#define XADV_HEAT_RECT 0
#define XADV_HEAT_DISK 1
#define XADV_HEAT_GAUSS 2
#define XADV_HEAT_MAX_PRIMS 16
#define XADV_HEAT_MAX_RADIUS 63

typedef struct {
  int kind;              /* XADV_HEAT_* */
  int x0, x1, y0, y1;    /* RECT: cells x0 <= x < x1, y0 <= y < y1 take value */
  int cx, cy, r2;        /* DISK: cells with (x-cx)^2 + (y-cy)^2 < r2 take value */
  float value;
  double gx, gy;         /* GAUSS: adds amplitude * exp(-((x-gx)^2 + (y-gy)^2) / den), */
  double den, amplitude; /*        in FP64, rounded to FP32 after each addition */
} xadv_heat_prim;

typedef struct {
  int mx, my;                   /* map size (the core's x and y extent) */
  float base;                   /* initial value of every cell */
  int nprims;                   /* <= XADV_HEAT_MAX_PRIMS */
  const xadv_heat_prim* prims;  /* host, applied in order */
  int radius;                   /* Gaussian blur half-width (<= XADV_HEAT_MAX_RADIUS), 0 = none */
  const double* weights;        /* host, weights[j] for offsets +-j, j = 0..radius */
  int normalise;                /* 1: divide by the maximum when it is positive */
} xadv_heat_desc;

/* Write a heat field (nx, ny, nz) FP32 into a caller device buffer: zero
 * everywhere except the z = 0 plane at x in [x_offset, x_offset + mx),
 * y in [0, my), which holds the map painted from its primitives, blurred
 * along x then y with reflecting edges (FP64 sums, FP32 after each axis) and
 * divided by its maximum.  Restates the ten generators of
 * src/geometry.py:69-319 (their scipy.ndimage.gaussian_filter and
 * division by the maximum) and the placement of generate_dataset.py:402-410.
 * Runs on the batch's stream and synchronises it. */
int xadv_heat_map(xadv_batch* b, const xadv_heat_desc* h, int nx, int ny, int nz, int x_offset, float* heat);

/* xadv_instance_load with d->mask (int8) and d->heat (FP32, or NULL) in
 * device memory: the mask checks, cell flags, heat source and compact column
 * table are built on the device; the results are the same as
 * xadv_instance_load's for the same mask and heat map.  When the mask breaks
 * more than one rule, the message names the first of: a code outside 0..3,
 * an INLET off x = 0, an OUTLET off x = nx-1.  On error the slot is left as it
 * was.  No reference counterpart (the reference copies host arrays,
 * src/simulation.py:62-86). */
int xadv_instance_load_device(xadv_batch* b, int slot, const xadv_instance_desc* d);

/* Message of the last error on this thread ("" if none). */
const char* xadv_last_error(void);

/* Library version string. */
const char* xadv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* XADV_H */
