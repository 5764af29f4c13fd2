/*
 * tlbm.h -- C ABI of the MI355X D3Q19 lattice Boltzmann step library (libtlbm.so).
 *
 * The reference composes its step from Tack device classes and JIT-compiles one kernel per
 * combination of plugin types (lbm.py:1-7, README.md:6-9).  This library replaces that boundary
 * with plain C: a kernel is chosen by a tlbm_spec (collision x obstacle x diagnostics x lateral
 * side x storage), every combination is a pre-instantiated, branch-free HIP kernel for gfx950,
 * and the instance scalars the reference passes as kernel arguments travel in tlbm_params.
 *
 * Conventions:
 *   - every function returns TLBM_OK (0) or a nonzero status; tlbm_last_error() then names the
 *     failure (per host thread).  No C++ exception crosses the ABI.
 *   - pointers to lattice data are device (or managed) pointers; "stream" is a hipStream_t
 *     passed as void* (NULL = the null stream).  All launches are asynchronous.
 *   - layout is the reference's: population q of cell idx = (z*ny + y)*nx + x lives at
 *     f[q*qstride + idx] (lbm.py:386-388, 426, 456), with qstride >= nx*ny*nz.  A step can
 *     instead take populations in tiles of k = tile_rows rows (tlbm_params.tile_rows > 0, no
 *     counterpart in the reference): with row r = z*ny + y and T = k*nx, population q of cell
 *     (x, r) lives at f[(r/k)*19*T + q*T + (r%k)*nx + x], i.e. each tile of k rows holds its 19
 *     planes together; tlbm_population_elems gives the buffer length.  The temperature planes
 *     g follow the same layout with 7 planes: g[q*qstride + idx], or tiles of k rows holding
 *     their 7 planes of k*nx cells together.
 */
#ifndef TLBM_H
#define TLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLBM_OK 0
#define TLBM_ERR_INVALID 1      /* bad argument: sizes, null pointers, unknown spec       */
#define TLBM_ERR_HIP 2          /* a HIP runtime call failed (message has hipGetErrorString) */
#define TLBM_ERR_NO_DEVICE 3    /* no HIP device visible                                  */

/* collision models: lbm.py:544 COLLISIONS = {"bgk": BGK, "trt": TRT, "les": SmagorinskyLES} */
enum { TLBM_BGK = 0, TLBM_TRT = 1, TLBM_LES = 2 };
/* obstacles: lbm.py:144-227 (Cylinder, Sphere, Box) and lbm.py:231-268 (HeatSink: a finned
 * sink on the floor of a duct whose y = 0 and y = ny - 1 rows are walls).  TLBM_FIELD has no
 * counterpart in the reference: a per-cell solid fraction phi in [0, 1] (tlbm_params.phi), used
 * as the `solid` value of lbm.py:403 and 416 -- partial bounce-back
 * f = pulled + phi (bounced - pulled) on each link, the smoothed, differentiable obstacle of
 * the adjoint. */
enum { TLBM_CYLINDER = 0, TLBM_SPHERE = 1, TLBM_BOX = 2, TLBM_FIELD = 3, TLBM_HEATSINK = 4 };
/* diagnostics variants: lbm.py:325-339 (Quiet.enabled = 0, Report.enabled = 1) */
enum { TLBM_QUIET = 0, TLBM_REPORT = 1 };
/* lateral y sides: lbm.py:342-369 (PeriodicY, SlipY); TLBM_LID_Y has no counterpart in the
 * reference: a moving no-slip lid halfway outside y = ny-1 over a free-slip floor. */
enum { TLBM_PERIODIC_Y = 0, TLBM_SLIP_Y = 1, TLBM_LID_Y = 2 };
/* population storage: FP32 as the reference (lbm.py:405-406, 566); FP16/BF16 store the
 * deviation f_q - w_q from the rest equilibrium, arithmetic stays FP32. */
enum { TLBM_F32 = 0, TLBM_F16 = 1, TLBM_BF16 = 2 };

/* Number of FP64 sums a report step produces, in this order:
 * the 13 fields of lbm.py:545 FIELDS (ux uy uz p vmag r_rho r_ux r_uy r_uz r_ke fx fy fz),
 * then the total mass sum(rho). */
#define TLBM_NSUMS 14

typedef struct {
  int32_t collision; /* TLBM_BGK / TLBM_TRT / TLBM_LES           */
  int32_t obstacle;  /* TLBM_CYLINDER / TLBM_SPHERE / TLBM_BOX / TLBM_FIELD / TLBM_HEATSINK */
  int32_t diag;      /* TLBM_QUIET / TLBM_REPORT                  */
  int32_t side;      /* TLBM_PERIODIC_Y / TLBM_SLIP_Y / TLBM_LID_Y */
  int32_t storage;   /* TLBM_F32 / TLBM_F16 / TLBM_BF16           */
  int32_t heat;      /* 0: NoHeat; 1: Heat, the D3Q7 temperature variant (lbm.py:279-319,
                        457-492), TLBM_F32 storage only; such handles step with tlbm_step_heat */
} tlbm_spec;

/* Runtime arguments of one step (the reference's plugin instance attributes). */
typedef struct {
  int32_t nx, ny, nz;      /* D3Q19(nx, ny, nz), lbm.py:42-44                           */
  int32_t block_x;         /* threads per block of the step: 64, 128, 192 or 256; 0: the
                              handle's default (tlbm_set_launch, else 256)                */
  int64_t qstride;         /* elements between population (and temperature) planes,
                              >= nx*ny*nz; not used for the layout when tile_rows > 0     */
  int32_t tile_rows;       /* 0: q-major population planes; k > 0 (a power of two): tiles
                              of k rows, 19 planes of k*nx cells each (see the layout note) */
  int32_t xcd_group;       /* quiet steps: blocks of consecutive cells go to the XCDs in
                              runs of this many (a power of two up to 64; 1: plain order);
                              0: 2 for FP16/BF16 storage, 1 for FP32 and the temperature
                              variant.  Report steps keep plain order.                    */
  float omega;             /* BGK: 1 / (3 nu + 1/2), lbm.py:63                          */
  float tau0;              /* TRT / LES: 3 nu + 1/2, lbm.py:110, 131                    */
  float cs;                /* Smagorinsky constant (TRT.cs, SmagorinskyLES.cs = 0.16)    */
  float magic;             /* TRT magic parameter (TRT.magic = 0.1875), lbm.py:106       */
  float obs[8];            /* cylinder: x, y, r; sphere: x, y, z, r;
                              box: x, y, z, hx, hy, hz  (lbm.py:146, 174, 203-205);
                              heat sink: x0, x1, z0, z1, y_base, y_top, fin_thickness,
                              fin_pitch (lbm.py:239-246; integers except the y levels) */
  float u0;                /* Flow.u0, lbm.py:375                                        */
  float sponge_start;      /* Flow.sponge_start                                          */
  float sponge_strength;   /* Flow.sponge_strength                                       */
  float lid[3];            /* TLBM_LID_Y wall velocity                                   */
  const float* omega_rot;  /* device/managed pointer to one float, the cylinder surface
                              speed read by every step (Flow.omega_rot, lbm.py:378-379,
                              417); NULL means 0                                          */
  const float* phi;        /* TLBM_FIELD: device pointer to nx*ny*nz solid fractions      */
} tlbm_params;

/* The 13 report fields of lbm.py:545, each n = nx*ny*nz floats (device or managed). */
typedef struct {
  float *ux, *uy, *uz, *p, *vmag, *r_rho, *r_ux, *r_uy, *r_uz, *r_ke, *fx, *fy, *fz;
} tlbm_fields;

/* ---- errors, devices --------------------------------------------------------------------- */
const char* tlbm_last_error(void);
int tlbm_version(void);
int tlbm_device_count(int* count);
int tlbm_set_device(int device);

/* ---- memory (replaces ManagedArray, lbm.py:525-541, and tack.array/empty/zeros) ------------- */
/* hipMallocManaged, preferred location = current device, prefetched there, zero-filled. */
int tlbm_alloc_managed(size_t bytes, void** out);
int tlbm_alloc_device(size_t bytes, void** out);
int tlbm_alloc_host(size_t bytes, void** out); /* pinned host memory (hipHostMalloc) */
int tlbm_free(void* ptr);                      /* device or managed */
int tlbm_free_host(void* ptr);
int tlbm_memcpy(void* dst, const void* src, size_t bytes);                   /* synchronous  */
int tlbm_memcpy_async(void* dst, const void* src, size_t bytes, void* stream);
int tlbm_memset(void* ptr, int value, size_t bytes, void* stream);  /* asynchronous; refused
                                                                       under graph capture */

/* ---- streams, events (replaces tack.sync, lbm.py:651) ------------------------------------ */
int tlbm_stream_create(void** stream);
int tlbm_stream_destroy(void* stream);  /* waits for the stream, frees its report workspace */
int tlbm_stream_sync(void* stream);
int tlbm_device_sync(void);
int tlbm_event_create(void** event);
int tlbm_event_destroy(void* event);
int tlbm_event_record(void* event, void* stream);
int tlbm_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- kernels ----------------------------------------------------------------------------- */
/* Resolve a specialisation (Tack's type-keyed kernel cache, README.md:55-56).  The handle is
 * stable for the life of the process. */
int tlbm_kernel_lookup(const tlbm_spec* spec, int* handle);
/* Launch shape of a specialisation: threads per block (64, 128, 192 or 256; 0 restores the
 * default, 256).  Each block steps that many consecutive cells in index order. */
int tlbm_set_launch(int handle, int block_x);
/* Elements one population buffer must hold for `params` (nx, ny, nz, qstride, tile_rows):
 * 19*qstride for q-major planes, 19*k*nx*ceil(ny*nz/k) for tiles of k rows. */
int tlbm_population_elems(const tlbm_params* params, uint64_t* elems);

/* lbm_init (lbm.py:383-394): f = feq(1, u0 fluid, seed_amp u0 fluid sin(2 pi x/nx)
 * cos(2 pi z/nz), 0), plus, when noise_amp != 0, a seeded uniform U(-noise_amp, noise_amp)
 * velocity noise on fluid cells (counter-based splitmix64 of (noise_seed, 3 idx + component)).
 * Uses the obstacle and storage of the handle; collision/diag/side are ignored. */
int tlbm_init(int handle, void* f, const tlbm_params* params, float seed_amp, float noise_amp,
              uint64_t noise_seed, void* stream);

/* lbm_step (lbm.py:397-456, 493-513): one fused pull-stream + bounce-back + collide + sponge
 * step from f_in to f_out (distinct buffers).  For a TLBM_REPORT handle, `fields` holds the
 * 13 report arrays (old ux, uy, uz, p are read, all 13 are written) and, when `sums` is not
 * NULL, TLBM_NSUMS FP64 sums are written to sums[0..13] (device, managed or pinned host
 * memory) by the same stream.  A report step may pass fields = NULL when sums is not NULL:
 * it then writes only the sums, with the five residual sums zero.  `fields` and `sums` are
 * ignored by TLBM_QUIET handles.  Report steps on different streams may run concurrently:
 * the library keeps one reduction workspace per (device, stream).
 * Replaces the host FP64 numpy sums of Solver.monitors (lbm.py:649-661). */
int tlbm_step(int handle, const void* f_in, void* f_out, const tlbm_params* params,
              const tlbm_fields* fields, double* sums, void* stream);

/* tlbm_step without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as tlbm_step; a report step's per-block FP64
 * partials go to a fixed array the library holds on each device it runs on (set up at the
 * first call on that device) instead of a stream workspace, so report steps of lattices of
 * more than TLBM_NOALLOC_MAX_BLOCKS blocks (at the default 256 threads per block:
 * 33,554,432 cells; 14.7 MB per device) are refused with TLBM_ERR_INVALID.  Every call on one
 * device shares that array: report steps of overlapping calls (other streams, other threads)
 * race, so serialise them. */
#define TLBM_NOALLOC_MAX_BLOCKS 131072
int tlbm_step_noalloc(int handle, const void* f_in, void* f_out, const tlbm_params* params,
                      const tlbm_fields* fields, double* sums, void* stream);

/* A plain device-to-device copy kernel (16 bytes per lane; pointers and size 16-byte aligned).
 * No counterpart in the reference: it is the known-size streaming kernel against which HBM
 * counters and bandwidth of the step are calibrated. */
int tlbm_copy(void* dst, const void* src, size_t bytes, void* stream);

/* The step's data movement without its arithmetic: f_out[q*qstride + idx] =
 * f_in[q*qstride + src_q(idx)] for the 19 populations, src_q the pull source of lbm.py:410-415
 * with x, y and z periodic (collisionless streaming), launched and addressed exactly as the
 * quiet step's loads and stores, with the same cache policies (params: nx, ny, nz, qstride,
 * tile_rows, xcd_group; storage picks 4- or 2-byte elements, copied raw).  No counterpart in
 * the reference: the bandwidth it reaches is the ceiling of the step's 19-read/19-write access
 * pattern on the device. */
int tlbm_pull_copy(const void* f_in, void* f_out, const tlbm_params* params, int storage,
                   void* stream);

/* coarsen (lbm.py:516-522): dst[(k*my + j)*mx + i] = src[((k s)*ny + j s)*nx + i s]. */
int tlbm_coarsen(const float* src, float* dst, int nx, int ny, int nz, int s, int mx, int my,
                 int mz, void* stream);

/* ---- the temperature variant (lbm.py:279-319, 457-492) ------------------------------------
 * D3Q7 advection-diffusion of the temperature inside the same step: g is pulled periodically
 * in x, y and z; a link from a solid voxel bounces back (adiabatic) and one from a sink voxel
 * (the whole obstacle, or the heat sink's base and fins) anti-bounces to its t_solid
 * (lbm.py:468-473); BGK relaxation with omega_g = 1 / (4 alpha_eff + 1/2),
 * alpha_eff = alpha + max(0, (tau - 1/2) / 3 - nu) / pr_t, tau the step's relaxation time;
 * the flow's sponge blend applies to the temperature too. */
typedef struct {
  const float* t_solid; /* n: the temperature of each solid voxel (Solver.set_wall_temperature) */
  float* q_solid;       /* n: report steps add, atomically, fluid * temp of every fluid cell to
                           each sink voxel it faces (lbm.py:482-492); the caller clears it    */
  float* temp;          /* n: report steps write fluid * temp + solid * t_solid (lbm.py:483)  */
  float alpha;          /* molecular diffusivity, nu / prandtl (lbm.py:597)                   */
  float nu;             /* molecular viscosity                                                */
  float pr_t;           /* turbulent Prandtl number (Heat.pr_t = 0.85, lbm.py:286)            */
} tlbm_heat;

/* lbm_step with heat enabled (lbm.py:397-513): the flow step of tlbm_step (same arguments and
 * report behaviour) plus the temperature step from g_in to g_out (distinct buffers of 7 planes
 * of qstride floats, or with tile_rows = k > 0 tiled as the populations: 7 planes of k*nx cells
 * per tile of k rows, 7*k*nx*ceil(ny*nz/k) floats).  Needs a handle looked up with
 * spec.heat = 1. */
int tlbm_step_heat(int handle, const void* f_in, void* f_out, const float* g_in, float* g_out,
                   const tlbm_params* params, const tlbm_heat* heat, const tlbm_fields* fields,
                   double* sums, void* stream);

/* ---- derived fields of the output grid (derived.py:48-93) ----------------------------------
 * Grid of mx x my x mz points (x fastest), spacing h, inv_h = 1 / h; its cells are
 * (mx - 1)(my - 1)(mz - 1), indexed over their lower corner.  Arrays are device memory. */

/* The velocity gradient at every point (derived.py:52-89, the solver's gradient): central
 * differences, one-sided at the grid's faces, zero along an axis of one point.  jac holds 9
 * planes of mx*my*mz floats: jac[(3 c + d) * m + p] = d u_c / d x_d at point p. */
int tlbm_velocity_gradient(const float* ux, const float* uy, const float* uz, float* jac, int mx,
                           int my, int mz, float inv_h, void* stream);

/* |curl u| and the Q-criterion at every cell centre (the default derived fields,
 * derived.py:26-29), from the velocity in one pass.  basis = 0: the cell's gradient is the mean
 * of its 8 corners' point gradients (those of tlbm_velocity_gradient, formed on the fly);
 * basis = 1: the trilinear cell's gradient at its centre, the mean of the four edge differences
 * per axis.  Either output may be NULL; outputs hold (mx - 1)(my - 1)(mz - 1) floats. */
int tlbm_vortex_fields(const float* ux, const float* uy, const float* uz, float* vorticity,
                       float* qcrit, int mx, int my, int mz, float inv_h, int basis, void* stream);

/* ---- adjoint (no counterpart in the reference) --------------------------------------------
 * One reverse step of the discrete adjoint of the BGK step with a TLBM_FIELD obstacle, periodic
 * in x, y and z, FP32 populations; f_t, f_t1, mu_in and mu_out have plane stride params->qstride
 * (params as for that forward step; phi is the solid fraction).  The objective is J = sum_t lam_t * (x-momentum exchanged with the solid
 * at step t) (lbm.py:429-430 summed over cells).  Given mu_in = dJ/dg^{t+1} (g: the pulled and
 * bounced populations before collision) and the states f_t, f_t1 of steps t and t + 1, writes
 * mu_out = dJ/dg^t and adds dJ/dphi of the cells of the parameter block
 * [origin, origin + size) (inside the lattice) to gphi (FP64, x fastest).  lam_t, lam_t1: the objective's weights
 * of steps t and t + 1 (0 outside the averaging window; lam_t1 = 0 and mu_in = 0 for the last
 * step).  final_links != 0 adds only the terms of step t + 1's links (mu_in = dJ/dg^0,
 * f_t1 = the initial state): call it once after the reverse sweep.  Blocks may not overlap
 * in flight (gphi is updated in place). */
int tlbm_adjoint_step(const tlbm_params* params, const float* f_t, const float* f_t1,
                      const float* mu_in, float* mu_out, float lam_t, float lam_t1,
                      const int32_t* block_origin, const int32_t* block_size, double* gphi,
                      int final_links, void* stream);

/* ---- batched lattices (parameter sweeps) --------------------------------------------------
 * No counterpart in the reference, which steps one lattice per launch.  A batch is a set of
 * independent lattices that share one specialisation (collision, obstacle, side, storage of
 * `handle`; its diag is ignored, both variants are available) and differ in size and runtime
 * arguments (params[d]).  Their populations live in one buffer: domain d at element offset
 * elem_offset(d), laid out as a single lattice of params[d] (tlbm_population_elems): q-major
 * with plane stride params[d].qstride (a value below the domain's nx*ny*nz means nx*ny*nz:
 * packed), or tiles of params[d].tile_rows rows, the domains one after another; packed
 * per-cell arrays (report fields) hold domain d at cell_offset(d).  One launch steps every
 * domain: a persistent grid walks chunks of block_x consecutive cells of one domain. */
typedef struct tlbm_batch tlbm_batch;
typedef struct tlbm_graph tlbm_graph;

int tlbm_batch_create(int handle, const tlbm_params* params, int ndomains, tlbm_batch** out);
int tlbm_batch_destroy(tlbm_batch* batch);
/* Offsets of `domain` (domain == ndomains gives the totals) and the buffer sizes: `elements`
 * population elements, `cells` lattice cells over all domains.  Any output may be NULL. */
int tlbm_batch_layout(const tlbm_batch* batch, int domain, uint64_t* elem_offset,
                      uint64_t* cell_offset, uint64_t* elements, uint64_t* cells);
/* Threads per block (64, 128, 192, 256; 0 = 256) and persistent grid size in blocks (0 = one block
 * per chunk).  Changing block_x synchronises the device and rebuilds the chunk tables. */
int tlbm_batch_set_launch(tlbm_batch* batch, int block_x, int grid_blocks);
/* Chunk order of the batch's launches: group = 1 (the default) takes the chunks in plain
 * order, consecutive chunks on consecutive XCDs; group = g > 1 gives each XCD runs of g
 * consecutive chunks, so the cache lines neighbouring chunks share are fetched into one L2
 * (a persistent grid is then rounded down to a multiple of 8 blocks).  Results do not depend
 * on the order.  No counterpart in the reference (tuned by tlbm.autotune). */
int tlbm_batch_set_order(tlbm_batch* batch, int group);
/* lbm_init of every domain (see tlbm_init). */
int tlbm_batch_init(tlbm_batch* batch, void* f, float seed_amp, float noise_amp,
                    uint64_t noise_seed, void* stream);
/* One step of every domain.  report != 0: the report variant; `fields` (packed over all domains)
 * may be NULL (then only the sums of u, p, |u|, forces and mass are formed and the residual
 * sums are 0); `sums`, when not NULL, receives TLBM_NSUMS FP64 totals per domain
 * (sums[d*TLBM_NSUMS + k]). */
int tlbm_batch_step(tlbm_batch* batch, const void* f_in, void* f_out, int report,
                    const tlbm_fields* fields, double* sums, void* stream);
/* Capture `steps` (even) consecutive batch steps, f_a -> f_b -> f_a ..., with a report step
 * every `report_every` steps (0: none), on `stream` (not the null stream) into a hipGraph;
 * launching the graph runs them all and leaves the state in f_a.  The graph uses the batch's
 * own tables and partials: destroy it before tlbm_batch_set_launch or tlbm_batch_destroy. */
int tlbm_batch_graph(tlbm_batch* batch, void* f_a, void* f_b, int steps, int report_every,
                     const tlbm_fields* fields, double* sums, void* stream, tlbm_graph** out);
int tlbm_graph_launch(tlbm_graph* graph, void* stream);
int tlbm_graph_destroy(tlbm_graph* graph);
/* Capture the library calls this thread issues on `stream` (a created stream, not the null
 * stream) between tlbm_capture_begin and tlbm_capture_end into a graph (launch it with
 * tlbm_graph_launch, on any stream of the device).  Steps, inits, copies and coarsening can
 * be captured; nothing that allocates: the first report step with sums on a stream (or on a
 * larger lattice than before) sizes its reduction workspace and must run eagerly first, and
 * fails with TLBM_ERR_INVALID under capture; tlbm_memset is refused.  When a call fails under
 * capture, end the capture and destroy what tlbm_capture_end returns.  A graph keeps the
 * reduction workspaces its report steps use: they stay allocated, even when the stream needs a
 * larger one or is destroyed, until the last graph using them is destroyed.  A graph replays
 * the buffers it was captured with (a caller alternating two buffers captures one graph per
 * order). */
int tlbm_capture_begin(void* stream);
int tlbm_capture_end(void* stream, tlbm_graph** out);
/* Compute units and architecture name of the current device (sizing persistent grids, keying
 * tuned launch shapes). */
int tlbm_device_cus(int* cus);
int tlbm_device_name(char* buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* TLBM_H */
