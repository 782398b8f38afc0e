/*
 * lbx.h -- C ABI of the MI355X-native D3Q15 collide-stream library
 * (liblbx.so, built from lambrex_amd/csrc).
 *
 * The reference (otbrown/LAMBReX) keeps this path inside the C++ class
 * AmrSim on top of AMReX MultiFabs (include/AmrSim.h:23-155).  This ABI is
 * what that class's protected machinery becomes: every distribution array
 * lives on one GPU, owned by a context; the host passes plain integers,
 * doubles and host buffers.  The C++ AmrSim facade
 * (lambrex_amd/host/include/AmrSim.h) sits on top of it with the
 * reference's public API.
 *
 * Conventions
 *   - Every call returns LBX_OK (0) or a negative LBX_E* code and never
 *     aborts or throws; lbx_last_error() describes the last failure.
 *   - One host thread per context; each context owns one HIP stream
 *     (default priority).  Calls are stream-ordered; calls that hand data
 *     back to the host synchronise before returning.
 *   - Boxes are inclusive integer boxes, 6 int32 per box: lo x,y,z then
 *     hi x,y,z, in the level's own index space (fine = 2 x coarse).
 *   - Host field buffers are x fastest: rho[nz][ny][nx],
 *     u[3][nz][ny][nx], f[15][nz][ny][nx] over the box's valid cells, with
 *     populations in the reference order (include/d3q15_bgk.h:14-22).
 */
#ifndef LBX_H
#define LBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBX_OK 0
#define LBX_E_ARG (-1)      /* invalid argument */
#define LBX_E_STATE (-2)    /* call not valid in the current state */
#define LBX_E_HIP (-3)      /* HIP runtime error */
#define LBX_E_NOMEM (-4)    /* device or host allocation failed */
#define LBX_E_UNSUPPORTED (-5)

/* Levels 0 .. LBX_MAX_LEVELS-1, refinement ratio r between neighbours (2
 * unless lbx_set_ref_ratio sets 3 or 4). */
#define LBX_MAX_LEVELS 6

#define LBX_FP64 0          /* populations stored and collided in FP64 */
#define LBX_FP32_SHIFTED 1  /* stored as f_i - w_i*rho0 in FP32, collided in FP32; the
                               two-level average-down sums and all reductions in FP64 */

typedef struct lbx_ctx lbx_ctx;
typedef struct lbx_sim lbx_sim;

/* Process / device set-up.  Replaces lambrexInit()/lambrexFinalise()
 * (include/lambrex.h:6-7, src/lambrex.cpp:4-14): creates the device
 * context (stream, scratch) that all simulations on that GPU share. */
int lbx_ctx_create(int device, lbx_ctx** out);
int lbx_ctx_destroy(lbx_ctx* ctx);

/* Last error message of this context (ctx may be NULL for errors raised
 * before a context existed).  Replaces amrex::Abort messages
 * (src/AmrSim.cpp:211, 291, 399, 690, 795). */
const char* lbx_last_error(const lbx_ctx* ctx);

/* Wait for all work queued on the context's stream. */
int lbx_sync(lbx_ctx* ctx);

/* Free and total device memory (bytes) as seen by the context's device. */
int lbx_mem_info(lbx_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* A simulation on the domain [0,dims).  Replaces the AmrSim constructor's
 * geometry and boundary set-up (src/AmrSim.cpp:753-802).  periodic[d] != 0
 * wraps direction d; periodic[d] == 0 puts halfway bounce-back walls on
 * both faces normal to d (the reference aborts instead, :794-796; walls are
 * supported on single-level simulations, and refined levels are then
 * refused with LBX_E_UNSUPPORTED).  rho0 is the reference density of LBX_FP32_SHIFTED
 * storage (ignored for LBX_FP64). */
int lbx_sim_create(lbx_ctx* ctx, const int32_t dims[3], const int32_t periodic[3],
                   int precision, double rho0, lbx_sim** out);
int lbx_sim_destroy(lbx_sim* sim);

/* Define the boxes of a level and allocate its device storage.  Level 0's
 * boxes must tile the domain exactly.  Replaces
 * MakeNewLevelFromScratch/MakeNewLevelFromCoarse's MultiFab definitions
 * (src/AmrSim.cpp:665-713, include/multilevel.h:62-67). */
int lbx_level_define(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi);

/* Level l >= 1 (refinement ratio r, index space [0, r^l * dims)): level
 * l-1 must be defined; boxes must lie in the level's domain, be aligned to
 * the ratio (lo and hi + 1 multiples of r), not overlap, and be properly
 * nested: the parent of every cell of each box grown by 2 ghost cells that
 * no box of level l holds (periodic images included), and of every valid
 * cell, must be a valid cell of level l-1 (LBX_E_ARG otherwise; AMReX builds its grids
 * that way).  Each box is stored with 2 ghost layers like the reference's
 * FABs.  Defining a level releases every finer level and leaves level l-1
 * unmasked (lbx_set_level_mask sets its mask). */

/* Replace the boxes of the finest level l >= 1, keeping its data
 * (RemakeLevel, src/AmrSim.cpp:715-744, with DistFnFillPatch, :359-391):
 * every valid cell of the new boxes takes the old populations where an old
 * box covered it, else its parent's (piecewise-constant interpolation).
 * Box rules as for lbx_level_define; relaxation rates carry over; the
 * parent's mask is not changed.  LBX_E_STATE if level l has finer levels
 * (clear them first). */
int lbx_level_remake(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi);

/* ErrorEst on device data.  The reference tags only a static box
 * (TagCell, src/AmrSim.cpp:413-417, via ErrorEst :633-663); these tags come
 * from the state.  tags[(z*ny + y)*nx + x] = 1 for the cells of level 0
 * where criterion LBX_TAG_DENSITY_GRADIENT: |grad rho|^2 > threshold^2
 * (central differences, periodic), or LBX_TAG_SPEED: |u|^2 > threshold^2,
 * with rho and u exactly as lbx_hydro returns them.  Synchronous. */
#define LBX_TAG_DENSITY_GRADIENT 0
#define LBX_TAG_SPEED 1
int lbx_tag_cells(lbx_sim* sim, int level, int criterion, double threshold, uint8_t* tags);

/* Cluster level-0 tags (as from lbx_tag_cells, domain dims) into fine boxes
 * for lbx_level_define/lbx_level_remake: tags grown by `buffer` cells
 * (periodic), tiles of `blocking` coarse cells refined when they hold a
 * grown tag, tagged tiles merged into non-overlapping boxes, refined by 2
 * and chopped at max_grid_size (even) with AMReX's BoxList::maxSize rule.
 * Replaces AMReX's clustering in AmrCore::regrid with a simpler,
 * deterministic one.  *n_boxes = number of boxes; with max_boxes = 0 only
 * the count is returned; LBX_E_ARG if lohi (6 per box) is too small.
 * Host only: needs no context or GPU. */
int lbx_cluster_tags(const int32_t dims[3], const uint8_t* tags, int32_t blocking, int32_t buffer,
                     int32_t max_grid_size, int32_t max_boxes, int32_t* lohi, int32_t* n_boxes);
/* The same over an index region [0, dims) of any level, with a mask: a tile
 * that holds a cell with allowed[...] == 0 is not refined (allowed NULL:
 * every cell allowed, as lbx_cluster_tags).  Callers pass the parent level's
 * cells whose 3x3x3 neighbourhood lies in the parent, so the new level is
 * properly nested.  Host only. */
int lbx_cluster_tags_allowed(const int32_t dims[3], const uint8_t* tags, const uint8_t* allowed,
                             int32_t blocking, int32_t buffer, int32_t max_grid_size, int32_t max_boxes,
                             int32_t* lohi, int32_t* n_boxes);

/* Release a level's device storage and that of every finer level
 * (ClearLevel, src/AmrSim.cpp:746-751); level-1, now the finest, is left
 * unmasked. */
int lbx_level_clear(lbx_sim* sim, int level);

/* Refinement ratio between consecutive levels, the same on every level: 2
 * (the default, the reference's constructor fixes it, src/AmrSim.cpp:766),
 * 3 or 4; set before level 1 is defined.  Level l's boxes are then given in
 * the index space ratio^l * dims, aligned to the ratio; their parents are
 * floor(pos / ratio), the mask boxes are coarsened by the ratio, the
 * average-down runs over ratio^3 blocks times 1/ratio^3 over each box
 * coarsened and grown by 2 / ratio coarse cells (AMReX's
 * sum_fine_to_coarse), a level between 0 and the finest takes `ratio` steps
 * per step of its parent (the REF_RATIO loop, src/AmrSim.cpp:453-455), and
 * the finest level `ratio` substeps per step of its parent (the reference
 * writes ratio 2's two substeps out, src/AmrSim.cpp:446-452, 621-624).  The
 * caller sets each level's tau as ComputeDt does at that ratio
 * (src/AmrSim.cpp:308-315). */
int lbx_set_ref_ratio(lbx_sim* sim, int ratio);

/* Relaxation rates of a level, omega = 1/(tau + 1/2) computed by the caller
 * exactly as src/AmrSim.cpp:125-126. */
int lbx_set_relax(lbx_sim* sim, int level, double omega_s, double omega_b);

/* Mask of level `level`'s collide when it has a finer level (MakeFineMask,
 * src/AmrSim.cpp:419-428, used by CoarseCollide :487-500): its valid cells
 * inside coarsen(box, r) of any of the given boxes (index space of level
 * level+1, r the refinement ratio) have their populations zeroed instead of collided in the Rohde
 * cycle.  The reference passes the level's own BoxArray as these boxes;
 * passing the finer level's BoxArray masks the refined region.  A level
 * starts unmasked. */
int lbx_set_level_mask(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi_fine);
/* lbx_set_level_mask of level 0. */
int lbx_set_fine_mask(lbx_sim* sim, int n_boxes, const int32_t* lohi_fine);

/* Fill level l's valid cells from level l-1 by piecewise-constant
 * interpolation (DistFnFillFromCoarse, src/AmrSim.cpp:393-411). */
int lbx_fill_from_coarse(lbx_sim* sim, int level);

/* Equilibrium initialisation of one box from host density and velocity
 * (CalcEquilibriumDist, src/AmrSim.cpp:845-936, same operation order). */
int lbx_init_equilibrium(lbx_sim* sim, int level, int box, const double* rho, const double* u);

/* Copy one box's populations in or out (FP64 on the host side whatever the
 * storage precision).  There is no public reference counterpart; the
 * reference's tests reach the MultiFabs through AmrTest
 * (tests/AmrTest.h:13-39). */
int lbx_upload_f(lbx_sim* sim, int level, int box, const double* f);
int lbx_download_f(lbx_sim* sim, int level, int box, double* f);

/* Advance nsteps level-0 steps (Iterate, src/AmrSim.cpp:981-993): the
 * single-level collide-stream step (src/AmrSim.cpp:124-135, 109-122,
 * 324-333) on every box, or, when finer levels are defined, one Rohde cycle
 * per step (src/AmrSim.cpp:430-631: masked collide and stream of each
 * level, r steps of each level per step of its parent, r substeps of the
 * finest level per step of its parent, average-down and add).  A level l
 * advances r^l steps per call step.  Stream-ordered: may return before
 * the GPU finishes; lbx_sync() waits. */
int lbx_iterate(lbx_sim* sim, int nsteps);

/* Density and velocity of one box's valid cells (CalcHydroVars,
 * src/AmrSim.cpp:938-979).  Synchronous. */
int lbx_hydro(lbx_sim* sim, int level, int box, double* rho, double* u);

/* Density and velocity of every box of a level in one call (the bulk
 * readback CalcHydroVars needs, src/AmrSim.cpp:938-979): boxes in
 * definition order, each box's cells x fastest; rho holds sum(n_b) values,
 * u holds, per box, its [3][n_b] block.  Synchronous. */
int lbx_hydro_level(lbx_sim* sim, int level, double* rho, double* u);

/* FP64 global sums over a level's valid cells: out[0] = sum rho,
 * out[1..3] = sum rho*u.  No reference counterpart (the per-level mass at
 * src/AmrSim.cpp:312,318 is never used); deterministic order. */
int lbx_conservation(lbx_sim* sim, int level, double out[4]);

/* The same sums as lbx_conservation, bit for bit, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * allocates nothing: the partial sums live in a fixed device array of the
 * library (65536 partials of 4 doubles, 2 MiB per GPU, held until the
 * process ends), and a level that needs more partials (a patched level of
 * more than 65536 boxes) is rejected with LBX_E_ARG.  d_out is device memory
 * of at least 4 doubles owned by the caller; the sums are written there in
 * stream order on the context's stream, and the call returns without
 * waiting (lbx_sync waits).  Every context on one GPU shares that array:
 * calls that overlap on one GPU (from two contexts, or two threads) race. */
int lbx_conservation_noalloc(lbx_sim* sim, int level, double* d_out);

/* Timing of the stream-collide kernel with HIP events recorded on the
 * context's stream around every launch while enabled.  *ms is the summed
 * kernel time, *launches the number of timed launches since the last
 * reset. */
int lbx_profile_enable(lbx_ctx* ctx, int enable);
int lbx_profile_read(lbx_ctx* ctx, double* ms, int64_t* launches);

/* Autotuning of the stream-collide kernel's launch shape: workgroup extent
 * (bx, by; 256 or 512 lanes) and the cap on resident workgroups per CU
 * (0 = none, else 1 to 6; enforced by reserving LDS the kernel does not use).
 * While enabled, the first launch for a new (nx, ny, precision) times every
 * candidate on a small proxy lattice with the same rows and keeps the
 * fastest for the context's lifetime; shapes never change results.  No
 * reference counterpart.  lbx_tuned_shape reports the shape launches of
 * that geometry use: the tuned choice and its time per proxy launch, or the
 * default shape (time 0) if the geometry was never tuned. */
int lbx_set_autotune(lbx_ctx* ctx, int enable);
int lbx_tuned_shape(lbx_ctx* ctx, int nx, int ny, int precision, int32_t* bx, int32_t* by,
                    int32_t* wg_per_cu, float* ms_per_launch);

/* Achievable read+write bandwidth of this device: a plain streaming copy
 * of `bytes` (read) into a second buffer (write), repeated `reps` times;
 * *gbs = 2*bytes*reps / time.  No reference counterpart: it is the measured
 * ceiling reported beside the stream-collide kernel's bandwidth, and the
 * calibration kernel for counter-based traffic figures. */
int lbx_bench_copy(lbx_ctx* ctx, size_t bytes, int reps, double* gbs);

/* ---------------------------------------------------------------------
 * Batched serving of independent runs.  No reference counterpart: the
 * reference runs one AmrSim at a time.  A request is a periodic run with
 * its own tau, initial state and step count: a single-level run (its FP64
 * populations must fit the server's 256 MB staging buffer: at most
 * 2,236,962 cells; 128^3 fits), or, with fine boxes, a two-level run
 * (ratio 2) stepped by the Rohde cycle (src/AmrSim.cpp:430-469).
 * Single-level requests are admitted while the server's device arena has
 * room and advance together, one patch-list kernel launch per step over the
 * work items of every active run; a refined run is admitted while the
 * device has memory for its levels, is held in a simulation of its own on
 * the context (one captured level-0 step per server step, on the context's
 * stream, overlapping the patch-list launches on the server's stream) and
 * is released when delivered.  Each run computes exactly what lbx_iterate
 * computes for it alone, bit for bit. */

#define LBX_INIT_PLANE_PULSE 0  /* rho = 1 + amplitude on plane `plane` normal to `axis`, u = 0 */
#define LBX_INIT_FIELDS 1       /* rho [nz][ny][nx] and u [3][nz][ny][nx] given (copied at submit) */

typedef struct lbx_server lbx_server;

typedef struct {
  int32_t dims[3];
  int32_t steps;           /* >= 1 */
  double tau_s, tau_b;     /* omega = 1/(tau + 1/2) */
  int32_t init_kind;
  int32_t axis, plane;     /* LBX_INIT_PLANE_PULSE */
  double amplitude;
  const double* rho;       /* LBX_INIT_FIELDS */
  const double* u;
  double* out_rho;         /* optional, filled at completion: [nz][ny][nx] */
  double* out_u;           /* optional: [3][nz][ny][nx] */
  double* out_f;           /* optional: FP64 populations [15][nz][ny][nx] */
  /* Refined runs (n_fine_boxes > 0): level 1 at refinement ratio r
   * (ref_ratio: 0 or 2 for the reference's 2, or 3 or 4) over fine_lohi, 6
   * int32 per box (inclusive lo, hi in the fine index space r*dims), aligned
   * to the ratio; tau of level 1 rescaled as ComputeDt does
   * (src/AmrSim.cpp:313-314); level 0 masked under level 1 (the fine-grid
   * mask).  Level 0 is box_lohi's n_boxes boxes (one box over the domain if
   * n_boxes = 0; single-level runs take none).  mass/momentum and the out_*
   * fields are level 0's.  Box lists are copied at submit. */
  int32_t n_boxes;
  const int32_t* box_lohi;
  int32_t n_fine_boxes;
  const int32_t* fine_lohi;
  int32_t ref_ratio;
} lbx_request;

typedef struct {
  uint64_t id;
  int32_t status;          /* LBX_OK, or why a refined run could not be set up */
  int32_t steps;
  int64_t cells;
  double mass;             /* FP64 sum of rho at completion */
  double momentum[3];
  double t_submit, t_admit, t_done;  /* host steady-clock seconds */
} lbx_result;

/* arena_bytes = 0 takes half of the free device memory; max_active <= 44. */
int lbx_serve_create(lbx_ctx* ctx, int precision, double rho0, size_t arena_bytes, int max_active,
                     lbx_server** out);
int lbx_serve_destroy(lbx_server* srv);
const char* lbx_serve_last_error(const lbx_server* srv);
/* Admission policy.  LBX_SERVE_FIFO (default): arrival order, a request that
 * does not fit the arena yet holds back those behind it.
 * LBX_SERVE_SHORTEST_FIRST: queued requests in order of their work
 * (cells x steps, ties by arrival), each admitted if it fits; requests that
 * have waited longer than max_wait_s seconds since submission come first,
 * in arrival order (max_wait_s = 0: no limit, and under sustained overload a
 * large request can wait until the small ones are served).  max_wait_s is
 * ignored by LBX_SERVE_FIFO. */
#define LBX_SERVE_FIFO 0
#define LBX_SERVE_SHORTEST_FIRST 1
int lbx_serve_set_policy(lbx_server* srv, int policy, double max_wait_s);
int lbx_serve_submit(lbx_server* srv, const lbx_request* req, uint64_t* id);
/* Retire finished runs, admit queued ones, launch up to max_steps batched
 * steps (at most 4 ahead of the GPU). */
int lbx_serve_progress(lbx_server* srv, int max_steps, int* active, int* queued);
/* Progress until every submitted request has completed. */
int lbx_serve_drain(lbx_server* srv);
/* Move up to max completed results into out; *n = number returned. */
int lbx_serve_poll(lbx_server* srv, lbx_result* out, int max, int* n);
int lbx_serve_stats(lbx_server* srv, int64_t* lattice_updates, int64_t* steps_launched,
                    size_t* arena_free_largest);

#ifdef __cplusplus
}
#endif
#endif
