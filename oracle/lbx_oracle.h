/*
 * lbx_oracle.h -- CPU restatement of LAMBReX's D3Q15 collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP product
 * path in lambrex_amd/ is compared against.  Only tests/, the smoke entry
 * point and bench.py's CPU-baseline leg may load it; the product never does.
 *
 * It follows the reference (otbrown/LAMBReX) line by line where the
 * arithmetic matters:
 *   - mode tables            src/AmrSim.cpp:1033-1073 (derived here from
 *                            exact rationals, see lbx_oracle.c)
 *   - Collide                src/AmrSim.cpp:25-107 (dense 15x15 loops)
 *   - Stream/PropagatePoint  src/AmrSim.cpp:109-122, include/component.h:22-29
 *   - CollideLevel           src/AmrSim.cpp:124-135
 *   - IterateLevel/Iterate   src/AmrSim.cpp:324-333, 981-993,
 *                            include/AmrSim.h:89-94
 *   - ghost fills            src/AmrSim.cpp:359-372 (FillPatchSingleLevel),
 *                            :19-23 and :132 (FillBoundary), periodic only
 *   - equilibrium            src/AmrSim.cpp:845-936
 *   - hydro fields           src/AmrSim.cpp:938-979
 *   - C-order init layout    src/AmrSim.cpp:138-295, include/AmrSim.h:79-83
 *   - Rohde cycle            src/AmrSim.cpp:419-631, recursive over levels
 *   - loop drivers           include/amr_help.h:79-167
 *
 * Storage mirrors AMReX FABs: one array per box grown by 2 ghost layers,
 * component-major, x fastest: data[n][k][j][i].  Every function here was
 * written to be obviously equal to the reference, not fast.
 */
#ifndef LBX_ORACLE_H
#define LBX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NV 15
#define ORC_HALO 2
#define ORC_MAX_LEVELS 4

/* Tables: velocity set (include/d3q15_bgk.h:14-22) and mode matrices. */
void orc_tables(double M[15][15], double MI[15][15], int C[15][3]);

/* Per-cell kernels, literal restatements. */
void orc_collide_cell(double f[15], double omega_s, double omega_b);
void orc_equilibrium_cell(double rho, const double u[3], double f[15]);
void orc_hydro_cell(const double f[15], double* rho, double u[3]);

/* Whole-lattice helpers on a single periodic domain stored x fastest
 * ([15][nz][ny][nx]); used by the permutation and shear-wave checks. */
void orc_stream_periodic(const double* src, double* dst, int nx, int ny, int nz);

typedef struct orc_sim orc_sim;

/* A simulation on the periodic domain [0,nx)x[0,ny)x[0,nz).  lohi holds 6
 * inclusive ints per level-0 box (lo x,y,z, hi x,y,z); the boxes must tile
 * the domain.  tau values are the reference's (omega = 1/(tau+1/2),
 * src/AmrSim.cpp:125-126).  Returns NULL on invalid input. */
orc_sim* orc_sim_create(int nx, int ny, int nz, int nboxes, const int* lohi,
                        double tau_s, double tau_b);
void orc_sim_destroy(orc_sim* s);

/* Initial density (nx*ny*nz) and velocity (3*nx*ny*nz, component fastest),
 * both in the reference's C order (i slowest, k fastest), then the
 * equilibrium distribution on level 0.  src/AmrSim.cpp:138-295, 845-936. */
int orc_sim_init(orc_sim* s, const double* rho_c, const double* u_c);

/* Add a new finest level (refinement ratio 2 unless set below, at most ORC_MAX_LEVELS
 * levels) made of the given boxes in its own index space, filled from the
 * current finest level by piecewise-constant interpolation
 * (src/AmrSim.cpp:689-713).  mask_lohi lists the boxes (new level's index
 * space) whose coarsened (by the ratio) cover is masked in the parent's collide; the
 * reference passes the parent's own BoxArray (src/AmrSim.cpp:419-428).
 * Returns -3 if a grown new box needs a parent cell outside the parent
 * level's valid region (not properly nested). */
int orc_sim_add_fine(orc_sim* s, int nboxes, const int* lohi_fine,
                     int nmask, const int* mask_lohi_fine);
/* Refinement ratio (2, 3 or 4) of the levels added after this call; the
 * reference's constructor fixes 2 (src/AmrSim.cpp:766).  The new level's
 * boxes, domain, parents, dt, tau, mask and average-down use it, and it
 * takes `ratio` substeps per step of its parent (2 written out in the
 * reference, src/AmrSim.cpp:446-452, 621-624).  Returns -1 otherwise. */
int orc_sim_set_ratio(orc_sim* s, int ratio);

void orc_sim_iterate(orc_sim* s, int nsteps);

/* Level data accessors.  f is [15][nz][ny][nx] over the level's bounding
 * box of valid cells (cells outside every box are left untouched);
 * rho [nz][ny][nx]; u [3][nz][ny][nx].  Returns the bounding box in bb[6]. */
void orc_sim_bbox(const orc_sim* s, int level, int bb[6]);
void orc_sim_get_f(const orc_sim* s, int level, double* f);
void orc_sim_set_f(orc_sim* s, int level, const double* f);
void orc_sim_hydro(const orc_sim* s, int level, double* rho, double* u);
double orc_sim_time(const orc_sim* s, int level);
int orc_sim_step(const orc_sim* s, int level);
double orc_sim_tau(const orc_sim* s, int level, int which /*0=s,1=b*/);

/* Test hooks.  A box's FAB grown by ORC_HALO, [15][gz][gy][gx] x fastest,
 * of the `now` (which = 0) or `next` (which = 1) state; and single stages of
 * the Rohde cycle on a level's `next` state, for the pins of the two-level
 * pieces:
 *   ORC_STAGE_INIT_POST_COLLISION  InitPostCollision(level), :471-485
 *   ORC_STAGE_ZERO_INVALID         ZeroInvalidComponents(level), :604-617
 *   ORC_STAGE_SUM_FROM_FINE        SumFromFine(level) from level+1's now,
 *                                  :592-602
 * Returns -1 for an invalid level or stage. */
#define ORC_STAGE_INIT_POST_COLLISION 0
#define ORC_STAGE_ZERO_INVALID 1
#define ORC_STAGE_SUM_FROM_FINE 2
int orc_sim_nbox(const orc_sim* s, int level);
void orc_sim_box(const orc_sim* s, int level, int box, int lohi[6]);
void orc_sim_get_grown(const orc_sim* s, int level, int which, int box, double* out);
void orc_sim_set_grown(orc_sim* s, int level, int which, int box, const double* in);
int orc_stage(orc_sim* s, int level, int stage);

#ifdef __cplusplus
}
#endif
#endif
