/*
 * lbx_oracle.c -- CPU restatement of LAMBReX's D3Q15 path (test
 * infrastructure only; see lbx_oracle.h for the file:line map).
 *
 * Build: gcc -O2 -ffp-contract=off (no -march=native, no fast-math), so that
 * every double operation is the plain IEEE one in the reference's order.
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   - tables: exact rational identities M*MI = I, MI[:,0] = w, sum_i w_i
 *     M[m][i] = delta_m0, checked against an independent Fraction inverse;
 *   - collide: exact-rational restatement on random cells; rest state is a
 *     fixed point; mass and momentum conserved; shear-wave decay at
 *     nu = tau/3 (reference tau);
 *   - stream: a permutation (checked against numpy.roll);
 *   - whole path: the reference's pulse goldens (tests/pulseRegression.h)
 *     under Catch2 Approx rules, and decomposition invariance (bitwise);
 *   - two-level cycle: no reference test pins it.  Its pieces are pinned
 *     one at a time against what AMReX's documented routines and the
 *     reference's own loops fix (tests/test_oracle_pins_twolevel.py):
 *     FillPatchTwoLevels + PC interpolation + the outer comp-0 zero on coded
 *     fields, ZeroInvalidComponents by brute-force enumeration,
 *     sum_fine_to_coarse by amrex_avgdown's closed form with ADD, and one
 *     whole cycle on a fully refined domain by the coarse cells' closed form;
 *     the ADD order of overlapping contributions stays unpinned;
 *   - deeper hierarchies: the recursion is pinned by an exact reduction (a
 *     level 1 covering the whole domain makes levels 1..n evolve bit for bit
 *     like a hierarchy one level shallower whose level 0 is that level 1);
 *     the level-0/level-1 coupling inside it is the two-level code above.
 */
#include "lbx_oracle.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Tables                                                                    */
/* ------------------------------------------------------------------------ */

/* D3Q15 velocities in the reference's population order
 * (include/d3q15_bgk.h:14-22): rest, +-x, +-y, +-z, then the 8 diagonals
 * with x slowest-varying sign. */
static const int CV[15][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1},
    {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1},
};

typedef struct { int64_t n, d; } rat;

static int64_t gcd64(int64_t a, int64_t b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) { int64_t t = a % b; a = b; b = t; }
  return a;
}
static rat rmk(int64_t n, int64_t d) {
  if (d < 0) { n = -n; d = -d; }
  int64_t g = gcd64(n, d);
  if (g == 0) g = 1;
  rat r = {n / g, d / g};
  if (r.n == 0) r.d = 1;
  return r;
}
static rat rsub(rat a, rat b) { return rmk(a.n * b.d - b.n * a.d, a.d * b.d); }
static rat rmul(rat a, rat b) { return rmk(a.n * b.n, a.d * b.d); }
static rat rdiv(rat a, rat b) { return rmk(a.n * b.d, a.d * b.n); }

/* Mode basis (src/AmrSim.cpp:1037-1054): density, momentum, the six
 * second moments c_a c_b - delta_ab/3 in the order xx,xy,xz,yy,yz,zz, and
 * five ghost modes g, c_x g, c_y g, c_z g, c_x c_y c_z with g = -2 for
 * |c|^2 in {0,3} and g = 1 for |c|^2 = 1. */
static rat mode_entry(int m, int p) {
  const int* c = CV[p];
  const int c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  const int g = (c2 == 1) ? 1 : -2;
  static const int pair[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  if (m == 0) return rmk(1, 1);
  if (m <= 3) return rmk(c[m - 1], 1);
  if (m <= 9) {
    const int a = pair[m - 4][0], b = pair[m - 4][1];
    return rmk(3 * c[a] * c[b] - (a == b ? 1 : 0), 3);
  }
  if (m == 10) return rmk(g, 1);
  if (m <= 13) return rmk(c[m - 11] * g, 1);
  return rmk(c[0] * c[1] * c[2], 1);
}

static double MM[15][15], MMI[15][15];
static int tables_ready = 0;

/* Exact Gauss-Jordan inverse over the rationals, then one rounding per
 * entry (a/b with small integers, as the reference's literals 1./24. etc.). */
static void build_tables(void) {
  if (tables_ready) return;
  rat A[15][30];
  for (int m = 0; m < 15; ++m)
    for (int p = 0; p < 30; ++p)
      A[m][p] = (p < 15) ? mode_entry(m, p) : rmk(p - 15 == m ? 1 : 0, 1);
  for (int col = 0; col < 15; ++col) {
    int piv = col;
    while (A[piv][col].n == 0) ++piv;
    if (piv != col)
      for (int p = 0; p < 30; ++p) { rat t = A[col][p]; A[col][p] = A[piv][p]; A[piv][p] = t; }
    rat inv = A[col][col];
    for (int p = 0; p < 30; ++p) A[col][p] = rdiv(A[col][p], inv);
    for (int r = 0; r < 15; ++r) {
      if (r == col || A[r][col].n == 0) continue;
      rat fac = A[r][col];
      for (int p = 0; p < 30; ++p) A[r][p] = rsub(A[r][p], rmul(fac, A[col][p]));
    }
  }
  for (int m = 0; m < 15; ++m)
    for (int p = 0; p < 15; ++p) {
      rat e = mode_entry(m, p);
      MM[m][p] = (double)e.n / (double)e.d;
      MMI[m][p] = (double)A[m][15 + p].n / (double)A[m][15 + p].d;
    }
  tables_ready = 1;
}

void orc_tables(double M[15][15], double MI[15][15], int C[15][3]) {
  build_tables();
  memcpy(M, MM, sizeof(MM));
  memcpy(MI, MMI, sizeof(MMI));
  memcpy(C, CV, sizeof(CV));
}

/* ------------------------------------------------------------------------ */
/* Per-cell kernels                                                          */
/* ------------------------------------------------------------------------ */

#define NDIMS 3
#define NMODES 15
/* DELTA[a][b] = 1/NMODES on the diagonal (src/AmrSim.cpp:1033-1035). */
static const double DELTA[3][3] = {{1.0 / NMODES, 0.0, 0.0},
                                   {0.0, 1.0 / NMODES, 0.0},
                                   {0.0, 0.0, 1.0 / NMODES}};
static const double CS2 = 1.0 / 3.0; /* include/AmrSim.h:34 */

/* src/AmrSim.cpp:28-104, the per-cell body of Collide. */
void orc_collide_cell(double f[15], double omega_s, double omega_b) {
  build_tables();
  double mode[15] = {0};
  for (int m = 0; m < NMODES; ++m)
    for (int p = 0; p < NMODES; ++p) mode[m] += f[p] * MM[m][p];

  const double density = mode[0];
  double velocity[NDIMS];
  double usq = 0.0;
  for (int a = 0; a < NDIMS; ++a) {
    velocity[a] = mode[a + 1] / density;
    usq += velocity[a] * velocity[a];
  }
  double stress[NDIMS][NDIMS] = {{mode[4], mode[5], mode[6]},
                                 {mode[5], mode[7], mode[8]},
                                 {mode[6], mode[8], mode[9]}};
  double TrS = 0.0;
  for (int a = 0; a < NDIMS; ++a) TrS += stress[a][a];
  for (int a = 0; a < NDIMS; ++a) stress[a][a] -= (TrS / NDIMS);
  TrS -= omega_b * (TrS - density * usq);
  for (int a = 0; a < NDIMS; ++a) {
    for (int b = 0; b < NDIMS; ++b)
      stress[a][b] -= omega_s * (stress[a][b] -
                                 density * (velocity[a] * velocity[b] - usq * DELTA[a][b]));
    stress[a][a] += (TrS / NDIMS);
  }
  mode[4] = stress[0][0];
  mode[5] = stress[0][1];
  mode[6] = stress[0][2];
  mode[7] = stress[1][1];
  mode[8] = stress[1][2];
  mode[9] = stress[2][2];
  for (int m = 10; m < 15; ++m) mode[m] = 0.0;

  for (int p = 0; p < NMODES; ++p) {
    double fp = 0;
    for (int m = 0; m < NMODES; ++m) fp += mode[m] * MMI[p][m];
    f[p] = fp;
  }
}

/* src/AmrSim.cpp:873-927, the per-cell body of CalcEquilibriumDist. */
void orc_equilibrium_cell(double rho, const double u[3], double f[15]) {
  double rho_w[3], u2[3], u_cs2[3], u2_2cs4[3];
  rho_w[0] = rho * 2.0 / 9.0;
  rho_w[1] = rho / 9.0;
  rho_w[2] = rho / 72.0;
  for (int a = 0; a < 3; ++a) u2[a] = u[a] * u[a];
  for (int a = 0; a < 3; ++a) u_cs2[a] = u[a] / CS2;
  for (int a = 0; a < 3; ++a) u2_2cs4[a] = u2[a] / (2.0 * CS2 * CS2);
  const double uv_cs4 = u_cs2[0] * u_cs2[1];
  const double vw_cs4 = u_cs2[1] * u_cs2[2];
  const double uw_cs4 = u_cs2[0] * u_cs2[2];
  const double mod_sq = (u2[0] + u2[1] + u2[2]) / (2.0 * CS2);
  const double mod_sq_2 = (u2[0] + u2[1] + u2[2]) * (1 - CS2) / (2.0 * CS2 * CS2);

  f[0] = rho_w[0] * (1.0 - mod_sq);
  /* axis populations: (1 - mod_sq) +- u_a/cs2 + u_a^2/(2 cs4) */
  for (int p = 1; p <= 6; ++p) {
    const int a = (p - 1) / 2;
    const double s = (p % 2) ? u_cs2[a] : -u_cs2[a];
    f[p] = rho_w[1] * (1.0 - mod_sq + s + u2_2cs4[a]);
  }
  /* diagonals: 1 + sum_a c_a u_a/cs2 + c_x c_y uv + c_y c_z vw + c_x c_z uw
   * + mod_sq_2, summed left to right as written in the reference. */
  for (int p = 7; p < 15; ++p) {
    const int* c = CV[p];
    double acc = 1.0;
    acc = (c[0] > 0) ? acc + u_cs2[0] : acc - u_cs2[0];
    acc = (c[1] > 0) ? acc + u_cs2[1] : acc - u_cs2[1];
    acc = (c[2] > 0) ? acc + u_cs2[2] : acc - u_cs2[2];
    acc = (c[0] * c[1] > 0) ? acc + uv_cs4 : acc - uv_cs4;
    acc = (c[1] * c[2] > 0) ? acc + vw_cs4 : acc - vw_cs4;
    acc = (c[0] * c[2] > 0) ? acc + uw_cs4 : acc - uw_cs4;
    acc = acc + mod_sq_2;
    f[p] = rho_w[2] * acc;
  }
}

/* src/AmrSim.cpp:957-971. */
void orc_hydro_cell(const double f[15], double* rho, double u[3]) {
  build_tables();
  double mode[15];
  for (int m = 0; m < NMODES; ++m) {
    mode[m] = 0.0;
    for (int p = 0; p < NMODES; ++p) mode[m] += f[p] * MM[m][p];
  }
  *rho = mode[0];
  for (int a = 0; a < NDIMS; ++a) u[a] = mode[a + 1] / mode[0];
}

static int wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

void orc_stream_periodic(const double* src, double* dst, int nx, int ny, int nz) {
  const size_t N = (size_t)nx * ny * nz;
  for (int q = 0; q < 15; ++q)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const int si = wrap(i - CV[q][0], nx), sj = wrap(j - CV[q][1], ny),
                    sk = wrap(k - CV[q][2], nz);
          dst[q * N + ((size_t)k * ny + j) * nx + i] =
              src[q * N + ((size_t)sk * ny + sj) * nx + si];
        }
}

/* ------------------------------------------------------------------------ */
/* Box-structured levels (AMReX FAB semantics, restated)                    */
/* ------------------------------------------------------------------------ */

typedef struct { int lo[3], hi[3]; } box3;

typedef struct {
  int nbox;
  box3* vb;        /* valid boxes */
  double** now;    /* per box, grown by ORC_HALO */
  double** next;
  double** tmp;
  int dom[3];      /* periodic level domain [0,dom) */
  int ratio;       /* refinement ratio to the parent level (level 0: 1) */
  double tau_s, tau_b, dt, time;
  int step;
  int* mask;       /* per coarse box valid cell: 1 = covered (FINE_VAL) */
  int* mask_off;
  int* owner;      /* per domain cell: index of the valid box holding it, or -1 */
} orc_level;

struct orc_sim {
  int n[3];
  int nlev;
  int ratio;       /* ratio of every level added next (refRatio, src/AmrSim.cpp:766: 2) */
  orc_level lev[ORC_MAX_LEVELS];
};

static int bnx(const box3* b, int d) { return b->hi[d] - b->lo[d] + 1; }
static size_t fab_cells(const box3* b) {
  return (size_t)(bnx(b, 0) + 2 * ORC_HALO) * (bnx(b, 1) + 2 * ORC_HALO) *
         (bnx(b, 2) + 2 * ORC_HALO);
}
/* index into a FAB grown by ORC_HALO at absolute position (i,j,k), comp n */
static size_t fidx(const box3* b, int i, int j, int k, int n) {
  const size_t gx = bnx(b, 0) + 2 * ORC_HALO, gy = bnx(b, 1) + 2 * ORC_HALO,
               gz = bnx(b, 2) + 2 * ORC_HALO;
  return (((size_t)n * gz + (k - b->lo[2] + ORC_HALO)) * gy + (j - b->lo[1] + ORC_HALO)) * gx +
         (i - b->lo[0] + ORC_HALO);
}
static int in_box(const box3* b, int i, int j, int k) {
  return i >= b->lo[0] && i <= b->hi[0] && j >= b->lo[1] && j <= b->hi[1] &&
         k >= b->lo[2] && k <= b->hi[2];
}

static void level_alloc(orc_level* L) {
  const size_t ncell = (size_t)L->dom[0] * L->dom[1] * L->dom[2];
  L->owner = (int*)malloc(ncell * sizeof(int));
  for (size_t c = 0; c < ncell; ++c) L->owner[c] = -1;
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int k = B->lo[2]; k <= B->hi[2]; ++k)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int i = B->lo[0]; i <= B->hi[0]; ++i)
          L->owner[((size_t)k * L->dom[1] + j) * L->dom[0] + i] = b;
  }
  L->now = (double**)calloc(L->nbox, sizeof(double*));
  L->next = (double**)calloc(L->nbox, sizeof(double*));
  L->tmp = (double**)calloc(L->nbox, sizeof(double*));
  for (int b = 0; b < L->nbox; ++b) {
    const size_t n = fab_cells(&L->vb[b]) * ORC_NV;
    L->now[b] = (double*)calloc(n, sizeof(double));
    L->next[b] = (double*)calloc(n, sizeof(double));
    L->tmp[b] = (double*)calloc(n, sizeof(double));
  }
}

static void level_free(orc_level* L) {
  for (int b = 0; b < L->nbox; ++b) { free(L->now[b]); free(L->next[b]); free(L->tmp[b]); }
  free(L->now); free(L->next); free(L->tmp); free(L->vb);
  free(L->mask); free(L->mask_off); free(L->owner);
  memset(L, 0, sizeof(*L));
}

/* Find the valid box of level L holding the periodic image of (i,j,k) and
 * the image itself; -1 if none. */
static int locate(const orc_level* L, int i, int j, int k, int* pi, int* pj, int* pk) {
  const int wi = wrap(i, L->dom[0]), wj = wrap(j, L->dom[1]), wk = wrap(k, L->dom[2]);
  *pi = wi; *pj = wj; *pk = wk;
  return L->owner[((size_t)wk * L->dom[1] + wj) * L->dom[0] + wi];
}

/* FillPatchSingleLevel(dst <- src) for ghost depth ng: every cell of each
 * grown dst box takes the periodic image's valid value
 * (src/AmrSim.cpp:364-372; AMReX ParallelCopy with periodicity).  With
 * ghosts_only, valid cells are left alone (FillBoundary, :132). */
static void fill_single_level(orc_level* L, double** dst, double** src, int ng, int ghosts_only) {
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int i = B->lo[0] - ng; i <= B->hi[0] + ng; ++i)
      for (int j = B->lo[1] - ng; j <= B->hi[1] + ng; ++j)
        for (int k = B->lo[2] - ng; k <= B->hi[2] + ng; ++k) {
          if (ghosts_only && in_box(B, i, j, k)) continue;
          int si, sj, sk;
          const int s = locate(L, i, j, k, &si, &sj, &sk);
          if (s < 0) continue;
          for (int n = 0; n < ORC_NV; ++n)
            dst[b][fidx(B, i, j, k, n)] = src[s][fidx(&L->vb[s], si, sj, sk, n)];
        }
  }
}

/* for_point_in over valid cells, i outer / k inner (include/amr_help.h:79-94). */
static void collide_level(orc_level* L, double** f, int masked) {
  const double OMEGA_S = 1.0 / (L->tau_s + 0.5);
  const double OMEGA_B = 1.0 / (L->tau_b + 0.5);
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int i = B->lo[0]; i <= B->hi[0]; ++i)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int k = B->lo[2]; k <= B->hi[2]; ++k) {
          double cell[15];
          if (masked && L->mask) {
            const size_t mi = L->mask_off[b] +
                              (((size_t)(k - B->lo[2]) * bnx(B, 1) + (j - B->lo[1])) * bnx(B, 0) +
                               (i - B->lo[0]));
            if (L->mask[mi] == 1) { /* src/AmrSim.cpp:499-500 */
              for (int m = 0; m < 15; ++m) f[b][fidx(B, i, j, k, m)] = 0.0;
              continue;
            }
          }
          for (int n = 0; n < 15; ++n) cell[n] = f[b][fidx(B, i, j, k, n)];
          orc_collide_cell(cell, OMEGA_S, OMEGA_B);
          for (int n = 0; n < 15; ++n) f[b][fidx(B, i, j, k, n)] = cell[n];
        }
  }
}

/* Stream (src/AmrSim.cpp:109-122): a fresh MultiFab is pulled over valid + 1
 * ghost layer; its outer ghost layer is never written by the reference
 * (uninitialised memory there); this restatement defines it as zero.  The
 * fresh array then replaces `f`. */
static void stream_level(orc_level* L, double*** f) {
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    double* in = (*f)[b];
    double* out = L->tmp[b];
    memset(out, 0, fab_cells(B) * ORC_NV * sizeof(double));
    for (int i = B->lo[0] - 1; i <= B->hi[0] + 1; ++i)
      for (int j = B->lo[1] - 1; j <= B->hi[1] + 1; ++j)
        for (int k = B->lo[2] - 1; k <= B->hi[2] + 1; ++k)
          for (int q = 0; q < ORC_NV; ++q)
            out[fidx(B, i, j, k, q)] = in[fidx(B, i - CV[q][0], j - CV[q][1], k - CV[q][2], q)];
  }
  double** t = *f;
  *f = L->tmp;
  L->tmp = t;
}

static void swap_now_next(orc_level* L) {
  double** t = L->now;
  L->now = L->next;
  L->next = t;
}

/* ------------------------------------------------------------------------ */
/* Two-level pieces (Rohde cycle, src/AmrSim.cpp:430-631)                    */
/* ------------------------------------------------------------------------ */

/* ORC_MUTANT plants one plausible misreading of the two-level pieces, for the
 * test that shows each pin catches it (built only by that test, never by the
 * Makefile):
 *   1 sum_fine_to_coarse without the 1/8      2 ... over valid children only
 *   3 ... overwrites instead of ADDing        4 ... children in i,j,k order
 *   5 ghosts from coarse even with a fine image
 *   6 parent index by truncation, not floor   7 comp 0 zeroed on both layers
 *   8 every component zeroed on the outer layer
 *   9 ZeroInvalid keeps pos + 2c in valid    10 ... keeps pos - c in valid
 *  11 average-down over the valid boxes only (nGrow 0) at every ratio
 *  12 average-down volume fraction 1/8 at every ratio
 *  13 the finest level takes 2 substeps at every ratio
 *  14 a level between 0 and the finest takes 2 steps per parent step at
 *     every ratio */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

static int floor_div(int i, int r) { return (i >= 0) ? i / r : -((-i + r - 1) / r); }
static int parent_of(int i, int r) { return ORC_MUTANT == 6 ? i / r : floor_div(i, r); }

/* FillPatchTwoLevels with PCInterp (src/AmrSim.cpp:373-388): every cell of
 * each grown fine box takes the fine valid value of its periodic image when
 * some fine box holds it, else the coarse `now` value of the coarse cell
 * that contains it (piecewise-constant over the refinement ratio, coarse
 * periodic wrap). */
static void fill_two_levels(orc_level* C, orc_level* F, double** dst) {
  const int r = F->ratio;
  for (int b = 0; b < F->nbox; ++b) {
    const box3* B = &F->vb[b];
    for (int i = B->lo[0] - ORC_HALO; i <= B->hi[0] + ORC_HALO; ++i)
      for (int j = B->lo[1] - ORC_HALO; j <= B->hi[1] + ORC_HALO; ++j)
        for (int k = B->lo[2] - ORC_HALO; k <= B->hi[2] + ORC_HALO; ++k) {
          int si, sj, sk;
          const int s = locate(F, i, j, k, &si, &sj, &sk);
          if (s >= 0 && (ORC_MUTANT != 5 || in_box(B, i, j, k))) {
            for (int n = 0; n < ORC_NV; ++n)
              dst[b][fidx(B, i, j, k, n)] = F->now[s][fidx(&F->vb[s], si, sj, sk, n)];
          } else {
            int ci, cj, ck;
            const int c = locate(C, parent_of(i, r), parent_of(j, r), parent_of(k, r), &ci, &cj, &ck);
            for (int n = 0; n < ORC_NV; ++n)
              dst[b][fidx(B, i, j, k, n)] = C->now[c][fidx(&C->vb[c], ci, cj, ck, n)];
          }
        }
  }
}

/* for_point_in_boundary(f, depth) (include/amr_help.h:99-167): every cell
 * of the grown box within `depth` of its outer surface, each once. */
static int in_outer_shell(const box3* B, int i, int j, int k, int depth) {
  const int g = ORC_HALO;
  const int lo[3] = {B->lo[0] - g, B->lo[1] - g, B->lo[2] - g};
  const int hi[3] = {B->hi[0] + g, B->hi[1] + g, B->hi[2] + g};
  const int p[3] = {i, j, k};
  for (int d = 0; d < 3; ++d)
    if (p[d] < lo[d] + depth || p[d] > hi[d] - depth) return 1;
  return 0;
}

/* InitPostCollision on a fine level (src/AmrSim.cpp:471-485): after the
 * two-level fill, component 0 only is zeroed on the outermost ghost layer. */
static void zero_outer_comp0(orc_level* F, double** f) {
  for (int b = 0; b < F->nbox; ++b) {
    const box3* B = &F->vb[b];
    for (int i = B->lo[0] - ORC_HALO; i <= B->hi[0] + ORC_HALO; ++i)
      for (int j = B->lo[1] - ORC_HALO; j <= B->hi[1] + ORC_HALO; ++j)
        for (int k = B->lo[2] - ORC_HALO; k <= B->hi[2] + ORC_HALO; ++k)
          if (in_outer_shell(B, i, j, k, ORC_MUTANT == 7 ? 2 : 1)) {
            f[b][fidx(B, i, j, k, 0)] = 0.0;
            if (ORC_MUTANT == 8)
              for (int n = 1; n < ORC_NV; ++n) f[b][fidx(B, i, j, k, n)] = 0.0;
          }
  }
}

/* ZeroInvalidComponents (src/AmrSim.cpp:604-617): in the 2-deep ghost shell
 * keep population m only if pos - 2 c_m lies in the box's valid region. */
static void zero_invalid(orc_level* L, double** f) {
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int i = B->lo[0] - ORC_HALO; i <= B->hi[0] + ORC_HALO; ++i)
      for (int j = B->lo[1] - ORC_HALO; j <= B->hi[1] + ORC_HALO; ++j)
        for (int k = B->lo[2] - ORC_HALO; k <= B->hi[2] + ORC_HALO; ++k) {
          if (!in_outer_shell(B, i, j, k, ORC_HALO)) continue;
          for (int m = 0; m < ORC_NV; ++m) {
            const int a = ORC_MUTANT == 9 ? -2 : ORC_MUTANT == 10 ? 1 : 2;
            if (!in_box(B, i - a * CV[m][0], j - a * CV[m][1], k - a * CV[m][2]))
              f[b][fidx(B, i, j, k, m)] = 0.0;
          }
        }
  }
}

/* sum_fine_to_coarse (src/AmrSim.cpp:592-602, AMReX's sum_fine_to_coarse
 * with ratio r): each fine box grown by its 2 fine ghosts is averaged over
 * r x r x r blocks (amrex_avgdown: sum in k,j,i order of the fine children,
 * then times the volume fraction 1/r^3) onto its coarsened box grown by
 * nGrow = 2 / r coarse cells (1 for ratio 2, 0 above), and the result is
 * ADDED into every coarse valid cell that holds the periodic image of that
 * coarse position.  Boxes are added in box order. */
static void sum_fine_to_coarse(orc_level* F, orc_level* C, double** cdst) {
  const int r = F->ratio, ng = ORC_MUTANT == 11 ? 0 : ORC_HALO / r;
  const double volfrac = ORC_MUTANT == 12 ? 0.125 : 1.0 / (double)(r * r * r);
  for (int b = 0; b < F->nbox; ++b) {
    const box3* B = &F->vb[b];
    const int clo[3] = {floor_div(B->lo[0], r) - ng, floor_div(B->lo[1], r) - ng, floor_div(B->lo[2], r) - ng};
    const int chi[3] = {floor_div(B->hi[0], r) + ng, floor_div(B->hi[1], r) + ng, floor_div(B->hi[2], r) + ng};
    for (int ci = clo[0]; ci <= chi[0]; ++ci)
      for (int cj = clo[1]; cj <= chi[1]; ++cj)
        for (int ck = clo[2]; ck <= chi[2]; ++ck) {
          int ti, tj, tk;
          const int c = locate(C, ci, cj, ck, &ti, &tj, &tk);
          if (c < 0) continue;
          for (int n = 0; n < ORC_NV; ++n) {
            double s = 0.0;
            for (int dk = 0; dk < r; ++dk)
              for (int dj = 0; dj < r; ++dj)
                for (int di = 0; di < r; ++di) {
                  int ii = r * ci + di, jj = r * cj + dj, kk = r * ck + dk;
                  if (ORC_MUTANT == 4) { ii = r * ci + dk; kk = r * ck + di; }
                  if (ORC_MUTANT == 2 && !in_box(B, ii, jj, kk)) continue;
                  s += F->now[b][fidx(B, ii, jj, kk, n)];
                }
            if (ORC_MUTANT != 1) s *= volfrac;
            if (ORC_MUTANT == 3)
              cdst[c][fidx(&C->vb[c], ti, tj, tk, n)] = s;
            else
              cdst[c][fidx(&C->vb[c], ti, tj, tk, n)] += s;
          }
        }
  }
}

/* ------------------------------------------------------------------------ */
/* Simulation                                                                */
/* ------------------------------------------------------------------------ */

orc_sim* orc_sim_create(int nx, int ny, int nz, int nboxes, const int* lohi,
                        double tau_s, double tau_b) {
  if (nx < 1 || ny < 1 || nz < 1 || nboxes < 1 || !lohi) return NULL;
  build_tables();
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  s->n[0] = nx; s->n[1] = ny; s->n[2] = nz;
  s->nlev = 1;
  orc_level* L = &s->lev[0];
  L->nbox = nboxes;
  L->vb = (box3*)calloc(nboxes, sizeof(box3));
  size_t cells = 0;
  for (int b = 0; b < nboxes; ++b) {
    for (int d = 0; d < 3; ++d) {
      L->vb[b].lo[d] = lohi[6 * b + d];
      L->vb[b].hi[d] = lohi[6 * b + 3 + d];
      if (L->vb[b].lo[d] < 0 || L->vb[b].hi[d] >= s->n[d] || L->vb[b].hi[d] < L->vb[b].lo[d]) {
        free(L->vb); free(s); return NULL;
      }
    }
    cells += (size_t)bnx(&L->vb[b], 0) * bnx(&L->vb[b], 1) * bnx(&L->vb[b], 2);
  }
  if (cells != (size_t)nx * ny * nz) { free(L->vb); free(s); return NULL; }
  for (int d = 0; d < 3; ++d) L->dom[d] = s->n[d];
  L->tau_s = tau_s;
  L->tau_b = tau_b;
  L->dt = 1.0; /* ComputeDt level 0, src/AmrSim.cpp:316-319 */
  L->ratio = 1;
  s->ratio = 2;
  level_alloc(L);
  return s;
}

void orc_sim_destroy(orc_sim* s) {
  if (!s) return;
  for (int l = 0; l < s->nlev; ++l) level_free(&s->lev[l]);
  free(s);
}

int orc_sim_init(orc_sim* s, const double* rho_c, const double* u_c) {
  orc_level* L = &s->lev[0];
  const int NY = s->n[1], NZ = s->n[2];
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int k = B->lo[2]; k <= B->hi[2]; ++k)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
          /* CLindex (include/AmrSim.h:79-83): i slowest, k fastest. */
          const size_t c = ((size_t)i * NY + j) * NZ + k;
          const double u[3] = {u_c[3 * c + 0], u_c[3 * c + 1], u_c[3 * c + 2]};
          double f[15];
          orc_equilibrium_cell(rho_c[c], u, f);
          for (int n = 0; n < 15; ++n) L->now[b][fidx(B, i, j, k, n)] = f[n];
        }
  }
  L->time = 0.0;
  L->step = 0;
  return 0;
}

static void iterate_single(orc_sim* s) {
  orc_level* L = &s->lev[0];
  /* CollideLevel (src/AmrSim.cpp:124-135) */
  fill_single_level(L, L->next, L->now, ORC_HALO, 0);
  collide_level(L, L->next, 0);
  fill_single_level(L, L->next, L->next, ORC_HALO, 1);
  /* Stream (src/AmrSim.cpp:109-122) */
  stream_level(L, &L->next);
  /* UpdateNow + time (include/multilevel.h:75-79, src/AmrSim.cpp:328-330) */
  swap_now_next(L);
  L->time += L->dt;
  ++L->step;
}

/* The finest level's part of RohdeCycle (src/AmrSim.cpp:445-452):
 * InitPostCollision (two-level fill, comp 0 zeroed on the outer shell), then
 * FineCollide + Stream once per fine substep with no ghost refresh between,
 * ZeroInvalidComponents, and UpdateDistribution's steps and swap
 * (src/AmrSim.cpp:619-631).  The reference writes the two substeps and
 * UpdateDistribution's 2 out for its ratio 2; at ratio r the finest level
 * takes r substeps (dt_fine = dt_coarse / r), which is the same for r = 2. */
static void finest_pair(orc_level* C, orc_level* F) {
  fill_two_levels(C, F, F->next);
  zero_outer_comp0(F, F->next);
  for (int sub = 0; sub < (ORC_MUTANT == 13 ? 2 : F->ratio); ++sub) {
    collide_level(F, F->next, 0);
    stream_level(F, &F->next);
  }
  zero_invalid(F, F->next);
  F->time += F->ratio * F->dt;
  F->step += F->ratio;
  swap_now_next(F);
}

/* RohdeCycle(l) (src/AmrSim.cpp:430-469): one step of level l, which has a
 * finer level.  Level 0's InitPostCollision is a single-level fill; a finer
 * level's is the two-level fill from its parent's `now` plus comp 0 zeroed on
 * the outer ghost layer (:471-485).  The collide is masked by level l's fine
 * mask (:487-580).  If level l+1 is the finest it takes its two substeps
 * once, else RohdeCycle(l+1) runs refRatio = 2 times (:453-455).  Then
 * Stream(l), SumFromFine(l) from level l+1's `now`, ZeroInvalidComponents(l),
 * UpdateBoundaries on level 0 only, and UpdateDistribution(l): one step. */
static void rohde_cycle(orc_sim* s, int l) {
  orc_level* L = &s->lev[l];
  orc_level* F = &s->lev[l + 1];
  if (l == 0) {
    fill_single_level(L, L->next, L->now, ORC_HALO, 0);
  } else {
    fill_two_levels(&s->lev[l - 1], L, L->next);
    zero_outer_comp0(L, L->next);
  }
  collide_level(L, L->next, 1);
  if (l + 1 == s->nlev - 1) {
    finest_pair(L, F);
  } else {
    for (int iter = 0; iter < (ORC_MUTANT == 14 ? 2 : F->ratio); ++iter) rohde_cycle(s, l + 1);  /* REF_RATIO, :453-455 */
  }
  stream_level(L, &L->next);
  sum_fine_to_coarse(F, L, L->next);
  zero_invalid(L, L->next);
  if (l == 0) fill_single_level(L, L->next, L->next, ORC_HALO, 1);
  L->time += L->dt;
  L->step += 1;
  swap_now_next(L);
}

int orc_sim_set_ratio(orc_sim* s, int ratio) {
  if (ratio < 2 || ratio > 4) return -1;
  s->ratio = ratio;
  return 0;
}

void orc_sim_iterate(orc_sim* s, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    if (s->nlev == 1) iterate_single(s);
    else rohde_cycle(s, 0);
  }
}

int orc_sim_add_fine(orc_sim* s, int nboxes, const int* lohi_fine, int nmask,
                     const int* mask_lohi_fine) {
  if (s->nlev < 1 || s->nlev >= ORC_MAX_LEVELS || nboxes < 1) return -1;
  orc_level* C = &s->lev[s->nlev - 1];
  orc_level* F = &s->lev[s->nlev];
  memset(F, 0, sizeof(*F));
  F->nbox = nboxes;
  F->vb = (box3*)calloc(nboxes, sizeof(box3));
  const int r = s->ratio;
  F->ratio = r;
  for (int d = 0; d < 3; ++d) F->dom[d] = r * C->dom[d];
  for (int b = 0; b < nboxes; ++b)
    for (int d = 0; d < 3; ++d) {
      F->vb[b].lo[d] = lohi_fine[6 * b + d];
      F->vb[b].hi[d] = lohi_fine[6 * b + 3 + d];
      if (F->vb[b].lo[d] < 0 || F->vb[b].hi[d] >= F->dom[d] || F->vb[b].hi[d] < F->vb[b].lo[d]) {
        free(F->vb); memset(F, 0, sizeof(*F)); return -2;
      }
    }
  /* ComputeDt(1) (src/AmrSim.cpp:308-315) at ratio r */
  F->dt = C->dt / r;
  F->tau_s = r * (C->tau_s - 0.5) + 0.5;
  F->tau_b = r * (C->tau_b - 0.5) + 0.5;
  F->time = C->time;
  F->step = 0;
  level_alloc(F);
  /* Proper nesting: every cell of each grown fine box that no fine box
   * holds, and every valid fine cell, needs a parent in the coarse level's
   * valid region (AMReX guarantees it when it builds the grids; a parent
   * outside would leave FillPatchTwoLevels without data). */
  for (int b = 0; b < F->nbox; ++b) {
    const box3* B = &F->vb[b];
    for (int i = B->lo[0] - ORC_HALO; i <= B->hi[0] + ORC_HALO; ++i)
      for (int j = B->lo[1] - ORC_HALO; j <= B->hi[1] + ORC_HALO; ++j)
        for (int k = B->lo[2] - ORC_HALO; k <= B->hi[2] + ORC_HALO; ++k) {
          int si, sj, sk;
          if (!in_box(B, i, j, k) && locate(F, i, j, k, &si, &sj, &sk) >= 0) continue;
          if (locate(C, floor_div(i, r), floor_div(j, r), floor_div(k, r), &si, &sj, &sk) < 0) {
            level_free(F);
            return -3;
          }
        }
  }
  s->nlev += 1;
  /* DistFnFillFromCoarse (src/AmrSim.cpp:393-411): PC interpolation of the
   * coarse `now` into fine valid cells.  Fine ghosts are not read before the
   * first two-level fill overwrites them. */
  for (int b = 0; b < F->nbox; ++b) {
    const box3* B = &F->vb[b];
    for (int i = B->lo[0]; i <= B->hi[0]; ++i)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int k = B->lo[2]; k <= B->hi[2]; ++k) {
          int ci, cj, ck;
          const int c = locate(C, floor_div(i, r), floor_div(j, r), floor_div(k, r), &ci, &cj, &ck);
          for (int n = 0; n < ORC_NV; ++n)
            F->now[b][fidx(B, i, j, k, n)] = C->now[c][fidx(&C->vb[c], ci, cj, ck, n)];
        }
  }
  /* Fine mask on coarse valid cells: FINE_VAL where the coarse cell lies in
   * coarsen(mask box, r) for some mask box (src/AmrSim.cpp:419-428). */
  size_t total = 0;
  free(C->mask);
  free(C->mask_off);
  C->mask_off = (int*)calloc(C->nbox, sizeof(int));
  for (int b = 0; b < C->nbox; ++b) {
    C->mask_off[b] = (int)total;
    total += (size_t)bnx(&C->vb[b], 0) * bnx(&C->vb[b], 1) * bnx(&C->vb[b], 2);
  }
  C->mask = (int*)calloc(total, sizeof(int));
  for (int b = 0; b < C->nbox; ++b) {
    const box3* B = &C->vb[b];
    for (int k = B->lo[2]; k <= B->hi[2]; ++k)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
          int hit = 0;
          for (int m = 0; m < nmask && !hit; ++m) {
            const int* mb = &mask_lohi_fine[6 * m];
            if (i >= floor_div(mb[0], r) && i <= floor_div(mb[3], r) && j >= floor_div(mb[1], r) &&
                j <= floor_div(mb[4], r) && k >= floor_div(mb[2], r) && k <= floor_div(mb[5], r))
              hit = 1;
          }
          C->mask[C->mask_off[b] + (((size_t)(k - B->lo[2]) * bnx(B, 1) + (j - B->lo[1])) * bnx(B, 0) +
                                     (i - B->lo[0]))] = hit;
        }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Accessors                                                                 */
/* ------------------------------------------------------------------------ */

void orc_sim_bbox(const orc_sim* s, int level, int bb[6]) {
  const orc_level* L = &s->lev[level];
  for (int d = 0; d < 3; ++d) { bb[d] = 1 << 30; bb[3 + d] = -(1 << 30); }
  for (int b = 0; b < L->nbox; ++b)
    for (int d = 0; d < 3; ++d) {
      if (L->vb[b].lo[d] < bb[d]) bb[d] = L->vb[b].lo[d];
      if (L->vb[b].hi[d] > bb[3 + d]) bb[3 + d] = L->vb[b].hi[d];
    }
}

typedef void (*cell_fn)(const orc_level* L, int b, int i, int j, int k, size_t o, size_t N, void* ctx);

static void for_valid(const orc_sim* s, int level, cell_fn fn, void* ctx) {
  const orc_level* L = &s->lev[level];
  int bb[6];
  orc_sim_bbox(s, level, bb);
  const size_t nx = bb[3] - bb[0] + 1, ny = bb[4] - bb[1] + 1, nz = bb[5] - bb[2] + 1;
  const size_t N = nx * ny * nz;
  for (int b = 0; b < L->nbox; ++b) {
    const box3* B = &L->vb[b];
    for (int k = B->lo[2]; k <= B->hi[2]; ++k)
      for (int j = B->lo[1]; j <= B->hi[1]; ++j)
        for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
          const size_t o = ((size_t)(k - bb[2]) * ny + (j - bb[1])) * nx + (i - bb[0]);
          fn(L, b, i, j, k, o, N, ctx);
        }
  }
}

static void get_f_cell(const orc_level* L, int b, int i, int j, int k, size_t o, size_t N, void* ctx) {
  double* f = (double*)ctx;
  for (int n = 0; n < 15; ++n) f[n * N + o] = L->now[b][fidx(&L->vb[b], i, j, k, n)];
}
static void set_f_cell(const orc_level* L, int b, int i, int j, int k, size_t o, size_t N, void* ctx) {
  const double* f = (const double*)ctx;
  for (int n = 0; n < 15; ++n) L->now[b][fidx(&L->vb[b], i, j, k, n)] = f[n * N + o];
}
typedef struct { double* rho; double* u; } hydro_out;
static void hydro_cell(const orc_level* L, int b, int i, int j, int k, size_t o, size_t N, void* ctx) {
  hydro_out* h = (hydro_out*)ctx;
  double f[15], u[3];
  for (int n = 0; n < 15; ++n) f[n] = L->now[b][fidx(&L->vb[b], i, j, k, n)];
  orc_hydro_cell(f, &h->rho[o], u);
  for (int a = 0; a < 3; ++a) h->u[a * N + o] = u[a];
}

void orc_sim_get_f(const orc_sim* s, int level, double* f) { for_valid(s, level, get_f_cell, f); }
void orc_sim_set_f(orc_sim* s, int level, const double* f) { for_valid(s, level, set_f_cell, (void*)f); }
void orc_sim_hydro(const orc_sim* s, int level, double* rho, double* u) {
  hydro_out h = {rho, u};
  for_valid(s, level, hydro_cell, &h);
}
double orc_sim_time(const orc_sim* s, int level) { return s->lev[level].time; }
int orc_sim_step(const orc_sim* s, int level) { return s->lev[level].step; }
double orc_sim_tau(const orc_sim* s, int level, int which) {
  return which ? s->lev[level].tau_b : s->lev[level].tau_s;
}

/* ------------------------------------------------------------------------ */
/* Test hooks: whole FABs and single stages of the cycle                     */
/* ------------------------------------------------------------------------ */

int orc_sim_nbox(const orc_sim* s, int level) { return s->lev[level].nbox; }

void orc_sim_box(const orc_sim* s, int level, int box, int lohi[6]) {
  const box3* B = &s->lev[level].vb[box];
  for (int d = 0; d < 3; ++d) { lohi[d] = B->lo[d]; lohi[3 + d] = B->hi[d]; }
}

static double** fabs_of(orc_sim* s, int level, int which) {
  return which ? s->lev[level].next : s->lev[level].now;
}

void orc_sim_get_grown(const orc_sim* s, int level, int which, int box, double* out) {
  const box3* B = &s->lev[level].vb[box];
  memcpy(out, fabs_of((orc_sim*)s, level, which)[box], fab_cells(B) * ORC_NV * sizeof(double));
}

void orc_sim_set_grown(orc_sim* s, int level, int which, int box, const double* in) {
  const box3* B = &s->lev[level].vb[box];
  memcpy(fabs_of(s, level, which)[box], in, fab_cells(B) * ORC_NV * sizeof(double));
}

int orc_stage(orc_sim* s, int level, int stage) {
  if (level < 0 || level >= s->nlev) return -1;
  orc_level* L = &s->lev[level];
  switch (stage) {
    case ORC_STAGE_INIT_POST_COLLISION: /* src/AmrSim.cpp:471-485 */
      if (level == 0) {
        fill_single_level(L, L->next, L->now, ORC_HALO, 0);
      } else {
        fill_two_levels(&s->lev[level - 1], L, L->next);
        zero_outer_comp0(L, L->next);
      }
      return 0;
    case ORC_STAGE_ZERO_INVALID: /* :604-617 */
      zero_invalid(L, L->next);
      return 0;
    case ORC_STAGE_SUM_FROM_FINE: /* :592-602 */
      if (level + 1 >= s->nlev) return -1;
      sum_fine_to_coarse(&s->lev[level + 1], L, L->next);
      return 0;
    default:
      return -1;
  }
}
