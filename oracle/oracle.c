/*
 * CPU oracle for the MassCCS trajectory-method hot path -- TEST
 * INFRASTRUCTURE ONLY (see oracle.h).  Every function cites the reference
 * file:line it restates; arithmetic is written in the reference's operation
 * order so that a -O2 -fno-fast-math -ffp-contract=off build reproduces the
 * strict reference build.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/headers/Constants.h:25-66 */
#define AMU_TO_KG (1.66053904E-27)
#define BOLTZMANN_K (1.38064E-23)
#define ANG_TO_M (1.0E-10)
#define FS_TO_S (1.0E-15)
#define KCALMOLANGAMU_TO_ANGFS2 (1.0 / 48.88821291 / 48.88821291)
#define OMEGA_TO_FS_INV (2.45400505E8)
#define KCOUL (332.06348078020324)
#define AMUANGFS2_TO_KCAL_MOL (2390.07)
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct oracle_ctx {
  oracle_params p;
  int ncells;
  double corner[3];
  int *cell_count, *cell_first;
  int *nb1_start, *nb1; /* CSR first-shell lists */
  int *nb2_start, *nb2; /* CSR second-shell lists (long-range only) */
  /* probe geometry, src/GasBuffer.cpp:58-195 */
  int nsites;
  double site_z[3], site_q[3], site_m[3], gas_mass, d_bond, inertia;
};

/* ------------------------------------------------------------------ */
/* Linked cell: LinkedCell::calculateIndex, src/LinkedCell.cpp:117-129 */
static int cell_index(const oracle_ctx *c, const double pos[3]) {
  double xi = pos[0] - c->corner[0];
  int i = (int)floor(xi / c->p.cell);
  double yi = pos[1] - c->corner[1];
  int j = (int)floor(yi / c->p.cell);
  double zi = pos[2] - c->corner[2];
  int k = (int)floor(zi / c->p.cell);
  return k + c->p.nz * j + c->p.nz * c->p.ny * i;
}

typedef struct { int n; int (*o)[3]; } offs_t;
static void offs_push(offs_t *t, int i, int j, int k) {
  t->o = realloc(t->o, sizeof(int[3]) * (size_t)(t->n + 1));
  t->o[t->n][0] = i; t->o[t->n][1] = j; t->o[t->n][2] = k;
  t->n++;
}

static double sq(int v) { return (double)v * (double)v; }

/* Neighbour stencils, LinkedCell::calculateCellsNeighbors,
 * src/LinkedCell.cpp:228-440: offsets are classified in (i,j,k) loop order;
 * first shell = cells whose nearest point is closer than lj_cutoff, second
 * shell = the rest closer than coul_cutoff. */
static void build_stencils(oracle_ctx *c) {
  const double d = c->p.cell;
  const double R1 = c->p.lj_cutoff / d;
  const int next = c->p.long_range_flag == 1;
  const double R2 = next ? c->p.coul_cutoff / d : 0.0;
  const int n1 = (int)ceil(R1);
  const int nc = next ? (int)ceil(R2) : n1;
  offs_t s1 = {0, NULL}, s2 = {0, NULL};
  for (int i = -nc; i < nc + 1; i++)
    for (int j = -nc; j < nc + 1; j++)
      for (int k = -nc; k < nc + 1; k++) {
        int nz = (i != 0) + (j != 0) + (k != 0);
        if (nz == 0) {
          offs_push(&s1, i, j, k);
          continue;
        }
        if (nz == 1) {
          int a = abs(i + j + k);
          if (a <= n1) offs_push(&s1, i, j, k);
          else if (next) offs_push(&s2, i, j, k);
          continue;
        }
        /* plane (two non-zero) or general (three non-zero) */
        double near = 0.0, far = 0.0;
        if (i != 0) { near += sq(abs(i) - 1); far += sq(i); }
        if (j != 0) { near += sq(abs(j) - 1); far += sq(j); }
        if (k != 0) { near += sq(abs(k) - 1); far += sq(k); }
        if (next) {
          if (far < R2 * R2 || near < R2 * R2) {
            if (far < R1 * R1 || near < R1 * R1) offs_push(&s1, i, j, k);
            else offs_push(&s2, i, j, k);
          }
        } else {
          if (far < R1 * R1 || near < R1 * R1) offs_push(&s1, i, j, k);
        }
      }
  const int Nx = c->p.nx, Ny = c->p.ny, Nz = c->p.nz;
  c->nb1_start = calloc((size_t)c->ncells + 1, sizeof(int));
  c->nb2_start = calloc((size_t)c->ncells + 1, sizeof(int));
  size_t cap1 = (size_t)c->ncells * (size_t)s1.n + 1;
  size_t cap2 = (size_t)c->ncells * (size_t)s2.n + 1;
  c->nb1 = malloc(cap1 * sizeof(int));
  c->nb2 = malloc(cap2 * sizeof(int));
  int64_t w1 = 0, w2 = 0;
  for (int i = 0; i < Nx; i++)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) {
        int idx = k + Nz * j + Nz * Ny * i;
        c->nb1_start[idx] = (int)w1;
        for (int m = 0; m < s1.n; m++) {
          int a = i + s1.o[m][0], b = j + s1.o[m][1], e = k + s1.o[m][2];
          if (-1 < a && a < Nx && -1 < b && b < Ny && -1 < e && e < Nz) {
            int nb = e + Nz * b + Nz * Ny * a;
            if (c->cell_count[nb] != 0) c->nb1[w1++] = nb;
          }
        }
        c->nb2_start[idx] = (int)w2;
        if (next)
          for (int m = 0; m < s2.n; m++) {
            int a = i + s2.o[m][0], b = j + s2.o[m][1], e = k + s2.o[m][2];
            if (-1 < a && a < Nx && -1 < b && b < Ny && -1 < e && e < Nz) {
              int nb = e + Nz * b + Nz * Ny * a;
              if (c->cell_count[nb] != 0) c->nb2[w2++] = nb;
            }
          }
      }
  /* CSR end markers: cells are visited in idx order above */
  c->nb1_start[c->ncells] = (int)w1;
  c->nb2_start[c->ncells] = (int)w2;
  free(s1.o);
  free(s2.o);
}

void oracle_stencil_sums(const oracle_ctx *c, int64_t *s1, int64_t *s2) {
  *s1 = c->nb1_start[c->ncells];
  *s2 = c->nb2_start[c->ncells];
}

/* Probe constants, GasBuffer::gas_properties, src/GasBuffer.cpp:58-195 */
static void gas_init(oracle_ctx *c) {
  memset(c->site_z, 0, sizeof c->site_z);
  memset(c->site_q, 0, sizeof c->site_q);
  memset(c->site_m, 0, sizeof c->site_m);
  switch (c->p.gas) {
    case 2: /* N2: N, N, massless charged dummy at the centre */
      c->nsites = 3;
      c->site_z[0] = 0.5488; c->site_z[1] = -0.5488; c->site_z[2] = 0.0;
      c->site_q[0] = -0.4825; c->site_q[1] = -0.4825; c->site_q[2] = 0.965;
      c->site_m[0] = 14.007; c->site_m[1] = 14.007; c->site_m[2] = 0.0;
      c->gas_mass = c->site_m[0] + c->site_m[1];
      c->d_bond = fabs(c->site_z[0] - c->site_z[1]);
      break;
    case 3: /* CO2: O, C, O */
      c->nsites = 3;
      c->site_z[0] = -1.149; c->site_z[1] = 0.0; c->site_z[2] = 1.149;
      c->site_q[0] = -0.3256; c->site_q[1] = 0.6512; c->site_q[2] = -0.3256;
      c->site_m[0] = 15.999; c->site_m[1] = 12.011; c->site_m[2] = 15.999;
      c->gas_mass = c->site_m[0] + c->site_m[1] + c->site_m[2];
      c->d_bond = fabs(c->site_z[0] - c->site_z[2]);
      break;
    case 4: c->nsites = 1; c->site_m[0] = 39.948; c->gas_mass = 39.948; c->d_bond = 0.0; break;
    case 5: c->nsites = 1; c->site_m[0] = 44.0098; c->gas_mass = 44.0098; c->d_bond = 0.0; break;
    default: c->nsites = 1; c->site_m[0] = 4.0026; c->gas_mass = 4.0026; c->d_bond = 0.0; break;
  }
  /* src/System.cpp:109-114 */
  c->inertia = 0.0;
  for (int i = 0; i < c->nsites; i++) c->inertia += c->site_m[i] * pow(c->site_z[i], 2);
}

oracle_ctx *oracle_create(const oracle_params *p) {
  oracle_ctx *c = calloc(1, sizeof *c);
  c->p = *p;
  c->ncells = p->nx * p->ny * p->nz;
  /* src/LinkedCell.cpp:50-91: box rounded up to whole cells, corner = -box/2 */
  c->corner[0] = -0.5 * ((double)p->nx * p->cell);
  c->corner[1] = -0.5 * ((double)p->ny * p->cell);
  c->corner[2] = -0.5 * ((double)p->nz * p->cell);
  c->cell_count = calloc((size_t)c->ncells, sizeof(int));
  c->cell_first = calloc((size_t)c->ncells, sizeof(int));
  /* Atoms are already in cell order; recover per-cell counts by binning. */
  for (int a = 0; a < p->natoms; a++) {
    double pos[3] = {p->x[a], p->y[a], p->z[a]};
    int idx = cell_index(c, pos);
    c->cell_count[idx]++;
  }
  int run = 0;
  for (int i = 0; i < c->ncells; i++) {
    c->cell_first[i] = run;
    run += c->cell_count[i];
  }
  build_stencils(c);
  gas_init(c);
  return c;
}

void oracle_destroy(oracle_ctx *c) {
  if (!c) return;
  free(c->cell_count); free(c->cell_first);
  free(c->nb1_start); free(c->nb1); free(c->nb2_start); free(c->nb2);
  free(c);
}

/* ------------------------------------------------------------------ */
/* Force kernels, src/Force.cpp.  One generic body per family; each term
 * block keeps the reference's expression order, so the per-kernel sums are
 * those of the member functions named in oracle.h. */

typedef struct {
  double U, fx, fy, fz, Ex, Ey, Ez, Exx, Eyy, Ezz, Exy, Exz, Eyz;
} acc_t;

/* LJ block with cutoff shift, e.g. src/Force.cpp:84-103 */
static inline void lj_cut_term(const oracle_ctx *c, double eps, double sig,
                               double r2inv, double dx, double dy, double dz,
                               acc_t *a) {
  double r6inv = r2inv * r2inv * r2inv;
  double lj1 = 4.0 * eps * pow(sig, 6.0);
  double lj2 = lj1 * pow(sig, 6.0);
  double Ulj = r6inv * (lj2 * r6inv - lj1);
  double rc6inv = 1.0 / pow(c->p.lj_cutoff, 6);
  double Ulj_cut = rc6inv * (lj2 * rc6inv - lj1);
  double lj3 = 6.0 * lj1;
  double lj4 = 12.0 * lj2;
  double flj = r6inv * (lj4 * r6inv - lj3) * r2inv;
  a->U += Ulj - Ulj_cut;
  a->fx += flj * dx;
  a->fy += flj * dy;
  a->fz += flj * dz;
}

/* damped shifted Coulomb block, src/Force.cpp:671-683 */
static inline void coul_cut_term(const oracle_ctx *c, double qi, double qj,
                                 double r, double r2, double dx, double dy,
                                 double dz, acc_t *a) {
  const double rc = c->p.coul_cutoff;
  double rinv = 1.0 / r;
  double Ucoul = qi * qj * rinv * KCOUL;
  double Ucoul_shift = -1.5 * qi * qj * KCOUL / rc + 0.5 * qi * qj * KCOUL * r2 / pow(rc, 3);
  double r2inv = 1.0 / r2;
  double fcoul = Ucoul * r2inv * (1.0 - pow(r / rc, 3));
  a->U += Ucoul + Ucoul_shift;
  a->fx += fcoul * dx;
  a->fy += fcoul * dy;
  a->fz += fcoul * dz;
}

/* smoothed ion-induced-dipole field and gradient, src/Force.cpp:393-426 */
static inline void dip_cut_term(const oracle_ctx *c, double q, double r,
                                double r2, double dx, double dy, double dz,
                                acc_t *a) {
  double r2inv = 1.0 / r2;
  double r3inv = 1.0 / r * r2inv;
  double r5inv = r3inv * r2inv;
  double rc3inv = 1.0 / pow(c->p.coul_cutoff, 3);
  double smooth_factor = (1.0 - r * r2 * rc3inv);
  double qr3inv = q * r3inv * smooth_factor;
  double qr5inv = -3.0 * q * r5inv * smooth_factor;
  double qrc = -3.0 * q * rc3inv * r2inv;
  a->Ex += dx * qr3inv;
  a->Ey += dy * qr3inv;
  a->Ez += dz * qr3inv;
  a->Exx += qr3inv + dx * dx * qr5inv + dx * dx * qrc;
  a->Eyy += qr3inv + dy * dy * qr5inv + dy * dy * qrc;
  a->Ezz += qr3inv + dz * dz * qr5inv + dz * dz * qrc;
  a->Exy += dx * dy * qr5inv + dx * dy * qrc;
  a->Exz += dx * dz * qr5inv + dx * dz * qrc;
  a->Eyz += dy * dz * qr5inv + dy * dz * qrc;
}

/* epilogue F += alpha (E.grad)E, U -= alpha/2 |E|^2, src/Force.cpp:483-487 */
static inline void dip_epilogue(const oracle_ctx *c, const acc_t *a, double f[3],
                                double *U) {
  const double alpha = c->p.alpha;
  f[0] = a->fx + alpha * (a->Ex * a->Exx + a->Ey * a->Exy + a->Ez * a->Exz);
  f[1] = a->fy + alpha * (a->Ex * a->Exy + a->Ey * a->Eyy + a->Ez * a->Eyz);
  f[2] = a->fz + alpha * (a->Ex * a->Exz + a->Ey * a->Eyz + a->Ez * a->Ezz);
  *U = a->U - 0.5 * alpha * (a->Ex * a->Ex + a->Ey * a->Ey + a->Ez * a->Ez);
}

/* Linked-cell family: lennardjones_LC (:34), lennardjones_induced_dipole_LC
 * (:305), lennardjones_coulomb_LC (:598), coulomb_LC (:728),
 * coulomb_induced_dipole_iso_LC (:903) and the *_CO2 variants (:1088, :1224,
 * :1420). */
static void site_lc(const oracle_ctx *c, const double rp[3], double qi,
                    int use_lj, int central, int use_coul, int use_dip,
                    int second, double f[3], double *U) {
  acc_t a;
  memset(&a, 0, sizeof a);
  int index = cell_index(c, rp);
  /* Out of the grid: zero force (src/Force.cpp:58-66).  coulomb_LC's guard
   * (src/Force.cpp:753) can never fire and then reads out of bounds; that
   * undefined case is defined as zero here. */
  if (index >= c->ncells || index < 0) {
    f[0] = f[1] = f[2] = 0.0;
    *U = 0.0;
    return;
  }
  const oracle_params *p = &c->p;
  const double *E = central ? p->eps_c : p->eps;
  const double *S = central ? p->sig_c : p->sig;
  for (int pass = 0; pass < 2; pass++) {
    const int *lst, *st;
    if (pass == 0) { lst = c->nb1; st = c->nb1_start; }
    else {
      if (!second) break;
      lst = c->nb2; st = c->nb2_start;
    }
    for (int ci = st[index]; ci < st[index + 1]; ci++) {
      int cell = lst[ci];
      int b = c->cell_first[cell], e = b + c->cell_count[cell];
#ifdef ORACLE_SIMD
      /* Timing build only (liboracle_fast_*.so, the CPU baseline): the
       * per-cell atom loops vectorised as the reference's '#pragma omp simd'
       * loops are under its release flags (src/Force.cpp, e.g. :638-677 for
       * the first shell with LJ and :680-700 for the second shell without),
       * with the same terms; the checker build below keeps the sequential
       * order. */
      const double rlj = p->lj_cutoff, rc = p->coul_cutoff;
      const double rc6inv = 1.0 / pow(rlj, 6), rc3inv = 1.0 / pow(rc, 3);
      double sU = 0, sfx = 0, sfy = 0, sfz = 0, sEx = 0, sEy = 0, sEz = 0;
      double sExx = 0, sEyy = 0, sEzz = 0, sExy = 0, sExz = 0, sEyz = 0;
#define ORACLE_ELECTROSTATIC_TERMS                                                         \
  if (r < rc) {                                                                            \
    double qj = p->q[t];                                                                   \
    if (use_coul) {                                                                        \
      double Ucoul = qi * qj / r * KCOUL;                                                  \
      double fcoul = Ucoul * r2inv * (1.0 - pow(r / rc, 3));                               \
      sU += Ucoul - 1.5 * qi * qj * KCOUL / rc + 0.5 * qi * qj * KCOUL * r2 * rc3inv;      \
      sfx += fcoul * dx;                                                                   \
      sfy += fcoul * dy;                                                                   \
      sfz += fcoul * dz;                                                                   \
    }                                                                                      \
    if (use_dip) {                                                                         \
      double r3inv = 1.0 / r * r2inv, r5inv = r3inv * r2inv;                               \
      double sm = 1.0 - r * r2 * rc3inv;                                                   \
      double q3 = qj * r3inv * sm, q5 = -3.0 * qj * r5inv * sm, qr = -3.0 * qj * rc3inv * r2inv; \
      sEx += dx * q3;                                                                      \
      sEy += dy * q3;                                                                      \
      sEz += dz * q3;                                                                      \
      sExx += q3 + dx * dx * (q5 + qr);                                                    \
      sEyy += q3 + dy * dy * (q5 + qr);                                                    \
      sEzz += q3 + dz * dz * (q5 + qr);                                                    \
      sExy += dx * dy * (q5 + qr);                                                         \
      sExz += dx * dz * (q5 + qr);                                                         \
      sEyz += dy * dz * (q5 + qr);                                                         \
    }                                                                                      \
  }
      if (pass == 0 && use_lj) {
#pragma omp simd reduction(+ : sU, sfx, sfy, sfz, sEx, sEy, sEz, sExx, sEyy, sEzz, sExy, sExz, sEyz)
        for (int t = b; t < e; t++) {
          double dx = rp[0] - p->x[t], dy = rp[1] - p->y[t], dz = rp[2] - p->z[t];
          double r2 = dx * dx + dy * dy + dz * dz;
          double r = sqrt(r2);
          double r2inv = 1.0 / r2;
          if (r < rlj) {
            double r6inv = r2inv * r2inv * r2inv;
            double lj1 = 4.0 * E[t] * pow(S[t], 6.0);
            double lj2 = lj1 * pow(S[t], 6.0);
            double flj = r6inv * (12.0 * lj2 * r6inv - 6.0 * lj1) * r2inv;
            sU += r6inv * (lj2 * r6inv - lj1) - rc6inv * (lj2 * rc6inv - lj1);
            sfx += flj * dx;
            sfy += flj * dy;
            sfz += flj * dz;
          }
          ORACLE_ELECTROSTATIC_TERMS
        }
      } else {
#pragma omp simd reduction(+ : sU, sfx, sfy, sfz, sEx, sEy, sEz, sExx, sEyy, sEzz, sExy, sExz, sEyz)
        for (int t = b; t < e; t++) {
          double dx = rp[0] - p->x[t], dy = rp[1] - p->y[t], dz = rp[2] - p->z[t];
          double r2 = dx * dx + dy * dy + dz * dz;
          double r = sqrt(r2);
          double r2inv = 1.0 / r2;
          ORACLE_ELECTROSTATIC_TERMS
        }
      }
#undef ORACLE_ELECTROSTATIC_TERMS
      a.U += sU; a.fx += sfx; a.fy += sfy; a.fz += sfz;
      a.Ex += sEx; a.Ey += sEy; a.Ez += sEz;
      a.Exx += sExx; a.Eyy += sEyy; a.Ezz += sEzz; a.Exy += sExy; a.Exz += sExz; a.Eyz += sEyz;
#else
      for (int t = b; t < e; t++) {
        double dx = rp[0] - p->x[t];
        double dy = rp[1] - p->y[t];
        double dz = rp[2] - p->z[t];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        if (pass == 0 && use_lj && r < p->lj_cutoff)
          lj_cut_term(c, E[t], S[t], 1.0 / r2, dx, dy, dz, &a);
        if (r < p->coul_cutoff) {
          if (use_coul) coul_cut_term(c, qi, p->q[t], r, r2, dx, dy, dz, &a);
          if (use_dip) dip_cut_term(c, p->q[t], r, r2, dx, dy, dz, &a);
        }
      }
#endif
    }
  }
  if (use_dip) {
    dip_epilogue(c, &a, f, U);
  } else {
    f[0] = a.fx; f[1] = a.fy; f[2] = a.fz;
    *U = a.U;
  }
}

/* All-atom family: lennardjones (:117), lennardjones_coulomb (:176),
 * coulomb (:252), lennardjones_induced_dipole (:495),
 * coulomb_induced_dipole_iso (:836) and *_CO2 (:1169, :1352, :1637). */
static void site_all(const oracle_ctx *c, const double rp[3], double qi,
                     int use_lj, int central, int use_coul, int use_dip,
                     double f[3], double *U) {
  acc_t a;
  memset(&a, 0, sizeof a);
  const oracle_params *p = &c->p;
  const double *E = central ? p->eps_c : p->eps;
  const double *S = central ? p->sig_c : p->sig;
  for (int i = 0; i < p->natoms; i++) {
    double dx = rp[0] - p->x[i];
    double dy = rp[1] - p->y[i];
    double dz = rp[2] - p->z[i];
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double r2inv = 1.0 / r2;
    if (use_lj) {
      double r6inv = r2inv * r2inv * r2inv;
      double lj1 = 4.0 * E[i] * pow(S[i], 6.0);
      double lj2 = lj1 * pow(S[i], 6.0);
      a.U += r6inv * (lj2 * r6inv - lj1);
      double lj3 = 6.0 * lj1;
      double lj4 = 12.0 * lj2;
      double flj = r6inv * (lj4 * r6inv - lj3) * r2inv;
      a.fx += flj * dx;
      a.fy += flj * dy;
      a.fz += flj * dz;
    }
    if (use_coul) {
      double rinv = 1.0 / r;
      double Ucoul = qi * p->q[i] * rinv * KCOUL;
      double fcoul = Ucoul * r2inv;
      a.U += Ucoul;
      a.fx += fcoul * dx;
      a.fy += fcoul * dy;
      a.fz += fcoul * dz;
    }
    if (use_dip) {
      double q = p->q[i];
      double r3inv = 1.0 / r * r2inv;
      double r5inv = r3inv * r2inv;
      double qr3inv = q * r3inv;
      double qr5inv = q * r5inv;
      a.Ex += dx * qr3inv;
      a.Ey += dy * qr3inv;
      a.Ez += dz * qr3inv;
      a.Exx += qr3inv - 3.0 * dx * dx * qr5inv;
      a.Eyy += qr3inv - 3.0 * dy * dy * qr5inv;
      a.Ezz += qr3inv - 3.0 * dz * dz * qr5inv;
      a.Exy += -3.0 * dx * dy * qr5inv;
      a.Exz += -3.0 * dx * dz * qr5inv;
      a.Eyz += -3.0 * dy * dz * qr5inv;
    }
  }
  if (use_dip) {
    dip_epilogue(c, &a, f, U);
  } else {
    f[0] = a.fx; f[1] = a.fy; f[2] = a.fz;
    *U = a.U;
  }
}

void oracle_site_force(const oracle_ctx *c, int k, const double r[3], double qi,
                       double f[3], double *U) {
  switch (k) {
    case ORK_LJ: site_all(c, r, qi, 1, 0, 0, 0, f, U); break;
    case ORK_LJ_LC: site_lc(c, r, qi, 1, 0, 0, 0, 0, f, U); break;
    case ORK_LJ_COUL: site_all(c, r, qi, 1, 0, 1, 0, f, U); break;
    case ORK_LJ_COUL_LC: site_lc(c, r, qi, 1, 0, 1, 0, 1, f, U); break;
    case ORK_COUL: site_all(c, r, qi, 0, 0, 1, 0, f, U); break;
    case ORK_COUL_LC: site_lc(c, r, qi, 0, 0, 1, 0, 1, f, U); break;
    case ORK_LJ_DIP: site_all(c, r, qi, 1, 0, 0, 1, f, U); break;
    case ORK_LJ_DIP_LC: site_lc(c, r, qi, 1, 0, 0, 1, 1, f, U); break;
    case ORK_COUL_DIP: site_all(c, r, qi, 0, 0, 1, 1, f, U); break;
    case ORK_COUL_DIP_LC: site_lc(c, r, qi, 0, 0, 1, 1, 1, f, U); break;
    case ORK_LJ_CO2: site_all(c, r, qi, 1, 1, 0, 0, f, U); break;
    case ORK_LJ_LC_CO2: site_lc(c, r, qi, 1, 1, 0, 0, 0, f, U); break;
    case ORK_LJ_COUL_CO2: site_all(c, r, qi, 1, 1, 1, 0, f, U); break;
    case ORK_LJ_COUL_LC_CO2: site_lc(c, r, qi, 1, 1, 1, 0, 1, f, U); break;
    case ORK_LJ_COUL_DIP_CO2: site_all(c, r, qi, 1, 1, 1, 1, f, U); break;
    case ORK_LJ_COUL_DIP_LC_CO2: site_lc(c, r, qi, 1, 1, 1, 1, 1, f, U); break;
    default: f[0] = f[1] = f[2] = 0.0; *U = 0.0; break;
  }
}

/* Kernel choice per (gas, force type, site), src/System.cpp:371-379,
 * :593-650 (N2), :1059-1100 (CO2).  Returns -1 for "zero force". */
static int site_kernel(const oracle_ctx *c, int site) {
  const int ft = c->p.force_type;
  if (c->p.gas == 2) {
    if (site == 2) {
      switch (ft) {
        case 1: case 2: return -1;
        case 3: return ORK_COUL;
        case 4: return ORK_COUL_LC;
        case 5: return ORK_COUL_DIP;
        default: return ORK_COUL_DIP_LC;
      }
    }
    switch (ft) {
      case 1: return ORK_LJ;
      case 2: return ORK_LJ_LC;
      case 3: case 5: return ORK_LJ_COUL;
      default: return ORK_LJ_COUL_LC;
    }
  }
  if (c->p.gas == 3) {
    if (site == 1) {
      switch (ft) {
        case 1: return ORK_LJ_CO2;
        case 2: return ORK_LJ_LC_CO2;
        case 3: return ORK_LJ_COUL_CO2;
        case 4: return ORK_LJ_COUL_LC_CO2;
        case 5: return ORK_LJ_COUL_DIP_CO2;
        default: return ORK_LJ_COUL_DIP_LC_CO2;
      }
    }
    switch (ft) {
      case 1: return ORK_LJ;
      case 2: return ORK_LJ_LC;
      case 3: case 5: return ORK_LJ_COUL;
      default: return ORK_LJ_COUL_LC;
    }
  }
  switch (ft) { /* He, Ar, co2 */
    case 1: return ORK_LJ;
    case 2: return ORK_LJ_LC;
    case 3: return ORK_LJ_DIP;
    case 4: return ORK_LJ_DIP_LC;
    default: return -1;
  }
}

static void eval_site(const oracle_ctx *c, int site, const double r[3],
                      double f[3], double *U) {
  int k = site_kernel(c, site);
  if (k < 0) {
    f[0] = f[1] = f[2] = 0.0;
    *U = 0.0;
    return;
  }
  oracle_site_force(c, k, r, c->site_q[site], f, U);
}

/* ------------------------------------------------------------------ */
/* std::mt19937 (Matsumoto & Nishimura 1998), as used by
 * RandomNumber::getRandomNumber, src/RandomNumber.cpp:19-29. */
typedef struct { uint32_t s[624]; int i; } mt_t;
static void mt_seed(mt_t *m, uint32_t seed) {
  m->s[0] = seed;
  for (int i = 1; i < 624; i++)
    m->s[i] = 1812433253u * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
  m->i = 624;
}
static uint32_t mt_next(mt_t *m) {
  if (m->i >= 624) {
    for (int k = 0; k < 624; k++) {
      uint32_t y = (m->s[k] & 0x80000000u) | (m->s[(k + 1) % 624] & 0x7fffffffu);
      m->s[k] = m->s[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    m->i = 0;
  }
  uint32_t y = m->s[m->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
static double mt_uniform(mt_t *m) {
  return ((double)mt_next(m) - 0.0) / 4294967295.0;
}

void oracle_mt_raw(uint32_t seed, int64_t n, uint32_t *out) {
  mt_t m;
  mt_seed(&m, seed);
  for (int64_t i = 0; i < n; i++) out[i] = mt_next(&m);
}

/* Sample pre-generation, src/System.cpp:184-197 */
void oracle_samples(uint32_t seed, int multi, double bmax, int64_t n, double *u) {
  mt_t m;
  mt_seed(&m, seed);
  for (int64_t j = 0; j < n; j++) {
    double *o = u + 9 * j;
    o[0] = sqrt(bmax * bmax * mt_uniform(&m));
    for (int k = 1; k < 5; k++) o[k] = mt_uniform(&m);
    for (int k = 5; k < 9; k++) o[k] = multi ? mt_uniform(&m) : 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* System::velDistr / velGenerator, src/System.cpp:1523-1555 */
static double vel_distr(double v, double m, double T) {
  m = m * AMU_TO_KG;
  v = v * ANG_TO_M / FS_TO_S;
  return pow(v, 5.0) * pow((m / (2.0 * BOLTZMANN_K * T)), 3.0) *
         exp(-(m * pow(v, 2.0)) / (2.0 * BOLTZMANN_K * T));
}

double oracle_vel_generator(double m, double T, double sd) {
  double v = 0.0, dv = 1.0E-5;
  double sum = vel_distr(v, m, T);
  while (sum < sd) {
    v += dv;
    sum += vel_distr(v, m, T);
  }
  return v;
}

static inline double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  for (int i = 0; i < 3; i++) r += a[i] * b[i];
  return r;
}

/* System::KineticEnergy, src/System.cpp:1560-1564 */
static inline double kinetic(double m, const double v[3]) {
  return 0.5 * m * dot3(v, v) * AMUANGFS2_TO_KCAL_MOL;
}

/* System::anglevec, src/System.cpp:1566-1579 */
static double anglevec(const double vi[3], const double vf[3]) {
  double t1 = dot3(vi, vf);
  double t2 = sqrt(dot3(vi, vi));
  double t3 = sqrt(dot3(vf, vf));
  double arg = t1 / (t2 * t3);
  if (fabs(arg) > 1.0) return 0.0;
  return acos(arg);
}

/* System::rotate, src/System.cpp:1583-1619 (Rz(g) Ry(p) Rz(t)) */
static void rotate(double r[3], double v[3], const double ang[3]) {
  double ri[3], vi[3];
  memcpy(ri, r, sizeof ri); memcpy(vi, v, sizeof vi);
  r[0] = ri[0] * cos(ang[0]) - ri[1] * sin(ang[0]);
  r[1] = ri[0] * sin(ang[0]) + ri[1] * cos(ang[0]);
  r[2] = ri[2];
  v[0] = vi[0] * cos(ang[0]) - vi[1] * sin(ang[0]);
  v[1] = vi[0] * sin(ang[0]) + vi[1] * cos(ang[0]);
  v[2] = vi[2];
  memcpy(ri, r, sizeof ri); memcpy(vi, v, sizeof vi);
  r[0] = ri[0] * cos(ang[1]) + ri[2] * sin(ang[1]);
  r[1] = ri[1];
  r[2] = ri[0] * -sin(ang[1]) + ri[2] * cos(ang[1]);
  v[0] = vi[0] * cos(ang[1]) + vi[2] * sin(ang[1]);
  v[1] = vi[1];
  v[2] = vi[0] * -sin(ang[1]) + vi[2] * cos(ang[1]);
  memcpy(ri, r, sizeof ri); memcpy(vi, v, sizeof vi);
  r[0] = ri[0] * cos(ang[2]) - ri[1] * sin(ang[2]);
  r[1] = ri[0] * sin(ang[2]) + ri[1] * cos(ang[2]);
  r[2] = ri[2];
  v[0] = vi[0] * cos(ang[2]) - vi[1] * sin(ang[2]);
  v[1] = vi[0] * sin(ang[2]) + vi[1] * cos(ang[2]);
  v[2] = vi[2];
}

/* System::rotate_gas, src/System.cpp:1624-1638 (Ry(theta) then Rz(phi)) */
static void rotate_gas(double v[3], double theta, double phi) {
  double vi[3];
  memcpy(vi, v, sizeof vi);
  v[0] = vi[0] * cos(theta) + vi[2] * sin(theta);
  v[1] = vi[1];
  v[2] = vi[0] * -sin(theta) + vi[2] * cos(theta);
  memcpy(vi, v, sizeof vi);
  v[0] = vi[0] * cos(phi) - vi[1] * sin(phi);
  v[1] = vi[0] * sin(phi) + vi[1] * cos(phi);
  v[2] = vi[2];
}

static inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* System::setup, src/System.cpp:1397-1518 */
int oracle_setup(const oracle_ctx *c, const double u[9], double pos[9],
                 double vel[9], double rcm[3], double vcm[3]) {
  const oracle_params *p = &c->p;
  rcm[0] = u[0];
  rcm[1] = 0.0;
  rcm[2] = p->bmax;
  double vi = oracle_vel_generator(p->mu, p->temperature, u[1]);
  vcm[0] = 0.0;
  vcm[1] = 0.0;
  vcm[2] = -vi;
  double ang[3];
  ang[0] = 2.0 * M_PI * u[2];
  ang[1] = asin(2.0 * u[3] - 1.0) + 0.5 * M_PI;
  ang[2] = 2.0 * M_PI * u[4];
  rotate(rcm, vcm, ang);
  double ur[3] = {rcm[0] / p->a, rcm[1] / p->b, rcm[2] / p->c};
  double uv[3] = {vcm[0] / p->a, vcm[1] / p->b, vcm[2] / p->c};
  double cr[3];
  cross3(ur, uv, cr);
  double delta = -dot3(cr, cr) + dot3(uv, uv);
  if (delta < 0) return 0;
  double lambda = -dot3(ur, uv) / dot3(uv, uv) - sqrt(delta) / dot3(uv, uv);
  double d[3] = {lambda * vcm[0], lambda * vcm[1], lambda * vcm[2]};
  rcm[0] = rcm[0] + d[0];
  rcm[1] = rcm[1] + d[1];
  rcm[2] = rcm[2] + d[2];
  memset(pos, 0, 9 * sizeof(double));
  memset(vel, 0, 9 * sizeof(double));
  if (c->nsites == 1) {
    for (int k = 0; k < 3; k++) { pos[k] = rcm[k]; vel[k] = vcm[k]; }
    return 1;
  }
  double theta = asin(2.0 * u[5] - 1.0) + 0.5 * M_PI;
  double phi = 2.0 * M_PI * u[6];
  double psi = 2.0 * M_PI * u[7];
  double omega = sqrt(2.0 * BOLTZMANN_K * p->temperature / c->inertia *
                      log(1. / (1. - u[8]))) * OMEGA_TO_FS_INV;
  double w[3] = {omega * cos(psi), omega * sin(psi), 0.0};
  rotate_gas(w, theta, phi);
  for (int i = 0; i < c->nsites; i++) {
    double rg[3] = {0.0, 0.0, c->site_z[i]};
    rotate_gas(rg, theta, phi);
    double vg[3];
    cross3(w, rg, vg);
    for (int k = 0; k < 3; k++) { pos[3 * i + k] = rg[k]; vel[3 * i + k] = vg[k]; }
  }
  for (int i = 0; i < c->nsites; i++)
    for (int k = 0; k < 3; k++) {
      pos[3 * i + k] += rcm[k];
      vel[3 * i + k] += vcm[k];
    }
  return 1;
}

/* ------------------------------------------------------------------ */
/* RATTLE, src/System.cpp:1700-1739 */
void oracle_rattle_first(double ri[3], double rj[3], double vi[3], double vj[3],
                         const double fi[3], const double fj[3], double mi,
                         double mj, double dt, double d2ij) {
  double muij = mi * mj / (mi + mj);
  double rci[3], rcj[3], vci[3], vcj[3], rcij[3], rij[3];
  for (int i = 0; i < 3; i++) {
    vci[i] = vi[i] + 0.5 * dt / mi * fi[i] * KCALMOLANGAMU_TO_ANGFS2;
    vcj[i] = vj[i] + 0.5 * dt / mj * fj[i] * KCALMOLANGAMU_TO_ANGFS2;
    rci[i] = ri[i] + dt * vci[i];
    rcj[i] = rj[i] + dt * vcj[i];
    rij[i] = ri[i] - rj[i];
    rcij[i] = rci[i] - rcj[i];
  }
  double r2ij = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  double rc2ij = rcij[0] * rcij[0] + rcij[1] * rcij[1] + rcij[2] * rcij[2];
  double rcijrij = rcij[0] * rij[0] + rcij[1] * rij[1] + rcij[2] * rij[2];
  double determ = rcijrij * rcijrij - r2ij * (rc2ij - d2ij);
  if (determ < 0.0) determ = 0.0;
  double lambda = (rcijrij - sqrt(determ)) / r2ij;
  for (int i = 0; i < 3; i++) {
    vi[i] = vci[i] - muij / (mi * dt) * lambda * rij[i];
    vj[i] = vcj[i] + muij / (mj * dt) * lambda * rij[i];
    ri[i] = rci[i] - muij / mi * lambda * rij[i];
    rj[i] = rcj[i] + muij / mi * lambda * rij[i]; /* reference uses mi here */
  }
}

/* RATTLE, src/System.cpp:1741-1767 */
void oracle_rattle_second(const double ri[3], const double rj[3], double vi[3],
                          double vj[3], const double fi[3], const double fj[3],
                          double mi, double mj, double dt) {
  double muij = mi * mj / (mi + mj);
  double rij[3], vcij[3], vci[3], vcj[3];
  for (int i = 0; i < 3; i++) {
    vci[i] = vi[i] + 0.5 * dt / mi * fi[i] * KCALMOLANGAMU_TO_ANGFS2;
    vcj[i] = vj[i] + 0.5 * dt / mj * fj[i] * KCALMOLANGAMU_TO_ANGFS2;
    rij[i] = ri[i] - rj[i];
    vcij[i] = vci[i] - vcj[i];
  }
  double rijvcij = rij[0] * vcij[0] + rij[1] * vcij[1] + rij[2] * vcij[2];
  double r2ij = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
  double lambda = muij * rijvcij / r2ij;
  for (int i = 0; i < 3; i++) {
    vi[i] = vci[i] - lambda / mi * rij[i];
    vj[i] = vcj[i] + lambda / mj * rij[i];
  }
}

/* ------------------------------------------------------------------ */
static inline int outside_box(const oracle_params *p, const double r[3]) {
  return fabs(r[0]) >= p->half_box[0] || fabs(r[1]) >= p->half_box[1] ||
         fabs(r[2]) >= p->half_box[2];
}

static inline double ellipsoid_r(const oracle_params *p, const double r[3]) {
  double ur[3] = {r[0] / p->a, r[1] / p->b, r[2] / p->c};
  return dot3(ur, ur);
}

/* The reference loops forever for a probe that never leaves (zero speed or a
 * non-finite state); here such a trajectory is reported lost with tries = -1
 * once one try has run this many steps, as the engine does.  A positive
 * max_total_steps also ends a trajectory whose steps over all its tries reach
 * it (the same report). */
#define ORACLE_MAX_STEPS 4000000
#define CAPPED(sc, r, v)                                                       \
  ((sc) >= ORACLE_MAX_STEPS ||                                                  \
   (c->p.max_total_steps > 0 && o->steps >= c->p.max_total_steps) ||           \
   !isfinite((r)[0] + (r)[1] + (r)[2] + (v)[0] + (v)[1] + (v)[2]))

/* System::run_He, src/System.cpp:338-511 (also Ar and 1-site co2) */
static void run_atomic(const oracle_ctx *c, const double pos0[3],
                       const double vel0[3], oracle_record *o) {
  const oracle_params *p = &c->p;
  const double time_step = p->dt, mu = p->mu;
  double dt = time_step;
  int maxTries = 8, trajTries = 0, stepcount = 0;
  const double theta_max = 1.5 * M_PI;
  double dtheta = 0.0, Up = 0.0;
  double rcm[3], vcm[3], rcm_i[3], vcm_i[3], rcm_old[3], f[3], fi[3];
  memcpy(rcm, pos0, sizeof rcm);
  memcpy(vcm, vel0, sizeof vcm);
  memcpy(rcm_i, rcm, sizeof rcm);
  memcpy(vcm_i, vcm, sizeof vcm);
  memcpy(rcm_old, rcm, sizeof rcm);
  eval_site(c, 0, rcm, f, &Up);
  memcpy(fi, f, sizeof f);
  double Ek = kinetic(mu, vcm);
  double Ei = Up + Ek;
  o->steps = 0;
  while (trajTries < maxTries && maxTries < 10) {
    if (stepcount == 0) {
      memcpy(vcm, vcm_i, sizeof vcm);
      memcpy(rcm, rcm_i, sizeof rcm);
      memcpy(rcm_old, rcm_i, sizeof rcm);
      memcpy(f, fi, sizeof f);
      dtheta = 0.0;
    }
    for (int i = 0; i < 3; i++) {
      vcm[i] += 0.5 * dt / mu * f[i] * KCALMOLANGAMU_TO_ANGFS2;
      rcm[i] += dt * vcm[i];
    }
    dtheta += anglevec(rcm, rcm_old);
    memcpy(rcm_old, rcm, sizeof rcm);
    eval_site(c, 0, rcm, f, &Up);
    for (int i = 0; i < 3; i++) vcm[i] += 0.5 * dt / mu * f[i] * KCALMOLANGAMU_TO_ANGFS2;
    stepcount++;
    o->steps++;
    if (CAPPED(stepcount, rcm, vcm)) {
      o->success = 0;
      o->tries = -1;
      return;
    }
    if (dtheta > theta_max) {
      trajTries++;
      dt = time_step / (trajTries + 1);
      stepcount = 0;
      continue;
    }
    if (outside_box(p, rcm) || ellipsoid_r(p, rcm) > 1.0) {
      Ek = kinetic(mu, vcm);
      double E = Up + Ek;
      double dE = fabs(E - Ei) / Ei * 100.0;
      if (dE > 0.5) {
        trajTries++;
        maxTries++;
        dt = time_step / (trajTries + 1);
        stepcount = 0;
        continue;
      }
      o->chi = anglevec(vcm_i, vcm);
      o->success = 1;
      o->tries = trajTries;
      return;
    }
  }
  o->success = 0;
  o->tries = trajTries;
}

/* System::run_N2, src/System.cpp:514-945 */
static void run_n2(const oracle_ctx *c, const double pos0[9], const double vel0[9],
                   const double rcm0[3], const double vcm0[3], oracle_record *o) {
  const oracle_params *p = &c->p;
  const double time_step = p->dt;
  double dt = time_step;
  int maxTries = 8, trajTries = 0, stepcount = 0;
  const double theta_max = 1.5 * M_PI;
  double dtheta = 0.0;
  const int n = 3;
  double x[9], v[9], xi[9], vxi[9], fg[9], fig[9];
  memcpy(x, pos0, sizeof x);
  memcpy(v, vel0, sizeof v);
  memcpy(xi, x, sizeof x);
  memcpy(vxi, v, sizeof v);
  const double *m = c->site_m;
  double rcm[3], vcm[3], rcm_i[3], vcm_i[3], rcm_old[3];
  memcpy(rcm, rcm0, sizeof rcm);
  memcpy(vcm, vcm0, sizeof vcm);
  memcpy(rcm_i, rcm, sizeof rcm);
  memcpy(vcm_i, vcm, sizeof vcm);
  memcpy(rcm_old, rcm, sizeof rcm);
  double Up_gas = 0.0, Up;
  for (int s = 0; s < n; s++) {
    eval_site(c, s, &x[3 * s], &fg[3 * s], &Up);
    Up_gas += Up;
  }
  memcpy(fig, fg, sizeof fg);
  double Ek = 0.0;
  for (int s = 0; s < n; s++) Ek += kinetic(m[s], &v[3 * s]);
  double Ei = Up_gas + Ek;
  const double d2ij = c->d_bond * c->d_bond;
  o->steps = 0;
  while (trajTries < maxTries && maxTries < 10) {
    if (stepcount == 0) {
      memcpy(vcm, vcm_i, sizeof vcm);
      memcpy(rcm, rcm_i, sizeof rcm);
      memcpy(rcm_old, rcm_i, sizeof rcm);
      memcpy(x, xi, sizeof x);
      memcpy(v, vxi, sizeof v);
      memcpy(fg, fig, sizeof fg);
      dtheta = 0.0;
    }
    double ri[3], rj[3], vi[3], vj[3], fi[3], fj[3];
    for (int k = 0; k < 3; k++) {
      ri[k] = x[k]; vi[k] = v[k];
      rj[k] = x[3 + k]; vj[k] = v[3 + k];
      fi[k] = fg[k] + 0.5 * fg[6 + k];
      fj[k] = fg[3 + k] + 0.5 * fg[6 + k];
    }
    oracle_rattle_first(ri, rj, vi, vj, fi, fj, m[0], m[1], dt, d2ij);
    for (int k = 0; k < 3; k++) {
      x[k] = ri[k]; v[k] = vi[k];
      x[3 + k] = rj[k]; v[3 + k] = vj[k];
      rcm[k] = 0.5 * (ri[k] + rj[k]);
      vcm[k] = 0.5 * (vi[k] + vj[k]);
      x[6 + k] = rcm[k];
    }
    dtheta += anglevec(rcm, rcm_old);
    memcpy(rcm_old, rcm, sizeof rcm);
    Up_gas = 0.0;
    for (int s = 0; s < n; s++) {
      eval_site(c, s, &x[3 * s], &fg[3 * s], &Up);
      Up_gas += Up;
    }
    for (int k = 0; k < 3; k++) {
      ri[k] = x[k]; vi[k] = v[k];
      rj[k] = x[3 + k]; vj[k] = v[3 + k];
      fi[k] = fg[k] + 0.5 * fg[6 + k];
      fj[k] = fg[3 + k] + 0.5 * fg[6 + k];
    }
    oracle_rattle_second(ri, rj, vi, vj, fi, fj, m[0], m[1], dt);
    for (int k = 0; k < 3; k++) {
      v[k] = vi[k];
      v[3 + k] = vj[k];
      vcm[k] = 0.5 * (vi[k] + vj[k]);
    }
    stepcount++;
    o->steps++;
    if (CAPPED(stepcount, rcm, vcm)) {
      o->success = 0;
      o->tries = -1;
      return;
    }
    if (dtheta > theta_max) {
      trajTries++;
      dt = time_step / (trajTries + 1);
      stepcount = 0;
      continue;
    }
    if (outside_box(p, rcm) || ellipsoid_r(p, rcm) > 1.0) {
      Ek = 0.0;
      for (int s = 0; s < n - 1; s++) Ek += kinetic(m[s], &v[3 * s]);
      double E = Up_gas + Ek;
      double dE = fabs((E - Ei) / Ei) * 100.0;
      if (dE > 1.0) {
        trajTries++;
        maxTries++;
        dt = time_step / (trajTries + 1);
        stepcount = 0;
        continue;
      }
      o->chi = anglevec(vcm_i, vcm);
      o->success = 1;
      o->tries = trajTries;
      return;
    }
  }
  o->success = 0;
  o->tries = trajTries;
}

/* System::run_CO2, src/System.cpp:948-1392, including the reference's
 * initial-velocity typo vcm[0] = gasProbe->vcm[1] (:1012) unless
 * co2_vcm_fix is set. */
static void run_co2(const oracle_ctx *c, const double pos0[9], const double vel0[9],
                    const double rcm0[3], const double vcm0[3], oracle_record *o) {
  const oracle_params *p = &c->p;
  const double time_step = p->dt;
  double dt = time_step;
  int maxTries = 8, trajTries = 0, stepcount = 0;
  const double theta_max = 1.5 * M_PI;
  double dtheta = 0.0;
  const int n = 3;
  double x[9], v[9], xi[9], vxi[9], fg[9], fig[9];
  memcpy(x, pos0, sizeof x);
  memcpy(v, vel0, sizeof v);
  memcpy(xi, x, sizeof x);
  memcpy(vxi, v, sizeof v);
  const double *m = c->site_m;
  const double M = c->gas_mass;
  (void)p;
  double rcm[3], vcm[3], rcm_i[3], vcm_i[3], rcm_old[3];
  memcpy(rcm, rcm0, sizeof rcm);
  vcm[0] = c->p.co2_vcm_fix ? vcm0[0] : vcm0[1];
  vcm[1] = vcm0[1];
  vcm[2] = vcm0[2];
  memcpy(rcm_i, rcm, sizeof rcm);
  memcpy(vcm_i, vcm, sizeof vcm);
  memcpy(rcm_old, rcm, sizeof rcm);
  double Up_gas = 0.0, Up;
  for (int s = 0; s < n; s++) {
    eval_site(c, s, &x[3 * s], &fg[3 * s], &Up);
    Up_gas += Up;
  }
  memcpy(fig, fg, sizeof fg);
  double Ek = 0.0;
  for (int s = 0; s < n; s++) Ek += kinetic(m[s], &v[3 * s]);
  double Ei = Up_gas + Ek;
  const double d2ij = c->d_bond * c->d_bond;
  o->steps = 0;
  while (trajTries < maxTries && maxTries < 10) {
    if (stepcount == 0) {
      memcpy(vcm, vcm_i, sizeof vcm);
      memcpy(rcm, rcm_i, sizeof rcm);
      memcpy(rcm_old, rcm_i, sizeof rcm);
      memcpy(x, xi, sizeof x);
      memcpy(v, vxi, sizeof v);
      memcpy(fg, fig, sizeof fg);
      dtheta = 0.0;
    }
    double ri[3], rj[3], vi[3], vj[3], fi[3], fj[3];
    double m2M = m[1] / M;
    for (int k = 0; k < 3; k++) {
      fi[k] = (1.0 - 0.5 * m2M) * fg[k] + m[0] / M * fg[3 + k] - 0.5 * m2M * fg[6 + k];
      fj[k] = (1.0 - 0.5 * m2M) * fg[6 + k] + m[0] / M * fg[3 + k] - 0.5 * m2M * fg[k];
      ri[k] = x[k]; vi[k] = v[k];
      rj[k] = x[6 + k]; vj[k] = v[6 + k];
    }
    oracle_rattle_first(ri, rj, vi, vj, fi, fj, m[0], m[2], dt, d2ij);
    for (int k = 0; k < 3; k++) {
      x[k] = ri[k]; v[k] = vi[k];
      x[6 + k] = rj[k]; v[6 + k] = vj[k];
      rcm[k] = 0.5 * (ri[k] + rj[k]);
      vcm[k] = 0.5 * (vi[k] + vj[k]);
      x[3 + k] = rcm[k];
      v[3 + k] = vcm[k];
    }
    dtheta += anglevec(rcm, rcm_old);
    memcpy(rcm_old, rcm, sizeof rcm);
    Up_gas = 0.0;
    for (int s = 0; s < n; s++) {
      eval_site(c, s, &x[3 * s], &fg[3 * s], &Up);
      Up_gas += Up;
    }
    m2M = m[1] / M;
    for (int k = 0; k < 3; k++) {
      fi[k] = (1.0 - 0.5 * m2M) * fg[k] + m[0] / M * fg[3 + k] - 0.5 * m2M * fg[6 + k];
      fj[k] = (1.0 - 0.5 * m2M) * fg[6 + k] + m[0] / M * fg[3 + k] - 0.5 * m2M * fg[k];
      ri[k] = x[k]; vi[k] = v[k];
      rj[k] = x[6 + k]; vj[k] = v[6 + k];
    }
    oracle_rattle_second(ri, rj, vi, vj, fi, fj, m[0], m[2], dt);
    for (int k = 0; k < 3; k++) {
      v[k] = vi[k];
      v[6 + k] = vj[k];
      vcm[k] = 0.5 * (vi[k] + vj[k]);
      v[3 + k] = vcm[k];
    }
    stepcount++;
    o->steps++;
    if (CAPPED(stepcount, rcm, vcm)) {
      o->success = 0;
      o->tries = -1;
      return;
    }
    if (dtheta > theta_max) {
      trajTries++;
      dt = time_step / (trajTries + 1);
      stepcount = 0;
      continue;
    }
    if (outside_box(p, rcm) || ellipsoid_r(p, rcm) > 1.0) {
      Ek = 0.0;
      for (int s = 0; s < n; s++) Ek += kinetic(m[s], &v[3 * s]);
      double E = Up_gas + Ek;
      double dE = fabs((E - Ei) / Ei) * 100.0;
      if (dE > 20.0) {
        trajTries++;
        maxTries++;
        dt = time_step / (trajTries + 1);
        stepcount = 0;
        continue;
      }
      o->chi = anglevec(vcm_i, vcm);
      o->success = 1;
      o->tries = trajTries;
      return;
    }
  }
  o->success = 0;
  o->tries = trajTries;
}

/* One trajectory: the body of the OpenMP loop, src/System.cpp:206-230 */
static void run_one(const oracle_ctx *c, const double u[9], oracle_record *o) {
  memset(o, 0, sizeof *o);
  double pos[9], vel[9], rcm[3], vcm[3];
  o->hit = oracle_setup(c, u, pos, vel, rcm, vcm);
  if (!o->hit) return;
  if (c->p.gas == 2) run_n2(c, pos, vel, rcm, vcm, o);
  else if (c->p.gas == 3) run_co2(c, pos, vel, rcm, vcm, o);
  else run_atomic(c, pos, vel, o);
  if (o->success) o->domega = M_PI * (1.0 - cos(o->chi)) * pow(c->p.bmax, 2.0);
}

void oracle_run(const oracle_ctx *c, const double *u, int64_t n,
                oracle_record *out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int64_t j = 0; j < n; j++) run_one(c, u + 9 * j, out + j);
}

/* Tally, src/System.cpp:287-317 */
void oracle_tally(const oracle_record *rec, int64_t n_iter, int64_t n_traj,
                  double *omega_iter, int64_t *counts, double *ccs_mean,
                  double *ccs_err) {
  double Omega = 0.0, Omega2 = 0.0;
  int64_t count = 0;
  for (int64_t i = 0; i < n_iter; i++) {
    double dOmega = 0.0;
    int Nscatter = 0, Nfree = 0, Nlost = 0;
    for (int64_t j = 0; j < n_traj; j++) {
      const oracle_record *r = &rec[count++];
      dOmega += r->domega;
      if (!r->hit) Nfree++;
      else if (r->success) Nscatter++;
      else Nlost++;
    }
    double w = 1.0 / ((double)(float)(Nscatter + Nfree)) * dOmega;
    Omega += w;
    Omega2 += pow(w, 2.0);
    if (omega_iter) omega_iter[i] = w;
    if (counts) {
      counts[3 * i + 0] = Nfree;
      counts[3 * i + 1] = Nscatter;
      counts[3 * i + 2] = Nlost;
    }
  }
  double ave = Omega / n_iter;
  double sig2 = Omega2 / n_iter - pow(ave, 2.0);
  *ccs_mean = ave;
  *ccs_err = sqrt(sig2 / (double)(float)n_iter);
}
