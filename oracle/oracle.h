/*
 * CPU oracle for the MassCCS trajectory-method hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (cces-cepid/massccs, src/System.cpp, src/Force.cpp,
 * src/LinkedCell.cpp, src/GasBuffer.cpp, src/RandomNumber.cpp,
 * src/headers/Constants.h), written to follow the reference's operation order
 * so that, compiled with -O2 -fno-fast-math -ffp-contract=off, it reproduces
 * the strict reference build bit for bit.  It is used only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker; the product never links or calls it.
 *
 * Pinning (tests/test_oracle.py):
 *   - per-point site forces of every kernel against the reference harness
 *     (oracle/_ref/ref_harness force) and the committed fixtures;
 *   - per-trajectory (hit, success, chi) against the reference harness;
 *   - per-iteration Omega against the reference's published logs;
 *   - closed forms: F = -grad U by finite differences, dipole gradient tensor
 *     by finite differences of the field, brute-force all-pairs sums, RATTLE
 *     invariants, the Maxwell-Boltzmann table's normalisation.
 * Kernels whose reference reads uninitialised accumulators
 * (src/Force.cpp:517-522, :836-853) are defined here with zero-initialised
 * accumulators: parity unpinned for those (no configuration uses them).
 */
#ifndef MASSCCS_ORACLE_H
#define MASSCCS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Prepared target: atoms already centred, rotated and sorted in cell order
 * (src/LinkedCell.cpp:132-225), plus the grid and run parameters that the
 * reference System holds. */
typedef struct {
  int natoms;
  const double *x, *y, *z, *q, *eps, *sig;
  const double *eps_c, *sig_c; /* CO2 carbon-site parameters, may be NULL */
  int nx, ny, nz;              /* cells per axis */
  double cell;                 /* lj_cutoff + skin */
  double lj_cutoff, coul_cutoff;
  double alpha;                /* already multiplied by ALPHA_TO_KCAL_MOL */
  int long_range_flag;         /* second-shell stencil present */
  int gas;                     /* 1 He, 2 N2, 3 CO2, 4 Ar, 5 co2 */
  int force_type;              /* 1..6, src/System.cpp:49-87 */
  double mu;                   /* reduced mass (He) or gas mass (N2/CO2) */
  double temperature, dt;
  double a, b, c, bmax;        /* ellipsoid and maximal impact parameter */
  double half_box[3];          /* 0.5 * linked-cell box */
  int co2_vcm_fix;             /* CO2: 1 measures chi from the true initial velocity
                                  (0 = the reference, src/System.cpp:1012) */
  int max_total_steps;         /* 0: no limit; else a trajectory whose steps over all
                                  its tries reach this is reported lost with tries = -1
                                  (a budget for classification runs, not a reference
                                  behaviour) */
} oracle_params;

typedef struct oracle_ctx oracle_ctx;

typedef struct {
  int32_t hit, success, tries, steps; /* tries = trajTries at exit; steps =
                                         integration steps summed over tries */
  double chi;
  double domega;                      /* pi (1 - cos chi) bmax^2 if scattered */
} oracle_record;

oracle_ctx *oracle_create(const oracle_params *p);
void oracle_destroy(oracle_ctx *ctx);

/* Stencil statistics: total first/second-shell list lengths over all cells. */
void oracle_stencil_sums(const oracle_ctx *ctx, int64_t *s1, int64_t *s2);

/* Kernel ids (one per reference Force member). */
enum {
  ORK_LJ = 0,            /* Force::lennardjones            Force.cpp:117 */
  ORK_LJ_LC,             /* Force::lennardjones_LC         Force.cpp:34  */
  ORK_LJ_COUL,           /* Force::lennardjones_coulomb    Force.cpp:176 */
  ORK_LJ_COUL_LC,        /* Force::lennardjones_coulomb_LC Force.cpp:598 */
  ORK_COUL,              /* Force::coulomb                 Force.cpp:252 */
  ORK_COUL_LC,           /* Force::coulomb_LC              Force.cpp:728 */
  ORK_LJ_DIP,            /* Force::lennardjones_induced_dipole    :495 */
  ORK_LJ_DIP_LC,         /* Force::lennardjones_induced_dipole_LC :305 */
  ORK_COUL_DIP,          /* Force::coulomb_induced_dipole_iso     :836 */
  ORK_COUL_DIP_LC,       /* Force::coulomb_induced_dipole_iso_LC  :903 */
  ORK_LJ_CO2,            /* Force::lennardjones_CO2              :1169 */
  ORK_LJ_LC_CO2,         /* Force::lennardjones_LC_CO2           :1088 */
  ORK_LJ_COUL_CO2,       /* Force::lennardjones_coulomb_CO2      :1352 */
  ORK_LJ_COUL_LC_CO2,    /* Force::lennardjones_coulomb_LC_CO2   :1224 */
  ORK_LJ_COUL_DIP_CO2,   /* ..._induced_dipole_iso_CO2           :1637 */
  ORK_LJ_COUL_DIP_LC_CO2,/* ..._induced_dipole_iso_LC_CO2        :1420 */
  ORK_COUNT
};

/* One site evaluation: probe site at r with charge qi. */
void oracle_site_force(const oracle_ctx *ctx, int kernel, const double r[3],
                       double qi, double f[3], double *U);

/* mt19937 sample stream, src/System.cpp:184-197: for trajectory j, writes
 * u[9j+0] = sqrt(bmax^2 u1), u[9j+1..4] = u2..u5 and, for N2/CO2,
 * u[9j+5..8] = u6..u9 (zero otherwise). */
void oracle_samples(uint32_t seed, int multi_site, double bmax, int64_t n,
                    double *u);

/* Raw mt19937 outputs (for the standard's known-answer check). */
void oracle_mt_raw(uint32_t seed, int64_t n, uint32_t *out);

/* Maxwell-Boltzmann speed by the reference's sequential scan
 * (System::velGenerator, src/System.cpp:1537-1555). */
double oracle_vel_generator(double m, double temperature, double u);

/* Runs trajectories [0, n) over samples u (9 doubles each); OpenMP over
 * trajectories with the reference's schedule(dynamic). */
void oracle_run(const oracle_ctx *ctx, const double *u, int64_t n,
                oracle_record *out, int nthreads);

/* Initial state of one trajectory as System::setup leaves it
 * (src/System.cpp:1397-1518): hit flag, site positions and velocities (3 sites
 * x 3), centre of mass position and velocity. */
int oracle_setup(const oracle_ctx *ctx, const double u[9], double pos[9],
                 double vel[9], double rcm[3], double vcm[3]);

/* Tally, src/System.cpp:287-317.  omega_iter receives nIter values. */
void oracle_tally(const oracle_record *rec, int64_t n_iter, int64_t n_traj,
                  double *omega_iter, int64_t *counts /* 3*nIter: free,
                  scatter, lost */, double *ccs_mean, double *ccs_err);

/* RATTLE stages, src/System.cpp:1700-1767 (exposed for invariant tests). */
void oracle_rattle_first(double ri[3], double rj[3], double vi[3], double vj[3],
                         const double fi[3], const double fj[3], double mi,
                         double mj, double dt, double d2ij);
void oracle_rattle_second(const double ri[3], const double rj[3], double vi[3],
                          double vj[3], const double fi[3], const double fj[3],
                          double mi, double mj, double dt);

#ifdef __cplusplus
}
#endif
#endif
