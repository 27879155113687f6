/*
 * massccs_amd -- C ABI of the MI355X trajectory-method engine.
 *
 * The engine is a drop-in for the hot path of MassCCS (cces-cepid/massccs):
 * the Monte-Carlo trajectory loop that the reference runs inside
 * System::System (src/System.cpp:160-317), its per-trajectory integrators
 * System::run_He / run_N2 / run_CO2 (src/headers/System.h:92-94,
 * src/System.cpp:338-1392), their setup (src/System.cpp:1397-1518) and the
 * per-site force kernels Force::<kernel> (src/headers/Force.h:42-62).
 * The host-side preparation that feeds them (src/Input.cpp, src/MoleculeTarget.cpp,
 * src/GasBuffer.cpp, src/LinkedCell.cpp, System::geometric_ellipsoid) is
 * restated here so that the same input.json / .xyz / .pqr / .ff files drive it.
 *
 * Conventions: plain pointers and sizes, no C++ or torch types.  Every call
 * returns an mccs_status; on failure mccs_last_error() (thread-local) holds a
 * message.  Units are the reference's: Angstrom, fs, amu, kcal/mol, e.
 * Handles are owned by the caller and released with the matching *_destroy.
 */
#ifndef MASSCCS_H
#define MASSCCS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MCCS_OK = 0,
  MCCS_ERR_INVALID = 1,     /* bad argument or inconsistent input */
  MCCS_ERR_IO = 2,          /* file missing or unreadable */
  MCCS_ERR_PARSE = 3,       /* malformed json/xyz/pqr/ff */
  MCCS_ERR_UNSUPPORTED = 4, /* option outside what the engine implements */
  MCCS_ERR_DEVICE = 5,      /* HIP runtime error or no usable GPU */
  MCCS_ERR_NOMEM = 6
} mccs_status;

/* Message of the last failed call on this thread. */
const char *mccs_last_error(void);

/* Probe gases, src/Input.cpp:98-117 ("GasBuffer"). */
enum { MCCS_GAS_HE = 1, MCCS_GAS_N2 = 2, MCCS_GAS_CO2 = 3, MCCS_GAS_AR = 4, MCCS_GAS_CO2_1SITE = 5 };

/* Run configuration: the fields of the reference's Input
 * (src/headers/Input.h:49-67, defaults src/headers/Constants.h:68-83 and
 * src/Input.cpp:75-262).  Strings are NUL-terminated paths. */
typedef struct {
  char target_file[1024];
  char force_field[1024];   /* empty: built-in table (MoleculeTarget::defaultFF) */
  uint32_t n_probe, n_iter, seed, nthreads;
  double dt, temperature, skin, lj_cutoff, coul_cutoff, alpha;
  int32_t gas;              /* MCCS_GAS_* */
  int32_t equipotential;    /* 1: equipotential ellipsoid (src/Equipotential.cpp), computed on the GPU */
  int32_t short_range_cutoff, long_range_flag, long_range_cutoff, polarizability;
  int32_t user_ff;          /* 1 if force_field is set */
  int32_t alpha_set;        /* 1 if alpha carries a value */
  int32_t co2_vcm_fix;      /* CO2 probe: 0 (default) measures chi from the reference's initial
                               velocity with its x component replaced by y (src/System.cpp:1012);
                               1 from the true initial velocity (optional key "CO2 vcm fix": "yes";
                               no counterpart in the reference) */
} mccs_config;

/* Reference defaults (Input::readInputFile with an input holding only
 * targetFileName).  Replaces src/Input.cpp:49-263. */
mccs_status mccs_config_default(mccs_config *cfg);

/* Parse input.json with the reference's keys, defaults and error exits
 * turned into status codes.  Relative file names in the json are kept as
 * given (resolved against the process working directory, as the reference
 * does).  Replaces Input::Input (src/Input.cpp:23-263). */
mccs_status mccs_config_load_json(const char *path, mccs_config *cfg);

/* ------------------------------------------------------------------ */
/* Prepared system: target read, parameterised, centred and rotated onto its
 * principal axes (MoleculeTarget, src/MoleculeTarget.cpp:19-931), probe
 * constants (GasBuffer, src/GasBuffer.cpp:19-195), force type and reduced
 * mass (src/System.cpp:49-117), geometric ellipsoid (src/System.cpp:1640-1698)
 * and linked cells with atoms reordered by cell (src/LinkedCell.cpp:19-440). */
typedef struct mccs_prepared mccs_prepared;

mccs_status mccs_prepare(const mccs_config *cfg, mccs_prepared **out);

/* Same preparation from atoms held in memory (synthetic targets): n atoms
 * with element symbols (used against the force-field table), positions and
 * charges; charges ignored for force types 1/2 as for a 5-column .xyz. */
mccs_status mccs_prepare_atoms(const mccs_config *cfg, int64_t n,
                               const char *const *symbols, const double *xyz,
                               const double *q, mccs_prepared **out);

void mccs_prepared_destroy(mccs_prepared *p);

typedef struct {
  int32_t gas, force_type, nsites, natoms;
  int32_t nx, ny, nz, ncells;
  int32_t filled_cells, max_atoms_cell;
  double mass, charge, mu, alpha;       /* alpha in kcal/mol units */
  double a, b, c, bmax;
  double half_box[3];
  double cell;                           /* lj_cutoff + skin */
  double corner[3];
  double radius;
  int64_t stencil1_atoms, stencil2_atoms; /* atoms over all cells' stencils */
} mccs_prepared_info;

mccs_status mccs_prepared_get_info(const mccs_prepared *p, mccs_prepared_info *info);

/* Sorted target arrays (cell order), each of length natoms; any pointer may
 * be NULL.  eps_c/sig_c are the CO2 carbon-site pair parameters. */
mccs_status mccs_prepared_get_atoms(const mccs_prepared *p, double *x, double *y,
                                    double *z, double *q, double *m, double *eps,
                                    double *sig, double *eps_c, double *sig_c);

/* Text the reference prints from the start of the run up to and including
 * "linked-cell calculation time" (Input::printReadInput,
 * MoleculeTarget::printFF/print, src/System.cpp:91-144).  len receives the
 * full length; at most cap-1 bytes are copied (NUL-terminated). */
mccs_status mccs_prepared_report(const mccs_prepared *p, char *buf, size_t cap, size_t *len);

/* Monte-Carlo samples of trajectories [first, first+n) in the reference's
 * mt19937 stream order (src/System.cpp:184-197): 9 doubles per trajectory,
 * u[0] = impact parameter b, u[1..4] = speed, orientation draws, u[5..8] the
 * probe-orientation draws (N2/CO2; zero otherwise). */
mccs_status mccs_samples(const mccs_prepared *p, uint64_t first, uint64_t n, double *u);

/* ------------------------------------------------------------------ */
/* Device engine on one GPU: the prepared target, cell index and speed table
 * resident in HBM. */
typedef struct mccs_engine mccs_engine;

typedef struct {
  int32_t device;          /* HIP device ordinal */
  int32_t lanes;           /* lanes cooperating on one trajectory: 0 = auto, 16, 32 or 64 */
  int32_t block;           /* threads per block: 0 = auto (768 when the target's atoms are
                              staged in LDS, else 256), else a multiple of 64 from 128 up
                              to that */
  int32_t use_lds;         /* -1 auto, 0 global-memory atoms, 1 atoms staged in LDS */
  int32_t count_work;      /* 1: fill the work counters of mccs_traj_record */
  int32_t spec_below;      /* tries of a restarting trajectory run one after another; once
                              fewer than this many trajectories are unresolved and no new
                              one is left, idle groups also start its later tries side by
                              side (results are the same either way): -1 = auto, 0 = never */
  int32_t list_cap;        /* Verlet-list entries per trajectory: 0 = auto (what the block's
                              LDS allows), else at most this many; a list that outgrows it
                              is replaced by a walk over the cell columns, which differs only
                              in summation order.  No counterpart in the reference (it has
                              no lists). */
  int32_t order;           /* order in which the run starts its trajectories: 0 = likely-long
                              first (small impact parameter and slow speed: fewer groups sit
                              idle at the end of the run), 1 = index order.  Records are the
                              same either way.  No counterpart in the reference (its OpenMP
                              loop schedules in index order). */
} mccs_engine_opts;

mccs_status mccs_engine_create(const mccs_prepared *p, const mccs_engine_opts *opts,
                               mccs_engine **out);
void mccs_engine_destroy(mccs_engine *e);

/* One trajectory's result.  hit/success/chi follow the reference loop body
 * (src/System.cpp:206-230); tries is trajTries at exit, steps the integration
 * steps over all tries; domega = pi (1 - cos chi) bmax^2 when scattered.
 * The work counters (summed over every site evaluation of the trajectory) are
 * filled when mccs_engine_opts.count_work is set: visits = (site, atom) pairs
 * examined, lj = pairs inside the LJ cutoff, coul = Coulomb pairs, dip =
 * induced-dipole pairs. */
typedef struct {
  int32_t hit, success, tries, steps;
  double chi, domega;
  uint64_t visits, lj, coul, dip;
} mccs_traj_record;

/* Runs trajectories over host-resident samples (n x 9 doubles, as written by
 * mccs_samples) and writes host records.  Replaces the OpenMP loop of
 * src/System.cpp:204-285 and the System::setup/run_* calls in it. */
mccs_status mccs_run(mccs_engine *e, const double *samples, uint64_t n,
                     mccs_traj_record *records);

/* Same with device-resident samples and records (HBM pointers), enqueued on
 * a HIP stream (NULL = the engine's stream) without synchronising. */
mccs_status mccs_run_device(mccs_engine *e, const double *samples_dev, uint64_t n,
                            mccs_traj_record *records_dev, void *stream);

/* Same as mccs_run_device, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no counterpart in the reference.  The call allocates nothing;
 * the trajectory work queue (counters and per-trajectory try table) lives in
 * a statically sized device array that the library holds on each GPU until
 * the process ends (about 24 MiB: a 1.5 KiB counter block plus 364 bytes per
 * trajectory and 64 KiB of spare spawn slots).  n may be at most
 * MCCS_NOALLOC_MAX_TRAJECTORIES; larger n returns MCCS_ERR_INVALID (run them
 * in chunks).  Every call on one GPU shares that array and the run clears
 * what it reads first: calls that overlap on one GPU (other streams, other
 * engines, other threads) race and must be serialised by the caller.  The
 * engine (created beforehand) still owns its target data; samples and
 * records are the caller's device buffers as in mccs_run_device.
 * mccs_engine_last_work reports the run's executed work as for
 * mccs_run_device. */
#define MCCS_NOALLOC_MAX_TRAJECTORIES 65536
mccs_status mccs_run_device_noalloc(mccs_engine *e, const double *samples_dev, uint64_t n,
                                    mccs_traj_record *records_dev, void *stream);

/* Work the device executed in the last mccs_run / mccs_run_device on this
 * engine, every try included (a trajectory that restarts has its later tries
 * run concurrently; tries made moot by an earlier success are abandoned):
 * totals[0..4] = (site, atom) pairs examined, LJ pairs, Coulomb pairs,
 * induced-dipole pairs, integration steps.  Synchronises the stream of that
 * run.  Calls on one engine must not overlap (they share its work queue). */
mccs_status mccs_engine_last_work(mccs_engine *e, uint64_t totals[5]);

/* Site roles for mccs_eval_sites.  For N2: MCCS_SITE_N uses
 * Force::lennardjones_coulomb[_LC] (or lennardjones[_LC] for types 1/2) with
 * charge -0.4825; MCCS_SITE_DUMMY uses coulomb[_LC] or
 * coulomb_induced_dipole_iso[_LC] with charge +0.965.  For He/Ar/co2:
 * MCCS_SITE_ATOM uses lennardjones[_LC] / lennardjones_induced_dipole[_LC].
 * For CO2: MCCS_SITE_O (lennardjones_coulomb*) and MCCS_SITE_C (the *_CO2
 * kernels). */
enum { MCCS_SITE_ATOM = 0, MCCS_SITE_N = 1, MCCS_SITE_DUMMY = 2, MCCS_SITE_O = 3, MCCS_SITE_C = 4 };

/* Per-point site forces: for each of n host points (xyz, 3n doubles) the force
 * (3n) and potential energy (n) of the probe site of the given role.
 * Replaces Force::<kernel>(GasBuffer*, int iatom, vector<double>& f,
 * double& Up) (src/headers/Force.h:42-62). */
mccs_status mccs_eval_sites(mccs_engine *e, int32_t site, const double *xyz, uint64_t n,
                            double *f, double *U);

/* Tally over nIter x nTraj records in index order, src/System.cpp:287-317:
 * omega_iter[i] = sum_j domega / (Nscatter + Nfree); counts[3i..3i+2] = Nfree,
 * Nscatter, Nlost; mean and standard error of the nIter values. */
mccs_status mccs_tally(const mccs_traj_record *records, uint64_t n_iter, uint64_t n_traj,
                       double *omega_iter, int64_t *counts, double *ccs_mean,
                       double *ccs_err);

/* Same tally on the device over device-resident records; writes n_iter
 * omega values, 3 n_iter counts and (mean, err) into device buffers. */
mccs_status mccs_tally_device(mccs_engine *e, const mccs_traj_record *records_dev,
                              uint64_t n_iter, uint64_t n_traj, double *omega_dev,
                              int64_t *counts_dev, double *mean_err_dev, void *stream);

/* ------------------------------------------------------------------ */
/* Target preparation steps on the GPU, for very large targets.  mccs_prepare
 * (host) stays the authoritative path; these compute the same results from
 * device arrays. */

/* Linked-cell binning and the stable reorder of the atoms by cell
 * (LinkedCell::calculateAtomsInsideOfCell / sortingAtoms,
 * src/LinkedCell.cpp:93-225): for n atoms at xyz_dev (n x 3 device doubles,
 * the oriented target) and the grid of mccs_prepared_info (cell edge, corner,
 * nx, ny, nz), writes cell_start_dev (nx ny nz + 1 device ints: the first
 * sorted position of every cell, then n) and order_dev (n device ints: the
 * input index of the atom at each sorted position; atoms of one cell keep
 * their input order).  An atom outside the grid returns MCCS_ERR_INVALID, as
 * the host preparation does.  Synchronises the stream. */
mccs_status mccs_cells_device(int32_t device, int64_t n, const double *xyz_dev, double cell,
                              const double corner[3], int32_t nx, int32_t ny, int32_t nz,
                              int32_t *cell_start_dev, int32_t *order_dev, void *stream);

/* Centre of mass, total mass and inertia tensor about the centre of mass of
 * n atoms (MoleculeTarget::calculateCenterOfMass and setInertia,
 * src/MoleculeTarget.cpp:701-745), from device positions and masses;
 * inertia receives 9 doubles, row-major.  The FP64 sums are reduced in
 * another order than the host's.  Synchronises the stream. */
mccs_status mccs_inertia_device(int32_t device, int64_t n, const double *xyz_dev, const double *m_dev,
                                double com[3], double *mass, double inertia[9], void *stream);

/* ------------------------------------------------------------------ */
/* Batches: many prepared targets in one run (no counterpart in the
 * reference, which runs one target per process).  Targets are grouped by
 * probe kind and force type; each group runs as one persistent launch whose
 * device work queue spans all its targets' trajectories (larger targets
 * first), and the launches are captured in a hipGraph that later runs of the
 * same sizes replay.  Supported: He/Ar/co2 force types 2 and 4, N2 force
 * type 6.  Records are those mccs_run would produce for each target with the
 * same launch shape. */
typedef struct mccs_batch mccs_batch;

typedef struct {
  int32_t device;
  int32_t lanes;           /* 0: tuned (or 64 until mccs_batch_autotune runs), else 16/32/64 */
  int32_t count_work;      /* 1: fill the work counters of the records */
  int32_t use_graph;       /* 1: capture the launches in a hipGraph and replay it */
  double skin;             /* Verlet-list skin in A; 0: tuned (or 1 A) */
  int32_t order;           /* as mccs_engine_opts.order */
} mccs_batch_opts;

/* The targets' data is copied to the device: the mccs_prepared handles may be
 * destroyed once the batch exists. */
mccs_status mccs_batch_create(const mccs_prepared *const *targets, int32_t ntargets,
                              const mccs_batch_opts *opts, mccs_batch **out);
void mccs_batch_destroy(mccs_batch *b);

/* Runs n[t] trajectories of every target t from host samples[t] (n[t] x 9
 * doubles, as written by mccs_samples) into host records[t]. */
mccs_status mccs_batch_run(mccs_batch *b, const uint64_t *n, const double *const *samples,
                           mccs_traj_record *const *records);

/* Same with device-resident samples[t] and records[t] (HBM pointers),
 * enqueued on a HIP stream (NULL = the batch's own stream) without
 * synchronising: the batch's launches start after the work already on that
 * stream and finish before later work on it. */
mccs_status mccs_batch_run_device(mccs_batch *b, const uint64_t *n,
                                  const double *const *samples_dev,
                                  mccs_traj_record *const *records_dev, void *stream);

/* Autotuner: for every group, times trial runs of the first n_trial[t]
 * samples of each target over lanes per trajectory {16, 32, 64}, Verlet
 * skin {0.5, 1, 2} A and LDS tile (the Verlet-list capacity: half a CU per
 * block, two blocks per CU, or a whole CU), keeps the fastest, and caches it
 * for this process by (probe, force type, cutoffs, target sizes, far field).
 * report (may be NULL) receives the trials as JSON ("lanes", "skin", "tile":
 * 1 half, 2 whole, "ms"); len the full length. */
mccs_status mccs_batch_autotune(mccs_batch *b, const uint64_t *n_trial,
                                const double *const *samples, char *report, size_t cap,
                                size_t *len);

/* Launch shape chosen for every target (either array may be NULL). */
mccs_status mccs_batch_shapes(const mccs_batch *b, int32_t *lanes, double *skin);

/* Executed work of the last batch run, as mccs_engine_last_work. */
mccs_status mccs_batch_last_work(mccs_batch *b, uint64_t totals[5]);

/* ------------------------------------------------------------------ */
/* Far-field surrogate (no counterpart in the reference).  A Fourier-feature
 * MLP -- 64 features [sin(2 pi B r'), cos(2 pi B r')], r' = (r - centre) *
 * in_scale, then 3 x 128 SiLU, then 2 outputs (times out_scale) -- trained on
 * the far field psi that the exact linked-cell evaluator computes (both
 * outputs are fitted to psi; the far field uses output 0).  BF16 weights,
 * FP32 accumulation, all products on MFMA.  Parameters are a flat FP32 array
 * of MCCS_MLP_NPARAM values: W1[64][128], b1[128], W2[128][128], b2[128],
 * W3[128][128], b3[128], W4[128][2], b4[2] (row-major, W[in][out]). */
#define MCCS_MLP_NPARAM 41602

typedef struct {
  float freq[32][3];       /* B */
  double centre[3];
  double in_scale;         /* 1 / length scale, 1/A */
  double out_scale[2];     /* kcal/mol per output unit */
} mccs_mlp_spec;

/* Outputs y (n x 2, kcal/mol) and, if dy is not NULL, their gradients
 * (n x 2 x 3, kcal/mol/A) at n points xyz (n x 3). */
mccs_status mccs_mlp_eval(int32_t device, const mccs_mlp_spec *spec, const float *params,
                          const double *xyz, uint64_t n, double *y, double *dy);

/* Mean squared error of the scaled outputs against e (n x 2, kcal/mol) over
 * one batch (n <= 65536) and its gradient with respect to the parameters. */
mccs_status mccs_mlp_loss_grad(int32_t device, const mccs_mlp_spec *spec, const float *params,
                               const double *xyz, const double *e, uint64_t n, double *loss,
                               float *grad);

/* Loss of values and directional derivatives over one batch (n <= 32768):
 * sum over points and both outputs of (y - e)^2 + w (dy - d)^2 / 2n in the
 * scaled output units, y the outputs at xyz and dy their derivatives along
 * the unit vectors dir (n x 3), carried through the network in forward mode
 * (the analytic directional derivative); e (n x 2, kcal/mol) and d (n x 2,
 * kcal/mol/A) are the targets.  grad receives the loss gradient with respect
 * to the parameters; y_out and dy_out (may be NULL) the scaled y and dy. */
mccs_status mccs_mlp_jvp_loss_grad(int32_t device, const mccs_mlp_spec *spec, const float *params,
                                   const double *xyz, const double *dir, const double *e,
                                   const double *d, double w, uint64_t n, double *loss,
                                   float *grad, double *y_out, double *dy_out);

/* Adam training (cosine-decayed learning rate lr -> lr/20) of params in place
 * on n points; loss_history (may be NULL) receives one mean loss per epoch. */
mccs_status mccs_mlp_train(int32_t device, const mccs_mlp_spec *spec, float *params,
                           const double *xyz, const double *e, uint64_t n, int32_t epochs,
                           int32_t batch, double lr, uint32_t seed, double *loss_history);

typedef struct {
  uint64_t n_train;        /* points at which the exact far field is sampled to train the
                              network (default 2e6; table_source 0 only) */
  int32_t epochs, batch;   /* default 300, 8192 */
  double lr;               /* default 5e-3 (cosine decay to lr / 20) */
  double shell;            /* pairs nearer than this keep the exact terms (default 8 A) */
  double blend;            /* width of the switch to the far field (default 4 A): pair terms
                              reach zero at shell + blend, which must not exceed the Coulomb
                              cutoff */
  double grid;             /* spacing of the tabulated far field (default 0.5 A) */
  double feature_length;   /* Fourier frequencies ~ N(0, 1 / feature_length^2) per A; 0 = auto,
                              0.23 x (largest target extent + 2 x the Coulomb cutoff) */
  uint32_t seed;
  int32_t table_source;    /* 0: the trained network is tabulated; 1: the exact far field is
                              evaluated at the grid nodes (no network); 2: no far field (the
                              near terms alone: the ablation) */
} mccs_surrogate_opts;

typedef struct {
  double train_loss;                 /* last epoch, scaled units (network only) */
  /* held-out error and scale of psi (value, [0] and [1]) and of grad psi
   * (per component, 1/A): [0] the tabulated field the trajectories use,
   * [1] the network before tabulation (zero without a network); against the
   * exact far field at points 2 A or more from every atom */
  double val_rmse_u[2], val_rms_u[2];
  double val_rmse_f[2], val_rms_f[2];
  double seconds_sample, seconds_train, seconds_table;
  uint64_t n_train, grid_nodes;
  mccs_mlp_spec spec;
} mccs_surrogate_report;

/* Range-split far field of an engine whose probe carries an induced dipole
 * (N2 force type 6, He/Ar/co2 force type 4); no counterpart in the
 * reference.  Every electrostatic pair term of src/Force.cpp (the shifted
 * Coulomb potential and the smoothed ion-induced-dipole field) is split by a
 * smooth partition S(r) of the pair distance: S = 1 below `shell`, 0 beyond
 * shell + blend (quintic smoothstep between).  The trajectory kernel keeps
 * summing S times the exact terms over the atoms within shell + blend (with
 * the Lennard-Jones terms, whose cutoff the list covers); the remainder is
 * one smooth scalar field psi(x) = sum_j q_j (1 - S) u_c(r_j),
 * u_c = 1/r - 1.5/rc + r^2/(2 rc^3), whose value, gradient and Hessian give
 * every far-field term: a charged site's q K psi and -q K grad psi, the
 * dipole field's E = -grad psi and grad E = -Hess psi.  psi is tabulated as
 * interpolating cubic B-spline coefficients: from the Fourier-feature MLP
 * trained (BF16 MFMA) on psi sampled by the linked-cell evaluator
 * (table_source 0), or from that evaluator at every node (table_source 1).
 * params_out (may be NULL) receives the trained parameters. */
mccs_status mccs_engine_build_surrogate(mccs_engine *e, const mccs_surrogate_opts *opts,
                                        mccs_surrogate_report *report, float *params_out);

/* 1: later runs use the range-split far field (64 lanes per trajectory);
 * 0: the exact force everywhere. */
mccs_status mccs_engine_use_surrogate(mccs_engine *e, int32_t on);

/* The same for every target of a batch that supports it (the others stay
 * exact); reports (may be NULL) receives one report per target (zero for
 * the others). */
mccs_status mccs_batch_build_surrogate(mccs_batch *b, const mccs_surrogate_opts *opts,
                                       mccs_surrogate_report *reports);
mccs_status mccs_batch_use_surrogate(mccs_batch *b, int32_t on);

/* Library version string. */
const char *mccs_version(void);

#ifdef __cplusplus
}
#endif
#endif
