// Device engine interface (internal to libmassccs_amd).
#pragma once

#include <cstdint>
#include <string>

#include "host.hpp"

namespace mccs {

struct Engine;

Engine *engine_create(const Prepared &p, const mccs_engine_opts &opts);
void engine_destroy(Engine *e);

// Trajectories over device-resident samples/records, enqueued on stream.
void engine_run_device(Engine *e, const double *u_dev, uint64_t n, mccs_traj_record *rec_dev,
                       void *stream);
// Same with the work queue in the library's static device workspace (no
// device allocation; at most MCCS_NOALLOC_MAX_TRAJECTORIES trajectories).
void engine_run_device_noalloc(Engine *e, const double *u_dev, uint64_t n,
                               mccs_traj_record *rec_dev, void *stream);
// Same over host buffers (copies in and out, synchronous).
void engine_run_host(Engine *e, const double *u, uint64_t n, mccs_traj_record *rec);
// Per-point site forces over host buffers.
void engine_eval_sites(Engine *e, int site, const double *xyz, uint64_t n, double *f, double *U);
// Executed work of the last run (visits, lj, coul, dip, steps), all tries
// included; synchronises the stream of that run.
void engine_last_work(Engine *e, uint64_t out[5]);
// Tally over device-resident records.
void engine_tally_device(Engine *e, const mccs_traj_record *rec_dev, uint64_t n_iter,
                         uint64_t n_traj, double *omega_dev, int64_t *counts_dev,
                         double *mean_err_dev, void *stream);

// Range-split far field of an N2 type-6 or He/Ar/co2 type-4 engine (the
// exact field or the network trained on it, tabulated on the GPU).
void engine_build_surrogate(Engine *e, const mccs_surrogate_opts &o, mccs_surrogate_report &rep,
                            float *params_out);
void engine_use_surrogate(Engine *e, bool on);

// Several targets in one run (one persistent launch per probe kind / force
// type, captured in a hipGraph).
struct Batch;
Batch *batch_create(const Prepared *const *targets, int ntargets, const mccs_batch_opts &opts);
void batch_destroy(Batch *b);
void batch_run(Batch *b, const uint64_t *n, const double *const *samples,
               mccs_traj_record *const *records);
void batch_run_device(Batch *b, const uint64_t *n, const double *const *samples_dev,
                      mccs_traj_record *const *records_dev, void *stream);
std::string batch_autotune(Batch *b, const uint64_t *n_trial, const double *const *samples);
void batch_last_work(Batch *b, uint64_t out[5]);
void batch_shapes(const Batch *b, int32_t *lanes, double *skin);
void batch_build_surrogate(Batch *b, const mccs_surrogate_opts &o, mccs_surrogate_report *reports);
void batch_use_surrogate(Batch *b, bool on);

// Target preparation steps on the GPU (prep_device.hip): linked-cell binning
// with the stable reorder, and centre of mass / inertia tensor.
void device_cell_sort(int device, int64_t n, const double *xyz_dev, double cell, const double corner[3],
                      int nx, int ny, int nz, int *cell_start_dev, int *order_dev, void *stream);
void device_inertia(int device, int64_t n, const double *xyz_dev, const double *m_dev, double com[3],
                    double *mass, double inertia[3][3], void *stream);

}  // namespace mccs
