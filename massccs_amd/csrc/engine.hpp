// Device engine interface (internal to libmassccs_amd).
#pragma once

#include <cstdint>

#include "host.hpp"

namespace mccs {

struct Engine;

Engine *engine_create(const Prepared &p, const mccs_engine_opts &opts);
void engine_destroy(Engine *e);

// Trajectories over device-resident samples/records, enqueued on stream.
void engine_run_device(Engine *e, const double *u_dev, uint64_t n, mccs_traj_record *rec_dev,
                       void *stream);
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

}  // namespace mccs
