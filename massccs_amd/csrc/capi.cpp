// extern "C" boundary of libmassccs_amd: converts exceptions to status codes.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"
#include "mlp.hpp"
#include "host.hpp"
#include "massccs.h"

struct mccs_prepared {
  mccs::Prepared *p;
};
struct mccs_engine {
  mccs::Engine *e;
};
struct mccs_batch {
  mccs::Batch *b;
};

namespace {
thread_local std::string g_err;

template <class F>
mccs_status guard(F &&f) {
  try {
    f();
    return MCCS_OK;
  } catch (const mccs::Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "out of memory";
    return MCCS_ERR_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return MCCS_ERR_INVALID;
  }
}

void need(bool c, const char *what) {
  if (!c) throw mccs::Error(MCCS_ERR_INVALID, what);
}
}  // namespace

extern "C" {

const char *mccs_last_error(void) { return g_err.c_str(); }

const char *mccs_version(void) { return "massccs_amd 0.1 (gfx950)"; }

mccs_status mccs_config_default(mccs_config *cfg) {
  return guard([&] {
    need(cfg, "null config");
    *cfg = mccs::default_config();
  });
}

mccs_status mccs_config_load_json(const char *path, mccs_config *cfg) {
  return guard([&] {
    need(path && cfg, "null argument");
    std::string gas;
    mccs::load_json_config(path, *cfg, gas);
  });
}

mccs_status mccs_prepare(const mccs_config *cfg, mccs_prepared **out) {
  return guard([&] {
    need(cfg && out, "null argument");
    *out = nullptr;
    auto *p = mccs::prepare(*cfg, mccs::gas_name(cfg->gas), 0, nullptr, nullptr, nullptr);
    *out = new mccs_prepared{p};
  });
}

mccs_status mccs_prepare_atoms(const mccs_config *cfg, int64_t n, const char *const *symbols,
                               const double *xyz, const double *q, mccs_prepared **out) {
  return guard([&] {
    need(cfg && out && symbols && xyz && n > 0, "null argument or empty target");
    *out = nullptr;
    auto *p = mccs::prepare(*cfg, mccs::gas_name(cfg->gas), n, symbols, xyz, q);
    *out = new mccs_prepared{p};
  });
}

void mccs_prepared_destroy(mccs_prepared *p) {
  if (!p) return;
  delete p->p;
  delete p;
}

mccs_status mccs_prepared_get_info(const mccs_prepared *pp, mccs_prepared_info *info) {
  return guard([&] {
    need(pp && info, "null argument");
    const mccs::Prepared &P = *pp->p;
    std::memset(info, 0, sizeof *info);
    info->gas = P.cfg.gas;
    info->force_type = P.force_type;
    info->nsites = P.gas.nsites;
    info->natoms = (int32_t)P.natoms;
    info->nx = P.nx; info->ny = P.ny; info->nz = P.nz; info->ncells = P.ncells;
    int filled = 0, mx = 0;
    for (int c : P.cell_count) {
      if (c) filled++;
      if (c > mx) mx = c;
    }
    info->filled_cells = filled;
    info->max_atoms_cell = mx;
    info->mass = P.mass; info->charge = P.Q; info->mu = P.mu; info->alpha = P.alpha;
    info->a = P.a; info->b = P.b; info->c = P.c; info->bmax = P.bmax;
    for (int k = 0; k < 3; k++) {
      info->half_box[k] = P.half_box[k];
      info->corner[k] = P.corner[k];
    }
    info->cell = P.cell;
    info->radius = P.radius;
    info->stencil1_atoms = P.stencil1_atoms;
    info->stencil2_atoms = P.stencil2_atoms;
  });
}

mccs_status mccs_prepared_get_atoms(const mccs_prepared *pp, double *x, double *y, double *z,
                                    double *q, double *m, double *eps, double *sig, double *eps_c,
                                    double *sig_c) {
  return guard([&] {
    need(pp, "null argument");
    const mccs::Prepared &P = *pp->p;
    const size_t n = (size_t)P.natoms;
    auto cp = [&](double *dst, const std::vector<double> &src) {
      if (dst) std::memcpy(dst, src.data(), sizeof(double) * n);
    };
    cp(x, P.x); cp(y, P.y); cp(z, P.z); cp(q, P.q); cp(m, P.m);
    cp(eps, P.eps); cp(sig, P.sig); cp(eps_c, P.eps_c); cp(sig_c, P.sig_c);
  });
}

mccs_status mccs_prepared_report(const mccs_prepared *pp, char *buf, size_t cap, size_t *len) {
  return guard([&] {
    need(pp, "null argument");
    const std::string &r = pp->p->report;
    if (len) *len = r.size();
    if (buf && cap) {
      size_t k = std::min(cap - 1, r.size());
      std::memcpy(buf, r.data(), k);
      buf[k] = 0;
    }
  });
}

mccs_status mccs_samples(const mccs_prepared *pp, uint64_t first, uint64_t n, double *u) {
  return guard([&] {
    need(pp && (u || n == 0), "null argument");
    mccs::samples(*pp->p, first, n, u);
  });
}

mccs_status mccs_engine_create(const mccs_prepared *pp, const mccs_engine_opts *opts,
                               mccs_engine **out) {
  return guard([&] {
    need(pp && out, "null argument");
    *out = nullptr;
    mccs_engine_opts o{};
    o.use_lds = -1;
    o.spec_below = -1;
    if (opts) o = *opts;
    auto *e = mccs::engine_create(*pp->p, o);
    *out = new mccs_engine{e};
  });
}

void mccs_engine_destroy(mccs_engine *e) {
  if (!e) return;
  mccs::engine_destroy(e->e);
  delete e;
}

mccs_status mccs_run(mccs_engine *e, const double *samples, uint64_t n, mccs_traj_record *records) {
  return guard([&] {
    need(e && ((samples && records) || n == 0), "null argument");
    mccs::engine_run_host(e->e, samples, n, records);
  });
}

mccs_status mccs_run_device(mccs_engine *e, const double *samples_dev, uint64_t n,
                            mccs_traj_record *records_dev, void *stream) {
  return guard([&] {
    need(e && ((samples_dev && records_dev) || n == 0), "null argument");
    mccs::engine_run_device(e->e, samples_dev, n, records_dev, stream);
  });
}

mccs_status mccs_run_device_noalloc(mccs_engine *e, const double *samples_dev, uint64_t n,
                                    mccs_traj_record *records_dev, void *stream) {
  return guard([&] {
    need(e && ((samples_dev && records_dev) || n == 0), "null argument");
    mccs::engine_run_device_noalloc(e->e, samples_dev, n, records_dev, stream);
  });
}

mccs_status mccs_eval_sites(mccs_engine *e, int32_t site, const double *xyz, uint64_t n, double *f,
                            double *U) {
  return guard([&] {
    need(e && ((xyz && f && U) || n == 0), "null argument");
    mccs::engine_eval_sites(e->e, site, xyz, n, f, U);
  });
}

mccs_status mccs_engine_last_work(mccs_engine *e, uint64_t totals[5]) {
  return guard([&] {
    need(e && totals, "null argument");
    mccs::engine_last_work(e->e, totals);
  });
}

mccs_status mccs_batch_create(const mccs_prepared *const *targets, int32_t ntargets,
                              const mccs_batch_opts *opts, mccs_batch **out) {
  return guard([&] {
    need(targets && out && ntargets > 0, "null argument or empty batch");
    *out = nullptr;
    mccs_batch_opts o{};
    o.use_graph = 1;
    o.count_work = 1;
    if (opts) o = *opts;
    std::vector<const mccs::Prepared *> ps;
    for (int32_t t = 0; t < ntargets; t++) {
      need(targets[t] != nullptr, "null target");
      ps.push_back(targets[t]->p);
    }
    auto *b = mccs::batch_create(ps.data(), ntargets, o);
    *out = new mccs_batch{b};
  });
}

void mccs_batch_destroy(mccs_batch *b) {
  if (!b) return;
  mccs::batch_destroy(b->b);
  delete b;
}

mccs_status mccs_batch_run(mccs_batch *b, const uint64_t *n, const double *const *samples,
                           mccs_traj_record *const *records) {
  return guard([&] {
    need(b && n && samples && records, "null argument");
    mccs::batch_run(b->b, n, samples, records);
  });
}

mccs_status mccs_batch_autotune(mccs_batch *b, const uint64_t *n_trial,
                                const double *const *samples, char *report, size_t cap,
                                size_t *len) {
  return guard([&] {
    need(b && n_trial && samples, "null argument");
    const std::string r = mccs::batch_autotune(b->b, n_trial, samples);
    if (len) *len = r.size();
    if (report && cap) {
      const size_t k = std::min(cap - 1, r.size());
      std::memcpy(report, r.data(), k);
      report[k] = 0;
    }
  });
}

mccs_status mccs_batch_shapes(const mccs_batch *b, int32_t *lanes, double *skin) {
  return guard([&] {
    need(b, "null argument");
    mccs::batch_shapes(b->b, lanes, skin);
  });
}

mccs_status mccs_batch_last_work(mccs_batch *b, uint64_t totals[5]) {
  return guard([&] {
    need(b && totals, "null argument");
    mccs::batch_last_work(b->b, totals);
  });
}

namespace {
mccs::MlpSpec to_spec(const mccs_mlp_spec &s) {
  mccs::MlpSpec m{};
  static_assert(sizeof(m.freq) == sizeof(s.freq), "frequency table");
  std::memcpy(m.freq, s.freq, sizeof m.freq);
  for (int d = 0; d < 3; d++) m.centre[d] = s.centre[d];
  m.in_scale = s.in_scale;
  m.out_scale[0] = s.out_scale[0];
  m.out_scale[1] = s.out_scale[1];
  return m;
}
struct MlpHolder {
  mccs::Mlp *m;
  ~MlpHolder() { mccs::mlp_destroy(m); }
};
}  // namespace

mccs_status mccs_mlp_eval(int32_t device, const mccs_mlp_spec *spec, const float *params,
                          const double *xyz, uint64_t n, double *y, double *dy) {
  return guard([&] {
    need(spec && params && ((xyz && y) || n == 0), "null argument");
    if (n == 0) return;
    MlpHolder h{mccs::mlp_create(device, to_spec(*spec), params, 16384)};
    mccs::mlp_eval(h.m, xyz, n, y, dy);
  });
}

mccs_status mccs_mlp_loss_grad(int32_t device, const mccs_mlp_spec *spec, const float *params,
                               const double *xyz, const double *e, uint64_t n, double *loss,
                               float *grad) {
  return guard([&] {
    need(spec && params && xyz && e && loss && grad, "null argument");
    need(n > 0 && n <= 65536, "batch must hold 1..65536 points");
    MlpHolder h{mccs::mlp_create(device, to_spec(*spec), params, (int)n)};
    *loss = mccs::mlp_loss_grad(h.m, xyz, e, n, grad);
  });
}

mccs_status mccs_mlp_jvp_loss_grad(int32_t device, const mccs_mlp_spec *spec, const float *params,
                                   const double *xyz, const double *dir, const double *e,
                                   const double *d, double w, uint64_t n, double *loss,
                                   float *grad, double *y_out, double *dy_out) {
  return guard([&] {
    need(spec && params && xyz && dir && e && d && loss && grad, "null argument");
    need(n > 0 && n <= 32768, "batch must hold 1..32768 points");
    MlpHolder h{mccs::mlp_create(device, to_spec(*spec), params, (int)(2 * n))};
    *loss = mccs::mlp_jvp_loss_grad(h.m, xyz, dir, e, d, w, n, grad, y_out, dy_out);
  });
}

mccs_status mccs_mlp_train(int32_t device, const mccs_mlp_spec *spec, float *params,
                           const double *xyz, const double *e, uint64_t n, int32_t epochs,
                           int32_t batch, double lr, uint32_t seed, double *loss_history) {
  return guard([&] {
    need(spec && params && xyz && e && n > 0 && epochs > 0 && batch > 0, "null or empty argument");
    need(batch <= 65536, "batch must be at most 65536");
    const mccs::MlpSpec sp = to_spec(*spec);
    MlpHolder h{mccs::mlp_create(device, sp, params, batch)};
    mccs::TrainOpts o;
    o.epochs = epochs;
    o.batch = batch;
    o.lr = lr;
    o.lr_final = lr / 20.0;
    o.seed = seed;
    std::vector<double> hist;
    mccs::mlp_train_host(h.m, xyz, e, n, o, &hist);
    mccs::mlp_get_params(h.m, params);
    if (loss_history)
      for (size_t i = 0; i < hist.size(); i++) loss_history[i] = hist[i];
  });
}

mccs_status mccs_engine_build_surrogate(mccs_engine *e, const mccs_surrogate_opts *opts,
                                        mccs_surrogate_report *report, float *params_out) {
  return guard([&] {
    need(e, "null engine");
    mccs_surrogate_opts o{};
    if (opts) o = *opts;
    mccs_surrogate_report r{};
    mccs::engine_build_surrogate(e->e, o, r, params_out);
    if (report) *report = r;
  });
}

mccs_status mccs_engine_use_surrogate(mccs_engine *e, int32_t on) {
  return guard([&] {
    need(e, "null engine");
    mccs::engine_use_surrogate(e->e, on != 0);
  });
}

mccs_status mccs_batch_run_device(mccs_batch *b, const uint64_t *n,
                                  const double *const *samples_dev,
                                  mccs_traj_record *const *records_dev, void *stream) {
  return guard([&] {
    need(b && n && samples_dev && records_dev, "null argument");
    mccs::batch_run_device(b->b, n, samples_dev, records_dev, stream);
  });
}

mccs_status mccs_batch_build_surrogate(mccs_batch *b, const mccs_surrogate_opts *opts,
                                       mccs_surrogate_report *reports) {
  return guard([&] {
    need(b, "null batch");
    mccs_surrogate_opts o{};
    if (opts) o = *opts;
    mccs::batch_build_surrogate(b->b, o, reports);
  });
}

mccs_status mccs_batch_use_surrogate(mccs_batch *b, int32_t on) {
  return guard([&] {
    need(b, "null batch");
    mccs::batch_use_surrogate(b->b, on != 0);
  });
}

mccs_status mccs_cells_device(int32_t device, int64_t n, const double *xyz_dev, double cell,
                              const double corner[3], int32_t nx, int32_t ny, int32_t nz,
                              int32_t *cell_start_dev, int32_t *order_dev, void *stream) {
  return guard([&] {
    need(corner && cell_start_dev && (n == 0 || (xyz_dev && order_dev)), "null argument");
    need(cell > 0.0 && nx > 0 && ny > 0 && nz > 0, "invalid grid");
    mccs::device_cell_sort(device, n, xyz_dev, cell, corner, nx, ny, nz, cell_start_dev, order_dev,
                           stream);
  });
}

mccs_status mccs_inertia_device(int32_t device, int64_t n, const double *xyz_dev, const double *m_dev,
                                double com[3], double *mass, double inertia[9], void *stream) {
  return guard([&] {
    need(com && mass && inertia && (n == 0 || (xyz_dev && m_dev)), "null argument");
    double I[3][3];
    mccs::device_inertia(device, n, xyz_dev, m_dev, com, mass, I, stream);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) inertia[3 * a + b] = I[a][b];
  });
}

// src/System.cpp:287-317
mccs_status mccs_tally(const mccs_traj_record *rec, uint64_t n_iter, uint64_t n_traj,
                       double *omega_iter, int64_t *counts, double *ccs_mean, double *ccs_err) {
  return guard([&] {
    need(rec || n_iter * n_traj == 0, "null argument");
    need(n_iter > 0, "nIter must be positive");
    double Omega = 0.0, Omega2 = 0.0;
    uint64_t k = 0;
    for (uint64_t i = 0; i < n_iter; i++) {
      double dO = 0.0;
      int nf = 0, ns = 0, nl = 0;
      for (uint64_t j = 0; j < n_traj; j++) {
        const mccs_traj_record &r = rec[k++];
        dO += r.domega;
        if (!r.hit) nf++;
        else if (r.success) ns++;
        else nl++;
      }
      const double w = 1.0 / ((double)(float)(ns + nf)) * dO;
      Omega += w;
      Omega2 += std::pow(w, 2.0);
      if (omega_iter) omega_iter[i] = w;
      if (counts) {
        counts[3 * i] = nf;
        counts[3 * i + 1] = ns;
        counts[3 * i + 2] = nl;
      }
    }
    const double ave = Omega / n_iter;
    const double sig2 = Omega2 / n_iter - std::pow(ave, 2.0);
    if (ccs_mean) *ccs_mean = ave;
    if (ccs_err) *ccs_err = std::sqrt(sig2 / (double)(float)n_iter);
  });
}

mccs_status mccs_tally_device(mccs_engine *e, const mccs_traj_record *rec_dev, uint64_t n_iter,
                              uint64_t n_traj, double *omega_dev, int64_t *counts_dev,
                              double *mean_err_dev, void *stream) {
  return guard([&] {
    need(e && rec_dev && omega_dev && counts_dev && mean_err_dev, "null argument");
    mccs::engine_tally_device(e->e, rec_dev, n_iter, n_traj, omega_dev, counts_dev, mean_err_dev,
                              stream);
  });
}

}  // extern "C"
