// massccs <input.json>: drop-in command line for the reference program
// (src/main.cpp:20-34), running the trajectories on the GPU engine and
// printing the reference's report (src/System.cpp:19-326).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "massccs.h"

static void die(mccs_status s) {
  std::cerr << "Error: " << mccs_last_error() << std::endl;
  std::exit(s == MCCS_OK ? 0 : 1);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cout << "Usage: " << argv[0] << " input.json" << std::endl;
    return 1;
  }
  using clk = std::chrono::steady_clock;
  auto t_start = clk::now();
  auto secs = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  mccs_config cfg;
  mccs_status st = mccs_config_load_json(argv[1], &cfg);
  if (st != MCCS_OK) die(st);
  mccs_prepared *prep = nullptr;
  st = mccs_prepare(&cfg, &prep);
  if (st != MCCS_OK) die(st);
  size_t len = 0;
  mccs_prepared_report(prep, nullptr, 0, &len);
  std::string rep(len + 1, '\0');
  mccs_prepared_report(prep, &rep[0], rep.size(), &len);
  rep.resize(len);
  std::cout << rep;

  mccs_engine_opts opts{};
  opts.use_lds = -1;
  opts.spec_below = -1;
  if (const char *d = std::getenv("MASSCCS_DEVICE")) opts.device = std::atoi(d);
  mccs_engine *eng = nullptr;
  st = mccs_engine_create(prep, &opts, &eng);
  if (st != MCCS_OK) die(st);

  const uint64_t niter = cfg.n_iter, ntraj = cfg.n_probe, n = niter * ntraj;
  auto t_ccs = clk::now();
  std::vector<double> u(9 * n);
  st = mccs_samples(prep, 0, n, u.data());
  if (st != MCCS_OK) die(st);
  std::cout << "*********************************************************" << std::endl;
  std::cout << "Trajectory calculations " << std::endl;
  std::cout << "*********************************************************" << std::endl;
  std::vector<mccs_traj_record> rec(n);
  st = mccs_run(eng, u.data(), n, rec.data());
  if (st != MCCS_OK) die(st);
  std::vector<double> om(niter);
  std::vector<int64_t> cnt(3 * niter);
  double mean = 0, err = 0;
  if (niter > 0) {
    st = mccs_tally(rec.data(), niter, ntraj, om.data(), cnt.data(), &mean, &err);
    if (st != MCCS_OK) die(st);
  }
  for (uint64_t i = 0; i < niter; i++) {
    printf("Ntraj: %i\n", (int)ntraj);
    printf("Nfree: %i\n", (int)cnt[3 * i]);
    printf("Nscatter: %i\n", (int)cnt[3 * i + 1]);
    printf("Nlost: %i\n", (int)cnt[3 * i + 2]);
    printf("omega: %g\n", om[i]);
  }
  fflush(stdout);
  auto t_end_ccs = clk::now();
  std::cout << "CCS time: " << secs(t_ccs, t_end_ccs) << " s" << std::endl;
  std::cout << "*********************************************************" << std::endl;
  std::cout << "average value of CCS = " << mean << " Ang^2" << std::endl;
  std::cout << "error value of CCS = " << err << " Ang^2" << std::endl;
  std::cout << "Total time: " << secs(t_start, clk::now()) << " s" << std::endl;
  mccs_engine_destroy(eng);
  mccs_prepared_destroy(prep);
  std::cout << "Program finished..." << std::endl;
  // the reference's System destructor runs after main's last line
  if (cfg.equipotential) std::cout << "equipotential destroyed" << std::endl;
  return 0;
}
