// massccs <input.json>: drop-in command line for the reference program
// (src/main.cpp:20-34), running the trajectories on the GPU engine and
// printing the reference's report (src/System.cpp:19-326).
//
// massccs --batch <list> [table.csv] [--tune]: many inputs in one batched GPU
// run (one persistent launch per probe kind), reported as the table the
// reference's data set describes (data/README.md:13-17: index, id, file,
// natoms, mass in kDa, charge, ccs_He, err_He, ccs_N2, err_N2), sorted by mass.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "massccs.h"

static void die(mccs_status s) {
  std::cerr << "Error: " << mccs_last_error() << std::endl;
  std::exit(s == MCCS_OK ? 0 : 1);
}

static std::string base_name(const std::string &path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? path : path.substr(k + 1);
}

static int run_batch(const char *list, const char *csv, bool tune) {
  std::ifstream in(list);
  if (!in) {
    std::cerr << "Error: cannot open " << list << std::endl;
    return 1;
  }
  std::vector<std::string> inputs;
  for (std::string ln; std::getline(in, ln);) {
    ln.erase(0, ln.find_first_not_of(" \t"));
    ln.erase(ln.find_last_not_of(" \t\r") + 1);
    if (!ln.empty() && ln[0] != '#') inputs.push_back(ln);
  }
  if (inputs.empty()) {
    std::cerr << "Error: no inputs in " << list << std::endl;
    return 1;
  }
  std::vector<mccs_config> cfgs(inputs.size());
  std::vector<mccs_prepared *> preps(inputs.size(), nullptr);
  for (size_t t = 0; t < inputs.size(); t++) {
    mccs_status st = mccs_config_load_json(inputs[t].c_str(), &cfgs[t]);
    if (st == MCCS_OK) st = mccs_prepare(&cfgs[t], &preps[t]);
    if (st != MCCS_OK) {
      std::cerr << inputs[t] << ": ";
      die(st);
    }
  }
  mccs_batch_opts bo{};
  bo.use_graph = 1;
  mccs_batch *b = nullptr;
  mccs_status st = mccs_batch_create(preps.data(), (int32_t)preps.size(), &bo, &b);
  if (st != MCCS_OK) die(st);
  std::vector<uint64_t> n(inputs.size());
  std::vector<std::vector<double>> u(inputs.size());
  std::vector<std::vector<mccs_traj_record>> rec(inputs.size());
  std::vector<const double *> up(inputs.size());
  std::vector<mccs_traj_record *> rp(inputs.size());
  for (size_t t = 0; t < inputs.size(); t++) {
    n[t] = (uint64_t)cfgs[t].n_iter * cfgs[t].n_probe;
    u[t].resize(9 * n[t]);
    rec[t].resize(n[t]);
    if (mccs_samples(preps[t], 0, n[t], u[t].data()) != MCCS_OK) die(MCCS_ERR_INVALID);
    up[t] = u[t].data();
    rp[t] = rec[t].data();
  }
  if (tune) {
    std::vector<uint64_t> nt(n);
    for (auto &x : nt) x = std::max<uint64_t>(std::min<uint64_t>(x, 64), x / 16);
    size_t len = 0;
    std::string rep(1 << 16, '\0');
    if (mccs_batch_autotune(b, nt.data(), up.data(), &rep[0], rep.size(), &len) != MCCS_OK) die(MCCS_ERR_DEVICE);
    rep.resize(std::min(len, rep.size() - 1));
    std::cout << "autotune: " << rep << std::endl;
  }
  auto t0 = std::chrono::steady_clock::now();
  st = mccs_batch_run(b, n.data(), up.data(), rp.data());
  if (st != MCCS_OK) die(st);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  struct Row {
    std::string id, file;
    int natoms = 0;
    double mass = 0, charge = 0;
    std::map<std::string, std::pair<double, double>> ccs;
  };
  std::map<std::string, Row> rows;  // by target file
  uint64_t total = 0;
  for (size_t t = 0; t < inputs.size(); t++) {
    double mean = 0, err = 0;
    std::vector<double> om(cfgs[t].n_iter);
    std::vector<int64_t> cnt(3 * cfgs[t].n_iter);
    if (mccs_tally(rec[t].data(), cfgs[t].n_iter, cfgs[t].n_probe, om.data(), cnt.data(), &mean, &err) != MCCS_OK)
      die(MCCS_ERR_INVALID);
    mccs_prepared_info info;
    mccs_prepared_get_info(preps[t], &info);
    const std::string gas = info.gas == MCCS_GAS_N2 ? "N2" : (info.gas == MCCS_GAS_HE ? "He" :
                            (info.gas == MCCS_GAS_AR ? "Ar" : "CO2"));
    const std::string file = base_name(cfgs[t].target_file);
    Row &r = rows[cfgs[t].target_file];
    r.file = file;
    r.id = file.substr(0, file.find_last_of('.'));
    r.natoms = info.natoms;
    r.mass = info.mass / 1000.0;
    r.charge = info.charge;
    r.ccs[gas] = {mean, err};
    total += n[t];
    std::cout << inputs[t] << ": " << gas << " CCS = " << mean << " +- " << err << " Ang^2" << std::endl;
  }
  std::vector<Row> sorted;
  for (auto &kv : rows) sorted.push_back(kv.second);
  std::stable_sort(sorted.begin(), sorted.end(), [](const Row &a, const Row &b) { return a.mass < b.mass; });
  std::ofstream out;
  std::ostream *os = &std::cout;
  if (csv) {
    out.open(csv);
    if (!out) {
      std::cerr << "Error: cannot write " << csv << std::endl;
      return 1;
    }
    os = &out;
  }
  *os << "index,id,file,natoms,mass,charge,ccs_He,err_He,ccs_N2,err_N2" << std::endl;
  for (size_t i = 0; i < sorted.size(); i++) {
    const Row &r = sorted[i];
    *os << (i + 1) << "," << r.id << "," << r.file << "," << r.natoms << "," << r.mass << "," << r.charge;
    for (const char *g : {"He", "N2"}) {
      auto it = r.ccs.find(g);
      if (it == r.ccs.end()) *os << ",,";
      else *os << "," << it->second.first << "," << it->second.second;
    }
    *os << std::endl;
  }
  std::cout << "batch: " << inputs.size() << " inputs, " << total << " trajectories in " << secs << " s"
            << std::endl;
  mccs_batch_destroy(b);
  for (auto *p : preps) mccs_prepared_destroy(p);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 3 && std::string(argv[1]) == "--batch") {
    bool tune = false;
    const char *csv = nullptr;
    for (int i = 3; i < argc; i++) {
      if (std::string(argv[i]) == "--tune") tune = true;
      else csv = argv[i];
    }
    return run_batch(argv[2], csv, tune);
  }
  if (argc < 2) {
    std::cout << "Usage: " << argv[0] << " input.json" << std::endl;
    std::cout << "       " << argv[0] << " --batch inputs.txt [table.csv] [--tune]" << std::endl;
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
