// Test harness that drives the reference MassCCS sources directly.
//
// This file is test infrastructure only: it is compiled against the
// reference sources where they lie (see oracle/Makefile) into oracle/_ref/,
// and is used to pin the CPU oracle and to write the golden fixtures under
// tests/golden/.  Nothing in the product links it.
//
// Usage:
//   ref_harness prep  <input.json>                      dump the prepared target
//   ref_harness force <input.json> <points.txt>         per-point site forces
//   ref_harness traj  <input.json> <first> <count>      per-trajectory records
//
// <input.json> must carry "numberProbe": 0 so that the reference constructor
// (src/System.cpp:19-326) runs its whole preamble (input, target orientation,
// ellipsoid, linked cell) and no trajectories.  The harness then replays the
// sample stream exactly as src/System.cpp:184-197 draws it and calls the
// reference's own setup()/run_He()/run_N2()/run_CO2() per trajectory.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>
#include <unistd.h>
#include "../rapidjson/document.h"

#define private public
#include "System.h"
#undef private

static void dump_prep(System &s) {
  MoleculeTarget *t = s.moleculeTarget;
  LinkedCell *lc = s.linkedcell;
  printf("force_type %u\n", s.force_type);
  printf("natoms %u\n", t->natoms);
  printf("mass %.17g\nQ %.17g\n", t->mass, t->Q);
  printf("abc %.17g %.17g %.17g\n", s.a, s.b, s.c);
  printf("bmax %.17g\nmu %.17g\nalpha %.17g\n", s.bmax, s.mu, s.alpha);
  printf("halfbox %.17g %.17g %.17g\n", s.lx, s.ly, s.lz);
  printf("grid %d %d %d %d\n", lc->Nx, lc->Ny, lc->Nz, lc->Ncells);
  printf("maxXYZ %.17g %.17g %.17g\n", t->maxX, t->maxY, t->maxZ);
  for (unsigned i = 0; i < t->natoms; i++) {
    printf("atom %.17g %.17g %.17g %.17g %.17g %.17g %.17g", t->x[i], t->y[i],
           t->z[i], t->q[i], t->m[i], t->eps[i], t->sig[i]);
    if (s.gas_buffer_flag == 3)
      printf(" %.17g %.17g", t->eps_central[i], t->sig_central[i]);
    printf("\n");
  }
  for (int c = 0; c < lc->Ncells; c++) {
    if (!lc->atoms_inside_cell[c]) continue;
    printf("cell %d %d %d\n", c, lc->atoms_inside_cell[c], lc->atoms_ids[c][0]);
  }
  // stencil sizes of every cell (first shell, second shell)
  long s1 = 0, s2 = 0;
  for (int c = 0; c < lc->Ncells; c++) {
    s1 += lc->neighbors1_cells[c];
    if (lc->next_neighbor == 1) s2 += lc->neighbors2_cells[c];
  }
  printf("stencil_sums %ld %ld\n", s1, s2);
}

// Kernel ids understood by "force": name -> member function.
typedef void (Force::*ForceFn)(GasBuffer *, int, vector<double> &, double &);
static const std::map<std::string, ForceFn> kKernels = {
    {"lennardjones", &Force::lennardjones},
    {"lennardjones_LC", &Force::lennardjones_LC},
    {"lennardjones_coulomb", &Force::lennardjones_coulomb},
    {"lennardjones_coulomb_LC", &Force::lennardjones_coulomb_LC},
    {"coulomb", &Force::coulomb},
    {"coulomb_LC", &Force::coulomb_LC},
    {"lennardjones_induced_dipole_LC", &Force::lennardjones_induced_dipole_LC},
    {"coulomb_induced_dipole_iso_LC", &Force::coulomb_induced_dipole_iso_LC},
    {"lennardjones_LC_CO2", &Force::lennardjones_LC_CO2},
    {"lennardjones_CO2", &Force::lennardjones_CO2},
    {"lennardjones_coulomb_CO2", &Force::lennardjones_coulomb_CO2},
    {"lennardjones_coulomb_LC_CO2", &Force::lennardjones_coulomb_LC_CO2},
    {"lennardjones_coulomb_induced_dipole_iso_LC_CO2",
     &Force::lennardjones_coulomb_induced_dipole_iso_LC_CO2},
};

// points file: lines "<kernel> <site> <x> <y> <z>"
static void run_force(System &s, const char *points) {
  std::ifstream in(points);
  std::string line;
  Force force(s.moleculeTarget, s.linkedcell, s.lj_cutoff, s.alpha, s.coul_cutoff);
  GasBuffer gas(s.gas_buffer_flag);
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string kname;
    int site;
    double x, y, z;
    if (!(ss >> kname >> site >> x >> y >> z)) continue;
    auto it = kKernels.find(kname);
    if (it == kKernels.end()) {
      fprintf(stderr, "unknown kernel %s\n", kname.c_str());
      exit(2);
    }
    gas.x[site] = x;
    gas.y[site] = y;
    gas.z[site] = z;
    vector<double> f(3, 0.0);
    double U = 0.0;
    (force.*(it->second))(&gas, site, f, U);
    printf("%s %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", kname.c_str(), site,
           x, y, z, f[0], f[1], f[2], U);
  }
}

static void run_traj(System &s, long first, long count) {
  const bool multi = (s.gas_buffer_flag == 2 || s.gas_buffer_flag == 3);
  const long total = first + count;
  std::vector<double> r(9 * total, 0.0);
  RandomNumber mt(s.seed);
  for (long j = 0; j < total; j++) {
    r[9 * j + 0] = sqrt(s.bmax * s.bmax * mt.getRandomNumber());
    for (int k = 1; k < 5; k++) r[9 * j + k] = mt.getRandomNumber();
    if (multi)
      for (int k = 5; k < 9; k++) r[9 * j + k] = mt.getRandomNumber();
  }
  std::vector<int> hit(count), ok(count);
  std::vector<double> chi(count, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (long jj = 0; jj < count; jj++) {
    const double *u = &r[9 * (first + jj)];
    Force force(s.moleculeTarget, s.linkedcell, s.lj_cutoff, s.alpha, s.coul_cutoff);
    GasBuffer probe(s.gas_buffer_flag);
    bool h = false, succ = false;
    double c = 0.0;
    s.setup(&probe, h, u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8]);
    if (h) {
      if (s.gas_buffer_flag == 2)
        s.run_N2(&probe, succ, c, s.dt, &force);
      else if (s.gas_buffer_flag == 3)
        s.run_CO2(&probe, succ, c, s.dt, &force);
      else
        s.run_He(&probe, succ, c, s.dt, &force);
    }
    hit[jj] = h;
    ok[jj] = succ;
    chi[jj] = c;
  }
  for (long jj = 0; jj < count; jj++)
    printf("traj %ld %d %d %.17g\n", first + jj, hit[jj], ok[jj], chi[jj]);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s prep|force|traj <input.json> [...]\n", argv[0]);
    return 1;
  }
  std::string mode = argv[1];
  // The reference constructor prints its report to stdout; keep it out of
  // the harness output by sending it to stderr while the preamble runs.
  std::streambuf *saved = std::cout.rdbuf(std::cerr.rdbuf());
  fflush(stdout);
  int saved_fd = dup(1);
  dup2(2, 1);
  System s(argv[2]);
  fflush(stdout);
  dup2(saved_fd, 1);
  close(saved_fd);
  std::cout.rdbuf(saved);
  if (mode == "prep") {
    dump_prep(s);
  } else if (mode == "force" && argc >= 4) {
    run_force(s, argv[3]);
  } else if (mode == "traj" && argc >= 5) {
    run_traj(s, atol(argv[3]), atol(argv[4]));
  } else {
    fprintf(stderr, "bad arguments\n");
    return 1;
  }
  return 0;
}
