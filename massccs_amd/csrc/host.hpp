// Host-side preparation of a trajectory-method run: input, target, probe,
// ellipsoid and linked cells.  Internal header of libmassccs_amd.
#pragma once

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "massccs.h"

namespace mccs {

// Error carrying an mccs_status; thrown inside the library, converted to a
// status code at the C boundary.
struct Error : std::runtime_error {
  mccs_status code;
  Error(mccs_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Reference constants, src/headers/Constants.h:17-84.
constexpr double AMU_TO_KG = 1.66053904E-27;
constexpr double BOLTZMANN_K = 1.38064E-23;
constexpr double ANG_TO_M = 1.0E-10;
constexpr double FS_TO_S = 1.0E-15;
constexpr double KCALMOLANGAMU_TO_ANGFS2 = 1.0 / 48.88821291 / 48.88821291;
constexpr double ALPHA_TO_KCAL_MOL = 331.842254885;
constexpr double OMEGA_TO_FS_INV = 2.45400505E8;
constexpr double KCOUL = 332.06348078020324;
constexpr double AMUANGFS2_TO_KCAL_MOL = 2390.07;

// Probe constants, src/GasBuffer.cpp:58-195.
struct Gas {
  int flag = 1;
  int nsites = 1;
  std::string name;
  double z[3] = {0, 0, 0}, q[3] = {0, 0, 0}, m[3] = {0, 0, 0};
  double mass = 0, d = 0;
};
Gas make_gas(int flag);

// Force-field table (MoleculeTarget::readUserFF / defaultFF).
struct ForceField {
  std::vector<std::string> name;
  std::vector<double> m, eps, sig;
};

struct Prepared {
  mccs_config cfg{};
  std::string gas_str;  // as printed ("He", "N2", ...)
  Gas gas;
  int force_type = 0;
  ForceField ff;
  // target in cell order
  int64_t natoms = 0;
  std::vector<double> x, y, z, q, m, eps, sig, eps_c, sig_c;
  double mass = 0, Q = 0, radius = 0;
  double inertia[3][3]{}, vectors[3][3]{};
  bool diagonal = false;
  double maxX = 0, maxY = 0, maxZ = 0;
  // run constants
  double mu = 0, alpha = 0, gas_inertia = 0, d_bond = 0;
  double a = 0, b = 0, c = 0, bmax = 0;
  double a0 = 0, b0 = 0, c0 = 0;  // initial ellipsoid axes (printed)
  // linked cell
  int nx = 0, ny = 0, nz = 0, ncells = 0;
  double cell = 0, box[3]{}, corner[3]{}, half_box[3]{};
  std::vector<int32_t> cell_count, cell_start;  // cell_start has ncells+1 entries
  int64_t stencil1_atoms = 0, stencil2_atoms = 0;
  // speed table: v_n and S_n of System::velGenerator's scan
  std::vector<double> vtab_v, vtab_s;
  // report text and timings
  std::string report;
};

mccs_config default_config();
void load_json_config(const std::string &path, mccs_config &cfg, std::string &gas_str);
std::string format_config(const mccs_config &cfg, const std::string &gas_str);

// Builds a Prepared from files (symbols == nullptr) or from memory.
Prepared *prepare(const mccs_config &cfg, const std::string &gas_str, int64_t n_mem,
                  const char *const *symbols, const double *xyz, const double *q);

int select_force_type(const mccs_config &cfg);
std::string gas_name(int flag);

// Stencil offsets of LinkedCell::calculateCellsNeighbors in reference order:
// (di, dj, dk) triples for the first and second shells.
void stencil_offsets(double lj_cutoff, double coul_cutoff, double cell, bool long_range,
                     std::vector<int> &first, std::vector<int> &second);

void samples(const Prepared &p, uint64_t first, uint64_t n, double *u);

}  // namespace mccs
