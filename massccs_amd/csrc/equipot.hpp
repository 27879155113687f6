// Equipotential ellipsoid (internal to libmassccs_amd).
#pragma once

#include <array>
#include <cstddef>
#include <string>
#include <vector>

namespace mccs {

struct EquiInput {
  int gas = 1, polarizability = 0;
  double temperature = 298.0, mu = 1.0, alpha = 0.0;  // alpha in kcal/mol units
  double maxX = 0, maxY = 0, maxZ = 0;                 // signed largest extents
  // oriented target in file order
  std::vector<double> x, y, z, q, eps, sig, eps_c, sig_c;
};

struct EquiResult {
  double a = 0, b = 0, c = 0;
  size_t points = 0;
  std::string report;  // the reference's "Equipotential Ellipsoid" block
};

// Runs on the current HIP device; throws Error(MCCS_ERR_DEVICE) without one.
EquiResult equipotential_ellipsoid(const EquiInput &in);

}  // namespace mccs
