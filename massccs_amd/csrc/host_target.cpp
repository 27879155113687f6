// Target and probe preparation on the host: restates MoleculeTarget
// (src/MoleculeTarget.cpp), GasBuffer (src/GasBuffer.cpp), the reduced-mass and
// ellipsoid steps of System (src/System.cpp:89-144, :1640-1698) and
// LinkedCell (src/LinkedCell.cpp).  Arithmetic keeps the reference's operation
// order: these coordinates feed the trajectories, so they must come out bit
// for bit as the reference computes them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>

#include "equipot.hpp"
#include "host.hpp"

namespace mccs {

Gas make_gas(int flag) {
  Gas g;
  g.flag = flag;
  g.name = gas_name(flag);
  switch (flag) {
    case MCCS_GAS_N2:  // src/GasBuffer.cpp:77-110
      g.nsites = 3;
      g.z[0] = 0.5488; g.z[1] = -0.5488; g.z[2] = 0.0;
      g.q[0] = -0.4825; g.q[1] = -0.4825; g.q[2] = 0.965;
      g.m[0] = 14.007; g.m[1] = 14.007; g.m[2] = 0.0;
      g.mass = g.m[0] + g.m[1];
      g.d = std::abs(g.z[0] - g.z[1]);
      break;
    case MCCS_GAS_CO2:  // src/GasBuffer.cpp:111-165
      g.nsites = 3;
      g.z[0] = -1.149; g.z[1] = 0.0; g.z[2] = 1.149;
      g.q[0] = -0.3256; g.q[1] = 0.6512; g.q[2] = -0.3256;
      g.m[0] = 15.999; g.m[1] = 12.011; g.m[2] = 15.999;
      g.mass = g.m[0] + g.m[1] + g.m[2];
      g.d = std::abs(g.z[0] - g.z[2]);
      break;
    case MCCS_GAS_AR:  // src/GasBuffer.cpp:166-179
      g.m[0] = 39.948; g.mass = g.m[0];
      break;
    case MCCS_GAS_CO2_1SITE:  // src/GasBuffer.cpp:180-193
      g.m[0] = 44.0098; g.mass = g.m[0];
      break;
    default:  // He, src/GasBuffer.cpp:58-76
      g.m[0] = 4.0026; g.mass = g.m[0];
      break;
  }
  return g;
}

namespace {

// MoleculeTarget::defaultFF, src/MoleculeTarget.cpp:372-538.
ForceField default_ff(int gas) {
  ForceField f;
  auto add = [&](const char *n, double m, double e, double s) {
    f.name.push_back(n); f.m.push_back(m); f.eps.push_back(e); f.sig.push_back(s);
  };
  if (gas == MCCS_GAS_HE) {
    add("C", 12.011, 0.0309, 3.043);
    add("N", 14.007, 0.0309, 3.043);
    add("H", 1.008, 0.0150, 2.38);
    add("O", 15.999, 0.0309, 3.043);
    add("S", 32.06, 0.0311, 3.5);
  } else if (gas == MCCS_GAS_N2) {
    add("C", 12.011, 0.0824736, 3.2255);
    add("N", 14.007, 0.0758527, 3.5719);
    add("H", 1.008, 0.0362711, 1.8986);
    add("O", 15.999, 0.062323, 3.0750);
    add("S", 32.06, 0.138032, 3.4237);
    add("P", 30.9738, 0.145372, 3.47);
    add("F", 18.9984, 0.04649, 3.1285);
  } else if (gas == MCCS_GAS_CO2 || gas == MCCS_GAS_AR) {
    add("C", 12.011, 0.10499, 3.43085);
    add("N", 14.007, 0.06899, 3.26069);
    add("H", 1.008, 0.04399, 2.57113);
    add("O", 15.999, 0.05999, 3.11815);
    add("S", 32.06, 0.27399, 3.59478);
    add("P", 30.9738, 0.30499, 3.69456);
  } else if (gas == MCCS_GAS_CO2_1SITE) {
    add("C", 12.011, 0.09759, 3.58140);
    add("N", 14.007, 0.05579, 3.25500);
    add("H", 1.008, 0.01892, 1.24090);
    add("O", 15.999, 0.08279, 4.39200);
    add("S", 32.06, 0.05999, 3.50000);
    add("P", 30.9738, 0.05999, 3.50000);
    add("K", 39.0983, 0.05999, 3.50000);
  }
  return f;
}

// MoleculeTarget::readUserFF, src/MoleculeTarget.cpp:321-370.
ForceField read_ff(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw Error(MCCS_ERR_IO, "Error: reading force-field parameters file " + path);
  ForceField f;
  std::string line;
  std::getline(in, line);
  std::stringstream ss(line);
  unsigned int n;
  if (!(ss >> n)) throw Error(MCCS_ERR_PARSE, "Error: reading force-field parameters file");
  std::getline(in, line);  // comment
  for (unsigned i = 0; i < n; i++) {
    std::getline(in, line);
    std::stringstream ls(line);
    std::string t;
    double m, e, s;
    if (!(ls >> t >> m >> e >> s))
      throw Error(MCCS_ERR_PARSE, "Error: reading force-field parameters file");
    f.name.push_back(t); f.m.push_back(m); f.eps.push_back(e); f.sig.push_back(s);
  }
  return f;
}

struct Params { double m, eps, sig, eps_c, sig_c; };

// MoleculeTarget::assignedParameter, src/MoleculeTarget.cpp:540-570.
Params assign(const ForceField &ff, int gas, bool user_ff, const std::string &sym) {
  for (size_t i = 0; i < ff.name.size(); i++) {
    if (sym == ff.name[i]) {
      Params p{ff.m[i], 0, 0, 0, 0};
      if (gas == MCCS_GAS_CO2) {
        p.eps = std::sqrt(ff.eps[i] * 0.159);
        p.sig = 0.5 * (ff.sig[i] + 3.033);
        p.eps_c = std::sqrt(ff.eps[i] * 0.055);
        p.sig_c = 0.5 * (ff.sig[i] + 2.757);
      } else if (gas == MCCS_GAS_AR) {
        p.eps = std::sqrt(ff.eps[i] * 0.185);
        p.sig = 0.5 * (ff.sig[i] + 3.446);
      } else {
        p.eps = ff.eps[i];
        p.sig = ff.sig[i];
      }
      return p;
    }
  }
  throw Error(MCCS_ERR_INVALID, user_ff ? "Atom type not found in the file user force field: " + sym
                                        : "Atom type not found in the default data base: " + sym);
}

struct RawTarget {
  std::vector<std::string> name;
  std::vector<double> x, y, z, q, m, eps, sig, eps_c, sig_c;
  double mass = 0, Q = 0;
};

void push_atom(RawTarget &t, const ForceField &ff, int gas, bool user_ff,
               const std::string &sym, double x, double y, double z, double q) {
  Params p = assign(ff, gas, user_ff, sym);
  t.name.push_back(sym);
  t.x.push_back(x); t.y.push_back(y); t.z.push_back(z); t.q.push_back(q);
  t.m.push_back(p.m); t.eps.push_back(p.eps); t.sig.push_back(p.sig);
  t.eps_c.push_back(p.eps_c); t.sig_c.push_back(p.sig_c);
  t.mass += p.m;
  t.Q += q;
}

// MoleculeTarget::readPQRfile, src/MoleculeTarget.cpp:60-146.  Fields are
// extracted with the same stream sequence; as in the reference, a field that
// fails to parse keeps its previous value and the record name of a blank line
// is the previous line's.
void read_pqr(const std::string &path, const ForceField &ff, int gas, bool user_ff,
              RawTarget &t) {
  std::ifstream in(path);
  if (!in) throw Error(MCCS_ERR_IO, "cannot open target file " + path);
  std::string line, rec, aname, rname;
  int serial = 0, chain = 0;
  double x = 0, y = 0, z = 0, q = 0, r = 0;
  while (std::getline(in, line)) {
    std::istringstream s(line);
    s >> rec >> serial >> aname >> rname >> chain >> x >> y >> z >> q >> r;
    if (rec == "ATOM") {
      std::string sym = user_ff ? aname : aname.substr(0, 1);
      push_atom(t, ff, gas, user_ff, sym, x, y, z, q);
    }
  }
}

// MoleculeTarget::readXYZfile, src/MoleculeTarget.cpp:148-319.
void read_xyz(const std::string &path, const ForceField &ff, int gas, bool user_ff,
              int force_type, RawTarget &t) {
  std::ifstream in(path);
  if (!in) throw Error(MCCS_ERR_IO, "cannot open target file " + path);
  std::string line;
  std::getline(in, line);
  unsigned int natoms;
  {
    std::stringstream ss(line);
    if (!(ss >> natoms)) throw Error(MCCS_ERR_PARSE, "Error: missing number ot atoms in xyz file");
  }
  std::getline(in, line);  // comment
  std::vector<std::string> rows;
  while (std::getline(in, line)) rows.push_back(line);
  std::vector<int> items;
  for (auto &l : rows) {
    std::istringstream iss(l);
    std::string it;
    int n = 0;
    while (iss >> it) n++;
    items.push_back(n);
  }
  if (items.empty()) throw Error(MCCS_ERR_PARSE, "Error: numbers of lines and natoms are differens in the xyz file");
  for (size_t i = 1; i < items.size(); i++)
    if (items[i] != items[0])
      throw Error(MCCS_ERR_PARSE, "Error: numbers of items by lines in the xyz file");
  const bool with_q = items[0] != 4;  // src/MoleculeTarget.cpp:212-218
  if (rows.size() != natoms)
    throw Error(MCCS_ERR_PARSE, "Error: numbers of lines and natoms are differens in the xyz file");
  std::string sym;
  double x = 0, y = 0, z = 0, q = 0;
  for (unsigned i = 0; i < natoms; i++) {
    std::stringstream ss(rows[i]);
    if (!with_q) {
      ss >> sym >> x >> y >> z;
      push_atom(t, ff, gas, user_ff, sym, x, y, z, 0.);
    } else {
      ss >> sym >> x >> y >> z >> q;
      const bool neutral = (force_type == 1 || force_type == 2);
      push_atom(t, ff, gas, user_ff, sym, x, y, z, neutral ? 0. : q);
    }
  }
}

inline double kdelta(int a, int b) { return a == b ? 1.0 : 0.0; }

// MoleculeTarget::getEigenvalues, src/MoleculeTarget.cpp:812-850.
void eigenvalues(const double A[3][3], double ev[3]) {
  double a0 = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] -
              A[1][1] * A[0][2] * A[2][0] - A[2][2] * A[0][1] * A[1][0] +
              A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[2][1] * A[1][0];
  double a1 = A[0][1] * A[1][0] + A[0][2] * A[2][0] + A[1][2] * A[2][1] - A[0][0] * A[1][1] -
              A[1][1] * A[2][2] - A[2][2] * A[0][0];
  double a2 = A[0][0] + A[1][1] + A[2][2];
  a0 = -a0; a1 = -a1; a2 = -a2;
  double p = (3 * a1 - a2 * a2) / 3;
  double q = (9 * a1 * a2 - 27 * a0 - 2 * a2 * a2 * a2) / 27;
  double Q = p / 3, R = q / 2, D = Q * Q * Q + R * R;
  double a, b;
  if (D < 0) { a = R; b = std::sqrt(-D); }
  else { a = R + std::sqrt(D); b = 0; }
  double phase = std::atan2(b, a);
  double mag = std::pow(a * a + b * b, 1.0 / 6.0);
  ev[0] = -(a2 / 3) + mag * 2 * std::cos(phase / 3.0);
  ev[1] = -(a2 / 3) - mag * (std::cos(phase / 3.0) + std::sqrt(3.0) * std::sin(phase / 3.0));
  ev[2] = -(a2 / 3) - mag * (std::cos(phase / 3.0) - std::sqrt(3.0) * std::sin(phase / 3.0));
}

// MoleculeTarget::getEigenvectors, src/MoleculeTarget.cpp:852-878.
void eigenvectors(const double A[3][3], const double l[3], double E[3][3]) {
  double det = l[0] * l[0] - l[0] * A[1][1] - A[1][2] * A[1][2] - l[0] * A[2][2] + A[1][1] * A[2][2];
  double y = (l[0] * A[1][0] - A[2][2] * A[1][0] + A[1][2] * A[2][0]) / det;
  double z = (A[2][1] * A[1][0] + l[0] * A[2][0] - A[1][1] * A[2][0]) / det;
  double nf = std::sqrt(1 + y * y + z * z);
  E[0][0] = 1.0 / nf; E[0][1] = y / nf; E[0][2] = z / nf;
  det = l[1] * l[1] - l[1] * A[0][0] - A[0][2] * A[0][2] - l[1] * A[2][2] + A[0][0] * A[2][2];
  double x = (l[1] * A[0][1] + A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  z = (A[0][1] * A[0][2] + (l[1] - A[0][0]) * A[2][1]) / det;
  nf = std::sqrt(1 + x * x + z * z);
  E[1][0] = x / nf; E[1][1] = 1 / nf; E[1][2] = z / nf;
  E[2][0] = E[0][1] * E[1][2] - E[0][2] * E[1][1];
  E[2][1] = E[0][2] * E[1][0] - E[0][0] * E[1][2];
  E[2][2] = E[0][0] * E[1][1] - E[0][1] * E[1][0];
}

// MoleculeTarget::orientateMolecule, src/MoleculeTarget.cpp:623-698.  theta is
// zero-initialised here (the reference leaves it unset when the farthest atom
// lies on a coordinate plane).
void orientate(RawTarget &t) {
  const size_t n = t.x.size();
  double rmax = 0.0, theta = 0.0, phi, gamma;
  size_t id = 0;
  for (size_t i = 0; i < n; i++) {
    double r = std::sqrt(t.x[i] * t.x[i] + t.y[i] * t.y[i] + t.z[i] * t.z[i]);
    if (r > rmax) { id = i; rmax = r; }
  }
  double xi = t.x[id], yi = t.y[id], zi = t.z[id];
  if ((xi != 0.0) & (yi != 0.0)) theta = std::atan2(yi, xi);
  phi = std::acos(zi / rmax);
  theta *= -1.0;
  for (size_t i = 0; i < n; i++) {
    xi = t.x[i]; yi = t.y[i];
    t.x[i] = std::cos(theta) * xi - std::sin(theta) * yi;
    t.y[i] = std::sin(theta) * xi + std::cos(theta) * yi;
  }
  phi *= -1.0;
  for (size_t i = 0; i < n; i++) {
    xi = t.x[i]; zi = t.z[i];
    t.x[i] = std::sin(phi) * zi + std::cos(phi) * xi;
    t.z[i] = std::cos(phi) * zi - std::sin(phi) * xi;
  }
  rmax = 0.0;
  for (size_t i = 0; i < n; i++) {
    double r = std::sqrt(t.x[i] * t.x[i] + t.y[i] * t.y[i]);
    if (r > rmax) { id = i; rmax = r; }
  }
  xi = t.x[id]; yi = t.y[id];
  gamma = std::atan2(yi, xi);
  gamma *= -1.0;
  for (size_t i = 0; i < n; i++) {
    xi = t.x[i]; yi = t.y[i];
    t.x[i] = std::cos(gamma) * xi - std::sin(gamma) * yi;
    t.y[i] = std::sin(gamma) * xi + std::cos(gamma) * yi;
  }
}

double dot3(const double *a, const double *b) {
  double r{};
  for (int i = 0; i < 3; i++) r += a[i] * b[i];
  return r;
}

// Centre of mass, radius, principal axes, longest extent on z:
// src/MoleculeTarget.cpp:54-58, :596-621, :701-805, :880-931.
void orient_target(RawTarget &t, Prepared &P) {
  const size_t n = t.x.size();
  double rcm[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < n; i++) {
    rcm[0] += t.m[i] * t.x[i];
    rcm[1] += t.m[i] * t.y[i];
    rcm[2] += t.m[i] * t.z[i];
  }
  rcm[0] /= t.mass; rcm[1] /= t.mass; rcm[2] /= t.mass;
  for (size_t i = 0; i < n; i++) {
    t.x[i] -= rcm[0]; t.y[i] -= rcm[1]; t.z[i] -= rcm[2];
  }
  {  // calculateMoleculeRadius
    double rmax = 0.0;
    size_t id = 0;
    for (size_t i = 0; i < n; i++) {
      double r = std::sqrt(t.x[i] * t.x[i] + t.y[i] * t.y[i] + t.z[i] * t.z[i]);
      if (r > rmax) { id = i; rmax = r; }
    }
    P.radius = n ? std::sqrt(t.x[id] * t.x[id] + t.y[id] * t.y[id] + t.z[id] * t.z[id]) : 0.0;
  }
  double (&I)[3][3] = P.inertia;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) I[a][b] = 0.0;
  for (size_t i = 0; i < n; i++) {
    double r[3] = {t.x[i], t.y[i], t.z[i]};
    double mi = t.m[i];
    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) I[a][b] += mi * (r2 * kdelta(a, b) - r[a] * r[b]);
  }
  P.diagonal = false;
  if (std::abs(I[0][1]) < 1.e-12 && std::abs(I[1][0]) < 1.e-12 && std::abs(I[1][2]) < 1.e-12 &&
      std::abs(I[2][1]) < 1.e-12 && std::abs(I[2][0]) < 1.e-12 && std::abs(I[0][2]) < 1.e-12) {
    orientate(t);
    P.diagonal = true;
  } else {
    double ev[3];
    eigenvalues(I, ev);
    eigenvectors(I, ev, P.vectors);
    double (&V)[3][3] = P.vectors;
    double dp = std::sqrt(std::pow(dot3(V[0], V[1]), 2.0));
    dp = std::max(std::sqrt(std::pow(dot3(V[0], V[2]), 2.0)), dp);
    dp = std::max(std::sqrt(std::pow(dot3(V[1], V[2]), 2.0)), dp);
    if (dp > 1e-6) {
      orientate(t);
    } else {
      for (size_t i = 0; i < n; i++) {
        double xi = t.x[i], yi = t.y[i], zi = t.z[i];
        t.x[i] = xi * V[0][0] + yi * V[0][1] + zi * V[0][2];
        t.y[i] = xi * V[1][0] + yi * V[1][1] + zi * V[1][2];
        t.z[i] = xi * V[2][0] + yi * V[2][1] + zi * V[2][2];
      }
    }
  }
  double mx = 0, my = 0, mz = 0;
  for (size_t i = 0; i < n; i++) {
    if (std::abs(t.x[i]) > std::abs(mx)) mx = std::abs(t.x[i]);
    if (std::abs(t.y[i]) > std::abs(my)) my = std::abs(t.y[i]);
    if (std::abs(t.z[i]) > std::abs(mz)) mz = std::abs(t.z[i]);
  }
  if (mx > my && mx > mz) {
    for (size_t i = 0; i < n; i++) {
      double xi = t.x[i], zi = t.z[i];
      t.x[i] = -zi; t.z[i] = xi;
    }
  } else if (my > mx && my > mz) {
    for (size_t i = 0; i < n; i++) {
      double yi = t.y[i], zi = t.z[i];
      t.y[i] = zi; t.z[i] = -yi;
    }
  }
  mx = my = mz = 0;
  for (size_t i = 0; i < n; i++) {
    if (std::abs(t.x[i]) > std::abs(mx)) mx = t.x[i];
    if (std::abs(t.y[i]) > std::abs(my)) my = t.y[i];
    if (std::abs(t.z[i]) > std::abs(mz)) mz = t.z[i];
  }
  P.maxX = mx; P.maxY = my; P.maxZ = mz;
}

// System::geometric_ellipsoid, src/System.cpp:1640-1698.
void ellipsoid(const RawTarget &t, Prepared &P) {
  const double mX = std::abs(P.maxX), mY = std::abs(P.maxY), mZ = std::abs(P.maxZ);
  const double d = P.gas.d;
  const double radius = P.cfg.long_range_flag == 1 ? 2.0 * P.cfg.coul_cutoff : 2.0 * P.cfg.lj_cutoff;
  double a = mX + radius + P.cfg.skin + d;
  double b = mY + radius + P.cfg.skin + d;
  double c = mZ + radius + P.cfg.skin + d;
  P.a0 = a; P.b0 = b; P.c0 = c;
  const double delta = 0.01;
  for (size_t i = 0; i < t.x.size(); i++) {
    double xi = t.x[i], yi = t.y[i], zi = t.z[i];
    double ri = std::sqrt(xi * xi + yi * yi + zi * zi);
    double ux = xi / ri, uy = yi / ri, uz = zi / ri;
    double x = xi + (radius + d) * ux;
    double y = yi + (radius + d) * uy;
    double z = zi + (radius + d) * uz;
    for (;;) {
      double e = std::pow(x / a, 2) + std::pow(y / b, 2) + std::pow(z / c, 2);
      if (e > 1.0) { a += delta; b += delta; c += delta; }
      else break;
    }
  }
  P.a = a; P.b = b; P.c = c;
  P.bmax = std::max(std::max(a, b), c);
}

// LinkedCell, src/LinkedCell.cpp:19-129 (grid, binning) and :132-225 (stable
// reorder of the atoms by cell index).
void linked_cell(RawTarget &t, Prepared &P) {
  const double cell = P.cfg.lj_cutoff + P.cfg.skin;
  P.cell = cell;
  double lx = 2.0 * P.a, ly = 2.0 * P.b, lz = 2.0 * P.c;
  P.nx = (int)std::ceil(lx / cell);
  P.ny = (int)std::ceil(ly / cell);
  P.nz = (int)std::ceil(lz / cell);
  P.box[0] = ((double)P.nx) * cell;
  P.box[1] = ((double)P.ny) * cell;
  P.box[2] = ((double)P.nz) * cell;
  P.ncells = P.nx * P.ny * P.nz;
  for (int k = 0; k < 3; k++) {
    P.corner[k] = -0.5 * P.box[k];
    P.half_box[k] = 0.5 * P.box[k];
  }
  const size_t n = t.x.size();
  std::vector<int> idx(n);
  P.cell_count.assign(P.ncells, 0);
  for (size_t i = 0; i < n; i++) {
    int ii = (int)std::floor((t.x[i] - P.corner[0]) / cell);
    int jj = (int)std::floor((t.y[i] - P.corner[1]) / cell);
    int kk = (int)std::floor((t.z[i] - P.corner[2]) / cell);
    int c = kk + P.nz * jj + P.nz * P.ny * ii;
    if (c < 0 || c >= P.ncells)
      throw Error(MCCS_ERR_INVALID, "target atom outside the linked-cell grid");
    idx[i] = c;
    P.cell_count[c]++;
  }
  P.cell_start.assign(P.ncells + 1, 0);
  for (int c = 0; c < P.ncells; c++) P.cell_start[c + 1] = P.cell_start[c] + P.cell_count[c];
  std::vector<int> fill(P.cell_start.begin(), P.cell_start.end() - 1);
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[fill[idx[i]]++] = i;
  auto perm = [&](const std::vector<double> &src) {
    std::vector<double> o(n);
    for (size_t i = 0; i < n; i++) o[i] = src[order[i]];
    return o;
  };
  P.natoms = (int64_t)n;
  P.x = perm(t.x); P.y = perm(t.y); P.z = perm(t.z); P.q = perm(t.q);
  P.m = perm(t.m); P.eps = perm(t.eps); P.sig = perm(t.sig);
  P.eps_c = perm(t.eps_c); P.sig_c = perm(t.sig_c);
}

std::string fmt_time(double s) {
  std::ostringstream o;
  o << s;
  return o.str();
}

// velGenerator table: the reference's scan (src/System.cpp:1523-1555) visits
// v_n = v_{n-1} + dv with S_n = S_{n-1} + f(v_n) and returns the first v_n
// with S_n >= u.  The table holds (v_n, S_n) until S stops growing.
void speed_table(Prepared &P) {
  const double m = P.mu * AMU_TO_KG;
  const double T = P.cfg.temperature;
  // velDistr takes the mass in amu and converts it itself; keep its order.
  auto vel_distr = [&](double v) {
    double mm = P.mu;
    mm = mm * AMU_TO_KG;
    double vv = v * ANG_TO_M / FS_TO_S;
    return std::pow(vv, 5.0) * std::pow((mm / (2.0 * BOLTZMANN_K * T)), 3.0) *
           std::exp(-(mm * std::pow(vv, 2.0)) / (2.0 * BOLTZMANN_K * T));
  };
  double v = 0.0;
  const double dv = 1.0E-5;
  double s = vel_distr(v);
  P.vtab_v.assign(1, v);
  P.vtab_s.assign(1, s);
  // most probable speed of v^5 exp(-m v^2 / 2kT), in A/fs
  const double vpeak = std::sqrt(5.0 * BOLTZMANN_K * T / m) * FS_TO_S / ANG_TO_M;
  int flat = 0;
  while (s < 1.0 && P.vtab_v.size() < 50000000) {
    v += dv;
    double ns = s + vel_distr(v);
    P.vtab_v.push_back(v);
    P.vtab_s.push_back(ns);
    if (ns == s && v > 2.0 * vpeak) {
      if (++flat > 1000) break;
    } else {
      flat = 0;
    }
    s = ns;
  }
}

}  // namespace

void stencil_offsets(double lj_cutoff, double coul_cutoff, double d, bool next,
                     std::vector<int> &s1, std::vector<int> &s2) {
  // LinkedCell::calculateCellsNeighbors, src/LinkedCell.cpp:228-411
  s1.clear();
  s2.clear();
  const double R1 = lj_cutoff / d;
  const double R2 = next ? coul_cutoff / d : 0.0;
  const int n1 = (int)std::ceil(R1);
  const int nc = next ? (int)std::ceil(R2) : n1;
  auto push = [](std::vector<int> &v, int i, int j, int k) {
    v.push_back(i); v.push_back(j); v.push_back(k);
  };
  auto sq = [](int a) { return (double)a * (double)a; };
  for (int i = -nc; i < nc + 1; i++)
    for (int j = -nc; j < nc + 1; j++)
      for (int k = -nc; k < nc + 1; k++) {
        int nz = (i != 0) + (j != 0) + (k != 0);
        if (nz == 0) { push(s1, i, j, k); continue; }
        if (nz == 1) {
          int a = std::abs(i + j + k);
          if (a <= n1) push(s1, i, j, k);
          else if (next) push(s2, i, j, k);
          continue;
        }
        double nearv = 0.0, farv = 0.0;
        if (i) { nearv += sq(std::abs(i) - 1); farv += sq(i); }
        if (j) { nearv += sq(std::abs(j) - 1); farv += sq(j); }
        if (k) { nearv += sq(std::abs(k) - 1); farv += sq(k); }
        if (next) {
          if (farv < R2 * R2 || nearv < R2 * R2) {
            if (farv < R1 * R1 || nearv < R1 * R1) push(s1, i, j, k);
            else push(s2, i, j, k);
          }
        } else if (farv < R1 * R1 || nearv < R1 * R1) {
          push(s1, i, j, k);
        }
      }
}

Prepared *prepare(const mccs_config &cfg, const std::string &gas_str, int64_t n_mem,
                  const char *const *symbols, const double *xyz, const double *q) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  if (cfg.gas < 1 || cfg.gas > 5) throw Error(MCCS_ERR_INVALID, "unknown gas buffer");
  if (!(cfg.dt > 0) || !(cfg.temperature > 0) || !(cfg.lj_cutoff > 0) || !(cfg.skin >= 0) ||
      !(cfg.coul_cutoff > 0))
    throw Error(MCCS_ERR_INVALID, "dt, Temp, LJ-cutoff and Coul-cutoff must be positive, skin non-negative");
  std::unique_ptr<Prepared> P(new Prepared);
  P->cfg = cfg;
  P->gas_str = gas_str;
  P->gas = make_gas(cfg.gas);
  P->force_type = select_force_type(cfg);
  std::ostringstream rep;
  rep << format_config(cfg, gas_str);

  auto t0 = clk::now();
  const bool user_ff = cfg.user_ff != 0;
  P->ff = user_ff ? read_ff(cfg.force_field) : default_ff(cfg.gas);
  // MoleculeTarget::printFF, src/MoleculeTarget.cpp:572-594
  rep << "*********************************************************\n";
  rep << "Force Field parameters: \n";
  rep << "*********************************************************\n";
  rep << "Symbol  mass (amu)  epsilon (kcal/mol)  sigma(Angstroms) \n";
  if (cfg.gas == MCCS_GAS_CO2) {
    for (size_t i = 0; i < P->ff.name.size(); i++)
      rep << P->ff.name[i] << "-C  " << P->ff.m[i] << "   " << std::sqrt(P->ff.eps[i] * 0.055)
          << "   " << 0.5 * (P->ff.sig[i] + 2.757) << "\n";
    for (size_t i = 0; i < P->ff.name.size(); i++)
      rep << P->ff.name[i] << "-O  " << P->ff.m[i] << "   " << std::sqrt(P->ff.eps[i] * 0.159)
          << "   " << 0.5 * (P->ff.sig[i] + 3.033) << "\n";
  } else {
    for (size_t i = 0; i < P->ff.name.size(); i++)
      rep << P->ff.name[i] << "   " << P->ff.m[i] << "   " << P->ff.eps[i] << "   " << P->ff.sig[i] << "\n";
  }
  RawTarget t;
  if (symbols) {
    const bool neutral = (P->force_type == 1 || P->force_type == 2);
    for (int64_t i = 0; i < n_mem; i++)
      push_atom(t, P->ff, cfg.gas, user_ff, symbols[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                (q && !neutral) ? q[i] : 0.0);
  } else {
    std::string fn = cfg.target_file;
    std::string ext = fn.substr(fn.find_last_of('.') + 1);
    if (ext == "pqr") read_pqr(fn, P->ff, cfg.gas, user_ff, t);
    else if (ext == "xyz") read_xyz(fn, P->ff, cfg.gas, user_ff, P->force_type, t);
    else throw Error(MCCS_ERR_INVALID, "Error: only acceptable PQR or XYZ or XYZ-Q format");
  }
  if (t.x.empty()) throw Error(MCCS_ERR_INVALID, "target has no atoms");
  P->mass = t.mass;
  P->Q = t.Q;
  orient_target(t, *P);
  // MoleculeTarget::print, src/MoleculeTarget.cpp:887-903
  rep << "*********************************************************\n";
  rep << "MOLECULE:: orientation around the inertia principal axis \n";
  rep << "*********************************************************\n";
  rep << "Inertia matrix a.u. Ang^2: \n";
  for (int a = 0; a < 3; a++)
    rep << "{  " << P->inertia[a][0] << "  " << P->inertia[a][1] << "  " << P->inertia[a][2] << "  }\n";
  if (!P->diagonal) {
    rep << "Rotation matrix applied to molecule: \n";
    for (int a = 0; a < 3; a++)
      rep << "{  " << P->vectors[a][0] << "  " << P->vectors[a][1] << "  " << P->vectors[a][2] << "  }\n";
  }
  rep << "Molecule radius: " << P->radius << " Ang \n";
  auto t1 = clk::now();
  rep << "orientation time of molecule target: " << fmt_time(secs(t0, t1)) << " s\n";

  // reduced mass and probe inertia, src/System.cpp:97-117
  const Gas &g = P->gas;
  if (g.nsites == 1) P->mu = P->mass * g.mass / (P->mass + g.mass);
  else {
    P->mu = g.mass;
    P->gas_inertia = 0.0;
    for (int i = 0; i < g.nsites; i++) P->gas_inertia += g.m[i] * std::pow(g.z[i], 2);
    P->d_bond = g.d;
  }
  rep << "target mass: " << P->mass << " amu\n";
  rep << "gas mass: " << g.mass << " amu\n";
  rep << "reduce mass: " << P->mu << " amu\n";
  rep << "charge state: " << P->Q << " e\n";
  P->alpha = (cfg.alpha_set ? cfg.alpha : 0.0) * ALPHA_TO_KCAL_MOL;

  auto t2 = clk::now();
  if (cfg.equipotential) {  // src/System.cpp:120-125
    EquiInput in;
    in.gas = cfg.gas;
    in.polarizability = cfg.polarizability;
    in.temperature = cfg.temperature;
    in.mu = P->mu;
    in.alpha = P->alpha;
    in.maxX = P->maxX; in.maxY = P->maxY; in.maxZ = P->maxZ;
    in.x = t.x; in.y = t.y; in.z = t.z; in.q = t.q;
    in.eps = t.eps; in.sig = t.sig; in.eps_c = t.eps_c; in.sig_c = t.sig_c;
    const EquiResult r = equipotential_ellipsoid(in);
    P->a = r.a; P->b = r.b; P->c = r.c;
    P->a0 = r.a; P->b0 = r.b; P->c0 = r.c;
    P->bmax = std::max(std::max(r.a, r.b), r.c);
    rep << r.report;
    rep << "ellipsoid axis length: " << P->a << "  " << P->b << "  " << P->c << "  Ang\n";
  } else {
    ellipsoid(t, *P);
    rep << "*********************************************************\n";
    rep << "Geometric Ellipsoid: \n";
    rep << "*********************************************************\n";
    rep << "maximal distances: " << std::abs(P->maxX) << "  " << std::abs(P->maxY) << "  "
        << std::abs(P->maxZ) << "  Ang\n";
    rep << "Initial axis length: " << P->a0 << "  " << P->b0 << "  " << P->c0 << "  Ang\n";
    rep << "Ellipsoid axis length: " << P->a << "  " << P->b << "  " << P->c << "  Ang\n";
  }
  auto t3 = clk::now();
  rep << "ellipsoid calculation time: " << fmt_time(secs(t2, t3)) << " s\n";
  rep << "maximal impact parameter: " << P->bmax << " Ang\n";

  auto t4 = clk::now();
  linked_cell(t, *P);
  {  // stencil totals (reference visits per cell) and LinkedCell::print
    std::vector<int> s1, s2;
    stencil_offsets(cfg.lj_cutoff, cfg.coul_cutoff, P->cell, cfg.long_range_flag == 1, s1, s2);
    int64_t v1 = 0, v2 = 0;
    for (int i = 0; i < P->nx; i++)
      for (int j = 0; j < P->ny; j++)
        for (int k = 0; k < P->nz; k++) {
          auto acc = [&](const std::vector<int> &s, int64_t &v) {
            for (size_t m = 0; m < s.size(); m += 3) {
              int a = i + s[m], b = j + s[m + 1], c = k + s[m + 2];
              if (a < 0 || a >= P->nx || b < 0 || b >= P->ny || c < 0 || c >= P->nz) continue;
              v += P->cell_count[c + P->nz * b + P->nz * P->ny * a];
            }
          };
          acc(s1, v1);
          acc(s2, v2);
        }
    P->stencil1_atoms = v1;
    P->stencil2_atoms = v2;
  }
  int filled = 0, empty = 0, sum = 0, mx = 0, mn = (int)P->natoms;
  for (int c = 0; c < P->ncells; c++) {
    int n = P->cell_count[c];
    if (n == 0) { empty++; continue; }
    sum += n;
    filled++;
    mx = std::max(mx, n);
    mn = std::min(mn, n);
  }
  rep << "*********************************************************\n";
  rep << "Linked-cell: \n";
  rep << "Numbers of cells: " << P->ncells << "\n";
  rep << "Nx: " << P->nx << " Ny: " << P->ny << " Nz: " << P->nz << "\n";
  rep << "Filled cells: " << filled << "\n";
  rep << "Empty cells: " << empty << "\n";
  rep << "Average atoms per cell: " << ((float)sum) / ((float)filled) << "\n";
  rep << "Maximum atoms per cell: " << mx << "\n";
  rep << "Minimum atoms per cell: " << mn << "\n";
  rep << "Simulation box: \n";
  rep << "lx: " << P->box[0] << " Ang\n";
  rep << "ly: " << P->box[1] << " Ang\n";
  rep << "lz: " << P->box[2] << " Ang\n";
  auto t5 = clk::now();
  rep << "linked-cell calculation time: " << fmt_time(secs(t4, t5)) << " s\n";
  speed_table(*P);
  P->report = rep.str();
  return P.release();
}

// Sample stream, src/System.cpp:184-197 with RandomNumber
// (src/RandomNumber.cpp:19-29): u = (mt() - min) / max.
void samples(const Prepared &P, uint64_t first, uint64_t n, double *u) {
  const bool multi = P.gas.nsites == 3;
  const uint64_t per = multi ? 9 : 5;
  std::mt19937 mt((std::mt19937::result_type)P.cfg.seed);
  mt.discard(first * per);
  const double mn = (double)std::mt19937::min(), mxv = (double)std::mt19937::max();
  auto rnd = [&]() { return ((double)mt() - mn) / mxv; };
  for (uint64_t j = 0; j < n; j++) {
    double *o = u + 9 * j;
    o[0] = std::sqrt(P.bmax * P.bmax * rnd());
    for (int k = 1; k < 5; k++) o[k] = rnd();
    for (int k = 5; k < 9; k++) o[k] = multi ? rnd() : 0.0;
  }
}

}  // namespace mccs
