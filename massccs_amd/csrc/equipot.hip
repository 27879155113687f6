// Equipotential ellipsoid (System's alternative to the geometric ellipsoid,
// src/Equipotential.cpp:19-713) on the GPU.
//
// The reference marches 20 x (levels + 1) rays towards the z axis in 0.1 A
// steps, evaluating an all-atom probe potential at every step until the
// potential energy falls below -1 % of the kinetic energy of a slow probe;
// the crossing points are fitted to an axis-aligned ellipsoid by least
// squares, which is then grown in 0.01 A steps until every point lies inside.
// Here one 64-lane group marches one ray: the atoms are strided over the
// lanes and the site sums reduced across the group, so a step costs one pass
// over the target; the rays march side by side.  The fit and the growth are
// the reference's formulas on the host.  The potentials' sums run in another
// order than the reference's OpenMP reductions, so a crossing can move by one
// step where the energy sits within rounding of the threshold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <sstream>
#include <string>
#include <vector>

#include "equipot.hpp"
#include "host.hpp"

#define EQ_OK(x)                                                                             \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw ::mccs::Error(MCCS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

namespace mccs {
namespace eqdev {

constexpr int L = 64;
constexpr double KB = 1.38064E-23, JEV = 6.242E18, EVKCAL = 23.061, KCOULD = 332.06348078020324;

struct EqT {
  const double4 *at;    // x, y, z, q (file order, oriented)
  const double2 *lj;    // 4 eps sig^6, (4 eps sig^6) sig^6
  const double2 *ljc;   // the same with the CO2 carbon-site parameters
  int n, kind, pol;     // kind: 1 atomic, 2 N2, 3 CO2
  double alpha, ekmin;  // alpha in kcal/mol units; 1 % of the slow probe's kinetic energy
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int d = 1; d < L; d <<= 1) v += __shfl_xor(v, d, L);
  return v;
}

// Boltzmann average at 500 K over the three axis-aligned orientations,
// src/Equipotential.cpp:322-344.
__device__ double orient_average(const double U[3]) {
  const double umin = fmin(U[0], fmin(U[1], U[2]));
  const double kbt = KB * 500.0 * JEV * EVKCAL;
  double Z = 0.0;
  for (int i = 0; i < 3; i++) Z += exp(-(U[i] - umin) / kbt);
  double pot = 0.0;
  for (int i = 0; i < 3; i++) pot += exp(-(U[i] - umin) / kbt) / Z * U[i];
  return pot;
}

// Equipotential::potential_He / potential_N2 / potential_CO2
// (src/Equipotential.cpp:196-239, :241-347, :350-468), group-cooperative.
__device__ double potential(const EqT &T, int lane, const double pos[3]) {
  if (T.kind == 1) {
    double ulj = 0, ex = 0, ey = 0, ez = 0;
    for (int i = lane; i < T.n; i += L) {
      const double4 a = T.at[i];
      const double2 p = T.lj[i];
      const double dx = pos[0] - a.x, dy = pos[1] - a.y, dz = pos[2] - a.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r2inv = 1.0 / r2;
      const double r6inv = r2inv * r2inv * r2inv;
      ulj += r6inv * (p.y * r6inv - p.x);
      if (T.pol) {
        const double rinv = sqrt(r2inv);
        const double r3inv = rinv * r2inv;
        ex += a.w * dx * r3inv;
        ey += a.w * dy * r3inv;
        ez += a.w * dz * r3inv;
      }
    }
    ulj = wsum(ulj);
    double pot = ulj;
    if (T.pol) {
      ex = wsum(ex);
      ey = wsum(ey);
      ez = wsum(ez);
      pot -= 0.5 * T.alpha * (ex * ex + ey * ey + ez * ez);
    }
    return pot;
  }
  const bool n2 = T.kind == 2;
  const double qc = n2 ? 0.965 : 0.6512, qs = -0.5 * qc, d = n2 ? 0.5488 : 1.149;
  double site[6][3];
  for (int k = 0; k < 6; k++)
    for (int c = 0; c < 3; c++) site[k][c] = pos[c];
  for (int ax = 0; ax < 3; ax++) {
    site[2 * ax][ax] = pos[ax] + d;
    site[2 * ax + 1][ax] = pos[ax] - d;
  }
  double ulj[6] = {0, 0, 0, 0, 0, 0}, uc[6] = {0, 0, 0, 0, 0, 0};
  double ucc = 0, uljc = 0, ex = 0, ey = 0, ez = 0;
  for (int i = lane; i < T.n; i += L) {
    const double4 a = T.at[i];
    const double2 p = T.lj[i];
    for (int k = 0; k < 6; k++) {
      const double dx = site[k][0] - a.x, dy = site[k][1] - a.y, dz = site[k][2] - a.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double r2inv = 1.0 / r2;
      const double r6inv = r2inv * r2inv * r2inv;
      ulj[k] += r6inv * (p.y * r6inv - p.x);
      if (T.pol) {
        const double rinv = sqrt(r2inv);
        uc[k] += qs * a.w * rinv * KCOULD;
      }
    }
    const double dx = pos[0] - a.x, dy = pos[1] - a.y, dz = pos[2] - a.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!n2) {  // the CO2 carbon site has LJ with the mixed parameters
      const double2 c = T.ljc[i];
      const double r2inv = 1.0 / r2;
      const double r6inv = r2inv * r2inv * r2inv;
      uljc += r6inv * (c.y * r6inv - c.x);
    }
    if (T.pol) {
      const double r = sqrt(r2);
      const double rinv = 1.0 / r;
      ucc += qc * a.w * rinv * KCOULD;
      const double r2inv = 1.0 / r2;
      const double r3inv = rinv * r2inv;
      ex += dx * a.w * r3inv;
      ey += dy * a.w * r3inv;
      ez += dz * a.w * r3inv;
    }
  }
  for (int k = 0; k < 6; k++) {
    ulj[k] = wsum(ulj[k]);
    uc[k] = wsum(uc[k]);
  }
  ucc = wsum(ucc);
  uljc = wsum(uljc);
  ex = wsum(ex);
  ey = wsum(ey);
  ez = wsum(ez);
  const double uind = -0.5 * T.alpha * (ex * ex + ey * ey + ez * ez);
  double U[3];
  for (int o = 0; o < 3; o++) {
    U[o] = ulj[2 * o] + ulj[2 * o + 1] + uc[2 * o] + uc[2 * o + 1] + uind;
    if (!n2) U[o] += uljc;
    U[o] += ucc;
  }
  return orient_average(U);
}

// Search down the z axis from |maxZ| + 20 A in 0.5 A steps (src/Equipotential.cpp:86-121).
// out: zMax, final previous position (3), status.
__global__ void zsearch_kernel(EqT T, double z0, long cap, double *out) {
  const int lane = threadIdx.x;
  double pos[3] = {0.0, 0.0, z0}, prev[3] = {0.0, 0.0, 0.0};
  double e = potential(T, lane, pos);
  const double slope = (e + T.ekmin < 0.0) ? -1.0 : 1.0;
  long it = 0;
  for (;;) {
    for (int c = 0; c < 3; c++) prev[c] = pos[c];
    pos[2] -= 0.5 * slope;
    e = potential(T, lane, pos);
    if ((e + T.ekmin) * slope < 0.0) break;
    if (++it > cap) {
      if (lane == 0) out[4] = 1.0;
      return;
    }
  }
  if (lane == 0) {
    out[0] = fmax(pos[2], prev[2]);
    out[1] = prev[0];
    out[2] = prev[1];
    out[3] = prev[2];
    out[4] = 0.0;
  }
}

// One ray per group (src/Equipotential.cpp:128-181).  Per ray: the position
// at the crossing and the previous position, whether the loop ended at its
// first test (then the reference compares with the previous ray's last
// position, resolved on the host), and the previous position at the end.
struct RayOut {
  double hit[3], prev[3];
  int status;  // 0 crossed the axis without a hit, 1 hit, 2 step cap reached
  int first;
};

__global__ void rays_kernel(EqT T, int nrot, double zmax, double zinc, double posmag, long cap,
                            int nrays, RayOut *out) {
  const int ray = blockIdx.x;
  if (ray >= nrays) return;
  const int lane = threadIdx.x;
  const int i = ray / nrot, n = ray % nrot;
  const double zpos = -zmax + i * zinc;
  const double theta = n * 2.0 * M_PI / ((double)nrot);
  const double sx = posmag * cos(theta), sy = posmag * sin(theta);
  double pos[3] = {sx, sy, zpos}, prev[3] = {0.0, 0.0, 0.0};
  double vel[3] = {-0.1 * cos(theta), -0.1 * sin(theta), 0.0};
  double e = potential(T, lane, pos);
  const double slope = (e + T.ekmin < 0.0) ? -1.0 : 1.0;
  for (int c = 0; c < 3; c++) vel[c] *= slope;
  int first = 1, status = 0;
  for (long it = 0;; it++) {
    e = potential(T, lane, pos);
    if ((e + T.ekmin) * slope < 0.0) {
      status = 1;
      break;
    }
    if ((sx == 0 && sy / pos[1] <= 0) || (sy == 0 && sx / pos[0] <= 0) ||
        (sx / pos[0] <= 0 && sy / pos[1] <= 0))
      break;
    if (it > cap) {
      status = 2;
      break;
    }
    for (int c = 0; c < 3; c++) {
      prev[c] = pos[c];
      pos[c] = pos[c] + vel[c];
    }
    first = 0;
  }
  if (lane == 0) {
    RayOut &o = out[ray];
    for (int c = 0; c < 3; c++) {
      o.hit[c] = pos[c];
      o.prev[c] = prev[c];
    }
    o.status = status;
    o.first = first;
  }
}

}  // namespace eqdev

using namespace eqdev;

namespace {
double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
}  // namespace

EquiResult equipotential_ellipsoid(const EquiInput &in) {
  const size_t n = in.x.size();
  if (n == 0) throw Error(MCCS_ERR_INVALID, "target has no atoms");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw Error(MCCS_ERR_DEVICE, "the equipotential ellipsoid is computed on the GPU; no HIP device available");
  int dev = 0;
  EQ_OK(hipGetDevice(&dev));
  // slow-probe kinetic energy, src/Equipotential.cpp:32-37 and :100
  const double v_mb = 2.0 * std::sqrt((2.0 * BOLTZMANN_K * in.temperature) / (M_PI * in.mu * AMU_TO_KG));
  const double v2_mb = 3.0 * BOLTZMANN_K * in.temperature / (in.mu * AMU_TO_KG);
  const double sig_mb = std::sqrt(v2_mb - v_mb * v_mb);
  const double v_min = v_mb - 2.0 * sig_mb;
  const double ekmin = 0.5 * in.mu * AMU_TO_KG * std::pow(v_min, 2.0) * 6.242E18 * 23.061;

  std::vector<double4> at(n);
  std::vector<double2> lj(n), ljc(n);
  for (size_t i = 0; i < n; i++) {
    at[i] = make_double4(in.x[i], in.y[i], in.z[i], in.q[i]);
    const double l1 = 4.0 * in.eps[i] * std::pow(in.sig[i], 6);
    lj[i] = make_double2(l1, l1 * std::pow(in.sig[i], 6));
    const double c1 = 4.0 * in.eps_c[i] * std::pow(in.sig_c[i], 6);
    ljc[i] = make_double2(c1, c1 * std::pow(in.sig_c[i], 6));
  }
  double4 *d_at = nullptr;
  double2 *d_lj = nullptr, *d_ljc = nullptr;
  double *d_z = nullptr;
  RayOut *d_rays = nullptr;
  auto release = [&]() {
    hipFree(d_at);
    hipFree(d_lj);
    hipFree(d_ljc);
    hipFree(d_z);
    hipFree(d_rays);
  };
  EquiResult R;
  try {
    EQ_OK(hipMalloc(&d_at, sizeof(double4) * n));
    EQ_OK(hipMalloc(&d_lj, sizeof(double2) * n));
    EQ_OK(hipMalloc(&d_ljc, sizeof(double2) * n));
    EQ_OK(hipMalloc(&d_z, sizeof(double) * 8));
    EQ_OK(hipMemcpy(d_at, at.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    EQ_OK(hipMemcpy(d_lj, lj.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    EQ_OK(hipMemcpy(d_ljc, ljc.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    EqT T;
    T.at = d_at;
    T.lj = d_lj;
    T.ljc = d_ljc;
    T.n = (int)n;
    T.kind = in.gas == MCCS_GAS_N2 ? 2 : (in.gas == MCCS_GAS_CO2 ? 3 : 1);
    T.pol = in.polarizability ? 1 : 0;
    T.alpha = in.alpha;
    T.ekmin = 0.01 * ekmin;

    // 1. extent along z
    const long cap = 2000000;
    hipLaunchKernelGGL(zsearch_kernel, dim3(1), dim3(L), 0, 0, T, std::abs(in.maxZ) + 20.0, cap, d_z);
    EQ_OK(hipGetLastError());
    double zo[5];
    EQ_OK(hipMemcpy(zo, d_z, sizeof zo, hipMemcpyDeviceToHost));
    if (zo[4] != 0.0) throw Error(MCCS_ERR_INVALID, "equipotential search along z did not end");
    double zmax = zo[0] * 1.3;
    double carry[3] = {zo[1], zo[2], zo[3]};

    // 2. rays
    const int nrot = 20;
    double posmag = std::abs(std::max(std::abs(in.maxX), std::abs(in.maxY)));
    posmag += std::abs(zmax - std::abs(in.maxZ));
    const int levels = (int)std::ceil((double)std::max(25, (int)std::ceil(zmax)));
    const double zinc = 2 * zmax / (levels);
    const int nrays = (levels + 1) * nrot;
    EQ_OK(hipMalloc(&d_rays, sizeof(RayOut) * nrays));
    const long rcap = (long)(2.0 * posmag / 0.1) + 1000;
    hipLaunchKernelGGL(rays_kernel, dim3(nrays), dim3(L), 0, 0, T, nrot, zmax, zinc, posmag, rcap, nrays,
                       d_rays);
    EQ_OK(hipGetLastError());
    std::vector<RayOut> ro(nrays);
    EQ_OK(hipMemcpy(ro.data(), d_rays, sizeof(RayOut) * nrays, hipMemcpyDeviceToHost));
    std::vector<std::array<double, 3>> pts;
    for (int r = 0; r < nrays; r++) {
      const RayOut &o = ro[r];
      if (o.status == 2) throw Error(MCCS_ERR_INVALID, "equipotential ray did not end");
      const double *prev = o.first ? carry : o.prev;
      if (o.status == 1) {
        const double *p = norm3(prev) > norm3(o.hit) ? prev : o.hit;
        if (!(p[0] == 0 && p[1] == 0 && p[2] == 0)) pts.push_back({p[0], p[1], p[2]});
      }
      if (!o.first)
        for (int c = 0; c < 3; c++) carry[c] = o.prev[c];
    }
    release();
    d_at = nullptr; d_lj = nullptr; d_ljc = nullptr; d_z = nullptr; d_rays = nullptr;
    if (pts.size() < 4) throw Error(MCCS_ERR_INVALID, "too few equipotential boundary points to fit an ellipsoid");

    // 3. least-squares ellipsoid x^2 A + y^2 B + z^2 C = 1 (src/Equipotential.cpp:471-560)
    double mu = 0, nu = 0, eps = 0, delta = 0, sigma = 0, rho = 0, al = 0, be = 0, ga = 0;
    for (auto &p : pts) {
      const double x2 = p[0] * p[0], y2 = p[1] * p[1], z2 = p[2] * p[2];
      mu += x2 * x2; nu += y2 * y2; eps += z2 * z2;
      delta += x2 * y2; sigma += x2 * z2; rho += y2 * z2;
      al += x2; be += y2; ga += z2;
    }
    const double det = eps * mu * nu - eps * delta * delta - mu * rho * rho + 2 * delta * rho * sigma -
                       nu * sigma * sigma;
    const double A = (al * (nu * eps - rho * rho) + delta * (rho * ga - be * eps) + sigma * (be * rho - nu * ga)) / det;
    const double B = (mu * (be * eps - ga * rho) + al * (rho * sigma - delta * eps) + sigma * (delta * ga - be * sigma)) / det;
    const double C = (mu * (nu * ga - be * rho) + delta * (be * sigma - delta * ga) + al * (delta * rho - nu * sigma)) / det;
    double a = 1.0 / std::sqrt(A), b = 1.0 / std::sqrt(B), c = 1.0 / std::sqrt(C);
    const double np = (double)pts.size();
    const double rmsd = std::sqrt((A * A * mu + B * B * nu + C * C * eps +
                                   2 * (A * B * delta + B * C * rho + A * C * sigma - A * al - B * be - C * ga) + np) /
                                  (np - 1));
    const double cov_aa = 2 * rmsd * rmsd * (eps * nu - rho * rho) * a * a / (4 * A * A * det);
    const double cov_bb = 2 * rmsd * rmsd * (eps * mu - sigma * sigma) * b * b / (4 * B * B * det);
    const double cov_cc = 2 * rmsd * rmsd * (mu * nu - delta * delta) * c * c / (4 * C * C * det);
    const double cov_ab = 2 * rmsd * rmsd * (rho * sigma - delta * eps) * a * b / (4 * A * B * det);
    const double cov_ac = 2 * rmsd * rmsd * (delta * rho - nu * sigma) * a * c / (4 * A * C * det);
    const double cov_bc = 2 * rmsd * rmsd * (delta * sigma - mu * rho) * b * c / (4 * B * C * det);
    const double aU = (a / A) * std::sqrt((eps * nu - rho * rho) * 2 / det) * rmsd / 2;
    const double bU = (b / B) * std::sqrt((eps * mu - sigma * sigma) * 2 / det) * rmsd / 2;
    const double cU = (c / C) * std::sqrt((mu * nu - delta * delta) * 2 / det) * rmsd / 2;
    const double avg = (a + b + c) / 3.0;
    const double avgU = std::sqrt(cov_aa + cov_bb + cov_cc + 2 * (cov_ab + cov_ac + cov_bc)) / 3.0;
    const double pp = 1.6075;
    const double f = (std::pow(a * b, pp) + std::pow(a * c, pp) + std::pow(b * c, pp)) / 3;
    const double area = 4 * M_PI * std::pow(f, 1.0 / pp);
    const double da = area * std::pow(a, pp - 1) * (std::pow(b, pp) + std::pow(c, pp)) / (3 * f);
    const double db = area * std::pow(b, pp - 1) * (std::pow(a, pp) + std::pow(c, pp)) / (3 * f);
    const double dc = area * std::pow(c, pp - 1) * (std::pow(a, pp) + std::pow(b, pp)) / (3 * f);
    const double areaU = std::sqrt(da * da * cov_aa + db * db * cov_bb + dc * dc * cov_cc +
                                   2 * (da * db * cov_ab + da * dc * cov_ac + db * dc * cov_bc));
    std::ostringstream o;  // Equipotential::print, src/Equipotential.cpp:562-585
    o << "*********************************************************\n";
    o << "Equipotential Ellipsoid: \n";
    o << "*********************************************************\n";
    o << "Axis Length: " << a << "  " << b << "  " << c << "  Ang\n";
    o << "Uncertainy: " << aU << "  " << bU << "  " << cU << "\n";
    o << "Percent Uncert: " << aU * 100 / a << "  " << bU * 100 / b << "  " << cU * 100 / b << "\n";
    o << "RMSD: " << rmsd << "\n";
    o << "Covariance matrix: " << rmsd << "\n";
    o << "{  " << cov_aa << "  " << cov_ab << "  " << cov_ac << "  }\n";
    o << "{  " << cov_ab << "  " << cov_bb << "  " << cov_bc << "  }\n";
    o << "{  " << cov_ac << "  " << cov_bc << "  " << cov_cc << "  }\n";
    o << "Average ellipsoid axes: " << avg << " +/- " << avgU << " Ang\n";
    o << "Ellipsoid surface area: " << area << " +/- " << areaU << " Ang^2\n";
    // 4. grow by 0.01 A until every point is inside (src/Equipotential.cpp:588-631)
    const double aOld = a;
    for (;;) {
      bool outside = false;
      if (a > 3 * aOld) break;
      for (auto &p : pts) {
        const double m = p[0] * p[0] / (a * a) + p[1] * p[1] / (b * b) + p[2] * p[2] / (c * c);
        if (m > 1) {
          outside = true;
          break;
        }
      }
      if (!outside) break;
      a += 0.01;
      b += 0.01;
      c += 0.01;
    }
    o << "Ellipsoid axes further expanded by: " << a - aOld << " Ang\n";
    R.a = a;
    R.b = b;
    R.c = c;
    R.points = pts.size();
    R.report = o.str();
  } catch (...) {
    release();
    throw;
  }
  return R;
}

}  // namespace mccs
