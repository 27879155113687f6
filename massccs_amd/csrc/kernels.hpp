// MI355X (gfx950) trajectory engine: device code shared by the kernel
// instantiation units (traj_gk*.hip) and the launcher (engine.hip).
//
// Work decomposition.  One trajectory is owned by a group of L lanes
// (L = 16, 32 or 64) of one wavefront.  The group integrates the probe with
// group-uniform FP64 state kept in LDS, and its L lanes cooperate on every
// force evaluation: the candidate atoms of the probe's Verlet list are strided
// over the lanes, each lane accumulating every site term in registers; a DPP
// butterfly (and scalar readlanes for 64-lane groups) then leaves the same sum
// in every lane.  Blocks are persistent and pull trajectories from a device
// counter, so trajectories of very different length share the GPU evenly.
//
// Reference behaviour restated here (cces-cepid/massccs):
//   setup                     src/System.cpp:1397-1518
//   run_He / run_N2 / run_CO2 src/System.cpp:338-1392
//   RATTLE                    src/System.cpp:1700-1767
//   site kernels              src/Force.cpp (LJ :34-111, dipole :305-489,
//                             Coulomb :598-723, dummy :903-1081, CO2 :1088-1725)
//   cell index / stencils     src/LinkedCell.cpp:117-129, :228-440
//   tally                     src/System.cpp:287-317
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "engine.hpp"
#include "massccs.h"

#define HIP_OK(x)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess)                                                                     \
      throw ::mccs::Error(MCCS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace mccs {
namespace dev {

constexpr double KCONV = 1.0 / 48.88821291 / 48.88821291;  // KCALMOLANGAMU_TO_ANGFS2
constexpr double KE_CONV = 2390.07;                          // amuAngfs2_to_KCAL_MOL
constexpr double KC = 332.06348078020324;                    // KCOUL
constexpr double PI = 3.14159265358979323846;
constexpr double FAR = 1.0e30;                               // parks an inactive site
constexpr int MAX_STEPS = 4000000;                           // per try

// Launch shape of the group kernel: threads per block (the largest block the
// launcher may use) and wavefronts per SIMD the register budget is sized for.
// Atoms staged in LDS: 768 threads, one block per CU, so that its 12
// wavefronts share one copy of the atoms, and 3 waves per SIMD (168 VGPRs).
// Atoms read through L2: 256 threads at 2 waves per SIMD (231 VGPRs), which
// leaves each group's Verlet list the LDS of a CU shared by 8 groups.
__host__ __device__ constexpr int traj_block(bool lds) { return lds ? 768 : 256; }
__host__ __device__ constexpr int traj_waves(bool lds) { return lds ? 3 : 2; }

// ST_HALF: a step whose first half was made at the end of the previous one
enum : int { ST_NEED = 0, ST_INIT = 1, ST_RESTART = 2, ST_STEP = 3, ST_DONE = 4, ST_HALF = 5 };

// Kernel argument: target, grid and run constants.
struct Tgt {
  const double4 *atoms;     // x, y, z, q in cell order
  const double2 *ljp;       // lj1 = 4 eps sig^6, lj2 = lj1 sig^6 per atom
  const double2 *ljc;       // the same with the CO2 carbon-site parameters
  const int *cstart;        // ncells + 1 prefix of atoms per cell
  const double *vs, *vv;    // speed table: cumulative S_n and speeds v_n
  const int *stencil;       // (n1 + n2) * 3 offsets, first shell then second
  int natoms, nx, ny, nz, ncells, vn, n1, n2;
  int list_cap;             // Verlet-list capacity per group
  double cell, cx, cy, cz;  // cell edge and grid corner
  double hx, hy, hz;        // half box
  double ea, eb, ec, bmax;  // ellipsoid semi-axes, maximal impact parameter
  double iea, ieb, iec;     // their reciprocals (a pre-test of the exit condition only)
  double rlj, rc, rlj2cut, rc2cut;  // sqrt(r2) < rlj exactly when r2 < rlj2cut; same for rc
  double irc3, cc1, cc2;    // 1/rc^3, -1.5/rc, 0.5/rc^3
  double rc6inv;            // 1/rlj^6 (LJ energy shift)
  double alpha, mu, dt, temp, ginertia, d2;
  double inv_cell;          // 1 / cell (fast-path cell arithmetic only)
  double blo[3], bhi[3];    // bounding box of the target atoms
  double rfar2;             // no term reaches beyond this squared distance
  double rlong;             // longest cutoff of the run's terms
  double skin;              // Verlet-list skin (A)
  int co2fix;               // CO2: chi from the true initial velocity (not src/System.cpp:1012)
  // Range-split far field (null when not in use).  Every electrostatic pair
  // term (shifted Coulomb and the smoothed dipole field) is split by a
  // partition of unity S(r) of the pair distance: S = 1 below sw1, 0 beyond
  // sw2, a quintic smoothstep between.  The pair loop sums S times the
  // reference's term over the atoms within sw2 (every pair nearer than sw1
  // exact); the remainder, (1 - S) times the same term summed over every
  // atom, is one smooth scalar field psi(x) = sum_j q_j (1 - S) u_c(r_j) with
  // u_c = 1/r - 1.5/rc + r^2/(2 rc^3), whose value, gradient and Hessian give
  // the Coulomb energy and force of a charged site (q K psi, -q K grad psi)
  // and the far part of the dipole field (E = -grad psi, grad E = -Hess psi).
  // psi is read from cubic B-spline coefficients on a regular grid (FP32:
  // their rounding, 6e-8 of psi, stays far below the interpolation error).
  const float *psi;
  double pox, poy, poz, pih;  // grid origin and 1 / spacing
  int pnx, pny, pnz;
  double sw1, sw2, isw;       // switch start and end, 1 / (sw2 - sw1)
  double sw2cut;              // pairs with r2 < sw2cut carry a near term
};

// ---------------------------------------------------------------- probes
// Which terms each site carries, per gas kind GK (1 atomic He/Ar/co2,
// 2 N2, 3 CO2) and force type FT (src/System.cpp:49-87, :371-379,
// :593-650, :1059-1100).
template <int GK, int FT>
struct Spec {
  static constexpr int NS = GK == 1 ? 1 : 3;
  static constexpr bool CUT = GK == 1 ? (FT == 2 || FT == 4) : (FT == 2 || FT == 4 || FT == 6);
  static constexpr bool lj(int s) { return GK == 2 ? s != 2 : true; }
  static constexpr bool cen(int s) { return GK == 3 && s == 1; }
  static constexpr bool coul(int s) { return GK != 1 && FT >= 3; }
  static constexpr bool dip(int s) {
    return GK == 1 ? (FT == 3 || FT == 4)
                   : ((GK == 2 ? s == 2 : s == 1) && (FT == 5 || FT == 6));
  }
  static constexpr double q(int s) {
    return GK == 2 ? (s == 2 ? 0.965 : -0.4825) : (GK == 3 ? (s == 1 ? 0.6512 : -0.3256) : 0.0);
  }
  static constexpr double m(int s) {
    return GK == 2 ? (s == 2 ? 0.0 : 14.007) : (GK == 3 ? (s == 1 ? 12.011 : 15.999) : 0.0);
  }
  static constexpr double z(int s) {
    return GK == 2 ? (s == 0 ? 0.5488 : (s == 1 ? -0.5488 : 0.0))
                   : (GK == 3 ? (s == 0 ? -1.149 : (s == 1 ? 0.0 : 1.149)) : 0.0);
  }
  // the site the others are symmetric about (N2 dummy, CO2 carbon)
  static constexpr int centre() { return GK == 2 ? 2 : (GK == 3 ? 1 : 0); }
};

// One site of a given role, for per-point evaluation.
template <bool LJ, bool CEN, bool COUL, bool DIP, bool CUTOFF>
struct OneSite {
  static constexpr int NS = 1;
  static constexpr bool CUT = CUTOFF;
  static constexpr bool lj(int) { return LJ; }
  static constexpr bool cen(int) { return CEN; }
  static constexpr bool coul(int) { return COUL; }
  static constexpr bool dip(int) { return DIP; }
  static constexpr int centre() { return 0; }
};

template <class S>
__device__ constexpr bool any_long() {
  bool r = false;
  for (int s = 0; s < S::NS; s++) r = r || S::coul(s) || S::dip(s);
  return r;
}

template <class S>
__device__ constexpr bool any_central() {
  bool r = false;
  for (int s = 0; s < S::NS; s++) r = r || (S::lj(s) && S::cen(s));
  return r;
}

// ---------------------------------------------------------------- lanes
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double swz16(double v) {  // lane ^ 16 inside 32 lanes
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
  hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Butterfly sum over the L lanes of a group; every lane receives the same
// bits (each step adds the same two operands in every lane).
template <int L>
__device__ __forceinline__ double gsum(double v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror: every lane of a row holds the row sum
  if (L == 32) v += swz16(v);
  if (L == 64)  // the four row sums through scalar registers: wave-uniform by construction
    v = (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
  return v;
}

template <int L>
__device__ __forceinline__ uint32_t gsum_u(uint32_t v);

// A work counter's total over the group (see count_of).
template <int L>
__device__ __forceinline__ uint32_t wtotal(uint32_t v) {
  if constexpr (L == 64) return (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  else return gsum_u<L>(v);
}

template <int L>
__device__ __forceinline__ uint32_t gsum_u(uint32_t v) {
#pragma unroll
  for (int d = 1; d < L; d <<= 1) v += (uint32_t)__shfl_xor((int)v, d, L);
  return v;
}

template <int L>
__device__ __forceinline__ int gscan_incl(int v, int lane) {
#pragma unroll
  for (int d = 1; d < L; d <<= 1) {
    int t = __shfl_up(v, d, L);
    if (lane >= d) v += t;
  }
  return v;
}

// Bits of a wave ballot that belong to this lane's group, and the count of
// set bits below this lane.
template <int L>
__device__ __forceinline__ void group_ballot(bool flag, int lane, int &below, int &total) {
  const unsigned long long m = __ballot(flag);
  const int gbase = (int)(threadIdx.x % 64) - lane;
  const unsigned long long gm = (L == 64) ? m : ((m >> gbase) & ((1ull << L) - 1ull));
  total = __popcll(gm);
  // a whole wavefront counts with mbcnt (no lane mask held in registers)
  below = (L == 64) ? (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))
                    : __popcll(gm & ((1ull << lane) - 1ull));
}

// Bits of a wave ballot that belong to this lane's group, bit i = group lane i.
template <int L>
__device__ __forceinline__ unsigned long long group_mask(bool flag, int lane) {
  const unsigned long long m = __ballot(flag);
  const int gbase = (int)(threadIdx.x % 64) - lane;
  return (L == 64) ? m : ((m >> gbase) & ((1ull << L) - 1ull));
}

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- forces
struct Acc {
  double U, fx, fy, fz, Ex, Ey, Ez, Exx, Eyy, Ezz, Exy, Exz, Eyz;
};

__device__ __forceinline__ void acc_zero(Acc &a) {
  a.U = a.fx = a.fy = a.fz = 0.0;
  a.Ex = a.Ey = a.Ez = a.Exx = a.Eyy = a.Ezz = a.Exy = a.Exz = a.Eyz = 0.0;
}

struct Work {
  uint32_t visits, lj, coul, dip;
};

struct GState;

struct Shm {
  GState *state;          // the group's trajectory state
  void *list;             // the group's Verlet list (uint16 indices when atoms are staged, else int)
  int *incl;              // per-lane inclusive run ends of the group
  int *base;              // per-lane (run start - run offset) * 2 + first-shell flag
  double *red;            // 64-lane groups: reduction rows (shares incl/base's storage)
  const double4 *atoms;   // LDS copies when staged
  const double2 *ljp, *ljc;
};

template <bool LDS, bool CEN>
__device__ __forceinline__ void fetch(const Tgt &T, const Shm &sh, int a, double4 &at, double2 &lp,
                                      double2 &lc) {
  if (LDS) {
    at = sh.atoms[a];
    lp = sh.ljp[a];
    if (CEN) lc = sh.ljc[a];
  } else {
    at = T.atoms[a];
    lp = T.ljp[a];
    if (CEN) lc = T.ljc[a];
  }
}

// Pair terms of one candidate atom against every site.  allow_lj is 0 for
// second-shell cells of an out-of-grid site (the reference applies LJ on the
// first shell only).
// 1/sqrt(x) for x > 0 (a pair distance squared): the hardware estimate with a
// second-order Newton-Raphson correction, y = y0 + y0 e (1/2 + 3/8 e) with
// e = 1 - x y0^2 (full FP64 precision), without the special-value handling of
// the library rsqrt (x is never 0, infinite or NaN here).
__device__ __forceinline__ double rsq_pos(double x) {
  const double y0 = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y0, y0, 1.0);
  return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// Work counters: per lane (summed over the group at the end of a try), or,
// for whole-wavefront groups (UNI), the wavefront's count from a ballot,
// accumulated uniformly in scalar registers (lane 0 takes part in every
// iteration of the strided loops, so its count is the total).
template <bool UNI>
__device__ __forceinline__ uint32_t count_of(bool c) {
  return UNI ? (uint32_t)__popcll(__ballot(c)) : (uint32_t)c;
}

template <class S, bool COUNT, bool UNI = false, bool SPLIT = false>
__device__ __forceinline__ void pair_terms(const Tgt &T, const double (&p)[S::NS][3],
                                           const double (&qs)[S::NS], const double4 at,
                                           const double2 lp, const double2 lc, int allow_lj,
                                           Acc (&a)[S::NS], Work &w) {
  // per-atom factors shared by the sites
  const double qk = at.w * KC;       // Coulomb: q_j K (times the site charge below)
  const double m3q = -3.0 * at.w;    // field gradient: -3 q_j
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (!(S::lj(s) || S::coul(s) || S::dip(s))) continue;
    const double dx = p[s][0] - at.x, dy = p[s][1] - at.y, dz = p[s][2] - at.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (COUNT) w.visits += count_of<UNI>(true);
    const double lj1 = S::cen(s) ? lc.x : lp.x, lj2 = S::cen(s) ? lc.y : lp.y;
    if (S::CUT) {
      // r < cutoff as the reference tests it (sqrt(r2) < rc): with a correctly
      // rounded square root that holds exactly for r2 below the smallest
      // double whose square root reaches the cutoff (set on the host)
      bool in_lj = false, in_c = false;
      if (S::lj(s)) in_lj = (allow_lj != 0) & (r2 < T.rlj2cut);
      if (S::coul(s) || S::dip(s)) in_c = r2 < T.rc2cut;
      if (COUNT) {
        if (S::lj(s)) w.lj += count_of<UNI>(in_lj);
        if (!SPLIT && S::coul(s)) w.coul += count_of<UNI>(in_c);
        if (!SPLIT && S::dip(s)) w.dip += count_of<UNI>(in_c);
      }
      // Coulomb and dipole terms without a branch: most listed atoms are
      // inside their cutoff, and outside it the weights (qq, A, B) are zero,
      // so those lanes add zeros (the terms of the others are unchanged)
      const bool any_c = S::coul(s) || S::dip(s);
      if (any_c || (S::lj(s) && allow_lj)) {
        const double rinv = rsq_pos(r2);
        const double r2inv = rinv * rinv;
        if (S::lj(s) && allow_lj) {
          // list strides past the inner part skip this as a whole; inside it,
          // atoms beyond the LJ cutoff get zero coefficients (zero terms)
          const double l1 = in_lj ? lj1 : 0.0, l2 = in_lj ? lj2 : 0.0;
          const double r6 = r2inv * r2inv * r2inv;
          const double ucut = T.rc6inv * (l2 * T.rc6inv - l1);
          a[s].U += r6 * (l2 * r6 - l1) - ucut;
          const double flj = r6 * (12.0 * l2 * r6 - 6.0 * l1) * r2inv;
          a[s].fx += flj * dx;
          a[s].fy += flj * dy;
          a[s].fz += flj * dz;
        }
        if (any_c && SPLIT) {
          // near part of the range-split terms: the reference's shifted
          // Coulomb potential q u(r), u = 1/r - 1.5/rc + r^2/(2 rc^3), and its
          // smoothed dipole field (the gradient of u) times the partition S(r);
          // for S = 1 (r <= sw1) these are the reference's terms exactly
          const bool in_s = r2 < T.sw2cut;
          if (COUNT) {
            if (S::coul(s)) w.coul += count_of<UNI>(in_s);
            if (S::dip(s)) w.dip += count_of<UNI>(in_s);
          }
          const double r = r2 * rinv, r3 = rinv * r2inv;
          const double t = fmin(fmax((r - T.sw1) * T.isw, 0.0), 1.0), t2 = t * t, omt = 1.0 - t;
          const double Sv = 1.0 - t2 * t * fma(t, fma(6.0, t, -15.0), 10.0);
          const double S1 = -30.0 * t2 * omt * omt * T.isw;
          const double S2 = -60.0 * t * omt * (1.0 - 2.0 * t) * (T.isw * T.isw);
          const double uc = rinv + T.cc1 + T.cc2 * r2;  // u
          const double u1 = r * T.irc3 - r2inv;         // u'
          const double u2 = 2.0 * r3 + T.irc3;          // u''
          const double ph1 = S1 * uc + Sv * u1;
          const double ph2 = S2 * uc + 2.0 * S1 * u1 + Sv * u2;
          // per unit charge: potential, E = Ar d, grad E = Ar I + Br d d^T
          const double ph = in_s ? Sv * uc : 0.0;
          const double Ar = in_s ? -ph1 * rinv : 0.0;
          const double Br = in_s ? (ph1 * rinv - ph2) * r2inv : 0.0;
          if (S::coul(s)) {
            const double qq = qs[s] * qk;
            a[s].U += qq * ph;
            const double fc = qq * Ar;
            a[s].fx += fc * dx;
            a[s].fy += fc * dy;
            a[s].fz += fc * dz;
          }
          if (S::dip(s)) {
            const double A = at.w * Ar, B = at.w * Br;
            a[s].Ex += dx * A;
            a[s].Ey += dy * A;
            a[s].Ez += dz * A;
            const double bx = B * dx, by = B * dy, bz = B * dz;
            a[s].Exx += A + bx * dx;
            a[s].Eyy += A + by * dy;
            a[s].Ezz += A + bz * dz;
            a[s].Exy += bx * dy;
            a[s].Exz += bx * dz;
            a[s].Eyz += by * dz;
          }
        } else if (any_c) {
          // shifted force r^-3 (1 - r^3/rc^3) = r^-3 - rc^-3 (src/Force.cpp:671-683 and
          // the smoothed field of :393-426)
          const double r3 = rinv * r2inv;
          const double r3s = r3 - T.irc3;
          if (S::coul(s)) {
            const double qq = in_c ? qs[s] * qk : 0.0;
            a[s].U += qq * (rinv + T.cc1 + T.cc2 * r2);
            const double fc = qq * r3s;
            a[s].fx += fc * dx;
            a[s].fy += fc * dy;
            a[s].fz += fc * dz;
          }
          if (S::dip(s)) {
            const double A = in_c ? at.w * r3s : 0.0;
            const double B = in_c ? m3q * r3 * r2inv : 0.0;
            a[s].Ex += dx * A;
            a[s].Ey += dy * A;
            a[s].Ez += dz * A;
            const double bx = B * dx, by = B * dy, bz = B * dz;
            a[s].Exx += A + bx * dx;
            a[s].Eyy += A + by * dy;
            a[s].Ezz += A + bz * dz;
            a[s].Exy += bx * dy;
            a[s].Exz += bx * dz;
            a[s].Eyz += by * dz;
          }
        }
      }
    } else {
      // all-atom kernels (no cutoff, no shifts): src/Force.cpp:117-300, :495-574, :836-897
      if (COUNT) {
        const uint32_t c = count_of<UNI>(true);
        if (S::lj(s)) w.lj += c;
        if (S::coul(s)) w.coul += c;
        if (S::dip(s)) w.dip += c;
      }
      const double rinv = rsqrt(r2);
      const double r2inv = rinv * rinv;
      if (S::lj(s)) {
        const double r6 = r2inv * r2inv * r2inv;
        a[s].U += r6 * (lj2 * r6 - lj1);
        const double flj = r6 * (12.0 * lj2 * r6 - 6.0 * lj1) * r2inv;
        a[s].fx += flj * dx;
        a[s].fy += flj * dy;
        a[s].fz += flj * dz;
      }
      if (S::coul(s)) {
        const double uc = qs[s] * at.w * KC * rinv;
        a[s].U += uc;
        const double fc = uc * r2inv;
        a[s].fx += fc * dx;
        a[s].fy += fc * dy;
        a[s].fz += fc * dz;
      }
      if (S::dip(s)) {
        const double r3 = rinv * r2inv;
        const double A = at.w * r3;
        const double B = -3.0 * at.w * r3 * r2inv;
        a[s].Ex += dx * A;
        a[s].Ey += dy * A;
        a[s].Ez += dz * A;
        a[s].Exx += A + B * dx * dx;
        a[s].Eyy += A + B * dy * dy;
        a[s].Ezz += A + B * dz * dz;
        a[s].Exy += B * dx * dy;
        a[s].Exz += B * dx * dz;
        a[s].Eyz += B * dy * dz;
      }
    }
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Candidate runs.  Mode 0: one run per cell column (i, j) that the sphere of
// radius sqrt(R2) around c reaches, covering the z-range of cells it cuts;
// atoms are stored in cell order with k fastest, so a column range is
// contiguous.  Mode 1 (site outside the grid): the reference stencils of the
// cell the reference's flat index aliases to.  Mode 2: all atoms
// (no-cutoff force types).
struct RunGen {
  int mode;
  double c[3], R2;
  int ix0, iy0, nys;       // mode 0
  int ci, cj, ck, nfirst;  // mode 1
  __device__ __forceinline__ void get(const Tgt &T, int r, int &start, int &len, int &first) const {
    start = 0;
    len = 0;
    first = 1;
    if (mode == 0) {
      const int i = ix0 + r / nys, j = iy0 + r % nys;
      const double lox = T.cx + i * T.cell, loy = T.cy + j * T.cell;
      const double dxm = fmax(0.0, fmax(lox - c[0], c[0] - (lox + T.cell)));
      const double dym = fmax(0.0, fmax(loy - c[1], c[1] - (loy + T.cell)));
      const double d2 = dxm * dxm + dym * dym;
      if (d2 >= R2) return;
      const double h = sqrt(R2 - d2);
      const int k0 = clampi((int)floor((c[2] - h - T.cz) * T.inv_cell), 0, T.nz - 1);
      const int k1 = clampi((int)floor((c[2] + h - T.cz) * T.inv_cell), 0, T.nz - 1);
      const int b = T.nz * (j + T.ny * i);
      start = T.cstart[b + k0];
      len = T.cstart[b + k1 + 1] - start;
    } else if (mode == 1) {
      const int *o = T.stencil + 3 * r;
      const int a = ci + o[0], b = cj + o[1], e = ck + o[2];
      first = r < nfirst;
      if (a < 0 || a >= T.nx || b < 0 || b >= T.ny || e < 0 || e >= T.nz) return;
      const int cell = e + T.nz * b + T.nz * T.ny * a;
      start = T.cstart[cell];
      len = T.cstart[cell + 1] - start;
    } else {
      start = 0;
      len = T.natoms;
    }
  }
};

// Walks the runs of one pass, flattening them over the group's lanes, and
// hands every candidate atom index (with its first-shell flag) to visit().
template <int L, class Visit>
__device__ void walk_runs(const Tgt &T, const Shm &sh, int lane, const RunGen &g, int nruns,
                          Visit &&visit) {
  for (int r0 = 0; r0 < nruns; r0 += L) {
    int start = 0, len = 0, first = 1;
    if (r0 + lane < nruns) g.get(T, r0 + lane, start, len, first);
    const int incl = gscan_incl<L>(len, lane);
    const int total = __shfl(incl, L - 1, L);
    sh.incl[lane] = incl;
    sh.base[lane] = (start - (incl - len)) * 2 + first;
    lds_fence();
    int cur = 0;
    for (int t = lane; t < total; t += L) {
      while (cur < L - 1 && t >= sh.incl[cur]) cur++;
      const int bw = sh.base[cur];
      visit((bw >> 1) + t, bw & 1);
    }
    lds_fence();
  }
}

// Sums of 64 lanes through LDS, four values at a time: row i of the group's
// scratch holds value i of every lane; lane 16r+q adds the four entries
// [4q, 4q+4) of row r, the 16 lanes of a row combine their partial sums with
// four DPP steps, and the totals are read from lanes 0, 16, 32, 48 into
// scalar registers (every lane receives the same bits; the order of every
// addition is fixed).
constexpr int RROWS = 4, RSTRIDE = 66;  // rows per pass, doubles per row (64 + pad)
template <int K>
__device__ __forceinline__ void sum64(double *const (&v)[K], double *red, int lane) {
#pragma unroll
  for (int c0 = 0; c0 < K; c0 += RROWS) {
    lds_fence();
#pragma unroll
    for (int i = 0; i < RROWS; i++)
      if (c0 + i < K) red[i * RSTRIDE + lane] = *v[c0 + i];
    lds_fence();
    const int r = lane >> 4, q = lane & 15;
    double x = 0.0;
    if (c0 + r < K) {
      const double2 *row = reinterpret_cast<const double2 *>(red + r * RSTRIDE + 4 * q);
      const double2 a0 = row[0], a1 = row[1];
      x = (a0.x + a0.y) + (a1.x + a1.y);
    }
    x += dpp<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp<0x141>(x);  // row_half_mirror
    x += dpp<0x140>(x);  // row_mirror: the 16 lanes of a row hold its sum
#pragma unroll
    for (int i = 0; i < RROWS; i++)
      if (c0 + i < K) *v[c0 + i] = readlane_d(x, 16 * i);
  }
  lds_fence();
}

template <class S, int L>
__device__ __forceinline__ void reduce(Acc (&a)[S::NS], double *red) {
  if constexpr (L == 64) {
    constexpr int K = [] {
      int k = 0;
      for (int s = 0; s < S::NS; s++) {
        if (S::lj(s) || S::coul(s) || S::dip(s)) k += 4;
        if (S::dip(s)) k += 9;
      }
      return k;
    }();
    double *v[K];
    int k = 0;
#pragma unroll
    for (int s = 0; s < S::NS; s++) {
      if (S::lj(s) || S::coul(s) || S::dip(s)) {
        v[k++] = &a[s].U; v[k++] = &a[s].fx; v[k++] = &a[s].fy; v[k++] = &a[s].fz;
      }
      if (S::dip(s)) {
        v[k++] = &a[s].Ex; v[k++] = &a[s].Ey; v[k++] = &a[s].Ez;
        v[k++] = &a[s].Exx; v[k++] = &a[s].Eyy; v[k++] = &a[s].Ezz;
        v[k++] = &a[s].Exy; v[k++] = &a[s].Exz; v[k++] = &a[s].Eyz;
      }
    }
    sum64<K>(v, red, threadIdx.x & 63);
    return;
  }
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (S::lj(s) || S::coul(s) || S::dip(s)) {
      a[s].U = gsum<L>(a[s].U);
      a[s].fx = gsum<L>(a[s].fx);
      a[s].fy = gsum<L>(a[s].fy);
      a[s].fz = gsum<L>(a[s].fz);
    }
    if (S::dip(s)) {
      a[s].Ex = gsum<L>(a[s].Ex);
      a[s].Ey = gsum<L>(a[s].Ey);
      a[s].Ez = gsum<L>(a[s].Ez);
      a[s].Exx = gsum<L>(a[s].Exx);
      a[s].Eyy = gsum<L>(a[s].Eyy);
      a[s].Ezz = gsum<L>(a[s].Ezz);
      a[s].Exy = gsum<L>(a[s].Exy);
      a[s].Exz = gsum<L>(a[s].Exz);
      a[s].Eyz = gsum<L>(a[s].Eyz);
    }
  }
}

template <class S>
__device__ __forceinline__ void finish(const Tgt &T, const Acc (&a)[S::NS], int only,
                                       double (&f)[S::NS][3], double (&U)[S::NS]) {
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (only >= 0 && s != only) continue;
    if (S::dip(s)) {
      f[s][0] = a[s].fx + T.alpha * (a[s].Ex * a[s].Exx + a[s].Ey * a[s].Exy + a[s].Ez * a[s].Exz);
      f[s][1] = a[s].fy + T.alpha * (a[s].Ex * a[s].Exy + a[s].Ey * a[s].Eyy + a[s].Ez * a[s].Eyz);
      f[s][2] = a[s].fz + T.alpha * (a[s].Ex * a[s].Exz + a[s].Ey * a[s].Eyz + a[s].Ez * a[s].Ezz);
      U[s] = a[s].U - 0.5 * T.alpha * (a[s].Ex * a[s].Ex + a[s].Ey * a[s].Ey + a[s].Ez * a[s].Ez);
    } else {
      f[s][0] = a[s].fx;
      f[s][1] = a[s].fy;
      f[s][2] = a[s].fz;
      U[s] = a[s].U;
    }
  }
}

// Verlet list of one group: every atom within R of the centre c0 it was built
// around.  It serves while |c - c0| + off <= R - rlong, i.e. while every atom
// that can be inside the longest cutoff of any site is on the list.
// The list holds first the atoms within Rin (the LJ cutoff plus the same
// margin) and then the others: only the first nin entries can be inside the
// LJ cutoff of a site while the list serves.
struct VList {
  double c0[3];
  double R;
  int n, nin;
  int valid;  // -1 none, 0 overflowed its capacity (walk the columns instead), 1 usable
};

template <bool LDS>
using LIdx = typename std::conditional<LDS, uint16_t, int>::type;

template <class S, int L, bool LDS>
__device__ void build_list(const Tgt &T, const Shm &sh, int lane, VList &vl, const double c[3],
                           double R, double Rin) {
  LIdx<LDS> *lst = static_cast<LIdx<LDS> *>(sh.list);
  const int cap = T.list_cap;
  const double R2 = R * R, Rin2 = Rin * Rin;
  int nin = 0, nout = 0;
  // one pass: atoms within Rin fill the list from the front, the others
  // within R from the back (entry t >= nin is at cap - 1 - (t - nin))
  auto take = [&](bool valid, int a) {
    bool kin = false, kout = false;
    double4 at = make_double4(0.0, 0.0, 0.0, 0.0);
    if (valid) {
      at = LDS ? sh.atoms[a] : T.atoms[a];
      const double ex = at.x - c[0], ey = at.y - c[1], ez = at.z - c[2];
      const double d2 = ex * ex + ey * ey + ez * ez;
      kin = d2 < Rin2;
      kout = !kin && d2 < R2;
    }
    int bin, cin, bout, cout;
    group_ballot<L>(kin, lane, bin, cin);
    group_ballot<L>(kout, lane, bout, cout);
    if (kin && nin + bin < cap) lst[nin + bin] = (LIdx<LDS>)a;
    if (kout && nout + bout < cap) lst[cap - 1 - (nout + bout)] = (LIdx<LDS>)a;
    nin += cin;
    nout += cout;
  };
  if (LDS) {
    // a target staged in LDS (at most a few thousand atoms): every atom, in
    // index (cell) order; every lane of the group runs the same iterations
    for (int t0 = 0; t0 < T.natoms; t0 += L) take(t0 + lane < T.natoms, t0 + lane);
  } else {
    // the cell columns that the sphere of radius R around c reaches
    RunGen g;
    g.mode = 0;
    g.c[0] = c[0];
    g.c[1] = c[1];
    g.c[2] = c[2];
    g.R2 = R2;
    const int ix0 = clampi((int)floor((c[0] - R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int ix1 = clampi((int)floor((c[0] + R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int iy0 = clampi((int)floor((c[1] - R - T.cy) * T.inv_cell), 0, T.ny - 1);
    const int iy1 = clampi((int)floor((c[1] + R - T.cy) * T.inv_cell), 0, T.ny - 1);
    g.ix0 = ix0;
    g.iy0 = iy0;
    g.nys = iy1 - iy0 + 1;
    // every lane of the group runs the same number of visits per chunk of L
    // candidates, so the ballots see the whole group
    for (int r0 = 0; r0 < (ix1 - ix0 + 1) * g.nys; r0 += L) {
      int start = 0, len = 0, first = 1;
      if (r0 + lane < (ix1 - ix0 + 1) * g.nys) g.get(T, r0 + lane, start, len, first);
      const int incl = gscan_incl<L>(len, lane);
      const int total = __shfl(incl, L - 1, L);
      sh.incl[lane] = incl;
      sh.base[lane] = start - (incl - len);
      lds_fence();
      int cur = 0;
      for (int t0 = 0; t0 < total; t0 += L) {
        const int t = t0 + lane;
        int a = 0;
        if (t < total) {
          while (cur < L - 1 && t >= sh.incl[cur]) cur++;
          a = sh.base[cur] + t;
        }
        take(t < total, a);
      }
      lds_fence();
    }
  }
  vl.nin = nin;
  vl.c0[0] = c[0];
  vl.c0[1] = c[1];
  vl.c0[2] = c[2];
  vl.R = R;
  vl.n = nin + nout;
  vl.valid = nin + nout <= cap;
}

// Uniform cubic B-spline basis piece i (0..3) at t in [0, 1), with its first
// and second derivatives in t.  Branch-free (the lanes of a wavefront take
// different pieces): every piece is formed and the lane's is selected.
__device__ __forceinline__ void bspline_piece(int i, double t, double &w, double &dw, double &ddw) {
  const double u = 1.0 - t, t2 = t * t, t3 = t2 * t;
  const double w0 = u * u * u * (1.0 / 6.0), w1 = (3.0 * t3 - 6.0 * t2 + 4.0) * (1.0 / 6.0);
  const double w2 = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) * (1.0 / 6.0), w3 = t3 * (1.0 / 6.0);
  const double d0 = -0.5 * u * u, d1 = 1.5 * t2 - 2.0 * t, d2 = -1.5 * t2 + t + 0.5, d3 = 0.5 * t2;
  const double e0 = u, e1 = 3.0 * t - 2.0, e2 = 1.0 - 3.0 * t, e3 = t;
  w = i == 0 ? w0 : (i == 1 ? w1 : (i == 2 ? w2 : w3));
  dw = i == 0 ? d0 : (i == 1 ? d1 : (i == 2 ? d2 : d3));
  ddw = i == 0 ? e0 : (i == 1 ? e1 : (i == 2 ? e2 : e3));
}

// Piece i (0..3) of the uniform cubic B-spline in FP32 (the far field's
// interpolation, whose own error is ~1e-3 of the field, runs in single
// precision), as the power-basis coefficients w(t) = c0 + c1 t + c2 t^2 +
// c3 t^3: (1 - t)^3 / 6, (3 t^3 - 6 t^2 + 4) / 6, (-3 t^3 + 3 t^2 + 3 t + 1) / 6,
// t^3 / 6.  A lane's piece index is fixed, so the coefficients are selected
// once per evaluation and each site costs six FMAs per axis.
struct BPiece {
  float c0, c1, c2, c3;
};
__device__ __forceinline__ BPiece bspline_coef_f(int i) {
  BPiece b;
  b.c0 = i == 0 ? 1.0f / 6.0f : (i == 1 ? 4.0f / 6.0f : (i == 2 ? 1.0f / 6.0f : 0.0f));
  b.c1 = i == 0 ? -0.5f : (i == 2 ? 0.5f : 0.0f);
  b.c2 = i == 0 ? 0.5f : (i == 1 ? -1.0f : (i == 2 ? 0.5f : 0.0f));
  b.c3 = i == 0 ? -1.0f / 6.0f : (i == 1 ? 0.5f : (i == 2 ? -0.5f : 1.0f / 6.0f));
  return b;
}
// w(t), w'(t) and w''(t) of a piece.
__device__ __forceinline__ void bspline_eval_f(const BPiece &b, float t, float &w, float &dw, float &ddw) {
  w = fmaf(fmaf(fmaf(b.c3, t, b.c2), t, b.c1), t, b.c0);
  dw = fmaf(fmaf(3.0f * b.c3, t, 2.0f * b.c2), t, b.c1);
  ddw = fmaf(6.0f * b.c3, t, 2.0f * b.c2);
}

// Far part of the range-split field at the sites: the L lanes of a group
// take the 4 x 4 x 4 B-spline nodes around each site, 64 / L nodes per lane
// (node k = lane + L h: x offset k / 16, y (k / 4) % 4, z k % 4).  The
// coefficients are loaded first (far_load, issued before the pair loop so
// that the loads overlap it) and applied after it (far_apply): each lane adds
// its nodes' share of psi and its derivatives to the site's accumulators,
// which the group reduction then sums together with the near terms.  A
// charged site gets q K psi and -q K grad psi, a polarisable one
// E -= grad psi and grad E -= Hess psi.  Nodes outside the grid hold zero
// (the field vanishes there).
template <int NPL>
struct FarNode {
  float t[3];     // position within the cell of the site, in grid units
  float c[NPL];   // this lane's coefficients
  bool on;
};

// The coefficients a trajectory read at its last evaluation, by site: a probe
// moves about a tenth of a grid spacing per step, so most steps need the same
// 4 x 4 x 4 nodes again.  b[s][0] = INT_MIN marks an empty entry.
template <int NS, int NPL>
struct FarCache {
  int b[NS][3];
  float c[NS][NPL];
  __device__ void clear() {
    for (int s = 0; s < NS; s++) b[s][0] = INT_MIN;
  }
};

template <class S, int L>
__device__ __forceinline__ void far_load(const Tgt &T, int lane, const double (&p)[S::NS][3],
                                         FarNode<64 / L> (&fn)[S::NS], int only,
                                         FarCache<S::NS, 64 / L> *cache = nullptr) {
  constexpr int NPL = 64 / L;
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    fn[s].on = false;
#pragma unroll
    for (int h = 0; h < NPL; h++) fn[s].c[h] = 0.0f;
    if (!(S::coul(s) || S::dip(s))) continue;
    if (only >= 0 && s != only) continue;
    const double ux = (p[s][0] - T.pox) * T.pih, uy = (p[s][1] - T.poy) * T.pih,
                 uz = (p[s][2] - T.poz) * T.pih;
    if (!(fabs(ux) < 1e9 && fabs(uy) < 1e9 && fabs(uz) < 1e9)) continue;  // parked or non-finite
    const double fx = floor(ux), fy = floor(uy), fz = floor(uz);
    fn[s].on = true;
    fn[s].t[0] = (float)(ux - fx);
    fn[s].t[1] = (float)(uy - fy);
    fn[s].t[2] = (float)(uz - fz);
    const int bx = (int)fx - 1, by = (int)fy - 1, bz = (int)fz - 1;
    if (cache && cache->b[s][0] == bx && cache->b[s][1] == by && cache->b[s][2] == bz) {
#pragma unroll
      for (int h = 0; h < NPL; h++) fn[s].c[h] = cache->c[s][h];
      continue;
    }
    const long iy = (long)by + ((lane >> 2) & 3), iz = (long)bz + (lane & 3);
#pragma unroll
    for (int h = 0; h < NPL; h++) {
      const long ix = (long)bx + ((lane + L * h) >> 4);
      if (ix >= 0 && ix < T.pnx && iy >= 0 && iy < T.pny && iz >= 0 && iz < T.pnz)
        fn[s].c[h] = T.psi[((size_t)ix * T.pny + (size_t)iy) * T.pnz + (size_t)iz];
    }
    if (cache) {
      cache->b[s][0] = bx;
      cache->b[s][1] = by;
      cache->b[s][2] = bz;
#pragma unroll
      for (int h = 0; h < NPL; h++) cache->c[s][h] = fn[s].c[h];
    }
  }
}

template <class S, int L>
__device__ __forceinline__ void far_apply(const Tgt &T, int lane, const FarNode<64 / L> (&fn)[S::NS],
                                          const double (&qs)[S::NS], Acc (&a)[S::NS]) {
  constexpr int NPL = 64 / L;
  const float h1 = (float)T.pih, h2 = h1 * h1;
  BPiece bx[NPL];
#pragma unroll
  for (int h = 0; h < NPL; h++) bx[h] = bspline_coef_f((lane + L * h) >> 4);
  const BPiece by = bspline_coef_f((lane >> 2) & 3), bz = bspline_coef_f(lane & 3);
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (!(S::coul(s) || S::dip(s))) continue;
    if (!fn[s].on) continue;
    float wy, dwy, ddwy, wz, dwz, ddwz;
    bspline_eval_f(by, fn[s].t[1], wy, dwy, ddwy);
    bspline_eval_f(bz, fn[s].t[2], wz, dwz, ddwz);
#pragma unroll
    for (int h = 0; h < NPL; h++) {
      const float c = fn[s].c[h];
      float wx, dwx, ddwx;
      bspline_eval_f(bx[h], fn[s].t[0], wx, dwx, ddwx);
      const float cyz = c * wy * wz, cxz = c * wx * wz, cxy = c * wx * wy;
      const double gx = (double)(dwx * cyz * h1), gy = (double)(dwy * cxz * h1),
                   gz = (double)(dwz * cxy * h1);
      if (S::coul(s)) {
        const double qk = qs[s] * KC;
        a[s].U += qk * (double)(wx * cyz);
        a[s].fx -= qk * gx;
        a[s].fy -= qk * gy;
        a[s].fz -= qk * gz;
      }
      if (S::dip(s)) {
        a[s].Ex -= gx;
        a[s].Ey -= gy;
        a[s].Ez -= gz;
        a[s].Exx -= (double)(ddwx * cyz * h2);
        a[s].Eyy -= (double)(ddwy * cxz * h2);
        a[s].Ezz -= (double)(ddwz * cxy * h2);
        a[s].Exy -= (double)(dwx * dwy * (c * wz) * h2);
        a[s].Exz -= (double)(dwx * dwz * (c * wy) * h2);
        a[s].Eyz -= (double)(dwy * dwz * (c * wx) * h2);
      }
    }
  }
}

// Forces and energies of all sites at positions p (group-uniform result).
template <class S, int L, bool LDS, bool COUNT, bool SPLIT = false>
__device__ void evaluate(const Tgt &T, const Shm &sh, int lane, const double (&p)[S::NS][3],
                         const double (&qs)[S::NS], double (&f)[S::NS][3], double (&U)[S::NS],
                         Work &w, VList *vl, FarCache<S::NS, 64 / L> *fcache = nullptr) {
  constexpr bool CEN = any_central<S>();
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    f[s][0] = f[s][1] = f[s][2] = 0.0;
    U[s] = 0.0;
  }
  Acc a[S::NS];
#pragma unroll
  for (int s = 0; s < S::NS; s++) acc_zero(a[s]);
  auto visit = [&](int ai, int first) {
    double4 at;
    double2 lp, lc;
    fetch<LDS, CEN>(T, sh, ai, at, lp, lc);
    pair_terms<S, COUNT, L == 64, SPLIT>(T, p, qs, at, lp, lc, first, a, w);
  };
  RunGen g;
  if (!S::CUT) {
    g.mode = 2;
    walk_runs<L>(T, sh, lane, g, 1, visit);
    reduce<S, L>(a, sh.red);
    finish<S>(T, a, -1, f, U);
    return;
  }
  // Exact zero: every site at least the longest cutoff away from the target's
  // bounding box, so no atom passes any cutoff test (also for aliased cells).
  bool near = false;
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double e = fmax(0.0, fmax(T.blo[d] - p[s][d], p[s][d] - T.bhi[d]));
      d2 += e * e;
    }
    near = near || d2 < T.rfar2;
  }
  if (!near) return;
  // every site's cell index inside the grid: 0 <= floor(u) < n exactly when
  // 0 <= u < n for an integer n (false for NaN), so no floor is needed
  bool inbox = true;
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    const double i = (p[s][0] - T.cx) * T.inv_cell;
    const double j = (p[s][1] - T.cy) * T.inv_cell;
    const double k = (p[s][2] - T.cz) * T.inv_cell;
    inbox = inbox && i >= 0.0 && i < T.nx && j >= 0.0 && j < T.ny && k >= 0.0 && k < T.nz;
  }
  if (inbox) {
    constexpr int cs = S::centre();
    double off2 = 0.0;
#pragma unroll
    for (int s = 0; s < S::NS; s++) {
      const double ex = p[s][0] - p[cs][0], ey = p[s][1] - p[cs][1], ez = p[s][2] - p[cs][2];
      off2 = fmax(off2, ex * ex + ey * ey + ez * ez);
    }
    const double off = S::NS == 1 ? 0.0 : sqrt(off2);  // one site is the centre (finite here)
    const double c[3] = {p[cs][0], p[cs][1], p[cs][2]};
    if (vl) {
      // rebuild when moved + off + rlong + 1e-9 > R, compared in squares
      const double e0 = c[0] - vl->c0[0], e1 = c[1] - vl->c0[1], e2 = c[2] - vl->c0[2];
      const double room = vl->R - (off + T.rlong + 1e-9);
      if (vl->valid < 0 || room < 0.0 || e0 * e0 + e1 * e1 + e2 * e2 > room * room) {
        VList nl;
        build_list<S, L, LDS>(T, sh, lane, nl, c, T.rlong + off + T.skin, T.rlj + off + T.skin);
        if (lane == 0) *vl = nl;
        lds_fence();
      }
      const int n = vl->n, nin = vl->nin;
      if (vl->valid > 0) {
        const LIdx<LDS> *lst = static_cast<const LIdx<LDS> *>(sh.list);
        // entry t of the list: the front part below nin, the back part after
        const int back = T.list_cap - 1 + nin;
        auto entry = [&](int t) { return (int)lst[t < nin ? t : back - t]; };
        // software-pipelined: the next atom's data is read while this one's
        // terms are computed (hides the LDS/L2 latency of the gather)
        double4 at = make_double4(0.0, 0.0, 0.0, 0.0);
        double2 lp = make_double2(0.0, 0.0), lc = make_double2(0.0, 0.0);
        // the far field's coefficients are loaded before the pair loop (their
        // latency overlaps it) and applied after it
        FarNode<64 / L> fn[S::NS];
        if constexpr (SPLIT) far_load<S, L>(T, lane, p, fn, -1, fcache);
        if (lane < n) fetch<LDS, CEN>(T, sh, entry(lane), at, lp, lc);
        for (int t = lane; t < n; t += L) {
          const double4 cat = at;
          const double2 clp = lp, clc = lc;
          if (t + L < n) fetch<LDS, CEN>(T, sh, entry(t + L), at, lp, lc);
          pair_terms<S, COUNT, L == 64, SPLIT>(T, p, qs, cat, clp, clc, t < nin, a, w);
        }
        if constexpr (SPLIT) far_apply<S, L>(T, lane, fn, qs, a);
        reduce<S, L>(a, sh.red);
        finish<S>(T, a, -1, f, U);
        return;
      }
    }
    // no (usable) list: walk the columns around the centre directly
    const double R = T.rlong + off + 1e-6;
    g.mode = 0;
    g.c[0] = c[0];
    g.c[1] = c[1];
    g.c[2] = c[2];
    g.R2 = R * R;
    const int ix0 = clampi((int)floor((c[0] - R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int ix1 = clampi((int)floor((c[0] + R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int iy0 = clampi((int)floor((c[1] - R - T.cy) * T.inv_cell), 0, T.ny - 1);
    const int iy1 = clampi((int)floor((c[1] + R - T.cy) * T.inv_cell), 0, T.ny - 1);
    g.ix0 = ix0;
    g.iy0 = iy0;
    g.nys = iy1 - iy0 + 1;
    walk_runs<L>(T, sh, lane, g, (ix1 - ix0 + 1) * g.nys, visit);
    if constexpr (SPLIT) {
      FarNode<64 / L> fn[S::NS];
      far_load<S, L>(T, lane, p, fn, -1);
      far_apply<S, L>(T, lane, fn, qs, a);
    }
    reduce<S, L>(a, sh.red);
    finish<S>(T, a, -1, f, U);
    return;
  }
  // A site left the grid (exit step only): evaluate each site alone with the
  // reference's stencils of its (possibly aliased) cell index.
#pragma unroll
  for (int ps = 0; ps < S::NS; ps++) {
    if (!(S::lj(ps) || S::coul(ps) || S::dip(ps))) continue;
    const int i = (int)floor((p[ps][0] - T.cx) / T.cell);
    const int j = (int)floor((p[ps][1] - T.cy) / T.cell);
    const int k = (int)floor((p[ps][2] - T.cz) / T.cell);
    const long long idx = (long long)k + (long long)T.nz * j + (long long)T.nz * T.ny * i;
    if (idx < 0 || idx >= T.ncells) continue;  // zero force, src/Force.cpp:58-66
    g.mode = 1;
    g.ci = (int)(idx / ((long long)T.ny * T.nz));
    g.cj = (int)((idx / T.nz) % T.ny);
    g.ck = (int)(idx % T.nz);
    g.nfirst = T.n1;
    const int nr = T.n1 + ((S::coul(ps) || S::dip(ps)) ? T.n2 : 0);
    double pp[S::NS][3];
#pragma unroll
    for (int s = 0; s < S::NS; s++)
#pragma unroll
      for (int d = 0; d < 3; d++) pp[s][d] = (s == ps) ? p[s][d] : FAR;
    Acc b[S::NS];
#pragma unroll
    for (int s = 0; s < S::NS; s++) acc_zero(b[s]);
    auto visit1 = [&](int ai, int first) {
      double4 at;
      double2 lp, lc;
      fetch<LDS, CEN>(T, sh, ai, at, lp, lc);
      pair_terms<S, COUNT, L == 64, SPLIT>(T, pp, qs, at, lp, lc, first, b, w);
    };
    walk_runs<L>(T, sh, lane, g, nr, visit1);
    if constexpr (SPLIT) {
      FarNode<64 / L> fn[S::NS];
      far_load<S, L>(T, lane, pp, fn, ps);
      far_apply<S, L>(T, lane, fn, qs, b);
    }
    reduce<S, L>(b, sh.red);
    finish<S>(T, b, ps, f, U);
  }
}

// ---------------------------------------------------------------- integration
__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// System::anglevec, src/System.cpp:1566-1579
__device__ __forceinline__ double anglevec(const double *a, const double *b) {
  const double arg = dot3(a, b) / (sqrt(dot3(a, a)) * sqrt(dot3(b, b)));
  if (fabs(arg) > 1.0) return 0.0;
  return acos(arg);
}
// 1/x for a positive finite x: the hardware estimate with two Newton-Raphson
// corrections (within an ulp).
__device__ __forceinline__ double rcp_pos(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(fma(-x, y, 1.0), y, y);
  return fma(fma(-x, y, 1.0), y, y);
}

// The angle between the centre positions of two consecutive steps (the
// reference's anglevec, src/System.cpp:1566-1579) with |b| given (the previous
// step's |a|) and |a| returned for the next step.  The square roots and the
// quotient come from the hardware estimates refined by Newton-Raphson (an ulp
// or so from the library's); small angles, the usual case, take
// acos x = 2 asin(sqrt((1 - x) / 2)) with the asin series to y^9 (y <= 0.005:
// truncation below 1e-24 relative; 1 - x is exact there).
__device__ __forceinline__ double anglevec_nb(const double *a, const double *b, double nb,
                                              double &na) {
  const double aa = dot3(a, a);
  na = aa * rsq_pos(aa);
  const double arg = dot3(a, b) * rcp_pos(na * nb);
  if (fabs(arg) > 1.0) return 0.0;
  if (arg > 0.99995) {
    const double t = (1.0 - arg) * 0.5;
    if (t == 0.0) return 0.0;
    const double y = t * rsq_pos(t), y2 = y * y;
    return 2.0 * (y + y * y2 * (1.0 / 6.0 + y2 * (3.0 / 40.0 + y2 * (5.0 / 112.0 + y2 * (35.0 / 1152.0)))));
  }
  return acos(arg);
}

// The reference's exit test (rx/a)^2 + (ry/b)^2 + (rz/c)^2 > 1
// (src/System.cpp:482-486), decided with reciprocals when the sum is not
// within 1e-12 of 1 (the quotients differ from the products by a few ulp).
__device__ __forceinline__ bool outside_ellipsoid(const Tgt &T, const double *r) {
  const double x = r[0] * T.iea, y = r[1] * T.ieb, z = r[2] * T.iec;
  const double q = x * x + y * y + z * z;
  if (q < 1.0 - 1e-12) return false;
  if (q > 1.0 + 1e-12) return true;
  return (r[0] / T.ea) * (r[0] / T.ea) + (r[1] / T.eb) * (r[1] / T.eb) + (r[2] / T.ec) * (r[2] / T.ec) >
         1.0;
}

__device__ __forceinline__ double kinetic(double m, const double *v) {
  return 0.5 * m * dot3(v, v) * KE_CONV;
}

// velGenerator through the prefix table: first v_n with S_n >= u.
__device__ double speed_of(const Tgt &T, double u) {
  int lo = 0, hi = T.vn - 1;
  if (T.vs[hi] < u) return T.vv[hi];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (T.vs[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  return T.vv[lo];
}

__device__ __forceinline__ void rot_z(double *r, double ang) {
  double s, c;
  sincos(ang, &s, &c);
  const double x = r[0], y = r[1];
  r[0] = x * c - y * s;
  r[1] = x * s + y * c;
}
__device__ __forceinline__ void rot_y(double *r, double ang) {
  double s, c;
  sincos(ang, &s, &c);
  const double x = r[0], z = r[2];
  r[0] = x * c + z * s;
  r[2] = x * -s + z * c;
}

// System::setup, src/System.cpp:1397-1518.  Returns the hit flag.
template <class S>
__device__ __attribute__((noinline)) bool setup(const Tgt &T, const double *u, double (&pos)[3][3], double (&vel)[3][3],
                      double (&rcm)[3], double (&vcm)[3]) {
  rcm[0] = u[0];
  rcm[1] = 0.0;
  rcm[2] = T.bmax;
  vcm[0] = 0.0;
  vcm[1] = 0.0;
  vcm[2] = -speed_of(T, u[1]);
  const double g = 2.0 * PI * u[2];
  const double ph = asin(2.0 * u[3] - 1.0) + 0.5 * PI;
  const double th = 2.0 * PI * u[4];
  rot_z(rcm, g); rot_z(vcm, g);
  rot_y(rcm, ph); rot_y(vcm, ph);
  rot_z(rcm, th); rot_z(vcm, th);
  const double ur[3] = {rcm[0] / T.ea, rcm[1] / T.eb, rcm[2] / T.ec};
  const double uv[3] = {vcm[0] / T.ea, vcm[1] / T.eb, vcm[2] / T.ec};
  const double cr[3] = {ur[1] * uv[2] - ur[2] * uv[1], ur[2] * uv[0] - ur[0] * uv[2],
                        ur[0] * uv[1] - ur[1] * uv[0]};
  const double uvuv = dot3(uv, uv);
  const double delta = -dot3(cr, cr) + uvuv;
  if (delta < 0) return false;
  const double lam = -dot3(ur, uv) / uvuv - sqrt(delta) / uvuv;
  for (int d = 0; d < 3; d++) rcm[d] = rcm[d] + lam * vcm[d];
  if (S::NS == 1) {
    for (int d = 0; d < 3; d++) {
      pos[0][d] = rcm[d];
      vel[0][d] = vcm[d];
    }
    return true;
  }
  const double theta = asin(2.0 * u[5] - 1.0) + 0.5 * PI;
  const double phi = 2.0 * PI * u[6];
  const double psi = 2.0 * PI * u[7];
  const double omega =
      sqrt(2.0 * 1.38064E-23 * T.temp / T.ginertia * log(1. / (1. - u[8]))) * 2.45400505E8;
  double w[3];
  sincos(psi, &w[1], &w[0]);
  w[0] *= omega;
  w[1] *= omega;
  w[2] = 0.0;
  rot_y(w, theta);
  rot_z(w, phi);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double rg[3] = {0.0, 0.0, S::z(s)};
    rot_y(rg, theta);
    rot_z(rg, phi);
    const double vg[3] = {w[1] * rg[2] - w[2] * rg[1], w[2] * rg[0] - w[0] * rg[2],
                          w[0] * rg[1] - w[1] * rg[0]};
    for (int d = 0; d < 3; d++) {
      pos[s][d] = rg[d] + rcm[d];
      vel[s][d] = vg[d] + vcm[d];
    }
  }
  return true;
}

// Per-try RATTLE coefficients (dt changes on every restart); each is the
// value the reference computes inline: 0.5*dt/m, muij/(m*dt), muij/mi.
struct Rat {
  double hi, hj, ci, cj, pij;
};
__device__ __forceinline__ Rat rattle_coeffs(double mi, double mj, double dt) {
  const double muij = mi * mj / (mi + mj);
  return Rat{0.5 * dt / mi, 0.5 * dt / mj, muij / (mi * dt), muij / (mj * dt), muij / mi};
}

// RATTLE position stage, src/System.cpp:1700-1739
__device__ __forceinline__ void rattle1(double *ri, double *rj, double *vi, double *vj,
                                        const double *fi, const double *fj, const Rat &c,
                                        double dt, double d2) {
  double rci[3], rcj[3], vci[3], vcj[3], rcij[3], rij[3];
  for (int k = 0; k < 3; k++) {
    vci[k] = vi[k] + c.hi * fi[k] * KCONV;
    vcj[k] = vj[k] + c.hj * fj[k] * KCONV;
    rci[k] = ri[k] + dt * vci[k];
    rcj[k] = rj[k] + dt * vcj[k];
    rij[k] = ri[k] - rj[k];
    rcij[k] = rci[k] - rcj[k];
  }
  const double r2 = dot3(rij, rij), rc2 = dot3(rcij, rcij), rr = dot3(rcij, rij);
  double det = rr * rr - r2 * (rc2 - d2);
  if (det < 0.0) det = 0.0;
  // the quotient by r2 (the bond length squared, positive) through the
  // refined hardware reciprocal: within an ulp of the division
  const double lam = (rr - sqrt(det)) * rcp_pos(r2);
  for (int k = 0; k < 3; k++) {
    vi[k] = vci[k] - c.ci * lam * rij[k];
    vj[k] = vcj[k] + c.cj * lam * rij[k];
    ri[k] = rci[k] - c.pij * lam * rij[k];
    rj[k] = rcj[k] + c.pij * lam * rij[k];  // the reference uses muij/mi for both sites
  }
}

// RATTLE velocity stage, src/System.cpp:1741-1767
__device__ __forceinline__ void rattle2(const double *ri, const double *rj, double *vi, double *vj,
                                        const double *fi, const double *fj, const Rat &c,
                                        double mi, double mj) {
  const double muij = mi * mj / (mi + mj);
  double rij[3], vcij[3], vci[3], vcj[3];
  for (int k = 0; k < 3; k++) {
    vci[k] = vi[k] + c.hi * fi[k] * KCONV;
    vcj[k] = vj[k] + c.hj * fj[k] * KCONV;
    rij[k] = ri[k] - rj[k];
    vcij[k] = vci[k] - vcj[k];
  }
  // quotients through reciprocals (within an ulp of the divisions; the
  // masses are constants of the probe)
  const double lam = muij * dot3(rij, vcij) * rcp_pos(dot3(rij, rij));
  const double li = lam * (1.0 / mi), lj = lam * (1.0 / mj);
  for (int k = 0; k < 3; k++) {
    vi[k] = vci[k] - li * rij[k];
    vj[k] = vcj[k] + lj * rij[k];
  }
}

// Forces on the two constrained sites: N2 dummy split (src/System.cpp:736-739)
// or CO2 Ciccotti redistribution (src/System.cpp:1154-1158).
template <class S, int GK>
__device__ __forceinline__ void pair_forces(const double (&fs)[3][3], double *fi, double *fj) {
  if (GK == 2) {
    for (int k = 0; k < 3; k++) {
      fi[k] = fs[0][k] + 0.5 * fs[2][k];
      fj[k] = fs[1][k] + 0.5 * fs[2][k];
    }
  } else {
    const double M = S::m(0) + S::m(1) + S::m(2);
    const double m2M = S::m(1) / M;
    for (int k = 0; k < 3; k++) {
      fi[k] = (1.0 - 0.5 * m2M) * fs[0][k] + S::m(0) / M * fs[1][k] - 0.5 * m2M * fs[2][k];
      fj[k] = (1.0 - 0.5 * m2M) * fs[2][k] + S::m(0) / M * fs[1][k] - 0.5 * m2M * fs[0][k];
    }
  }
}

template <int L>
__device__ __forceinline__ void write_record(mccs_traj_record *rec, uint64_t j, int lane, int hit,
                                             int success, int tries, int steps, double chi,
                                             double bmax, const Work &w, bool count) {
  uint32_t v = 0, l = 0, c = 0, d = 0;
  if (count) {
    v = wtotal<L>(w.visits);
    l = wtotal<L>(w.lj);
    c = wtotal<L>(w.coul);
    d = wtotal<L>(w.dip);
  }
  if (lane == 0) {
    mccs_traj_record r;
    r.hit = hit;
    r.success = success;
    r.tries = tries;
    r.steps = steps;
    r.chi = chi;
    r.domega = success ? PI * (1.0 - cos(chi)) * (bmax * bmax) : 0.0;
    r.visits = v;
    r.lj = l;
    r.coul = c;
    r.dip = d;
    rec[j] = r;
  }
}

// Trajectory state of one group, kept in LDS between the phases of a step so
// that only positions and force accumulators occupy registers during the
// force evaluation.  Every lane of the group computes the same values; lane 0
// stores them.
struct GState {
  double pos[3][3], vel[3][3], fs[3][3];
  double rcm[3], vcm[3], rold[3], vcm_i[3];
  double Ei, dt, dtheta, hdm, Utot;
  double rnorm;                 // |previous centre| (the angle sum's |b|)
  Rat rc;
  VList vl;
  unsigned long long j;         // trajectory (index over all targets of the launch)
  int t;                        // its target
  int k;                        // try (restart) index of the trajectory, 0..TRIES_MAX-1
  unsigned long long slot;      // spawn slot taken while its item is not yet written
  int steps;                    // steps of this try
  int cont;                     // 1: the group goes on with try k of trajectory j
  int has_slot;
};

// Outcome of one try of one trajectory.
enum : int { TRY_PENDING = 0, TRY_SUCCESS = 1, TRY_ANGLE = 2, TRY_ENERGY = 3, TRY_CAPPED = 4,
             TRY_ABANDON = 5 };  // TRY_ABANDON: moot, never stored
// The reference's retry loop (src/System.cpp:392, maxTries = 8 raised by one
// per energy failure while below 10) runs at most 9 tries: e.g. seven
// angular restarts, an energy failure, and a ninth try.
constexpr int TRIES_MAX = 9;
struct TryOut {
  int code, steps;
  double chi;
  uint32_t visits, lj, coul, dip;
};

// Device work queue.  Primary items are tries 0 of trajectories [0, n) in
// index order (ticket counter).  Each try k restarts from the trajectory's
// initial state with dt0 / (k + 1), exactly as the reference's loop does after
// k failed tries (src/System.cpp:453-458, :466-473), so the tries of one
// trajectory are independent of each other and may run concurrently; the
// reference's sequence of restarts is then folded over their outcomes in order.
// tries[j] counts the tries of j claimed so far (each index is claimed once).
// A group whose try fails claims the trajectory's next try and goes on with
// it.  Once fewer than spec_below trajectories are unresolved (the end of the
// run, when groups fall idle), the group running a trajectory that has
// already restarted, or whose first try has run long, claims all its
// remaining tries at once and spawns them as items that idle groups take, so
// the last trajectories' tries run side by side instead of one after another.
// A claimed try that the fold turns out not to need is abandoned at its next
// check.
struct Queue {
  unsigned long long *ticket;  // primary tickets handed out
  unsigned long long *shead;   // spawned items taken
  unsigned long long *stail;   // spawned items reserved
  long long *active;           // trajectories not yet resolved
  unsigned long long *work;    // executed totals: visits, lj, coul, dip, steps
  int *stalled;                // set by a group that gave up waiting
  unsigned long long *slots;   // spawned items j * 16 + k, ~0 until written
  unsigned long long nslots;
  int *tries;                  // tries claimed, per trajectory
  int *spawned;                // 1 once the trajectory's remaining tries are spawned
  int *resolved;               // 1 once the trajectory's record is written
  TryOut *tab;                 // TRIES_MAX per trajectory
  // Primary ticket t takes trajectory order[t] (null: trajectory t).  The
  // order only decides when a trajectory starts: every try is a pure
  // function of (trajectory, try), so the records are the same in any order.
  const unsigned *order;
  unsigned *bucket;            // ORDER_BUCKETS counters of the ordering pass
  unsigned long long wait_ticks;  // wall-clock ticks a group waits for work at most
  long long spec_below;        // spawning starts below this many unresolved trajectories
};

// Buckets of the start order (order_count / order_scan / order_scatter).
constexpr int ORDER_BUCKETS = 1024;

// A first try that has run this many steps is likely to orbit and restart.
constexpr int SPAWN_STEPS = 16384;
// A trajectory whose tries have failed this many times is one of the long
// orbiting chains (on the 50k ellipsoid: 0.4 % of the trajectories, each up
// to 3.5 x 10^5 steps over its nine tries, longer than a group's whole share
// of the launch): its remaining tries are spawned at once, whatever the
// stage of the launch, and groups take spawned tries before new trajectories,
// so that the chain's tries run side by side instead of one after another.
constexpr int SPAWN_FAILS = 4;

// The launch ends when its longest chain of tries does, and that chain runs
// at the latency of one step while the other waves of its SIMD compete for
// the same issue slots.  A whole-wave group raises its wave's issue priority
// with the try's index (a restarted trajectory is more likely one of the long
// chains; a spawned late try is one) and as the try grows long, so the long
// tries issue first and the short ones fill the remaining cycles; the
// priority is dropped when the try ends.  Only the order of issue changes.
constexpr int PRIO_STEPS1 = 4096, PRIO_STEPS2 = 16384, PRIO_STEPS3 = 65536;
template <int L>
__device__ __forceinline__ void set_prio(int level) {
  if constexpr (L == 64) {
    const int v = __builtin_amdgcn_readfirstlane(level);
    if (v <= 0) __builtin_amdgcn_s_setprio(0);
    else if (v == 1) __builtin_amdgcn_s_setprio(1);
    else if (v == 2) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(3);
  }
}
__device__ __forceinline__ int prio_of_try(int k) { return k >= SPAWN_FAILS ? 2 : (k >= 1 ? 1 : 0); }
template <int L>
__device__ __forceinline__ void prio_try_start(int k) {
  set_prio<L>(prio_of_try(k));
}
template <int L>
__device__ __forceinline__ void prio_step(int steps, int k) {
  if constexpr (L == 64) {
    const int s = __builtin_amdgcn_readfirstlane(steps);
    if (s == PRIO_STEPS1 || s == PRIO_STEPS2 || s == PRIO_STEPS3) {
      const int lv = s == PRIO_STEPS1 ? 1 : (s == PRIO_STEPS2 ? 2 : 3);
      const int base = prio_of_try(k);
      set_prio<L>(lv > base ? lv : base);
    }
  }
}
template <int L>
__device__ __forceinline__ void prio_try_end() {
  if constexpr (L == 64) __builtin_amdgcn_s_setprio(0);
}

__device__ __forceinline__ unsigned long long ld_acq(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_acq_i(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T_>
__device__ __forceinline__ T_ ld_rlx(const T_ *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Claims every try of trajectory j not claimed yet and spawns them (lane 0).
__device__ __attribute__((noinline)) void spawn_rest(const Queue &Q, uint64_t j) {
  if (atomicCAS(Q.spawned + j, 0, 1) != 0) return;
  const int c = atomicExch(Q.tries + j, TRIES_MAX);
  if (c >= TRIES_MAX) return;
  const unsigned long long r0 = atomicAdd(Q.stail, (unsigned long long)(TRIES_MAX - c));
  for (int kk = c; kk < TRIES_MAX; kk++)
    if (r0 + (kk - c) < Q.nslots)
      __hip_atomic_store(Q.slots + r0 + (kk - c), (unsigned long long)(16 * j + kk),
                         __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-group scratch: the run scratch of the column walks (2 ints per lane),
// which 64-lane groups share with their reduction rows.
__host__ __device__ constexpr size_t group_scratch(int lanes) {
  return lanes == 64 ? sizeof(double) * RROWS * RSTRIDE : sizeof(int) * 2 * (size_t)lanes;
}

// Block LDS layout: atoms and LJ pairs (when staged) | per-group scratch |
// one GState per group | one Verlet list per group.
__host__ __device__ inline size_t lds_bytes(int natoms, int list_cap, int block, int lanes,
                                            bool lds_atoms, bool central) {
  const size_t lidx = lds_atoms ? sizeof(uint16_t) : sizeof(int);
  size_t b = 0;
  if (lds_atoms)
    b += (sizeof(double4) + sizeof(double2) * (central ? 2 : 1)) * (size_t)natoms;
  b += group_scratch(lanes) * (size_t)(block / lanes);
  b += sizeof(GState) * (size_t)(block / lanes);
  b = (b + 31) / 32 * 32;
  b += lidx * (size_t)list_cap * (size_t)(block / lanes);
  return b + 64;
}

template <int L, bool LDS, bool CEN>
__device__ void stage(const Tgt &T, Shm &sh, char *smem) {
  char *p = smem;
  double4 *atoms = reinterpret_cast<double4 *>(p);
  if (LDS) p += sizeof(double4) * (size_t)T.natoms;
  double2 *ljp = reinterpret_cast<double2 *>(p);
  if (LDS) p += sizeof(double2) * (size_t)T.natoms;
  double2 *ljc = reinterpret_cast<double2 *>(p);
  if (LDS && CEN) p += sizeof(double2) * (size_t)T.natoms;
  char *scratch = p;
  p += group_scratch(L) * (blockDim.x / L);
  GState *states = reinterpret_cast<GState *>(p);
  p += sizeof(GState) * (blockDim.x / L);
  // the lists start on a 32-byte boundary (double4 entries of the copied lists)
  char *lists = smem + ((size_t)(p - smem) + 31) / 32 * 32;
  if (LDS) {
    for (int i = threadIdx.x; i < T.natoms; i += blockDim.x) {
      atoms[i] = T.atoms[i];
      ljp[i] = T.ljp[i];
      if (CEN) ljc[i] = T.ljc[i];
    }
  }
  __syncthreads();
  const int g = threadIdx.x / L;
  sh.atoms = atoms;
  sh.ljp = ljp;
  sh.ljc = ljc;
  sh.incl = reinterpret_cast<int *>(scratch + group_scratch(L) * g);
  sh.base = sh.incl + L;
  sh.red = reinterpret_cast<double *>(sh.incl);
  sh.state = states + g;
  sh.list = lists + sizeof(LIdx<LDS>) * (size_t)T.list_cap * g;
}

// ---------------------------------------------------------------- kernels
template <int L>
__device__ __forceinline__ unsigned long long bcast(unsigned long long v) {
  return __shfl(v, 0, L);
}

// Folds the tries of trajectory j in order, as the reference's retry loop
// would have run them (src/System.cpp:389-511): a try that succeeds ends the
// trajectory; an angular restart (dtheta > 1.5 pi) counts a try, an energy
// failure counts a try and raises the allowance; the loop stops when
// tries >= maxTries or maxTries reaches 10.  Writes the record once the
// outcome is determined (by whichever completing try determines it first).
// Returns true when the trajectory is resolved.  Called by lane 0 after the
// caller's own outcome is stored; the sequentially consistent fence makes sure
// that of two tries ending together at least one sees the other's outcome.
__device__ __attribute__((noinline)) bool resolve(const Queue &Q, mccs_traj_record *rec, uint64_t j, double bmax) {
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
  int tries = 0, max_tries = 8, steps = 0, success = 0, final = -1, capped = 0;
  uint32_t v = 0, l = 0, c = 0, d = 0;
  double chi = 0.0;
  for (int k = 0; k < TRIES_MAX; k++) {
    const int code = ld_acq_i(&Q.tab[TRIES_MAX * j + k].code);
    if (code == TRY_PENDING) return ld_acq_i(&Q.resolved[j]) != 0;  // not determined yet
    const TryOut t = Q.tab[TRIES_MAX * j + k];
    steps += t.steps;
    v += t.visits; l += t.lj; c += t.coul; d += t.dip;
    if (code == TRY_SUCCESS) {
      success = 1;
      chi = t.chi;
      final = k;
      break;
    }
    if (code == TRY_CAPPED) {
      capped = 1;
      final = k;
      break;
    }
    tries = k + 1;
    if (code == TRY_ENERGY) max_tries++;
    if (!(tries < max_tries && max_tries < 10) || k == TRIES_MAX - 1) {
      final = k;
      break;
    }
  }
  if (atomicCAS(&Q.resolved[j], 0, 1) != 0) return true;
  mccs_traj_record r;
  r.hit = 1;
  r.success = success;
  r.tries = capped ? -1 : (success ? final : tries);
  r.steps = steps;
  r.chi = chi;
  r.domega = success ? PI * (1.0 - cos(chi)) * (bmax * bmax) : 0.0;
  r.visits = v;
  r.lj = l;
  r.coul = c;
  r.dip = d;
  rec[j] = r;
  // records are read by the host after the launch: no fence before the count
  atomicAdd((unsigned long long *)Q.active, ~0ull);  // -1
  return true;
}

// Persistent trajectory kernel.  Each group of L lanes pulls work items
// (trajectory, try) until every trajectory is resolved; per loop iteration
// every working group performs exactly one force evaluation: the initial
// force of a try, or the force of one velocity-Verlet/RATTLE step.
template <int GK, int FT, int L, bool LDS, bool COUNT, bool MT, bool SPLIT = false>
__global__ __launch_bounds__(traj_block(LDS)) __attribute__((amdgpu_waves_per_eu(traj_waves(LDS)))) void traj_kernel(
    const Tgt *__restrict__ tg, const unsigned long long *__restrict__ toff, int ntg,
    const double *__restrict__ u, uint64_t n, mccs_traj_record *__restrict__ rec, Queue Q) {
  static_assert(!(MT && LDS), "targets staged in LDS need a launch of their own");
  using S = Spec<GK, FT>;
  constexpr bool CEN = any_central<S>();
  extern __shared__ __align__(32) char smem[];
  Shm sh;
  stage<L, LDS, CEN>(tg[0], sh, smem);
  const int lane = threadIdx.x % L;
  GState &st = *sh.state;
  constexpr int PA = 0, PB = GK == 2 ? 1 : 2;  // the constrained pair (N-N or O-O)
  constexpr double MA = S::NS == 1 ? 1.0 : S::m(PA), MB = S::NS == 1 ? 1.0 : S::m(PB);
  // target of trajectory j: the last t with toff[t] <= j (group-uniform)
  auto target_of = [&](uint64_t j) -> int {
    if (!MT) return 0;
    int lo = 0, hi = ntg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (toff[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    return L == 64 ? __builtin_amdgcn_readfirstlane(lo) : lo;
  };
  // the group's current target; for a whole-wave group the index is made
  // wave-uniform so that target fields are read with scalar loads
  auto cur = [&]() -> int {
    if (!MT) return 0;
    return L == 64 ? __builtin_amdgcn_readfirstlane(st.t) : st.t;
  };
  double qs[S::NS];
#pragma unroll
  for (int s = 0; s < S::NS; s++) qs[s] = S::q(s);

  Work w{0, 0, 0, 0};
  int status = ST_NEED;
  unsigned long long wait_since = 0;  // wall clock when the group began waiting, 0 if not
  FarCache<S::NS, 64 / L> fcache;  // far-field coefficients of the last evaluation (range split)
  fcache.clear();
  if (lane == 0) {
    st.vl.valid = -1;
    st.cont = 0;
    st.has_slot = 0;
    st.t = 0;
  }
  lds_fence();

  while (status != ST_DONE) {
    // ---- phase A: take a work item, or the first half-step
    if (status == ST_NEED) {
      double pos[3][3], vel[3][3], rcm[3], vcm[3];
      uint64_t j = 0;
      int k = 0, tj = 0;
      bool got = false, wait = false, done = false;
      if (st.cont) {  // the next try of the trajectory whose try just failed
        j = st.j;
        k = st.k;
        tj = cur();
        setup<S>(tg[tj], u + 9 * j, pos, vel, rcm, vcm);  // a trajectory with tries hits
        got = true;
      }
      while (!got && !wait && !done) {
        // 1. a spawned try (the slot taken earlier, or the next one): spawned
        // tries are the long chains', which start before new trajectories
        if (Q.spec_below > 0) {
          unsigned long long slot = ~0ull;
          if (lane == 0) {
            if (st.has_slot) {
              slot = st.slot;
            } else if (ld_rlx(Q.shead) < ld_rlx(Q.stail)) {
              slot = atomicAdd(Q.shead, 1ull);
            }
          }
          slot = bcast<L>(slot);
          if (slot != ~0ull) {
            unsigned long long item = ~0ull;
            int moot = 0;
            if (lane == 0) {
              item = slot < Q.nslots ? ld_acq(Q.slots + slot) : ~0ull;
              st.has_slot = item == ~0ull;  // written soon by its spawner, or never
              st.slot = slot;
              if (item != ~0ull) moot = ld_rlx(Q.resolved + (item >> 4));
            }
            item = bcast<L>(item);
            moot = __shfl(moot, 0, L);
            if (item != ~0ull) {
              if (moot) continue;  // the trajectory is resolved: take the next one
              j = item >> 4;
              k = (int)(item & 15);
              tj = target_of(j);
              setup<S>(tg[tj], u + 9 * j, pos, vel, rcm, vcm);  // a trajectory with tries hits
              got = true;
              break;
            }
          }
        }
        // 2. the next primary trajectory (the loop only repeats past misses)
        unsigned long long t = 0;
        if (lane == 0) {  // look before taking: idle groups must not hammer the counter
          t = __hip_atomic_load(Q.ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (t < n) t = atomicAdd(Q.ticket, 1ull);
        }
        t = bcast<L>(t);
        if (t < n) {
          j = Q.order ? Q.order[t] : t;
          k = 0;
          tj = target_of(j);
          w = Work{0, 0, 0, 0};
          if (setup<S>(tg[tj], u + 9 * j, pos, vel, rcm, vcm)) {
            got = true;
          } else {
            write_record<L>(rec, j, lane, 0, 0, 0, 0, 0.0, tg[tj].bmax, w, COUNT);
            if (lane == 0) atomicAdd((unsigned long long *)Q.active, ~0ull);
          }
          continue;
        }
        if (Q.spec_below <= 0) {  // nothing is ever spawned: this group is finished
          done = true;
          break;
        }
        // 3. nothing to take: finished once every trajectory is resolved
        long long act = 0;
        if (lane == 0) act = __hip_atomic_load(Q.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        act = (long long)bcast<L>((unsigned long long)act);
        if (act <= 0) done = true;
        else wait = true;
      }
      if (wait) {
        // a try in flight always ends (MAX_STEPS), so waiting ends; a group
        // that has waited far longer than any try takes gives up and flags it
        const unsigned long long now = wall_clock64();
        if (!wait_since) wait_since = now;
        if (now - wait_since > Q.wait_ticks) {
          if (lane == 0) __hip_atomic_store(Q.stalled, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          done = true;
        }
      }
      if (done) {
        status = ST_DONE;
        break;
      }
      if (wait) {
        // a whole-wave group sleeps ~24k cycles between looks; a smaller group
        // shares its wave with working groups and looks once per their step
        if (L == 64) {
          __builtin_amdgcn_s_sleep(127);
          __builtin_amdgcn_s_sleep(127);
          __builtin_amdgcn_s_sleep(127);
        }
        lds_fence();
        continue;
      }
      wait_since = 0;
      w = Work{0, 0, 0, 0};
      fcache.clear();  // a new try, possibly of another target
      if (lane == 0) {
        st.vl.valid = -1;  // each try builds its own neighbour list: results do not
                           // depend on what the group ran before
        for (int s = 0; s < S::NS; s++)
          for (int d = 0; d < 3; d++) {
            st.pos[s][d] = pos[s][d];
            st.vel[s][d] = vel[s][d];
          }
        for (int d = 0; d < 3; d++) {
          st.rcm[d] = rcm[d];
          st.vcm[d] = vcm[d];
        }
        st.j = j;
        st.t = tj;
        st.k = k;
        st.cont = 0;
      }
      prio_try_start<L>(k);
      status = ST_INIT;
    } else if (status == ST_STEP) {  // first half (the first step of a try)
      const Tgt &T = tg[cur()];
      const double dt = st.dt;
      double rcm[3];
      if (S::NS == 1) {
        const double hdm = st.hdm;
        double v[3], x[3];
        for (int d = 0; d < 3; d++) {
          v[d] = st.vel[0][d] + hdm * st.fs[0][d] * KCONV;
          x[d] = st.pos[0][d] + dt * v[d];
          rcm[d] = x[d];
        }
        if (lane == 0)
          for (int d = 0; d < 3; d++) {
            st.vel[0][d] = v[d];
            st.pos[0][d] = x[d];
          }
      } else {
        double fs[3][3], ri[3], rj[3], vi[3], vj[3], fi[3], fj[3];
        for (int s = 0; s < 3; s++)
          for (int d = 0; d < 3; d++) fs[s][d] = st.fs[s][d];
        for (int d = 0; d < 3; d++) {
          ri[d] = st.pos[PA][d]; rj[d] = st.pos[PB][d];
          vi[d] = st.vel[PA][d]; vj[d] = st.vel[PB][d];
        }
        pair_forces<S, GK>(fs, fi, fj);
        const Rat rc = st.rc;
        rattle1(ri, rj, vi, vj, fi, fj, rc, dt, T.d2);
        constexpr int mid = GK == 2 ? 2 : 1;
        double vcm[3];
        for (int d = 0; d < 3; d++) {
          rcm[d] = 0.5 * (ri[d] + rj[d]);
          vcm[d] = 0.5 * (vi[d] + vj[d]);
        }
        if (lane == 0)
          for (int d = 0; d < 3; d++) {
            st.pos[PA][d] = ri[d]; st.pos[PB][d] = rj[d];
            st.vel[PA][d] = vi[d]; st.vel[PB][d] = vj[d];
            st.pos[mid][d] = rcm[d];
            st.vcm[d] = vcm[d];
          }
      }
      double rold[3], na;
      for (int d = 0; d < 3; d++) rold[d] = st.rold[d];
      const double dth = st.dtheta + anglevec_nb(rcm, rold, st.rnorm, na);
      if (lane == 0) {
        st.rnorm = na;
        st.dtheta = dth;
        for (int d = 0; d < 3; d++) {
          st.rold[d] = rcm[d];
          st.rcm[d] = rcm[d];
        }
      }
    }
    lds_fence();
    const Tgt &T = tg[cur()];

    // ---- phase B: one force evaluation for every working group; positions,
    // forces and the total energy stay in registers for phase C (the group
    // holds them uniformly; lane 0 also stores forces and energy to LDS)
    double p[S::NS][3], f[S::NS][3], Ut = 0.0;
    {
      double U[S::NS];
#pragma unroll
      for (int s = 0; s < S::NS; s++)
        for (int d = 0; d < 3; d++) p[s][d] = st.pos[s][d];
      evaluate<S, L, LDS, COUNT, SPLIT>(T, sh, lane, p, qs, f, U, w, &st.vl, &fcache);
#pragma unroll
      for (int s = 0; s < S::NS; s++) Ut += U[s];
      if (lane == 0) {
#pragma unroll
        for (int s = 0; s < S::NS; s++)
          for (int d = 0; d < 3; d++) st.fs[s][d] = f[s][d];
        st.Utot = Ut;
      }
    }
    lds_fence();

    // ---- phase C: try initialisation, or second half-step and exit tests
    if (status == ST_INIT) {
      if (lane == 0) {
        double Ek = 0.0;
        if (S::NS == 1) Ek = kinetic(T.mu, st.vel[0]);
        else  // the N2 dummy is massless and adds nothing
          for (int s = 0; s < (GK == 2 ? 2 : 3); s++) Ek += kinetic(S::m(s), st.vel[s]);
        st.Ei = st.Utot + Ek;
        st.steps = 0;
        const double dt = T.dt / (st.k + 1);  // dt of try k, src/System.cpp:455
        st.dt = dt;
        st.hdm = 0.5 * dt / T.mu;
        st.rc = rattle_coeffs(MA, MB, dt);
        for (int d = 0; d < 3; d++) st.vcm_i[d] = st.vcm[d];
        if (GK == 3 && !T.co2fix) st.vcm_i[0] = st.vcm[1];  // src/System.cpp:1012 reads vcm[1] for x
        st.dtheta = 0.0;
        for (int d = 0; d < 3; d++) st.rold[d] = st.rcm[d];
        const double aa = dot3(st.rcm, st.rcm);
        st.rnorm = aa * rsq_pos(aa);  // as anglevec_nb computes |a|
      }
      status = ST_STEP;
      lds_fence();
      continue;
    }
    // ST_STEP / ST_HALF: second half
    double vcm[3], Ek = 0.0;
    // the integrated particles (the atom, or the constrained pair) and their
    // forces stay in registers for the next step's first half
    double ri[3], rj[3], vi[3], vj[3], fi[3], fj[3];
    const double dt = st.dt, hdm = st.hdm;
    const Rat rc = st.rc;
    if constexpr (S::NS == 1) {
      for (int d = 0; d < 3; d++) {
        fi[d] = f[0][d];
        ri[d] = p[0][d];
        vcm[d] = st.vel[0][d] + hdm * fi[d] * KCONV;
        vi[d] = vcm[d];
      }
      Ek = kinetic(T.mu, vcm);
    } else {
      for (int d = 0; d < 3; d++) {
        ri[d] = p[PA][d]; rj[d] = p[PB][d];
        vi[d] = st.vel[PA][d]; vj[d] = st.vel[PB][d];
      }
      pair_forces<S, GK>(f, fi, fj);
      rattle2(ri, rj, vi, vj, fi, fj, rc, MA, MB);
      for (int d = 0; d < 3; d++) vcm[d] = 0.5 * (vi[d] + vj[d]);
      // exit-test kinetic energy: N2 over its two atoms, CO2 over all three
      // sites with the carbon moving at vcm (src/System.cpp:880-886, :1323-1329)
      Ek = kinetic(MA, vi) + kinetic(MB, vj);
      if (GK == 3) Ek = kinetic(S::m(0), vi) + kinetic(S::m(1), vcm) + kinetic(S::m(2), vj);
    }
    double rcm[3];
    for (int d = 0; d < 3; d++) rcm[d] = st.rcm[d];
    const int steps = st.steps + 1;
    const uint64_t j = st.j;
    const int k = st.k;
    int code = TRY_PENDING;
    double chi = 0.0;
    if (steps >= MAX_STEPS || !isfinite(rcm[0] + rcm[1] + rcm[2] + vcm[0] + vcm[1] + vcm[2])) {
      // the reference would loop forever here (zero speed or non-finite
      // state); the trajectory is reported lost with tries = -1
      code = TRY_CAPPED;
    } else if (st.dtheta > 1.5 * PI) {
      code = TRY_ANGLE;
    } else if (fabs(rcm[0]) >= T.hx || fabs(rcm[1]) >= T.hy || fabs(rcm[2]) >= T.hz ||
               outside_ellipsoid(T, rcm)) {
      const double Ei = st.Ei;
      const double dE = S::NS == 1 ? fabs(Ut + Ek - Ei) / Ei * 100.0
                                   : fabs((Ut + Ek - Ei) / Ei) * 100.0;
      const double tol = GK == 1 ? 0.5 : (GK == 2 ? 1.0 : 20.0);
      if (dE > tol) {
        code = TRY_ENERGY;
      } else {
        double vi[3];
        for (int d = 0; d < 3; d++) vi[d] = st.vcm_i[d];
        chi = anglevec(vi, vcm);
        code = TRY_SUCCESS;
      }
    } else if ((steps & 255) == 0) {
      // a later try is moot once the trajectory is resolved or an earlier
      // try has succeeded; near the end of the run, a trajectory that has
      // restarted or runs long spawns its remaining tries
      int moot = 0;
      if (lane == 0) {
        if (k > 0) {
          moot = ld_acq_i(&Q.resolved[j]);
          for (int kk = 0; kk < k && !moot; kk++)
            moot = ld_acq_i(&Q.tab[TRIES_MAX * j + kk].code) == TRY_SUCCESS;
        }
        if (!moot && (k > 0 || steps >= SPAWN_STEPS) && ld_rlx(Q.spawned + j) == 0 &&
            ld_rlx(Q.active) < Q.spec_below)
          spawn_rest(Q, j);
      }
      if (__shfl(moot, 0, L)) code = TRY_ABANDON;
    }
    if (code == TRY_PENDING) {
      prio_step<L>(steps, k);
      // the first half of the next step, from the registers
      double rn[3];
      if (S::NS == 1) {
        for (int d = 0; d < 3; d++) {
          vi[d] = vi[d] + hdm * fi[d] * KCONV;
          ri[d] = ri[d] + dt * vi[d];
          rn[d] = ri[d];
        }
      } else {
        rattle1(ri, rj, vi, vj, fi, fj, rc, dt, T.d2);
        for (int d = 0; d < 3; d++) {
          rn[d] = 0.5 * (ri[d] + rj[d]);
          vcm[d] = 0.5 * (vi[d] + vj[d]);
        }
      }
      if constexpr (SPLIT) {
        // the far-field coefficients of the next evaluation's sites, loaded
        // into the cache now so that their latency (gathers that mostly
        // miss L2) passes during the state stores, the loop and the list
        // check instead of stalling the next pair loop
        double pn[S::NS][3];
        constexpr int mid = GK == 2 ? 2 : 1;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if (S::NS == 1) {
            pn[0][d] = ri[d];
          } else {
            pn[PA][d] = ri[d];
            pn[PB][d] = rj[d];
            pn[mid][d] = rn[d];
          }
        }
        FarNode<64 / L> fnn[S::NS];
        far_load<S, L>(T, lane, pn, fnn, -1, &fcache);
      }
      double na;
      const double dth = st.dtheta + anglevec_nb(rn, rcm, st.rnorm, na);
      lds_fence();
      if (lane == 0) {
        constexpr int mid = GK == 2 ? 2 : 1;
        if (S::NS == 1) {
          for (int d = 0; d < 3; d++) {
            st.vel[0][d] = vi[d];
            st.pos[0][d] = ri[d];
          }
        } else {
          for (int d = 0; d < 3; d++) {
            st.pos[PA][d] = ri[d]; st.pos[PB][d] = rj[d];
            st.vel[PA][d] = vi[d]; st.vel[PB][d] = vj[d];
            st.pos[mid][d] = rn[d];
            st.vcm[d] = vcm[d];
          }
        }
        st.rnorm = na;
        st.dtheta = dth;
        for (int d = 0; d < 3; d++) {
          st.rold[d] = rn[d];
          st.rcm[d] = rn[d];
        }
        st.steps = steps;
      }
      status = ST_HALF;
      lds_fence();
      continue;
    }
    lds_fence();
    // the try has ended
    prio_try_end<L>();
    uint32_t wv = 0, wl = 0, wc = 0, wd = 0;
    if (COUNT) {
      wv = wtotal<L>(w.visits);
      wl = wtotal<L>(w.lj);
      wc = wtotal<L>(w.coul);
      wd = wtotal<L>(w.dip);
    }
    if (lane == 0) {
      if (COUNT) {
        atomicAdd(Q.work + 0, (unsigned long long)wv);
        atomicAdd(Q.work + 1, (unsigned long long)wl);
        atomicAdd(Q.work + 2, (unsigned long long)wc);
        atomicAdd(Q.work + 3, (unsigned long long)wd);
      }
      atomicAdd(Q.work + 4, (unsigned long long)steps);
      if (k == 0 && (code == TRY_SUCCESS || code == TRY_CAPPED) && ld_rlx(Q.spawned + j) == 0) {
        // the common case: the first try decides (no retry bookkeeping; only
        // this group can have spawned later tries of j while try 0 ran)
        mccs_traj_record r;
        r.hit = 1;
        r.success = code == TRY_SUCCESS;
        r.tries = code == TRY_SUCCESS ? 0 : -1;
        r.steps = steps;
        r.chi = chi;
        r.domega = code == TRY_SUCCESS ? PI * (1.0 - cos(chi)) * (T.bmax * T.bmax) : 0.0;
        r.visits = wv;
        r.lj = wl;
        r.coul = wc;
        r.dip = wd;
        rec[j] = r;
        // read by the host after the launch: no fence (a release here writes
        // back the whole L2 of the XCD)
        atomicAdd((unsigned long long *)Q.active, ~0ull);
      } else if (code != TRY_ABANDON) {
        TryOut o;
        o.code = TRY_PENDING;
        o.steps = steps;
        o.chi = chi;
        o.visits = wv;
        o.lj = wl;
        o.coul = wc;
        o.dip = wd;
        Q.tab[TRIES_MAX * j + k] = o;
        // the release store publishes the outcome to the other XCDs' groups
        __hip_atomic_store(&Q.tab[TRIES_MAX * j + k].code, code, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
        // a failed try needs the next one unless the fold is already decided;
        // claim it unless another group has, and near the end of the run
        // spawn the ones after it
        if (!resolve(Q, rec, j, T.bmax) && code != TRY_SUCCESS && code != TRY_CAPPED &&
            k + 1 < TRIES_MAX) {
          if (atomicCAS(Q.tries + j, k + 1, k + 2) == k + 1) {
            st.k = k + 1;
            st.cont = 1;
          }
          if (ld_rlx(Q.active) < Q.spec_below || (Q.spec_below > 0 && k + 1 >= SPAWN_FAILS))
            spawn_rest(Q, j);
        }
      }
    }
    status = ST_NEED;
    lds_fence();
  }
}

static __global__ void queue_init(Queue Q, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t s = i; s < n; s += stride) {
    Q.tries[s] = 1;  // try 0 belongs to the primary ticket
    Q.spawned[s] = 0;
    Q.resolved[s] = 0;
  }
  for (uint64_t s = i; s < Q.nslots; s += stride) Q.slots[s] = ~0ull;
  for (uint64_t s = i; s < TRIES_MAX * n; s += stride) Q.tab[s].code = TRY_PENDING;
  if (i == 0) {
    *Q.ticket = 0;
    *Q.shead = 0;
    *Q.stail = 0;
    *Q.active = (long long)n;
    for (int k = 0; k < 5; k++) Q.work[k] = 0;
    *Q.stalled = 0;
  }
  if (Q.bucket)
    for (uint64_t s = i; s < (uint64_t)ORDER_BUCKETS; s += stride) Q.bucket[s] = 0;
}

// ---------------------------------------------------------------- start order
// Long trajectories started last make the launch's tail: the other groups
// sit idle while a few long tries finish.  Trajectories with a small impact
// parameter and a slow speed are the likely long ones (they reach the
// target's core and orbit), so primary tickets take them first: a counting
// sort of the trajectories by key = (b / bmax)^2 u_v (u_v the speed draw,
// whose speed grows with it) into ORDER_BUCKETS buckets, ascending.  Over
// several targets (a batch class, its targets largest first) the buckets
// are shared out over the targets, which keep their order: a launch works
// through one target's atoms at a time (mixing the targets measured 14 %
// slower on the 64-assembly batch, the atoms of many targets competing for
// L2), each target's trajectories in the start order.  The order inside a
// bucket is that of the atomics: any order gives the same records.

__device__ __forceinline__ int order_bucket(const Tgt *tg, const unsigned long long *toff, int ntg,
                                            const double *u, uint64_t j) {
  int lo = 0, hi = ntg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  const double x = u[9 * j] / tg[lo].bmax, v = u[9 * j + 1];
  const double key = x * x * v;  // in [0, 1] for b in [0, bmax]; NaN-safe below
  if (ntg > 1) {  // the target's share of the buckets, in target order
    const int per = ORDER_BUCKETS / ntg;
    if (per < 2) return (int)(((long long)lo * ORDER_BUCKETS) / ntg);
    const int b = (int)(key * per);
    return lo * per + (b < 0 ? 0 : (b >= per ? per - 1 : b));
  }
  const int b = (int)(key * ORDER_BUCKETS);
  return b < 0 ? 0 : (b >= ORDER_BUCKETS ? ORDER_BUCKETS - 1 : b);
}

static __global__ void order_count(const Tgt *tg, const unsigned long long *toff, int ntg,
                                   const double *u, uint64_t n, Queue Q) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) atomicAdd(Q.bucket + order_bucket(tg, toff, ntg, u, j), 1u);
}

// Exclusive prefix sum of the bucket counts in place (one block of
// ORDER_BUCKETS threads).
static __global__ __launch_bounds__(ORDER_BUCKETS) void order_scan(Queue Q) {
  __shared__ unsigned sc[ORDER_BUCKETS];
  const int i = threadIdx.x;
  const unsigned c = Q.bucket[i];
  sc[i] = c;
  __syncthreads();
  for (int off = 1; off < ORDER_BUCKETS; off <<= 1) {
    const unsigned add = i >= off ? sc[i - off] : 0u;
    __syncthreads();
    sc[i] += add;
    __syncthreads();
  }
  Q.bucket[i] = sc[i] - c;
}

static __global__ void order_scatter(const Tgt *tg, const unsigned long long *toff, int ntg,
                                     const double *u, uint64_t n, Queue Q) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned pos = atomicAdd(Q.bucket + order_bucket(tg, toff, ntg, u, j), 1u);
  const_cast<unsigned *>(Q.order)[pos] = (unsigned)j;
}

using TrajFn = void (*)(const Tgt *, const unsigned long long *, int, const double *, uint64_t,
                        mccs_traj_record *, Queue);

template <int GK, int FT, int L>
TrajFn pick_traj(bool lds, bool count) {
  if (lds)
    return count ? traj_kernel<GK, FT, L, true, true, false> : traj_kernel<GK, FT, L, true, false, false>;
  return count ? traj_kernel<GK, FT, L, false, true, false> : traj_kernel<GK, FT, L, false, false, false>;
}

// Tuned force types get every launch shape; the others only 64 lanes with
// work counters, which the launcher falls back to.
template <int GK, int FT>
TrajFn pick_traj_full(int L, bool lds, bool count) {
  switch (L) {
    case 16: return pick_traj<GK, FT, 16>(lds, count);
    case 32: return pick_traj<GK, FT, 32>(lds, count);
    default: return pick_traj<GK, FT, 64>(lds, count);
  }
}

// Range-split far field (force types with a dipole field: He/Ar/co2 type 4,
// N2 type 6): 64 lanes.
template <int GK, int FT>
TrajFn pick_traj_split(int L, bool lds, bool count) {
  switch (L) {
    case 16:
      if (lds) return count ? traj_kernel<GK, FT, 16, true, true, false, true> : traj_kernel<GK, FT, 16, true, false, false, true>;
      return count ? traj_kernel<GK, FT, 16, false, true, false, true> : traj_kernel<GK, FT, 16, false, false, false, true>;
    case 32:
      if (lds) return count ? traj_kernel<GK, FT, 32, true, true, false, true> : traj_kernel<GK, FT, 32, true, false, false, true>;
      return count ? traj_kernel<GK, FT, 32, false, true, false, true> : traj_kernel<GK, FT, 32, false, false, false, true>;
    default:
      if (lds) return count ? traj_kernel<GK, FT, 64, true, true, false, true> : traj_kernel<GK, FT, 64, true, false, false, true>;
      return count ? traj_kernel<GK, FT, 64, false, true, false, true> : traj_kernel<GK, FT, 64, false, false, false, true>;
  }
}

template <int GK, int FT>
TrajFn pick_traj_multi_split(int L, bool count) {
  switch (L) {
    case 16: return count ? traj_kernel<GK, FT, 16, false, true, true, true> : traj_kernel<GK, FT, 16, false, false, true, true>;
    case 32: return count ? traj_kernel<GK, FT, 32, false, true, true, true> : traj_kernel<GK, FT, 32, false, false, true, true>;
    default: return count ? traj_kernel<GK, FT, 64, false, true, true, true> : traj_kernel<GK, FT, 64, false, false, true, true>;
  }
}

template <int GK, int FT>
TrajFn pick_traj_basic(int L, bool lds, bool count) {
  if (L != 64 || !count) return nullptr;
  return lds ? traj_kernel<GK, FT, 64, true, true, false> : traj_kernel<GK, FT, 64, false, true, false>;
}

// Launches over several targets (atoms read from HBM/L2, never staged).
template <int GK, int FT>
TrajFn pick_traj_multi(int L, bool count) {
  switch (L) {
    case 16: return count ? traj_kernel<GK, FT, 16, false, true, true> : traj_kernel<GK, FT, 16, false, false, true>;
    case 32: return count ? traj_kernel<GK, FT, 32, false, true, true> : traj_kernel<GK, FT, 32, false, false, true>;
    default: return count ? traj_kernel<GK, FT, 64, false, true, true> : traj_kernel<GK, FT, 64, false, false, true>;
  }
}

TrajFn select_traj_gk1(int ft, int L, bool lds, bool count, bool split);
TrajFn select_traj_gk2(int ft, int L, bool lds, bool count, bool split);
TrajFn select_traj_gk3(int ft, int L, bool lds, bool count);
TrajFn select_traj_multi1(int ft, int L, bool count, bool split);
TrajFn select_traj_multi2(int ft, int L, bool count, bool split);

}  // namespace dev
}  // namespace mccs
