#pragma once
// MI355X (gfx950) trajectory engine: device code shared by the kernel
// instantiation units (traj_gk*.hip) and the launcher (engine.hip).
//
// Work decomposition.  One trajectory is owned by a group of L lanes
// (L = 16, 32 or 64) of one wavefront.  The group integrates the probe with
// wave-uniform FP64 state, and its L lanes cooperate on every force
// evaluation: the candidate atoms around the probe are flattened into one
// index space and strided over the lanes, each lane accumulating all site
// terms in registers; a DPP butterfly then leaves the identical sum in every
// lane.  Blocks are persistent and pull trajectories from a device counter,
// so trajectories of very different length share the GPU evenly.
//
// Reference behaviour restated here (cces-cepid/massccs):
//   setup                   src/System.cpp:1397-1518
//   run_He / run_N2 / run_CO2 src/System.cpp:338-1392
//   RATTLE                  src/System.cpp:1700-1767
//   site kernels            src/Force.cpp (LJ :34-111, dipole :305-489,
//                           Coulomb :598-723, dummy :903-1081, CO2 :1088-1725)
//   cell index / stencils   src/LinkedCell.cpp:117-129, :228-440
//   tally                   src/System.cpp:287-317
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "massccs.h"

#define HIP_OK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      throw ::mccs::Error(MCCS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace mccs {
namespace dev {

constexpr double KCONV = 1.0 / 48.88821291 / 48.88821291;  // KCALMOLANGAMU_TO_ANGFS2
constexpr double KE_CONV = 2390.07;                          // amuAngfs2_to_KCAL_MOL
constexpr double KC = 332.06348078020324;                    // KCOUL
constexpr double PI = 3.14159265358979323846;
constexpr double FAR = 1.0e30;                               // parks an inactive site
constexpr int MAX_STEPS = 4000000;                           // per trajectory, over all tries

enum : int { ST_NEED = 0, ST_INIT = 1, ST_RESTART = 2, ST_STEP = 3, ST_DONE = 4 };

// Kernel argument: target, grid and run constants.
struct Tgt {
  const double4 *atoms;     // x, y, z, q in cell order
  const uint8_t *type;      // LJ parameter row per atom
  const double *lj;         // per row: lj1 lj2 lj3 lj4 ucut | same for the CO2 carbon site
  const int *cstart;        // ncells + 1 prefix of atoms per cell
  const double *vs, *vv;    // speed table: cumulative S_n and speeds v_n
  const int *stencil;       // (n1 + n2) * 3 offsets, first shell then second
  int natoms, ntypes, nx, ny, nz, ncells, vn, n1, n2;
  double cell, cx, cy, cz;  // cell edge and grid corner
  double hx, hy, hz;        // half box
  double ea, eb, ec, bmax;  // ellipsoid semi-axes, maximal impact parameter
  double rlj, rc, rlj2lo, rlj2hi, rc2lo, rc2hi;
  double irc3, cc1, cc2;    // 1/rc^3, -1.5/rc, 0.5/rc^3
  double alpha, mu, dt, temp, ginertia, d2;
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
};

template <class S>
__device__ constexpr bool any_long() {
  bool r = false;
  for (int s = 0; s < S::NS; s++) r = r || S::coul(s) || S::dip(s);
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

// Butterfly sum over the L lanes of a group; every lane receives the same
// bits (each step adds the same two operands in every lane).
template <int L>
__device__ __forceinline__ double gsum(double v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror
  if (L >= 32) v += swz16(v);
  if (L >= 64) v += __shfl_xor(v, 32);
  return v;
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

struct Shm {
  int *incl;            // per-lane inclusive run ends of the group
  int *base;            // per-lane (run start - run offset) * 2 + first-shell flag
  const double *lj;     // LJ rows
  const double4 *atoms; // LDS copy when staged
  const uint8_t *type;
};

template <bool LDS>
__device__ __forceinline__ void fetch(const Tgt &T, const Shm &sh, int a, double4 &at, int &ty) {
  if (LDS) {
    at = sh.atoms[a];
    ty = sh.type[a];
  } else {
    at = T.atoms[a];
    ty = T.type[a];
  }
}

// Pair terms of one candidate atom against every site.  allow_lj is 0 for
// second-shell cells of an out-of-grid site (the reference applies LJ on the
// first shell only).
template <class S, bool COUNT>
__device__ __forceinline__ void pair_terms(const Tgt &T, const Shm &sh, const double (&p)[S::NS][3],
                                           const double (&qs)[S::NS], const double4 at, int ty,
                                           int allow_lj, Acc (&a)[S::NS], Work &w) {
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (!(S::lj(s) || S::coul(s) || S::dip(s))) continue;
    const double dx = p[s][0] - at.x, dy = p[s][1] - at.y, dz = p[s][2] - at.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (COUNT) w.visits++;
    if (S::CUT) {
      // r < cutoff as the reference tests it (sqrt(r2) < rc), decided on r2
      // except inside a relative band of 1e-15 around the cutoff
      bool in_lj = false, in_c = false;
      if (S::lj(s))
        in_lj = allow_lj && (r2 < T.rlj2lo || (r2 < T.rlj2hi && sqrt(r2) < T.rlj));
      if (S::coul(s) || S::dip(s)) in_c = (r2 < T.rc2lo || (r2 < T.rc2hi && sqrt(r2) < T.rc));
      if (in_lj || in_c) {
        const double rinv = rsqrt(r2);
        const double r2inv = rinv * rinv;
        if (S::lj(s) && in_lj) {
          const double *row = sh.lj + ty * 10 + (S::cen(s) ? 5 : 0);
          const double r6 = r2inv * r2inv * r2inv;
          a[s].U += r6 * (row[1] * r6 - row[0]) - row[4];
          const double flj = r6 * (row[3] * r6 - row[2]) * r2inv;
          a[s].fx += flj * dx;
          a[s].fy += flj * dy;
          a[s].fz += flj * dz;
          if (COUNT) w.lj++;
        }
        if (in_c) {
          const double r = r2 * rinv;
          const double sm = 1.0 - r * r2 * T.irc3;
          if (S::coul(s)) {
            const double qq = qs[s] * at.w * KC;
            a[s].U += qq * (rinv + T.cc1 + T.cc2 * r2);
            const double fc = qq * rinv * r2inv * sm;
            a[s].fx += fc * dx;
            a[s].fy += fc * dy;
            a[s].fz += fc * dz;
            if (COUNT) w.coul++;
          }
          if (S::dip(s)) {
            const double r3 = rinv * r2inv;
            const double A = at.w * (r3 - T.irc3);
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
            if (COUNT) w.dip++;
          }
        }
      }
    } else {
      // all-atom kernels (no cutoff, no shifts): src/Force.cpp:117-300, :495-574, :836-897
      const double rinv = rsqrt(r2);
      const double r2inv = rinv * rinv;
      if (S::lj(s)) {
        const double *row = sh.lj + ty * 10 + (S::cen(s) ? 5 : 0);
        const double r6 = r2inv * r2inv * r2inv;
        a[s].U += r6 * (row[1] * r6 - row[0]);
        const double flj = r6 * (row[3] * r6 - row[2]) * r2inv;
        a[s].fx += flj * dx;
        a[s].fy += flj * dy;
        a[s].fz += flj * dz;
        if (COUNT) w.lj++;
      }
      if (S::coul(s)) {
        const double uc = qs[s] * at.w * KC * rinv;
        a[s].U += uc;
        const double fc = uc * r2inv;
        a[s].fx += fc * dx;
        a[s].fy += fc * dy;
        a[s].fz += fc * dz;
        if (COUNT) w.coul++;
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
        if (COUNT) w.dip++;
      }
    }
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Candidate runs.  Mode 0 (probe inside the grid): one run per cell column
// (i, j) that the sphere of radius R around the centre reaches, covering the
// z-range of cells it cuts; atoms are stored in cell order with k fastest, so
// a column range is contiguous.  Mode 1 (site outside the grid): the
// reference stencils of the cell the reference's flat index aliases to.
// Mode 2: all atoms (no-cutoff force types).
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
      const int k0 = clampi((int)floor((c[2] - h - T.cz) / T.cell), 0, T.nz - 1);
      const int k1 = clampi((int)floor((c[2] + h - T.cz) / T.cell), 0, T.nz - 1);
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

// Flattens the runs of one pass over the group's lanes and accumulates.
template <class S, int L, bool LDS, bool COUNT>
__device__ void sweep(const Tgt &T, const Shm &sh, int lane, const RunGen &g, int nruns,
                      const double (&p)[S::NS][3], const double (&qs)[S::NS], Acc (&a)[S::NS],
                      Work &w) {
  for (int r0 = 0; r0 < nruns; r0 += L) {
    int start = 0, len = 0, first = 1;
    if (r0 + lane < nruns) g.get(T, r0 + lane, start, len, first);
    const int incl = gscan_incl<L>(len, lane);
    const int total = __shfl(incl, L - 1, L);
    sh.incl[lane] = incl;
    sh.base[lane] = (start - (incl - len)) * 2 + first;
    __builtin_amdgcn_wave_barrier();
    int cur = 0;
    for (int t = lane; t < total; t += L) {
      while (cur < L - 1 && t >= sh.incl[cur]) cur++;
      const int bw = sh.base[cur];
      const int at_i = (bw >> 1) + t;
      double4 at;
      int ty;
      fetch<LDS>(T, sh, at_i, at, ty);
      pair_terms<S, COUNT>(T, sh, p, qs, at, ty, bw & 1, a, w);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <class S, int L>
__device__ __forceinline__ void reduce(Acc (&a)[S::NS]) {
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

// Forces and energies of all sites at positions p (group-uniform result).
template <class S, int L, bool LDS, bool COUNT>
__device__ void evaluate(const Tgt &T, const Shm &sh, int lane, const double (&p)[S::NS][3],
                         const double (&qs)[S::NS], double (&f)[S::NS][3], double (&U)[S::NS],
                         Work &w) {
  Acc res[S::NS];
#pragma unroll
  for (int s = 0; s < S::NS; s++) acc_zero(res[s]);
  RunGen g;
  if (!S::CUT) {
    g.mode = 2;
    Acc a[S::NS];
#pragma unroll
    for (int s = 0; s < S::NS; s++) acc_zero(a[s]);
    sweep<S, L, LDS, COUNT>(T, sh, lane, g, 1, p, qs, a, w);
    reduce<S, L>(a);
#pragma unroll
    for (int s = 0; s < S::NS; s++) res[s] = a[s];
  } else {
    bool inbox = true;
#pragma unroll
    for (int s = 0; s < S::NS; s++) {
      const int i = (int)floor((p[s][0] - T.cx) / T.cell);
      const int j = (int)floor((p[s][1] - T.cy) / T.cell);
      const int k = (int)floor((p[s][2] - T.cz) / T.cell);
      inbox = inbox && i >= 0 && i < T.nx && j >= 0 && j < T.ny && k >= 0 && k < T.nz;
    }
    if (inbox) {
      // centre of the probe: the site the others are symmetric about
      const int cs = S::NS == 1 ? 0 : (S::NS == 3 && S::cen(1) ? 1 : S::NS - 1);
      double off2 = 0.0;
#pragma unroll
      for (int s = 0; s < S::NS; s++) {
        const double ex = p[s][0] - p[cs][0], ey = p[s][1] - p[cs][1], ez = p[s][2] - p[cs][2];
        off2 = fmax(off2, ex * ex + ey * ey + ez * ez);
      }
      const double R = (any_long<S>() ? T.rc : T.rlj) + sqrt(off2) + 1e-6;
      g.mode = 0;
      g.c[0] = p[cs][0];
      g.c[1] = p[cs][1];
      g.c[2] = p[cs][2];
      g.R2 = R * R;
      const int ix0 = clampi((int)floor((g.c[0] - R - T.cx) / T.cell), 0, T.nx - 1);
      const int ix1 = clampi((int)floor((g.c[0] + R - T.cx) / T.cell), 0, T.nx - 1);
      const int iy0 = clampi((int)floor((g.c[1] - R - T.cy) / T.cell), 0, T.ny - 1);
      const int iy1 = clampi((int)floor((g.c[1] + R - T.cy) / T.cell), 0, T.ny - 1);
      g.ix0 = ix0;
      g.iy0 = iy0;
      g.nys = iy1 - iy0 + 1;
      Acc a[S::NS];
#pragma unroll
      for (int s = 0; s < S::NS; s++) acc_zero(a[s]);
      sweep<S, L, LDS, COUNT>(T, sh, lane, g, (ix1 - ix0 + 1) * g.nys, p, qs, a, w);
      reduce<S, L>(a);
#pragma unroll
      for (int s = 0; s < S::NS; s++) res[s] = a[s];
    } else {
      // A site left the grid (exit step only): evaluate each site alone with
      // the reference's stencils of its (possibly aliased) cell index.
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
        Acc a[S::NS];
#pragma unroll
        for (int s = 0; s < S::NS; s++) acc_zero(a[s]);
        sweep<S, L, LDS, COUNT>(T, sh, lane, g, nr, pp, qs, a, w);
        reduce<S, L>(a);
        res[ps] = a[ps];
      }
    }
  }
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (S::dip(s)) {
      const Acc &a = res[s];
      f[s][0] = a.fx + T.alpha * (a.Ex * a.Exx + a.Ey * a.Exy + a.Ez * a.Exz);
      f[s][1] = a.fy + T.alpha * (a.Ex * a.Exy + a.Ey * a.Eyy + a.Ez * a.Eyz);
      f[s][2] = a.fz + T.alpha * (a.Ex * a.Exz + a.Ey * a.Eyz + a.Ez * a.Ezz);
      U[s] = a.U - 0.5 * T.alpha * (a.Ex * a.Ex + a.Ey * a.Ey + a.Ez * a.Ez);
    } else {
      f[s][0] = res[s].fx;
      f[s][1] = res[s].fy;
      f[s][2] = res[s].fz;
      U[s] = res[s].U;
    }
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
__device__ bool setup(const Tgt &T, const double *u, double (&pos)[3][3], double (&vel)[3][3],
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

// RATTLE position stage, src/System.cpp:1700-1739
__device__ __forceinline__ void rattle1(double *ri, double *rj, double *vi, double *vj,
                                        const double *fi, const double *fj, double mi, double mj,
                                        double dt, double d2) {
  const double muij = mi * mj / (mi + mj);
  double rci[3], rcj[3], vci[3], vcj[3], rcij[3], rij[3];
  for (int k = 0; k < 3; k++) {
    vci[k] = vi[k] + 0.5 * dt / mi * fi[k] * KCONV;
    vcj[k] = vj[k] + 0.5 * dt / mj * fj[k] * KCONV;
    rci[k] = ri[k] + dt * vci[k];
    rcj[k] = rj[k] + dt * vcj[k];
    rij[k] = ri[k] - rj[k];
    rcij[k] = rci[k] - rcj[k];
  }
  const double r2 = dot3(rij, rij), rc2 = dot3(rcij, rcij), rr = dot3(rcij, rij);
  double det = rr * rr - r2 * (rc2 - d2);
  if (det < 0.0) det = 0.0;
  const double lam = (rr - sqrt(det)) / r2;
  for (int k = 0; k < 3; k++) {
    vi[k] = vci[k] - muij / (mi * dt) * lam * rij[k];
    vj[k] = vcj[k] + muij / (mj * dt) * lam * rij[k];
    ri[k] = rci[k] - muij / mi * lam * rij[k];
    rj[k] = rcj[k] + muij / mi * lam * rij[k];  // the reference uses mi here too
  }
}

// RATTLE velocity stage, src/System.cpp:1741-1767
__device__ __forceinline__ void rattle2(const double *ri, const double *rj, double *vi, double *vj,
                                        const double *fi, const double *fj, double mi, double mj,
                                        double dt) {
  const double muij = mi * mj / (mi + mj);
  double rij[3], vcij[3], vci[3], vcj[3];
  for (int k = 0; k < 3; k++) {
    vci[k] = vi[k] + 0.5 * dt / mi * fi[k] * KCONV;
    vcj[k] = vj[k] + 0.5 * dt / mj * fj[k] * KCONV;
    rij[k] = ri[k] - rj[k];
    vcij[k] = vci[k] - vcj[k];
  }
  const double lam = muij * dot3(rij, vcij) / dot3(rij, rij);
  for (int k = 0; k < 3; k++) {
    vi[k] = vci[k] - lam / mi * rij[k];
    vj[k] = vcj[k] + lam / mj * rij[k];
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
    v = gsum_u<L>(w.visits);
    l = gsum_u<L>(w.lj);
    c = gsum_u<L>(w.coul);
    d = gsum_u<L>(w.dip);
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

__device__ void stage(const Tgt &T, Shm &sh, char *smem, bool lds_atoms) {
  // layout: atoms (double4) | lj rows | run scratch | types
  char *p = smem;
  double4 *atoms = reinterpret_cast<double4 *>(p);
  if (lds_atoms) p += sizeof(double4) * (size_t)T.natoms;
  double *lj = reinterpret_cast<double *>(p);
  p += sizeof(double) * 10 * (size_t)T.ntypes;
  int *scratch = reinterpret_cast<int *>(p);
  p += sizeof(int) * 2 * blockDim.x;
  uint8_t *type = reinterpret_cast<uint8_t *>(p);
  for (int i = threadIdx.x; i < 10 * T.ntypes; i += blockDim.x) lj[i] = T.lj[i];
  if (lds_atoms) {
    for (int i = threadIdx.x; i < T.natoms; i += blockDim.x) {
      atoms[i] = T.atoms[i];
      type[i] = T.type[i];
    }
  }
  __syncthreads();
  sh.lj = lj;
  sh.atoms = atoms;
  sh.type = type;
  sh.incl = scratch;
  sh.base = scratch;
}

template <int L>
__device__ __forceinline__ void group_scratch(Shm &sh, char *smem, const Tgt &T, bool lds_atoms) {
  char *p = smem;
  if (lds_atoms) p += sizeof(double4) * (size_t)T.natoms;
  p += sizeof(double) * 10 * (size_t)T.ntypes;
  int *scratch = reinterpret_cast<int *>(p);
  const int g = threadIdx.x / L;
  sh.incl = scratch + 2 * L * g;
  sh.base = scratch + 2 * L * g + L;
}

// ---------------------------------------------------------------- kernels
// Persistent trajectory kernel.  Each group of L lanes pulls trajectory
// indices from *queue until n is reached; per loop iteration every live
// group performs exactly one force evaluation: the initial force of a new
// (or restarted) trajectory, or the force of one velocity-Verlet/RATTLE step.
template <int GK, int FT, int L, bool LDS, bool COUNT>
__global__ __launch_bounds__(512) void traj_kernel(Tgt T, const double *__restrict__ u, uint64_t n,
                                                   mccs_traj_record *__restrict__ rec,
                                                   unsigned long long *__restrict__ queue) {
  using S = Spec<GK, FT>;
  extern __shared__ __align__(16) char smem[];
  Shm sh;
  stage(T, sh, smem, LDS);
  group_scratch<L>(sh, smem, T, LDS);
  const int lane = threadIdx.x % L;

  double qs[S::NS];
#pragma unroll
  for (int s = 0; s < S::NS; s++) qs[s] = S::q(s);

  int status = ST_NEED;
  uint64_t j = 0;
  int tries = 0, max_tries = 8, steps = 0;
  double dt = T.dt, dtheta = 0.0, Ei = 0.0;
  double pos[3][3], vel[3][3], fs[3][3], Us[3];
  double rcm[3], vcm[3], rold[3], vcm_i[3];
  Work w{0, 0, 0, 0};

  while (status != ST_DONE) {
    // ---- phase A: new trajectory, restart, or first half-step
    while (status == ST_NEED) {
      unsigned long long t = 0;
      if (lane == 0) t = atomicAdd(queue, 1ull);
      t = __shfl(t, 0, L);
      if (t >= n) {
        status = ST_DONE;
        break;
      }
      j = t;
      w = Work{0, 0, 0, 0};
      if (setup<S>(T, u + 9 * j, pos, vel, rcm, vcm)) {
        status = ST_INIT;
      } else {
        write_record<L>(rec, j, lane, 0, 0, 0, 0, 0.0, T.bmax, w, COUNT);
      }
    }
    if (status == ST_DONE) break;
    if (status == ST_RESTART) {
      setup<S>(T, u + 9 * j, pos, vel, rcm, vcm);
    } else if (status == ST_STEP) {
      if (S::NS == 1) {
        for (int d = 0; d < 3; d++) {
          vel[0][d] += 0.5 * dt / T.mu * fs[0][d] * KCONV;
          pos[0][d] += dt * vel[0][d];
          rcm[d] = pos[0][d];
        }
      } else {
        const int a = 0, b = GK == 2 ? 1 : 2;  // the constrained pair (N-N or O-O)
        double fi[3], fj[3];
        pair_forces<S, GK>(fs, fi, fj);
        rattle1(pos[a], pos[b], vel[a], vel[b], fi, fj, S::m(a), S::m(b), dt, T.d2);
        const int mid = GK == 2 ? 2 : 1;
        for (int d = 0; d < 3; d++) {
          rcm[d] = 0.5 * (pos[a][d] + pos[b][d]);
          vcm[d] = 0.5 * (vel[a][d] + vel[b][d]);
          pos[mid][d] = rcm[d];
        }
      }
      dtheta += anglevec(rcm, rold);
      for (int d = 0; d < 3; d++) rold[d] = rcm[d];
    }

    // ---- phase B: one force evaluation for every live group
    {
      double p[S::NS][3], f[S::NS][3], U[S::NS];
#pragma unroll
      for (int s = 0; s < S::NS; s++)
        for (int d = 0; d < 3; d++) p[s][d] = pos[s][d];
      evaluate<S, L, LDS, COUNT>(T, sh, lane, p, qs, f, U, w);
#pragma unroll
      for (int s = 0; s < S::NS; s++) {
        Us[s] = U[s];
        for (int d = 0; d < 3; d++) fs[s][d] = f[s][d];
      }
    }
    double Utot = 0.0;
#pragma unroll
    for (int s = 0; s < S::NS; s++) Utot += Us[s];

    // ---- phase C: bookkeeping, second half-step, exit tests
    if (status == ST_INIT || status == ST_RESTART) {
      if (status == ST_INIT) {
        double Ek = 0.0;
        if (S::NS == 1) Ek = kinetic(T.mu, vel[0]);
        else
          for (int s = 0; s < S::NS; s++) Ek += kinetic(S::m(s), vel[s]);
        Ei = Utot + Ek;
        tries = 0;
        max_tries = 8;
        steps = 0;
        dt = T.dt;
        for (int d = 0; d < 3; d++) vcm_i[d] = vcm[d];
        if (GK == 3) vcm_i[0] = vcm[1];  // src/System.cpp:1012 reads vcm[1] for x
      }
      dtheta = 0.0;
      for (int d = 0; d < 3; d++) rold[d] = rcm[d];
      if (GK == 3) {  // CO2 restarts from vcm_i (src/System.cpp:1134)
        for (int d = 0; d < 3; d++) vcm[d] = vcm_i[d];
      }
      status = ST_STEP;
      continue;
    }
    // ST_STEP: second half
    if (S::NS == 1) {
      for (int d = 0; d < 3; d++) {
        vel[0][d] += 0.5 * dt / T.mu * fs[0][d] * KCONV;
        vcm[d] = vel[0][d];
      }
    } else {
      const int a = 0, b = GK == 2 ? 1 : 2;
      double fi[3], fj[3];
      pair_forces<S, GK>(fs, fi, fj);
      rattle2(pos[a], pos[b], vel[a], vel[b], fi, fj, S::m(a), S::m(b), dt);
      for (int d = 0; d < 3; d++) vcm[d] = 0.5 * (vel[a][d] + vel[b][d]);
      if (GK == 3)
        for (int d = 0; d < 3; d++) vel[1][d] = vcm[d];
    }
    steps++;
    bool restart = false;
    if (steps >= MAX_STEPS || !isfinite(rcm[0] + rcm[1] + rcm[2] + vcm[0] + vcm[1] + vcm[2])) {
      // the reference would loop forever here (zero speed or non-finite
      // state); the trajectory is reported lost with tries = -1
      write_record<L>(rec, j, lane, 1, 0, -1, steps, 0.0, T.bmax, w, COUNT);
      status = ST_NEED;
      continue;
    }
    if (dtheta > 1.5 * PI) {
      tries++;
      dt = T.dt / (tries + 1);
      restart = true;
    } else if (fabs(rcm[0]) >= T.hx || fabs(rcm[1]) >= T.hy || fabs(rcm[2]) >= T.hz ||
               (rcm[0] / T.ea) * (rcm[0] / T.ea) + (rcm[1] / T.eb) * (rcm[1] / T.eb) +
                       (rcm[2] / T.ec) * (rcm[2] / T.ec) >
                   1.0) {
      double Ek = 0.0, dE;
      if (S::NS == 1) {
        Ek = kinetic(T.mu, vcm);
        dE = fabs(Utot + Ek - Ei) / Ei * 100.0;
      } else {
        const int ns_ke = GK == 2 ? 2 : 3;  // N2 excludes the massless dummy
        for (int s = 0; s < ns_ke; s++) Ek += kinetic(S::m(s), vel[s]);
        dE = fabs((Utot + Ek - Ei) / Ei) * 100.0;
      }
      const double tol = GK == 1 ? 0.5 : (GK == 2 ? 1.0 : 20.0);
      if (dE > tol) {
        tries++;
        max_tries++;
        dt = T.dt / (tries + 1);
        restart = true;
      } else {
        const double chi = anglevec(vcm_i, vcm);
        write_record<L>(rec, j, lane, 1, 1, tries, steps, chi, T.bmax, w, COUNT);
        status = ST_NEED;
        continue;
      }
    }
    if (restart) {
      if (tries < max_tries && max_tries < 10) {
        status = ST_RESTART;
      } else {
        write_record<L>(rec, j, lane, 1, 0, tries, steps, 0.0, T.bmax, w, COUNT);
        status = ST_NEED;
      }
    }
  }
}


using TrajFn = void (*)(Tgt, const double *, uint64_t, mccs_traj_record *, unsigned long long *);

template <int GK, int FT, int L>
TrajFn pick_traj(bool lds, bool count) {
  if (lds) return count ? traj_kernel<GK, FT, L, true, true> : traj_kernel<GK, FT, L, true, false>;
  return count ? traj_kernel<GK, FT, L, false, true> : traj_kernel<GK, FT, L, false, false>;
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

template <int GK, int FT>
TrajFn pick_traj_basic(int L, bool lds, bool count) {
  if (L != 64 || !count) return nullptr;
  return lds ? traj_kernel<GK, FT, 64, true, true> : traj_kernel<GK, FT, 64, false, true>;
}

TrajFn select_traj_gk1(int ft, int L, bool lds, bool count);
TrajFn select_traj_gk2(int ft, int L, bool lds, bool count);
TrajFn select_traj_gk3(int ft, int L, bool lds, bool count);

}  // namespace dev
}  // namespace mccs
