// MI355X (gfx950) trajectory engine, lane-per-trajectory form.
//
// Work decomposition.  Every lane of a wavefront owns one trajectory slot and
// integrates it alone: setup, velocity Verlet or RATTLE, the angle sum, the
// exit and energy tests, the retry rule and the device work queue all run
// SIMT, one trajectory per lane.  The site forces of the slots that need them
// are then evaluated by the whole wavefront, one slot after the other: the
// slot's site positions are broadcast through scalar registers, its Verlet
// list (one per slot, in HBM) is strided over the 64 lanes, and the lane
// partial sums are reduced through an LDS transpose whose totals come back
// through scalar registers and are written into the owner lane.  Slots whose
// sites are all farther than the longest cutoff from the target's bounding box
// (the reference's force is exactly zero there) cost no wavefront work at all.
//
// Compared to a group of lanes per trajectory (kernels.hpp, traj_kernel),
// the integration is done once per trajectory instead of once per lane, and
// a site evaluation's reduction is one transpose instead of a butterfly per
// accumulator.  The physics, the queue protocol and the results are the same:
// every try builds its own neighbour list, the list order and the reduction
// order depend only on the trajectory, so records do not depend on which lane
// or wavefront ran what.
//
// Reference behaviour restated (cces-cepid/massccs): setup
// src/System.cpp:1397-1518, run_He / run_N2 / run_CO2 src/System.cpp:338-1392,
// RATTLE src/System.cpp:1700-1767, site kernels src/Force.cpp, cell index and
// stencils src/LinkedCell.cpp:117-129, :228-440.
#pragma once
#include "kernels.hpp"

namespace mccs {
namespace dev {

constexpr int WBLOCK = 512;      // threads per block: 8 wavefronts, 2 per SIMD
constexpr int RED_ROWS = 8;      // accumulators reduced per transpose pass
constexpr int RED_STRIDE = 66;   // doubles per transpose row (64 lanes + pad)

// Per-slot Verlet lists of one launch.
struct WList {
  void *base;  // slots x cap indices (uint16 when atoms are staged in LDS)
  int cap;     // entries per slot
  int slots;   // trajectory slots per wavefront (16, 32 or 64)
  int period;  // ordinary tries step on every period-th iteration, urgent ones on every one
  int urgent;  // a first try becomes urgent after this many steps (later tries always are)
};

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int lane) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(v & 0xffffffffu), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// Global-memory writes of one lane visible to the loads of the others (the
// Verlet list a wavefront has just built).
__device__ __forceinline__ void glist_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Work of one site evaluation, counted with wavefront ballots (the value in
// lane 0 is the wavefront's total: lane 0 runs every iteration of the strided
// loops).
struct WorkU {
  uint32_t visits, lj, coul, dip;
};

// FP64 accumulators a site evaluation reduces: U and F for every site with
// terms, the field and its gradient for dipole sites.
template <class S>
__host__ __device__ constexpr int nacc() {
  int k = 0;
  for (int s = 0; s < S::NS; s++) {
    if (S::lj(s) || S::coul(s) || S::dip(s)) k += 4;
    if (S::dip(s)) k += 9;
  }
  return k;
}

template <class S>
__device__ __forceinline__ void pack(const Acc (&a)[S::NS], double (&v)[nacc<S>()]) {
  int k = 0;
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    if (S::lj(s) || S::coul(s) || S::dip(s)) {
      v[k++] = a[s].U; v[k++] = a[s].fx; v[k++] = a[s].fy; v[k++] = a[s].fz;
    }
    if (S::dip(s)) {
      v[k++] = a[s].Ex; v[k++] = a[s].Ey; v[k++] = a[s].Ez;
      v[k++] = a[s].Exx; v[k++] = a[s].Eyy; v[k++] = a[s].Ezz;
      v[k++] = a[s].Exy; v[k++] = a[s].Exz; v[k++] = a[s].Eyz;
    }
  }
}

template <class S>
__device__ __forceinline__ void unpack(const double (&v)[nacc<S>()], Acc (&a)[S::NS]) {
  int k = 0;
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    acc_zero(a[s]);
    if (S::lj(s) || S::coul(s) || S::dip(s)) {
      a[s].U = v[k++]; a[s].fx = v[k++]; a[s].fy = v[k++]; a[s].fz = v[k++];
    }
    if (S::dip(s)) {
      a[s].Ex = v[k++]; a[s].Ey = v[k++]; a[s].Ez = v[k++];
      a[s].Exx = v[k++]; a[s].Eyy = v[k++]; a[s].Ezz = v[k++];
      a[s].Exy = v[k++]; a[s].Exz = v[k++]; a[s].Eyz = v[k++];
    }
  }
}

// Sums each of the K per-lane values over the 64 lanes; every lane receives
// the same bits.  Eight values at a time are transposed through LDS (row i =
// value i of every lane); lane 8r+p then adds the eight entries [8p, 8p+8) of
// row r, three DPP steps combine the eight partial sums of a row, and the
// totals are read from lanes 0, 8, ..., 56.  The order of every addition is
// fixed, so the result does not depend on scheduling.
template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K], double *red, int lane) {
#pragma unroll
  for (int c0 = 0; c0 < K; c0 += RED_ROWS) {
#pragma unroll
    for (int i = 0; i < RED_ROWS; i++)
      if (c0 + i < K) red[i * RED_STRIDE + lane] = v[c0 + i];
    lds_fence();
    const int r = lane >> 3, p = lane & 7;
    double s = 0.0;
    if (c0 + r < K) {
      const double2 *q = reinterpret_cast<const double2 *>(red + r * RED_STRIDE + 8 * p);
      const double2 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3];
      s = ((a0.x + a0.y) + (a1.x + a1.y)) + ((a2.x + a2.y) + (a3.x + a3.y));
    }
    s += dpp<0xB1>(s);   // quad_perm [1,0,3,2]
    s += dpp<0x4E>(s);   // quad_perm [2,3,0,1]
    s += dpp<0x141>(s);  // row_half_mirror: the eight lanes of a half row hold its sum
#pragma unroll
    for (int i = 0; i < RED_ROWS; i++)
      if (c0 + i < K) v[c0 + i] = readlane_d(s, 8 * i);
    lds_fence();
  }
}

// Pair terms of one candidate atom against every site (the arithmetic of
// pair_terms, src/Force.cpp per term), with the work counted by ballots.
template <class S, bool COUNT>
__device__ __forceinline__ void pair_w(const Tgt &T, const double (&p)[S::NS][3],
                                       const double (&qs)[S::NS], const double4 at,
                                       const double2 lp, const double2 lc, int allow_lj,
                                       Acc (&a)[S::NS], WorkU &w) {
  Work lw{0, 0, 0, 0};
  pair_terms<S, false>(T, p, qs, at, lp, lc, allow_lj, a, lw);
  if (COUNT) {
#pragma unroll
    for (int s = 0; s < S::NS; s++) {
      if (!(S::lj(s) || S::coul(s) || S::dip(s))) continue;
      const double dx = p[s][0] - at.x, dy = p[s][1] - at.y, dz = p[s][2] - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      w.visits += (uint32_t)__popcll(__ballot(1));
      if (S::CUT) {
        bool in_lj = false, in_c = false;
        if (S::lj(s)) in_lj = allow_lj && (r2 < T.rlj2lo || (r2 < T.rlj2hi && sqrt(r2) < T.rlj));
        if (S::coul(s) || S::dip(s)) in_c = (r2 < T.rc2lo || (r2 < T.rc2hi && sqrt(r2) < T.rc));
        if (S::lj(s)) w.lj += (uint32_t)__popcll(__ballot(in_lj));
        if (S::coul(s)) w.coul += (uint32_t)__popcll(__ballot(in_c));
        if (S::dip(s)) w.dip += (uint32_t)__popcll(__ballot(in_c));
      } else {
        const uint32_t c = (uint32_t)__popcll(__ballot(1));
        if (S::lj(s)) w.lj += c;
        if (S::coul(s)) w.coul += c;
        if (S::dip(s)) w.dip += c;
      }
    }
  }
}

// Verlet list of one slot, built by the whole wavefront into the slot's HBM
// list: every atom within R of c, in the order of the cell columns (the same
// criterion and order as build_list: atoms within Rin first).  Returns the
// count (> cap: overflowed); nin receives the count within Rin.
template <class S, bool LDS>
__device__ int wbuild_list(const Tgt &T, const Shm &sh, int lane, LIdx<LDS> *__restrict__ out,
                           int cap, const double c[3], double R, double Rin, int &nin) {
  RunGen g;
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
  const double R2 = R * R, Rin2 = Rin * Rin;
  const int nruns = (ix1 - ix0 + 1) * g.nys;
  int n = 0;
  nin = 0;
  for (int pass = 0; pass < 2; pass++) {
  const double lo2 = pass ? Rin2 : -1.0, hi2 = pass ? R2 : Rin2;
  if (pass) nin = n;
  for (int r0 = 0; r0 < nruns; r0 += 64) {
    int start = 0, len = 0, first = 1;
    if (r0 + lane < nruns) g.get(T, r0 + lane, start, len, first);
    const int incl = gscan_incl<64>(len, lane);
    const int total = __shfl(incl, 63, 64);
    sh.incl[lane] = incl;
    sh.base[lane] = start - (incl - len);
    lds_fence();
    int cur = 0;
    for (int t0 = 0; t0 < total; t0 += 64) {
      const int t = t0 + lane;
      bool keep = false;
      int a = 0;
      if (t < total) {
        while (cur < 63 && t >= sh.incl[cur]) cur++;
        a = sh.base[cur] + t;
        const double4 at = LDS ? sh.atoms[a] : T.atoms[a];
        const double ex = at.x - c[0], ey = at.y - c[1], ez = at.z - c[2];
        const double d2 = ex * ex + ey * ey + ez * ez;
        keep = d2 >= lo2 && d2 < hi2;
      }
      const unsigned long long m = __ballot(keep);
      const int below = __popcll(m & ((1ull << lane) - 1ull));
      if (keep && n + below < cap) out[n + below] = (LIdx<LDS>)a;
      n += __popcll(m);
    }
    lds_fence();
  }
  }
  glist_fence();
  return n;
}

// Evaluation modes of a slot's sites.
enum : int { EV_NONE = 0, EV_LIST = 1, EV_STENCIL = 2, EV_ALL = 3 };

// Forces and energies of all sites at the (wavefront-uniform) positions p,
// evaluated by the whole wavefront.  EV_LIST: the slot's Verlet list (rebuilt
// first when build is set; vl is the slot's list state, updated in place),
// or the cell columns when the list overflowed; EV_STENCIL: a site outside the
// grid, each site alone with the reference's stencils of its aliased cell;
// EV_ALL: every atom (force types without cutoffs).
template <class S, bool LDS, bool COUNT>
__device__ void wave_force(const Tgt &T, const Shm &sh, double *red, int lane,
                           const double (&p)[S::NS][3], const double (&qs)[S::NS], int md,
                           double off, bool build, LIdx<LDS> *__restrict__ lst, int cap,
                           VList &vl, double (&f)[S::NS][3], double (&U)[S::NS], WorkU &wu) {
  constexpr bool CEN = any_central<S>();
  constexpr int K = nacc<S>();
  WorkU w{0, 0, 0, 0};
  Acc a[S::NS];
#pragma unroll
  for (int s = 0; s < S::NS; s++) acc_zero(a[s]);
#pragma unroll
  for (int s = 0; s < S::NS; s++) {
    f[s][0] = f[s][1] = f[s][2] = 0.0;
    U[s] = 0.0;
  }
  auto visit = [&](int ai, int first) {
    double4 at;
    double2 lp, lc;
    fetch<LDS, CEN>(T, sh, ai, at, lp, lc);
    pair_w<S, COUNT>(T, p, qs, at, lp, lc, first, a, w);
  };
  if (md == EV_STENCIL) {
    // a site left the grid (exit step only): each site alone, as evaluate()
#pragma unroll
    for (int ps = 0; ps < S::NS; ps++) {
      if (!(S::lj(ps) || S::coul(ps) || S::dip(ps))) continue;
      const int i = (int)floor((p[ps][0] - T.cx) / T.cell);
      const int j = (int)floor((p[ps][1] - T.cy) / T.cell);
      const int k = (int)floor((p[ps][2] - T.cz) / T.cell);
      const long long idx = (long long)k + (long long)T.nz * j + (long long)T.nz * T.ny * i;
      if (idx < 0 || idx >= T.ncells) continue;  // zero force, src/Force.cpp:58-66
      RunGen g;
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
        pair_w<S, COUNT>(T, pp, qs, at, lp, lc, first, b, w);
      };
      walk_runs<64>(T, sh, lane, g, nr, visit1);
      double v[K];
      pack<S>(b, v);
      wave_sum<K>(v, red, lane);
      unpack<S>(v, b);
      finish<S>(T, b, ps, f, U);
    }
  } else {
    if (md == EV_ALL) {
      RunGen g;
      g.mode = 2;
      walk_runs<64>(T, sh, lane, g, 1, visit);
    } else {
      constexpr int cs = S::centre();
      const double c[3] = {p[cs][0], p[cs][1], p[cs][2]};
      if (build) {
        const double R = T.rlong + off + T.skin;
        int nin = 0;
        const int n = wbuild_list<S, LDS>(T, sh, lane, lst, cap, c, R, T.rlj + off + T.skin, nin);
        vl.nin = nin;
        vl.c0[0] = c[0];
        vl.c0[1] = c[1];
        vl.c0[2] = c[2];
        vl.R = R;
        vl.n = n;
        vl.valid = n <= cap;
      }
      if (vl.valid > 0) {
        const int n = vl.n, nin = vl.nin;
        // two-stage pipeline: the index one chunk ahead is loaded from HBM
        // while the atom of the next chunk is read from LDS (or L2)
        double4 at = make_double4(0.0, 0.0, 0.0, 0.0);
        double2 lp = make_double2(0.0, 0.0), lc = make_double2(0.0, 0.0);
        int nxt = 0;
        if (lane < n) fetch<LDS, CEN>(T, sh, (int)lst[lane], at, lp, lc);
        if (lane + 64 < n) nxt = (int)lst[lane + 64];
        for (int t = lane; t < n; t += 64) {
          const double4 cat = at;
          const double2 clp = lp, clc = lc;
          if (t + 64 < n) fetch<LDS, CEN>(T, sh, nxt, at, lp, lc);
          if (t + 128 < n) nxt = (int)lst[t + 128];
          pair_w<S, COUNT>(T, p, qs, cat, clp, clc, t < nin, a, w);
        }
      } else {
        // the list overflowed its capacity: walk the columns around the centre
        const double R = T.rlong + off + 1e-6;
        RunGen g;
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
        walk_runs<64>(T, sh, lane, g, (ix1 - ix0 + 1) * g.nys, visit);
      }
    }
    double v[K];
    pack<S>(a, v);
    wave_sum<K>(v, red, lane);
    unpack<S>(v, a);
    finish<S>(T, a, -1, f, U);
  }
  if (COUNT) {
    wu.visits = (uint32_t)readlane_i((int)w.visits, 0);
    wu.lj = (uint32_t)readlane_i((int)w.lj, 0);
    wu.coul = (uint32_t)readlane_i((int)w.coul, 0);
    wu.dip = (uint32_t)readlane_i((int)w.dip, 0);
  }
}

// Block LDS of the wave kernel: atoms and LJ pairs (when staged) | per
// wavefront: transpose rows, run scratch (2 x 64 ints).
__host__ __device__ inline size_t wave_lds_bytes(int natoms, bool lds_atoms, bool central) {
  size_t b = 0;
  if (lds_atoms) b += (sizeof(double4) + sizeof(double2) * (central ? 2 : 1)) * (size_t)natoms;
  b += (sizeof(double) * RED_ROWS * RED_STRIDE + sizeof(int) * 128) * (WBLOCK / 64);
  return b + 64;
}

// Persistent lane-per-trajectory kernel.  Per iteration of the main loop:
//   A. lanes without work take a work item (primary tickets are handed out
//      per wavefront with one atomic) and set it up; stepping lanes make the
//      first half step;
//   B. the wavefront evaluates the forces of every slot whose sites are near
//      the target (the others get the exact zero force);
//   C. lanes that set up a try initialise it; stepping lanes make the second
//      half step, the exit and energy tests, and resolve an ended try.
// The queue protocol (tries, side-by-side spawned tries, resolve) is the one
// of traj_kernel, run by each lane for its own slot.
template <int GK, int FT, bool LDS, bool COUNT>
__global__ __launch_bounds__(WBLOCK) __attribute__((amdgpu_waves_per_eu(2))) void wave_kernel(
    const Tgt *__restrict__ tg, const double *__restrict__ u, uint64_t n,
    mccs_traj_record *__restrict__ rec, Queue Q, WList WL) {
  using S = Spec<GK, FT>;
  constexpr bool CEN = any_central<S>();
  constexpr int NS = S::NS;
  constexpr int NP = NS == 1 ? 1 : 2;          // integrated particles
  constexpr int PA = 0, PB = GK == 2 ? 1 : 2;  // sites of the constrained pair
  constexpr int CS = S::centre();
  constexpr double MA = NS == 1 ? 1.0 : S::m(PA), MB = NS == 1 ? 1.0 : S::m(PB);
  extern __shared__ __align__(16) char smem[];
  const Tgt &T = tg[0];

  // ---- LDS staging
  char *pc = smem;
  double4 *atoms = reinterpret_cast<double4 *>(pc);
  if (LDS) pc += sizeof(double4) * (size_t)T.natoms;
  double2 *ljp = reinterpret_cast<double2 *>(pc);
  if (LDS) pc += sizeof(double2) * (size_t)T.natoms;
  double2 *ljc = reinterpret_cast<double2 *>(pc);
  if (LDS && CEN) pc += sizeof(double2) * (size_t)T.natoms;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *red = reinterpret_cast<double *>(pc) + (size_t)wave * RED_ROWS * RED_STRIDE;
  pc += sizeof(double) * RED_ROWS * RED_STRIDE * (WBLOCK / 64);
  int *scratch = reinterpret_cast<int *>(pc) + 128 * wave;
  if (LDS) {
    for (int i = threadIdx.x; i < T.natoms; i += blockDim.x) {
      atoms[i] = T.atoms[i];
      ljp[i] = T.ljp[i];
      if (CEN) ljc[i] = T.ljc[i];
    }
  }
  __syncthreads();
  Shm sh;
  sh.state = nullptr;
  sh.list = nullptr;
  sh.atoms = atoms;
  sh.ljp = ljp;
  sh.ljc = ljc;
  sh.incl = scratch;
  sh.base = scratch + 64;
  sh.red = nullptr;
  using IDX = LIdx<LDS>;
  IDX *lists = static_cast<IDX *>(WL.base) +
               ((uint64_t)blockIdx.x * (WBLOCK / 64) + wave) * (uint64_t)WL.slots * WL.cap;

  double qs[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) qs[s] = S::q(s);

  // ---- per-lane trajectory state
  double r[NP][3], v[NP][3], fp[NP][3];  // positions, velocities, forces of the integrated particles
  double rc[3];                          // centre site (the previous centre during a step)
  double vci[3];                         // vcm at setup (the chi reference after initialisation)
  double Ei = 0.0, dt = 0.0, dth = 0.0, Ut = 0.0, hdm = 0.0;
  Rat rat{0.0, 0.0, 0.0, 0.0, 0.0};
  VList vl;
  vl.valid = -1;
  vl.n = 0;
  vl.nin = 0;
  vl.R = 0.0;
  vl.c0[0] = vl.c0[1] = vl.c0[2] = 0.0;
  for (int q = 0; q < NP; q++)
    for (int d = 0; d < 3; d++) r[q][d] = v[q][d] = fp[q][d] = 0.0;
  for (int d = 0; d < 3; d++) rc[d] = vci[d] = 0.0;
  uint64_t j = 0;
  int k = 0, steps = 0, cont = 0, has_slot = 0, nomore = 0;
  unsigned long long slot = 0, wait_since = 0;
  uint32_t wv = 0, wl = 0, wc = 0, wd = 0;
  int status = lane < WL.slots ? ST_NEED : ST_DONE;

  // Iterations.  A try's step rate is one step per iteration of its
  // wavefront, and an iteration costs one force evaluation per stepping slot.
  // Ordinary tries step on every period-th ("major") iteration only; urgent
  // tries (restarts, and first tries that have run long) step on every
  // iteration, so that the long restart chains, which bound the launch's
  // duration, advance at nearly the rate of a wavefront of their own.  Work
  // is taken and the queue polled on major iterations only.
  for (unsigned it = 0;; it++) {
    const bool major = it % (unsigned)WL.period == 0;
    // ================================================================ A
    if (__ballot(status == ST_NEED && (major || cont))) {
      double pos[3][3], vel[3][3], rcm[3], vcm[3];
      bool got = false;
      if (status == ST_NEED && cont) {  // the next try of the trajectory whose try failed
        setup<S>(T, u + 9 * j, pos, vel, rcm, vcm);  // a trajectory with tries hits
        got = true;
        cont = 0;
      }
      // primary tickets, handed out per wavefront
      while (major) {
        const bool want = status == ST_NEED && !got && !nomore;
        const unsigned long long m = __ballot(want);
        if (!m) break;
        const int leader = __builtin_ctzll(m);
        unsigned long long t0 = 0;
        if (lane == leader) {
          t0 = __hip_atomic_load(Q.ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (t0 < n) t0 = atomicAdd(Q.ticket, (unsigned long long)__popcll(m));
        }
        t0 = readlane_u64(t0, leader);
        if (want) {
          const uint64_t t = t0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
          if (t < n) {
            j = t;
            k = 0;
            if (setup<S>(T, u + 9 * j, pos, vel, rcm, vcm)) {
              got = true;
            } else {  // a miss: free trajectory, recorded at once
              mccs_traj_record rr;
              rr.hit = 0;
              rr.success = 0;
              rr.tries = 0;
              rr.steps = 0;
              rr.chi = 0.0;
              rr.domega = 0.0;
              rr.visits = rr.lj = rr.coul = rr.dip = 0;
              rec[j] = rr;
              __threadfence();
              atomicAdd((unsigned long long *)Q.active, ~0ull);
            }
          } else {
            nomore = 1;
          }
        }
      }
      // spawned tries, then the end test (lanes without a primary ticket)
      if (major && status == ST_NEED && !got) {
        bool done = false, wait = false;
        if (Q.spec_below <= 0) {
          done = true;  // nothing is ever spawned
        } else {
          unsigned long long sl = ~0ull;
          if (has_slot) sl = slot;
          else if (ld_rlx(Q.shead) < ld_rlx(Q.stail)) sl = atomicAdd(Q.shead, 1ull);
          if (sl != ~0ull) {
            const unsigned long long item = sl < Q.nslots ? ld_acq(Q.slots + sl) : ~0ull;
            has_slot = item == ~0ull;  // written soon by its spawner, or never
            slot = sl;
            if (item != ~0ull && !ld_rlx(Q.resolved + (item >> 4))) {
              j = item >> 4;
              k = (int)(item & 15);
              setup<S>(T, u + 9 * j, pos, vel, rcm, vcm);  // a trajectory with tries hits
              got = true;
            }
          }
          if (!got) {
            const long long act = __hip_atomic_load(Q.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (act <= 0) done = true;
            else wait = true;
          }
        }
        if (wait) {
          const unsigned long long now = wall_clock64();
          if (!wait_since) wait_since = now;
          if (now - wait_since > Q.wait_ticks) {
            __hip_atomic_store(Q.stalled, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            done = true;
          }
        }
        if (done) status = ST_DONE;
      }
      if (got) {
        wait_since = 0;
        wv = wl = wc = wd = 0;
        vl.valid = -1;  // each try builds its own neighbour list
        if (NS == 1) {
          for (int d = 0; d < 3; d++) {
            r[0][d] = pos[0][d];
            v[0][d] = vel[0][d];
          }
        } else {
          for (int d = 0; d < 3; d++) {
            r[0][d] = pos[PA][d];
            r[1][d] = pos[PB][d];
            v[0][d] = vel[PA][d];
            v[1][d] = vel[PB][d];
          }
        }
        for (int d = 0; d < 3; d++) {
          rc[d] = rcm[d];
          vci[d] = vcm[d];
        }
        steps = 0;
        status = ST_INIT;
      }
    }
    // the slots that make a step (or initialise a try) in this iteration
    const bool go = status == ST_INIT ||
                    (status == ST_STEP && (major || k > 0 || steps >= WL.urgent));
    if (status == ST_STEP && go) {  // first half step
      double rn[3];
      if (NS == 1) {
        for (int d = 0; d < 3; d++) {
          v[0][d] = v[0][d] + hdm * fp[0][d] * KCONV;
          r[0][d] = r[0][d] + dt * v[0][d];
          rn[d] = r[0][d];
        }
      } else {
        rattle1(r[0], r[1], v[0], v[1], fp[0], fp[1], rat, dt, T.d2);
        for (int d = 0; d < 3; d++) rn[d] = 0.5 * (r[0][d] + r[1][d]);
      }
      dth = dth + anglevec(rn, rc);
      for (int d = 0; d < 3; d++) rc[d] = rn[d];
    }

    // ================================================================ B
    {
      double p[NS][3];
      if constexpr (NS == 1) {
        for (int d = 0; d < 3; d++) p[0][d] = r[0][d];
      } else {
        for (int d = 0; d < 3; d++) {
          p[PA][d] = r[0][d];
          p[PB][d] = r[1][d];
          p[CS][d] = rc[d];
        }
      }
      int md = EV_NONE;
      bool nb = false;
      double off = 0.0;
      if (go) {
        if (!S::CUT) {
          md = EV_ALL;
        } else {
          bool near = false;
#pragma unroll
          for (int s = 0; s < NS; s++) {
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
              const double e = fmax(0.0, fmax(T.blo[d] - p[s][d], p[s][d] - T.bhi[d]));
              d2 += e * e;
            }
            near = near || d2 < T.rfar2;
          }
          if (near) {
            bool inbox = true;
#pragma unroll
            for (int s = 0; s < NS; s++) {
              const double i = floor((p[s][0] - T.cx) * T.inv_cell);
              const double jj = floor((p[s][1] - T.cy) * T.inv_cell);
              const double kk = floor((p[s][2] - T.cz) * T.inv_cell);
              inbox = inbox && i >= 0.0 && i < T.nx && jj >= 0.0 && jj < T.ny && kk >= 0.0 && kk < T.nz;
            }
            if (inbox) {
              md = EV_LIST;
              double off2 = 0.0;
#pragma unroll
              for (int s = 0; s < NS; s++) {
                const double ex = p[s][0] - p[CS][0], ey = p[s][1] - p[CS][1], ez = p[s][2] - p[CS][2];
                off2 = fmax(off2, ex * ex + ey * ey + ez * ez);
              }
              off = sqrt(off2);
              const double e0 = p[CS][0] - vl.c0[0], e1 = p[CS][1] - vl.c0[1], e2 = p[CS][2] - vl.c0[2];
              const double moved = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
              nb = vl.valid < 0 || moved + off + T.rlong + 1e-9 > vl.R;
            } else {
              md = EV_STENCIL;
            }
          }
        }
        if (md == EV_NONE) {  // exact zero: no atom within any cutoff of any site
          for (int q = 0; q < NP; q++)
            for (int d = 0; d < 3; d++) fp[q][d] = 0.0;
          Ut = 0.0;
        }
      }
      unsigned long long m = __ballot(md != EV_NONE);
      while (m) {
        const int t = __builtin_ctzll(m);
        m &= m - 1;
        const bool me = lane == t;  // the slot's owner receives the results
        double pt[NS][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if constexpr (NS == 1) {
            pt[0][d] = readlane_d(r[0][d], t);
          } else {
            pt[PA][d] = readlane_d(r[0][d], t);
            pt[PB][d] = readlane_d(r[NP - 1][d], t);
            pt[CS][d] = readlane_d(rc[d], t);
          }
        }
        const int mdt = readlane_i(md, t);
        const bool bt = readlane_i((int)nb, t) != 0;
        const double offt = readlane_d(off, t);
        VList lt;
        lt.c0[0] = readlane_d(vl.c0[0], t);
        lt.c0[1] = readlane_d(vl.c0[1], t);
        lt.c0[2] = readlane_d(vl.c0[2], t);
        lt.R = readlane_d(vl.R, t);
        lt.n = readlane_i(vl.n, t);
        lt.nin = readlane_i(vl.nin, t);
        lt.valid = readlane_i(vl.valid, t);
        double f[NS][3], Us[NS];
        WorkU wu{0, 0, 0, 0};
        wave_force<S, LDS, COUNT>(T, sh, red, lane, pt, qs, mdt, offt, bt,
                                  lists + (uint64_t)t * WL.cap, WL.cap, lt, f, Us, wu);
        if (bt) {
          vl.c0[0] = (me ? lt.c0[0] : vl.c0[0]);
          vl.c0[1] = (me ? lt.c0[1] : vl.c0[1]);
          vl.c0[2] = (me ? lt.c0[2] : vl.c0[2]);
          vl.R = (me ? lt.R : vl.R);
          vl.n = (int)(me ? ((uint32_t)lt.n) : (uint32_t)vl.n);
          vl.nin = (int)(me ? ((uint32_t)lt.nin) : (uint32_t)vl.nin);
          vl.valid = (int)(me ? ((uint32_t)lt.valid) : (uint32_t)vl.valid);
        }
        double Usum = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) Usum += Us[s];
        if constexpr (NS == 1) {
          for (int d = 0; d < 3; d++) fp[0][d] = (me ? f[0][d] : fp[0][d]);
        } else {
          double fi[3], fj[3];
          pair_forces<S, GK>(f, fi, fj);
          for (int d = 0; d < 3; d++) {
            fp[0][d] = (me ? fi[d] : fp[0][d]);
            fp[1][d] = (me ? fj[d] : fp[1][d]);
          }
        }
        Ut = (me ? Usum : Ut);
        if (COUNT) {
          wv = (me ? ((uint32_t)readlane_i((int)wv, t) + wu.visits) : wv);
          wl = (me ? ((uint32_t)readlane_i((int)wl, t) + wu.lj) : wl);
          wc = (me ? ((uint32_t)readlane_i((int)wc, t) + wu.coul) : wc);
          wd = (me ? ((uint32_t)readlane_i((int)wd, t) + wu.dip) : wd);
        }
      }
    }

    // ================================================================ C
    if (status == ST_INIT) {
      double Ek = 0.0;
      if (NS == 1) Ek = kinetic(T.mu, v[0]);
      else if (GK == 2) Ek = kinetic(S::m(0), v[0]) + kinetic(S::m(1), v[1]);  // the dummy is massless
      else Ek = kinetic(S::m(0), v[0]) + kinetic(S::m(1), vci) + kinetic(S::m(2), v[1]);
      Ei = Ut + Ek;
      dt = T.dt / (k + 1);  // dt of try k, src/System.cpp:455
      hdm = 0.5 * dt / T.mu;
      rat = rattle_coeffs(MA, MB, dt);
      if (GK == 3) vci[0] = vci[1];  // src/System.cpp:1012 reads vcm[1] for x
      dth = 0.0;
      steps = 0;
      status = ST_STEP;
    } else if (status == ST_STEP && go) {
      double vcm[3], Ek = 0.0;
      if (NS == 1) {
        for (int d = 0; d < 3; d++) {
          v[0][d] = v[0][d] + hdm * fp[0][d] * KCONV;
          vcm[d] = v[0][d];
        }
        Ek = kinetic(T.mu, vcm);
      } else {
        rattle2(r[0], r[1], v[0], v[1], fp[0], fp[1], rat, MA, MB);
        for (int d = 0; d < 3; d++) vcm[d] = 0.5 * (v[0][d] + v[1][d]);
        // exit-test kinetic energy: N2 over its two atoms, CO2 over all three
        // sites with the carbon moving at vcm (src/System.cpp:880-886, :1323-1329)
        Ek = kinetic(MA, v[0]) + kinetic(MB, v[1]);
        if (GK == 3) Ek = kinetic(S::m(0), v[0]) + kinetic(S::m(1), vcm) + kinetic(S::m(2), v[1]);
      }
      const int steps1 = steps + 1;
      int code = TRY_PENDING;
      double chi = 0.0;
      if (steps1 >= MAX_STEPS || !isfinite(rc[0] + rc[1] + rc[2] + vcm[0] + vcm[1] + vcm[2])) {
        // the reference would loop forever here (zero speed or non-finite
        // state); the trajectory is reported lost with tries = -1
        code = TRY_CAPPED;
      } else if (dth > 1.5 * PI) {
        code = TRY_ANGLE;
      } else if (fabs(rc[0]) >= T.hx || fabs(rc[1]) >= T.hy || fabs(rc[2]) >= T.hz ||
                 (rc[0] / T.ea) * (rc[0] / T.ea) + (rc[1] / T.eb) * (rc[1] / T.eb) +
                         (rc[2] / T.ec) * (rc[2] / T.ec) >
                     1.0) {
        const double dE = NS == 1 ? fabs(Ut + Ek - Ei) / Ei * 100.0 : fabs((Ut + Ek - Ei) / Ei) * 100.0;
        const double tol = GK == 1 ? 0.5 : (GK == 2 ? 1.0 : 20.0);
        if (dE > tol) {
          code = TRY_ENERGY;
        } else {
          chi = anglevec(vci, vcm);
          code = TRY_SUCCESS;
        }
      } else if ((steps1 & 255) == 0) {
        // a later try is moot once the trajectory is resolved or an earlier
        // try has succeeded; near the end of the run, a trajectory that has
        // restarted or runs long spawns its remaining tries
        int moot = 0;
        if (k > 0) {
          moot = ld_acq_i(&Q.resolved[j]);
          for (int kk = 0; kk < k && !moot; kk++)
            moot = ld_acq_i(&Q.tab[TRIES_MAX * j + kk].code) == TRY_SUCCESS;
        }
        if (!moot && (k > 0 || steps1 >= SPAWN_STEPS) && ld_rlx(Q.spawned + j) == 0 &&
            ld_rlx(Q.active) < Q.spec_below)
          spawn_rest(Q, j);
        if (moot) code = TRY_ABANDON;
      }
      if (code == TRY_PENDING) {
        steps = steps1;
      } else {
        // the try has ended
        if (COUNT) {
          atomicAdd(Q.work + 0, (unsigned long long)wv);
          atomicAdd(Q.work + 1, (unsigned long long)wl);
          atomicAdd(Q.work + 2, (unsigned long long)wc);
          atomicAdd(Q.work + 3, (unsigned long long)wd);
        }
        atomicAdd(Q.work + 4, (unsigned long long)steps1);
        if (k == 0 && (code == TRY_SUCCESS || code == TRY_CAPPED) && ld_rlx(Q.spawned + j) == 0) {
          // the common case: the first try decides (no retry bookkeeping; only
          // this lane can have spawned later tries of j while try 0 ran)
          mccs_traj_record rr;
          rr.hit = 1;
          rr.success = code == TRY_SUCCESS;
          rr.tries = code == TRY_SUCCESS ? 0 : -1;
          rr.steps = steps1;
          rr.chi = chi;
          rr.domega = code == TRY_SUCCESS ? PI * (1.0 - cos(chi)) * (T.bmax * T.bmax) : 0.0;
          rr.visits = wv;
          rr.lj = wl;
          rr.coul = wc;
          rr.dip = wd;
          rec[j] = rr;
          __threadfence();
          atomicAdd((unsigned long long *)Q.active, ~0ull);
        } else if (code != TRY_ABANDON) {
          TryOut o;
          o.code = TRY_PENDING;
          o.steps = steps1;
          o.chi = chi;
          o.visits = wv;
          o.lj = wl;
          o.coul = wc;
          o.dip = wd;
          Q.tab[TRIES_MAX * j + k] = o;
          __threadfence();
          __hip_atomic_store(&Q.tab[TRIES_MAX * j + k].code, code, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
          // a failed try needs the next one unless the fold is already decided;
          // claim it unless another lane has, and near the end of the run
          // spawn the ones after it
          if (!resolve(Q, rec, j, T.bmax) && code != TRY_SUCCESS && code != TRY_CAPPED &&
              k + 1 < TRIES_MAX) {
            if (atomicCAS(Q.tries + j, k + 1, k + 2) == k + 1) {
              k = k + 1;
              cont = 1;
            }
            if (ld_rlx(Q.active) < Q.spec_below) spawn_rest(Q, j);
          }
        }
        status = ST_NEED;
      }
    }

    if (!__ballot(status != ST_DONE)) break;
    if (!__ballot(status == ST_INIT || status == ST_STEP || (status == ST_NEED && cont))) {
      // every working slot of the wavefront waits for work: look again later
      __builtin_amdgcn_s_sleep(127);
      __builtin_amdgcn_s_sleep(127);
      __builtin_amdgcn_s_sleep(127);
    }
  }
}

using WaveFn = void (*)(const Tgt *, const double *, uint64_t, mccs_traj_record *, Queue, WList);

template <int GK, int FT>
WaveFn pick_wave(bool lds, bool count) {
  if (lds) return count ? wave_kernel<GK, FT, true, true> : wave_kernel<GK, FT, true, false>;
  return count ? wave_kernel<GK, FT, false, true> : wave_kernel<GK, FT, false, false>;
}

WaveFn select_wave_gk1(int ft, bool lds, bool count);
WaveFn select_wave_gk2(int ft, bool lds, bool count);
WaveFn select_wave_gk3(int ft, bool lds, bool count);

}  // namespace dev
}  // namespace mccs
