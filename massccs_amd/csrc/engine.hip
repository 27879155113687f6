// MI355X (gfx950) trajectory engine: per-point site kernel, tally kernels and
// the host-side launcher behind the C ABI.
#include "kernels.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <random>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>

namespace mccs {
namespace dev {

// Per-point site forces: one group of L lanes per point.
template <class S, int L, bool LDS, bool SPLIT = false>
__global__ __launch_bounds__(512) void eval_kernel(Tgt T, const double *__restrict__ xyz, uint64_t n,
                                                   double qsite, double *__restrict__ fout,
                                                   double *__restrict__ uout) {
  extern __shared__ __align__(32) char smem[];
  Shm sh;
  stage<L, LDS, any_central<S>()>(T, sh, smem);
  const int lane = threadIdx.x % L;
  const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / L);
  Work w{0, 0, 0, 0};
  for (uint64_t g = (uint64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; g < n; g += groups) {
    double p[1][3] = {{xyz[3 * g], xyz[3 * g + 1], xyz[3 * g + 2]}};
    double qs[1] = {qsite};
    double f[1][3], U[1];
    evaluate<S, L, LDS, false, SPLIT>(T, sh, lane, p, qs, f, U, w, nullptr);
    if (lane == 0) {
      fout[3 * g] = f[0][0];
      fout[3 * g + 1] = f[0][1];
      fout[3 * g + 2] = f[0][2];
      uout[g] = U[0];
    }
  }
}

// Tally: one wavefront per iteration sums its trajectories' dOmega in index
// order exactly as src/System.cpp:287-307 (a chain of additions in the
// order of the records, as the reference's sequential loop): the wavefront
// loads 64 records at a time and every lane adds the 64 values in lane
// order from scalar registers (readlane_d); the three counts come from
// ballots.

__global__ __launch_bounds__(64) void tally_iter_kernel(const mccs_traj_record *__restrict__ rec,
                                                        uint64_t n_traj, double *__restrict__ omega,
                                                        int64_t *__restrict__ counts) {
  const uint64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const mccs_traj_record *r = rec + i * n_traj;
  double dO = 0.0;
  int64_t nf = 0, ns = 0, nl = 0;
  for (uint64_t base = 0; base < n_traj; base += 64) {
    const uint64_t k = base + lane;
    const bool valid = k < n_traj;
    double d = 0.0;
    int hit = 0, suc = 0;
    if (valid) {
      d = r[k].domega;
      hit = r[k].hit;
      suc = r[k].success;
    }
    nf += __popcll(__ballot(valid && !hit));
    ns += __popcll(__ballot(valid && hit && suc));
    nl += __popcll(__ballot(valid && hit && !suc));
    const int m = (int)(n_traj - base < 64 ? n_traj - base : 64);
    for (int l = 0; l < m; l++) dO += readlane_d(d, l);
  }
  if (lane == 0) {
    omega[i] = 1.0 / ((double)(float)(ns + nf)) * dO;
    counts[3 * i] = nf;
    counts[3 * i + 1] = ns;
    counts[3 * i + 2] = nl;
  }
}

// Mean and standard error over iterations, src/System.cpp:295-317.
__global__ void tally_final_kernel(const double *__restrict__ omega, uint64_t n_iter,
                                   double *__restrict__ mean_err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double O = 0.0, O2 = 0.0;
  for (uint64_t i = 0; i < n_iter; i++) {
    O += omega[i];
    O2 += omega[i] * omega[i];
  }
  const double ave = O / n_iter;
  const double sig2 = O2 / n_iter - ave * ave;
  mean_err[0] = ave;
  mean_err[1] = sqrt(sig2 / (double)(float)n_iter);
}


// ---------------------------------------------------------------- dispatch
using EvalFn = void (*)(Tgt, const double *, uint64_t, double, double *, double *);

TrajFn select_traj(int gk, int ft, int L, bool lds, bool count, bool split) {
  switch (gk) {
    case 1: return select_traj_gk1(ft, L, lds, count, split);
    case 2: return select_traj_gk2(ft, L, lds, count, split);
    case 3: return split ? nullptr : select_traj_gk3(ft, L, lds, count);
  }
  return nullptr;
}
template <bool LJ, bool CEN, bool COUL, bool DIP, bool CUT>
EvalFn pick_eval(bool lds) {
  using S = OneSite<LJ, CEN, COUL, DIP, CUT>;
  return lds ? eval_kernel<S, 64, true> : eval_kernel<S, 64, false>;
}

template <bool LJ, bool CEN, bool COUL, bool DIP>
EvalFn pick_eval_c(bool cut, bool lds) {
  return cut ? pick_eval<LJ, CEN, COUL, DIP, true>(lds) : pick_eval<LJ, CEN, COUL, DIP, false>(lds);
}

// The site roles whose terms the range-split far field covers (N2 nitrogen
// and charge/dipole sites, the one-site dipole probe), with cutoffs, at L
// lanes per point (the trajectory kernel's group size: the far field's
// nodes are split over the lanes as there).
template <class S, int L>
EvalFn pick_eval_split(bool lds) {
  return lds ? eval_kernel<S, L, true, true> : eval_kernel<S, L, false, true>;
}
template <class S>
EvalFn pick_eval_split_l(int L, bool lds) {
  return L == 16 ? pick_eval_split<S, 16>(lds) : (L == 32 ? pick_eval_split<S, 32>(lds) : pick_eval_split<S, 64>(lds));
}
EvalFn select_eval_split(bool lj, bool coul, bool dip, bool lds, int L) {
  if (lj && coul && !dip) return pick_eval_split_l<OneSite<true, false, true, false, true>>(L, lds);
  if (!lj && coul && dip) return pick_eval_split_l<OneSite<false, false, true, true, true>>(L, lds);
  if (lj && !coul && dip) return pick_eval_split_l<OneSite<true, false, false, true, true>>(L, lds);
  return nullptr;
}

EvalFn select_eval(bool lj, bool cen, bool coul, bool dip, bool cut, bool lds) {
  const int key = (lj << 3) | (cen << 2) | (coul << 1) | (int)dip;
  switch (key) {
    case 0b1000: return pick_eval_c<true, false, false, false>(cut, lds);
    case 0b1001: return pick_eval_c<true, false, false, true>(cut, lds);
    case 0b1010: return pick_eval_c<true, false, true, false>(cut, lds);
    case 0b0010: return pick_eval_c<false, false, true, false>(cut, lds);
    case 0b0011: return pick_eval_c<false, false, true, true>(cut, lds);
    case 0b1100: return pick_eval_c<true, true, false, false>(cut, lds);
    case 0b1110: return pick_eval_c<true, true, true, false>(cut, lds);
    case 0b1111: return pick_eval_c<true, true, true, true>(cut, lds);
  }
  return nullptr;
}

}  // namespace dev

using namespace dev;

// ---------------------------------------------------------------- engine
constexpr int QLINE = 32;  // 64-bit words per counter line
constexpr int QCTL_WORDS = 6 * QLINE;
constexpr uint64_t SLOT_SLACK = 8192;  // spawn slots beyond 8 per trajectory

// Per trajectory: TRIES_MAX try outcomes, 8 spawn slots, claimed tries,
// spawned and resolved flags, its place in the start order.
constexpr size_t QUEUE_PER_TRAJ =
    TRIES_MAX * sizeof(TryOut) + 8 * sizeof(unsigned long long) + 3 * sizeof(int) + sizeof(unsigned);
// Workspace bytes of a queue over n trajectories (with the start order's
// bucket counters).
constexpr size_t queue_ws_bytes(uint64_t n) {
  return QUEUE_PER_TRAJ * n + sizeof(unsigned long long) * SLOT_SLACK + sizeof(unsigned) * ORDER_BUCKETS;
}

// Queue of n trajectories over a counter block ctl (QCTL_WORDS words) and a
// workspace ws of queue_ws_bytes(n) bytes, both device memory.
static Queue queue_over(unsigned long long *ctl, char *ws, uint64_t n, long long spec_below,
                        unsigned long long wait_ticks, bool ordered) {
  Queue Q;
  Q.ticket = ctl + 0 * QLINE;
  Q.shead = ctl + 1 * QLINE;
  Q.stail = ctl + 2 * QLINE;
  Q.active = reinterpret_cast<long long *>(ctl + 3 * QLINE);
  Q.work = ctl + 4 * QLINE;
  Q.stalled = reinterpret_cast<int *>(ctl + 5 * QLINE);
  Q.tab = reinterpret_cast<TryOut *>(ws);
  Q.slots = reinterpret_cast<unsigned long long *>(ws + TRIES_MAX * sizeof(TryOut) * n);
  Q.nslots = 8 * n + SLOT_SLACK;
  Q.tries = reinterpret_cast<int *>(Q.slots + Q.nslots);
  Q.spawned = Q.tries + n;
  Q.resolved = Q.spawned + n;
  unsigned *ord = reinterpret_cast<unsigned *>(Q.resolved + n);
  Q.order = ordered ? ord : nullptr;
  Q.bucket = ordered ? ord + n : nullptr;
  Q.wait_ticks = wait_ticks;
  Q.spec_below = spec_below;
  return Q;
}

// Throws when a group of the last run over this counter block gave up
// waiting for work (its records are then incomplete).  The run must have
// completed.
static void queue_check_stall(const unsigned long long *ctl) {
  unsigned long long c = 0;
  HIP_OK(hipMemcpy(&c, ctl + 5 * QLINE, sizeof c, hipMemcpyDeviceToHost));
  if ((int)(c & 0xffffffffu) != 0)
    throw Error(MCCS_ERR_DEVICE, "trajectory work queue stalled; records are incomplete");
}

static void queue_work(const unsigned long long *ctl, uint64_t out[5]) {
  HIP_OK(hipMemcpy(out, ctl + 4 * QLINE, sizeof(uint64_t) * 5, hipMemcpyDeviceToHost));
}

// A group waits for work at most 120 s (any single try takes far less).
static unsigned long long queue_wait_ticks(int dev) {
  int khz = 0;
  HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) khz = 100000;
  return 120ull * 1000ull * (unsigned long long)khz;
}

// Device work queue of one launch: counters, each polled one on its own
// 256-byte line (ticket, shead, stail, active, work[5], stalled), and a
// per-trajectory workspace grown on demand.
struct QueueWS {
  // owns device memory: never copied
  QueueWS() = default;
  QueueWS(const QueueWS &) = delete;
  QueueWS &operator=(const QueueWS &) = delete;
  unsigned long long *qctl = nullptr;
  char *qws = nullptr;
  uint64_t qcap = 0;
  unsigned long long wait_ticks = 0;
  int device = 0;

  void init(int dev) {
    device = dev;
    HIP_OK(hipMalloc(&qctl, sizeof(unsigned long long) * QCTL_WORDS));
    HIP_OK(hipMemset(qctl, 0, sizeof(unsigned long long) * QCTL_WORDS));
    wait_ticks = queue_wait_ticks(dev);
  }
  // Queue over n trajectories; spawning below spec_below unresolved ones;
  // primary trajectories taken in the start order when ordered.
  Queue make(uint64_t n, long long spec_below, bool ordered) {
    if (n > qcap) {
      if (qws) HIP_OK(hipFree(qws));
      qws = nullptr;
      qcap = 0;
      const uint64_t cap = std::max<uint64_t>(n, 1024);
      HIP_OK(hipMalloc(&qws, queue_ws_bytes(cap)));
      qcap = cap;
    }
    return queue_over(qctl, qws, n, spec_below, wait_ticks, ordered);
  }
  // Throws when a group of the last run gave up waiting for work (its
  // records are then incomplete).  The run must have completed.
  void check_stall() const { queue_check_stall(qctl); }
  void work(uint64_t out[5]) const { queue_work(qctl, out); }
  ~QueueWS() {
    if (qctl || qws) hipSetDevice(device);
    if (qctl) hipFree(qctl);
    if (qws) hipFree(qws);
  }
};

// The smallest double t with sqrt(t) >= rc, so that the reference's cutoff
// test sqrt(r2) < rc holds exactly when r2 < t: the square root is correctly
// rounded on the host and on the device, hence monotonic, and t lies within a
// few units in the last place of rc * rc.
static double sqrt_cut(double rc) {
  if (!(rc > 0.0)) return 0.0;                      // nothing is inside
  if (std::isinf(rc)) return rc;                    // everything finite is inside
  double t = rc * rc;
  for (int i = 0; i < 8 && std::sqrt(t) < rc; i++) t = std::nextafter(t, HUGE_VAL);
  for (int i = 0; i < 8; i++) {
    const double d = std::nextafter(t, 0.0);
    if (std::sqrt(d) >= rc) t = d;
    else break;
  }
  return t;
}

// Verlet-list size a probe is likely to need: the 99th percentile, over
// sampled points of the target's box that a probe's centre can reach (at least
// dmin from every atom) and that have atoms within R, of the number of atoms
// within R, counted through the linked cells.  The seed is fixed, so a target
// always gets the same answer.
static double list_need(const Prepared &P, double R, double dmin) {
  if (P.natoms == 0 || P.ncells == 0) return 0.0;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t a = 0; a < P.natoms; a++) {
    const double v[3] = {P.x[a], P.y[a], P.z[a]};
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], v[d]);
      hi[d] = std::max(hi[d], v[d]);
    }
  }
  const double R2 = R * R, d2min = dmin * dmin, ic = 1.0 / P.cell;
  const int n[3] = {P.nx, P.ny, P.nz};
  std::mt19937_64 rng(0x6c697374ull);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  std::vector<int> counts;
  for (int smp = 0; smp < 2048; smp++) {
    double p[3];
    int c0[3], c1[3];
    for (int d = 0; d < 3; d++) {
      p[d] = lo[d] - R + uni(rng) * (hi[d] - lo[d] + 2.0 * R);
      c0[d] = std::min(std::max((int)std::floor((p[d] - R - P.corner[d]) * ic), 0), n[d] - 1);
      c1[d] = std::min(std::max((int)std::floor((p[d] + R - P.corner[d]) * ic), 0), n[d] - 1);
    }
    int cnt = 0;
    double m2 = 1e300;
    for (int i = c0[0]; i <= c1[0]; i++)
      for (int j = c0[1]; j <= c1[1]; j++)
        for (int k = c0[2]; k <= c1[2]; k++) {
          const int64_t c = (int64_t)k + (int64_t)P.nz * j + (int64_t)P.nz * P.ny * i;
          for (int32_t a = P.cell_start[c]; a < P.cell_start[c + 1]; a++) {
            const double ex = P.x[a] - p[0], ey = P.y[a] - p[1], ez = P.z[a] - p[2];
            const double d2 = ex * ex + ey * ey + ez * ez;
            cnt += d2 < R2;
            m2 = std::min(m2, d2);
          }
        }
    if (cnt > 0 && m2 >= d2min) counts.push_back(cnt);
  }
  if (counts.empty()) return 0.0;
  std::sort(counts.begin(), counts.end());
  return counts[(size_t)(0.99 * (double)(counts.size() - 1))];
}

// LDS a block may use.  A CU has 160 KB: a target staged in LDS takes a whole
// CU per block; lists of a target read through L2 take half when they fit in
// it, so that two blocks (2 waves per SIMD) share the CU, else the whole.
constexpr size_t LDS_BLOCK_FULL = 150 * 1024, LDS_BLOCK_HALF = 78 * 1024;
// a list radius beyond the cutoff and skin: the largest site offset from the
// probe's centre (CO2 oxygens, 1.16 A) with margin
constexpr double LIST_OFF_MAX = 1.2;

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int gk = 1, ft = 2, lanes = 64, block = 256, cus = 0;
  bool lds = false, count = false;
  size_t shmem = 0;
  size_t shmem_split = 0;  // LDS per block of the range-split kernel
  Tgt T{};
  double site_q[5] = {0, -0.4825, 0.965, -0.3256, 0.6512};
  std::vector<void *> allocs;
  QueueWS q;
  Tgt *d_tg = nullptr;  // T in device memory (the kernel's target array)
  // far-field surrogate
  Mlp *mlp = nullptr;
  float *d_psi = nullptr;   // B-spline coefficients of the far field
  Tgt Ts{};                 // T with the range-split far field
  bool sur_built = false;
  bool split = false;       // runs use Ts
  int spec_below = -1;
  bool ordered = true;      // primary trajectories in the start order (likely-long first)
  hipStream_t last_stream = nullptr;
  const unsigned long long *last_ctl = nullptr;  // queue counters of the last run
  double list_need = 0.0;  // estimated largest Verlet list at skin 1 A (L2-fed targets)
  int nsites = 1;
  void *dalloc(size_t bytes) {
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    allocs.push_back(p);
    return p;
  }
  template <class V>
  const typename V::value_type *upload(const V &v) {
    void *p = dalloc(sizeof(typename V::value_type) * v.size());
    if (!v.empty())
      HIP_OK(hipMemcpy(p, v.data(), sizeof(typename V::value_type) * v.size(), hipMemcpyHostToDevice));
    return reinterpret_cast<const typename V::value_type *>(p);
  }
  ~Engine() {
    if (device >= 0) hipSetDevice(device);
    if (mlp) mlp_destroy(mlp);
    if (d_psi) hipFree(d_psi);
    for (void *p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
};

static void noalloc_setup(int device);

Engine *engine_create(const Prepared &P, const mccs_engine_opts &o) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw Error(MCCS_ERR_DEVICE, "no HIP device available: the trajectory engine needs an MI355X (gfx950) GPU");
  if (o.device < 0 || o.device >= ndev) throw Error(MCCS_ERR_INVALID, "device ordinal out of range");
  std::unique_ptr<Engine> E(new Engine);
  E->device = o.device;
  HIP_OK(hipSetDevice(o.device));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, o.device));
  E->cus = prop.multiProcessorCount;
  HIP_OK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
  const int gas = P.cfg.gas;
  E->gk = (gas == MCCS_GAS_N2) ? 2 : (gas == MCCS_GAS_CO2 ? 3 : 1);
  E->ft = P.force_type;
  E->nsites = P.gas.nsites;
  if (E->gk == 1 && E->ft > 4) E->ft = 2;
  E->count = o.count_work != 0;
  E->spec_below = o.spec_below;
  if (o.order < 0 || o.order > 1) throw Error(MCCS_ERR_INVALID, "order must be 0 or 1");
  E->ordered = o.order == 0;

  // per-atom LJ pairs in the reference's order (src/Force.cpp:84-85):
  // lj1 = 4 eps sig^6, lj2 = lj1 sig^6; the second pair for the CO2 carbon site
  std::vector<double2> ljp(P.natoms), ljc(P.natoms);
  for (int64_t i = 0; i < P.natoms; i++) {
    const double l1 = 4.0 * P.eps[i] * std::pow(P.sig[i], 6.0);
    ljp[i] = make_double2(l1, l1 * std::pow(P.sig[i], 6.0));
    const double c1 = 4.0 * P.eps_c[i] * std::pow(P.sig_c[i], 6.0);
    ljc[i] = make_double2(c1, c1 * std::pow(P.sig_c[i], 6.0));
  }
  std::vector<double4> atoms(P.natoms);
  for (int64_t i = 0; i < P.natoms; i++) atoms[i] = make_double4(P.x[i], P.y[i], P.z[i], P.q[i]);
  std::vector<int> s1, s2;
  stencil_offsets(P.cfg.lj_cutoff, P.cfg.coul_cutoff, P.cell, P.cfg.long_range_flag == 1, s1, s2);
  std::vector<int> st(s1);
  st.insert(st.end(), s2.begin(), s2.end());

  Tgt &T = E->T;
  T.atoms = E->upload(atoms);
  T.ljp = E->upload(ljp);
  T.ljc = E->upload(ljc);
  T.cstart = E->upload(P.cell_start);
  T.vs = E->upload(P.vtab_s);
  T.vv = E->upload(P.vtab_v);
  T.stencil = E->upload(st);
  T.natoms = (int)P.natoms;
  T.nx = P.nx; T.ny = P.ny; T.nz = P.nz; T.ncells = P.ncells;
  T.vn = (int)P.vtab_s.size();
  T.n1 = (int)s1.size() / 3;
  T.n2 = (int)s2.size() / 3;
  T.cell = P.cell;
  T.cx = P.corner[0]; T.cy = P.corner[1]; T.cz = P.corner[2];
  T.hx = P.half_box[0]; T.hy = P.half_box[1]; T.hz = P.half_box[2];
  T.ea = P.a; T.eb = P.b; T.ec = P.c; T.bmax = P.bmax;
  T.iea = 1.0 / P.a; T.ieb = 1.0 / P.b; T.iec = 1.0 / P.c;
  const double rlj = P.cfg.lj_cutoff, rc = P.cfg.coul_cutoff;
  T.rlj = rlj; T.rc = rc;
  T.rlj2cut = sqrt_cut(rlj);
  T.rc2cut = sqrt_cut(rc);
  T.irc3 = 1.0 / (rc * rc * rc);
  T.rc6inv = 1.0 / std::pow(rlj, 6);
  T.cc1 = -1.5 / rc;
  T.cc2 = 0.5 / (rc * rc * rc);
  T.alpha = P.alpha;
  T.mu = P.mu;
  T.dt = P.cfg.dt;
  T.temp = P.cfg.temperature;
  T.ginertia = P.gas_inertia;
  T.d2 = P.d_bond * P.d_bond;
  T.inv_cell = 1.0 / P.cell;
  for (int d = 0; d < 3; d++) {
    T.blo[d] = 1e300;
    T.bhi[d] = -1e300;
  }
  for (int64_t i = 0; i < P.natoms; i++) {
    const double v[3] = {P.x[i], P.y[i], P.z[i]};
    for (int d = 0; d < 3; d++) {
      T.blo[d] = std::min(T.blo[d], v[d]);
      T.bhi[d] = std::max(T.bhi[d], v[d]);
    }
  }
  {
    const int ft = P.force_type, gk = (gas == MCCS_GAS_N2) ? 2 : (gas == MCCS_GAS_CO2 ? 3 : 1);
    const bool longr = gk == 1 ? (ft == 3 || ft == 4) : ft >= 3;
    const double rf = std::max(longr ? rc : rlj, rlj);
    T.rfar2 = rf * rf * (1.0 + 1e-12);
    T.rlong = longr ? rc : rlj;
  }
  T.skin = 1.0;
  T.co2fix = P.cfg.co2_vcm_fix != 0;
  E->q.init(E->device);
  noalloc_setup(E->device);

  // launch shape: lanes per trajectory by target size unless given
  // launch shape: 64 lanes per trajectory unless given (measured fastest from
  // 60 to 10^6 atoms: the wavefront's 64 lanes share each force evaluation)
  int L = o.lanes ? o.lanes : 64;
  if (L != 16 && L != 32 && L != 64) throw Error(MCCS_ERR_INVALID, "lanes must be 0, 16, 32 or 64");
  E->lanes = L;
  // Atoms (and their LJ pairs) are staged in LDS when the block's share of a
  // CU's LDS (the CU holds 4 x traj_waves wavefronts) fits them together with
  // full-size Verlet lists, which then hold 16-bit indices.  The launch shape
  // follows: traj_block(true) threads when staged, else traj_block(false).
  if (o.block && (o.block % 64 != 0 || o.block < 128))
    throw Error(MCCS_ERR_INVALID, "block must be 0 or a multiple of 64, at least 128");
  const bool central = E->gk == 3;
  const int lds_block = o.block ? o.block : traj_block(true);
  const int blocks_per_cu = std::max(1, 4 * traj_waves(true) * 64 / lds_block);
  const size_t per_block = std::min<size_t>(150 * 1024, 160 * 1024 / blocks_per_cu);
  const bool fits = T.natoms <= 65535 && lds_block <= traj_block(true) &&
                    lds_bytes(T.natoms, T.natoms, lds_block, L, true, central) <= per_block;
  E->lds = o.use_lds < 0 ? fits : (o.use_lds != 0);
  if (E->lds && T.natoms > 65535) throw Error(MCCS_ERR_INVALID, "target too large to stage in LDS");
  E->block = o.block ? o.block : traj_block(E->lds);
  if (E->block > traj_block(E->lds))
    throw Error(MCCS_ERR_INVALID, "block must be at most " + std::to_string(traj_block(E->lds)) +
                                      (E->lds ? " with atoms in LDS" : " with atoms in global memory"));
  // list capacity: every atom for small targets, else what the block's LDS
  // allows: half a CU when the estimated lists fit in 90 % of it, else all
  const int groups = E->block / L;
  const size_t lidx = E->lds ? 2 : 4;
  const size_t base = lds_bytes(T.natoms, 0, E->block, L, E->lds, central);
  size_t budget = LDS_BLOCK_FULL;
  if (!E->lds) {
    E->list_need = list_need(P, T.rlong + T.skin + LIST_OFF_MAX, 2.5);
    const size_t half = base < LDS_BLOCK_HALF ? (LDS_BLOCK_HALF - base) / (lidx * groups) : 0;
    if (E->list_need <= 0.9 * (double)half) budget = LDS_BLOCK_HALF;
  }
  const size_t room = base < budget ? budget - base : 0;
  T.list_cap = (int)std::min<size_t>((size_t)T.natoms, room / (lidx * groups));
  if (o.list_cap < 0) throw Error(MCCS_ERR_INVALID, "list_cap must be >= 0 (0 = auto)");
  if (o.list_cap > 0) T.list_cap = std::min(T.list_cap, (int)o.list_cap);
  E->shmem = lds_bytes(T.natoms, T.list_cap, E->block, L, E->lds, central);
  if (E->shmem > 160 * 1024) throw Error(MCCS_ERR_INVALID, "target too large to stage in LDS");
  E->d_tg = reinterpret_cast<Tgt *>(E->dalloc(sizeof(Tgt)));
  HIP_OK(hipMemcpy(E->d_tg, &T, sizeof(Tgt), hipMemcpyHostToDevice));
  return E.release();
}

void engine_destroy(Engine *e) { delete e; }

// Clears the queue; the launch's own status is checked (not the thread's
// sticky last error, which unchecked teardown calls may have set).
static void launch_queue_init(Queue Q, uint64_t n, hipStream_t s) {
  void *args[] = {&Q, &n};
  HIP_OK(hipLaunchKernel((const void *)queue_init, dim3(256), dim3(256), args, 0, s));
}

// The start order of a queue that has one (after launch_queue_init, which
// clears its bucket counters): count, scan, scatter.
static void launch_order(Queue Q, const Tgt *tg, const unsigned long long *toff, int ntg,
                         const double *u, uint64_t n, hipStream_t s) {
  if (!Q.order || n == 0) return;
  if (n > 0xffffffffull) throw Error(MCCS_ERR_INVALID, "too many trajectories for one launch");
  const dim3 grid((unsigned)((n + 255) / 256));
  void *args[] = {&tg, &toff, &ntg, &u, &n, &Q};
  HIP_OK(hipLaunchKernel((const void *)order_count, grid, dim3(256), args, 0, s));
  void *sargs[] = {&Q};
  HIP_OK(hipLaunchKernel((const void *)order_scan, dim3(1), dim3(ORDER_BUCKETS), sargs, 0, s));
  HIP_OK(hipLaunchKernel((const void *)order_scatter, grid, dim3(256), args, 0, s));
}

static void set_shmem(const void *fn, size_t bytes) {
  if (bytes > 64 * 1024)
    HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Launch shape of a run of n trajectories: the kernel, its blocks and the
// spawning threshold (fewer trajectories than resident groups unresolved,
// i.e. when groups begin to fall idle, unless given).
struct TrajLaunch {
  TrajFn fn = nullptr;
  uint64_t blocks = 0;
  long long spec_below = 0;
  size_t shmem = 0;
};

static TrajLaunch traj_launch(Engine *E, uint64_t n) {
  TrajLaunch tl;
  tl.fn = select_traj(E->gk, E->ft, E->lanes, E->lds, E->count, E->split);
  if (!tl.fn && !E->split) {  // force types without tuned variants run 64 lanes with work counters
    E->lanes = 64;
    E->count = true;
    tl.fn = select_traj(E->gk, E->ft, 64, E->lds, true, false);
  }
  if (!tl.fn) throw Error(MCCS_ERR_UNSUPPORTED, "no trajectory kernel for this gas / force type");
  tl.shmem = E->split ? E->shmem_split : E->shmem;
  set_shmem((const void *)tl.fn, tl.shmem);
  int per_cu = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)tl.fn, E->block, tl.shmem));
  if (per_cu < 1) throw Error(MCCS_ERR_DEVICE, "trajectory kernel does not fit on a CU");
  const uint64_t groups_per_block = E->block / E->lanes;
  tl.blocks = std::min<uint64_t>((uint64_t)per_cu * E->cus,
                                 (n + groups_per_block - 1) / groups_per_block);
  tl.spec_below = E->spec_below >= 0 ? E->spec_below : (long long)(tl.blocks * groups_per_block);
  return tl;
}

static void launch_traj(Engine *E, const TrajLaunch &tl, Queue Q, const double *u_dev, uint64_t n,
                        mccs_traj_record *rec_dev, hipStream_t s) {
  launch_queue_init(Q, n, s);
  const Tgt *tg = E->d_tg;
  const unsigned long long *toff = nullptr;
  int ntg = 1;
  launch_order(Q, tg, toff, ntg, u_dev, n, s);
  void *args[] = {&tg, &toff, &ntg, &u_dev, &n, &rec_dev, &Q};
  HIP_OK(hipLaunchKernel((const void *)tl.fn, dim3((unsigned)tl.blocks), dim3(E->block), args,
                         tl.shmem, s));
  E->last_stream = s;
}

void engine_run_device(Engine *E, const double *u_dev, uint64_t n, mccs_traj_record *rec_dev,
                       void *stream) {
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  hipStream_t s = stream ? (hipStream_t)stream : E->stream;
  const TrajLaunch tl = traj_launch(E, n);
  Queue Q = E->q.make(n, tl.spec_below, E->ordered);
  launch_traj(E, tl, Q, u_dev, n, rec_dev, s);
  E->last_ctl = E->q.qctl;
}

// ---------------------------------------------------------------- no-allocation runs
// The work queue of mccs_run_device_noalloc lives in this statically sized
// device array (each GPU holds its own copy until the process ends): the
// counter block, then the per-trajectory workspace of at most
// MCCS_NOALLOC_MAX_TRAJECTORIES trajectories.  In host code the name is a
// stand-in, not a device address: hipGetSymbolAddress gives each device's
// address once (noalloc_setup, at engine creation with the device current),
// and only that address is handed to kernels.
constexpr size_t NOALLOC_CTL_BYTES = sizeof(unsigned long long) * QCTL_WORDS;
constexpr size_t NOALLOC_SCRATCH_BYTES =
    NOALLOC_CTL_BYTES + queue_ws_bytes(MCCS_NOALLOC_MAX_TRAJECTORIES);
static_assert(NOALLOC_CTL_BYTES % 256 == 0, "workspace must stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
struct ScratchDev {
  unsigned char *ptr = nullptr;  // the device's address of scratch
  unsigned long long wait_ticks = 0;
};
std::mutex g_scratch_mu;
std::map<int, ScratchDev> g_scratch;
}  // namespace

// Resolves the address of scratch on the current device (once per device).
static void noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch.count(device)) return;
  void *ptr = nullptr;
  HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw Error(MCCS_ERR_DEVICE, "no device address for the run workspace");
  ScratchDev d;
  d.ptr = static_cast<unsigned char *>(ptr);
  d.wait_ticks = queue_wait_ticks(device);
  g_scratch[device] = d;
}

void engine_run_device_noalloc(Engine *E, const double *u_dev, uint64_t n,
                               mccs_traj_record *rec_dev, void *stream) {
  if (n > MCCS_NOALLOC_MAX_TRAJECTORIES)
    throw Error(MCCS_ERR_INVALID, "mccs_run_device_noalloc runs at most " +
                                      std::to_string(MCCS_NOALLOC_MAX_TRAJECTORIES) +
                                      " trajectories per call (got " + std::to_string(n) + ")");
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  ScratchDev d;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(E->device);
    if (it == g_scratch.end()) throw Error(MCCS_ERR_DEVICE, "run workspace not set up on this device");
    d = it->second;
  }
  hipStream_t s = stream ? (hipStream_t)stream : E->stream;
  const TrajLaunch tl = traj_launch(E, n);
  unsigned long long *ctl = reinterpret_cast<unsigned long long *>(d.ptr);
  // queue_init (a kernel) clears the counters and every per-trajectory entry
  // the run reads before the trajectory kernel starts
  Queue Q = queue_over(ctl, reinterpret_cast<char *>(d.ptr + NOALLOC_CTL_BYTES), n, tl.spec_below,
                       d.wait_ticks, E->ordered);
  launch_traj(E, tl, Q, u_dev, n, rec_dev, s);
  E->last_ctl = ctl;
}

void engine_last_work(Engine *E, uint64_t out[5]) {
  HIP_OK(hipSetDevice(E->device));
  HIP_OK(hipStreamSynchronize(E->last_stream ? E->last_stream : E->stream));
  const unsigned long long *ctl = E->last_ctl ? E->last_ctl : E->q.qctl;
  queue_check_stall(ctl);
  queue_work(ctl, out);
}

void engine_run_host(Engine *E, const double *u, uint64_t n, mccs_traj_record *rec) {
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  double *du = nullptr;
  mccs_traj_record *dr = nullptr;
  HIP_OK(hipMalloc(&du, sizeof(double) * 9 * n));
  if (hipMalloc(&dr, sizeof(mccs_traj_record) * n) != hipSuccess) {
    hipFree(du);
    throw Error(MCCS_ERR_NOMEM, "device allocation of trajectory records failed");
  }
  try {
    HIP_OK(hipMemcpyAsync(du, u, sizeof(double) * 9 * n, hipMemcpyHostToDevice, E->stream));
    engine_run_device(E, du, n, dr, E->stream);
    HIP_OK(hipMemcpyAsync(rec, dr, sizeof(mccs_traj_record) * n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
    E->q.check_stall();
  } catch (...) {
    hipFree(du);
    hipFree(dr);
    throw;
  }
  hipFree(du);
  hipFree(dr);
}

// Kernel and probe charge of a site role; fn null for a site without terms.
static EvalFn site_kernel(const Engine *E, int site, double &q) {
  const int ft = E->ft, gk = E->gk;
  bool lj = false, cen = false, coul = false, dip = false, cut = false;
  q = 0.0;
  if (site == MCCS_SITE_ATOM) {
    if (gk != 1) throw Error(MCCS_ERR_INVALID, "site ATOM needs a one-site probe (He, Ar, co2)");
    lj = true; dip = (ft == 3 || ft == 4); cut = (ft == 2 || ft == 4);
  } else if (site == MCCS_SITE_N || site == MCCS_SITE_DUMMY) {
    if (gk != 2) throw Error(MCCS_ERR_INVALID, "sites N/DUMMY need the N2 probe");
    cut = (ft == 2 || ft == 4 || ft == 6);
    if (site == MCCS_SITE_N) { lj = true; coul = ft >= 3; q = -0.4825; }
    else { coul = ft >= 3; dip = (ft == 5 || ft == 6); q = 0.965; }
  } else if (site == MCCS_SITE_O || site == MCCS_SITE_C) {
    if (gk != 3) throw Error(MCCS_ERR_INVALID, "sites O/C need the CO2 probe");
    cut = (ft == 2 || ft == 4 || ft == 6);
    lj = true; coul = ft >= 3;
    if (site == MCCS_SITE_O) q = -0.3256;
    else { cen = true; dip = (ft == 5 || ft == 6); q = 0.6512; }
  } else {
    throw Error(MCCS_ERR_INVALID, "unknown site role");
  }
  if (!(lj || coul || dip)) return nullptr;  // N2 dummy of force types 1/2: no interaction
  EvalFn fn = E->split ? select_eval_split(lj, coul, dip, E->lds, E->lanes)
                       : select_eval(lj, cen, coul, dip, cut, E->lds);
  if (!fn) throw Error(MCCS_ERR_UNSUPPORTED, "no site kernel for this role");
  return fn;
}

// Site forces at device points (exact evaluator), enqueued on the engine's stream.
static void eval_sites_dev(Engine *E, int site, const double *dx, uint64_t n, double *df, double *du) {
  double q = 0.0;
  EvalFn fn = site_kernel(E, site, q);
  if (!fn) {
    HIP_OK(hipMemsetAsync(df, 0, sizeof(double) * 3 * n, E->stream));
    HIP_OK(hipMemsetAsync(du, 0, sizeof(double) * n, E->stream));
    return;
  }
  const int block = 256;
  const int L = E->split ? E->lanes : 64;  // lanes per point
  const size_t shmem = lds_bytes(E->T.natoms, E->T.list_cap, block, L, E->lds, E->gk == 3);
  set_shmem((const void *)fn, shmem);
  const uint64_t gpb = block / L;
  const uint64_t blocks = std::min<uint64_t>((n + gpb - 1) / gpb, (uint64_t)E->cus * 8);
  Tgt T = E->split ? E->Ts : E->T;  // with the far field when it is in use
  void *args[] = {&T, &dx, &n, &q, &df, &du};
  HIP_OK(hipLaunchKernel((const void *)fn, dim3((unsigned)blocks), dim3(block), args, shmem, E->stream));
}

void engine_eval_sites(Engine *E, int site, const double *xyz, uint64_t n, double *f, double *U) {
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  double *dx = nullptr, *df = nullptr, *du = nullptr;
  HIP_OK(hipMalloc(&dx, sizeof(double) * 3 * n));
  HIP_OK(hipMalloc(&df, sizeof(double) * 3 * n));
  HIP_OK(hipMalloc(&du, sizeof(double) * n));
  try {
    HIP_OK(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, E->stream));
    eval_sites_dev(E, site, dx, n, df, du);
    HIP_OK(hipMemcpyAsync(f, df, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipMemcpyAsync(U, du, sizeof(double) * n, hipMemcpyDeviceToHost, E->stream));
    HIP_OK(hipStreamSynchronize(E->stream));
  } catch (...) {
    hipFree(dx);
    hipFree(df);
    hipFree(du);
    throw;
  }
  hipFree(dx);
  hipFree(df);
  hipFree(du);
}

void engine_tally_device(Engine *E, const mccs_traj_record *rec_dev, uint64_t n_iter, uint64_t n_traj,
                         double *omega_dev, int64_t *counts_dev, double *mean_err_dev, void *stream) {
  if (n_iter == 0) return;
  HIP_OK(hipSetDevice(E->device));
  hipStream_t s = stream ? (hipStream_t)stream : E->stream;
  hipLaunchKernelGGL(tally_iter_kernel, dim3((unsigned)n_iter), dim3(64), 0, s, rec_dev, n_traj,
                     omega_dev, counts_dev);
  hipLaunchKernelGGL(tally_final_kernel, dim3(1), dim3(64), 0, s, omega_dev, n_iter, mean_err_dev);
  HIP_OK(hipGetLastError());
}



// ---------------------------------------------------------------- surrogate
// Range-split far field (see Tgt): psi(x) = sum_j q_j (1 - S(r_j)) u_c(r_j)
// over the atoms within the Coulomb cutoff, its exact evaluation through the
// linked cells, the network trained on it, and its tabulation as cubic
// B-spline coefficients.
namespace dev {

struct FarGrid {
  double ox, oy, oz, h;
  int nx, ny, nz;
};

// Partition of unity of the pair distance: 1 below sw1, 0 beyond sw2.
__device__ __forceinline__ void switch_fn(const Tgt &T, double r, double &S, double &dS) {
  const double t = fmin(fmax((r - T.sw1) * T.isw, 0.0), 1.0), t2 = t * t, u = 1.0 - t;
  S = 1.0 - t2 * t * fma(t, fma(6.0, t, -15.0), 10.0);
  dS = -30.0 * t2 * u * u * T.isw;
}

// psi and grad psi at n points (or at the nodes of a grid when g is given):
// one wavefront per point, the atoms of the cell columns within the cutoff
// strided over its lanes.  grad may be null.
template <bool GRID>
__global__ __launch_bounds__(256) void far_eval_kernel(Tgt T, const double *__restrict__ xyz,
                                                      FarGrid g, uint64_t n,
                                                      double *__restrict__ val,
                                                      double *__restrict__ grad) {
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t pi = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; pi < n;
       pi += waves) {
    double x, y, z;
    if (GRID) {
      const uint64_t iz = pi % g.nz, iy = (pi / g.nz) % g.ny, ix = pi / ((uint64_t)g.ny * g.nz);
      x = g.ox + g.h * ix;
      y = g.oy + g.h * iy;
      z = g.oz + g.h * iz;
    } else {
      x = xyz[3 * pi];
      y = xyz[3 * pi + 1];
      z = xyz[3 * pi + 2];
    }
    const double R = T.rc;
    const int ix0 = clampi((int)floor((x - R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int ix1 = clampi((int)floor((x + R - T.cx) * T.inv_cell), 0, T.nx - 1);
    const int iy0 = clampi((int)floor((y - R - T.cy) * T.inv_cell), 0, T.ny - 1);
    const int iy1 = clampi((int)floor((y + R - T.cy) * T.inv_cell), 0, T.ny - 1);
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    // nothing within the cutoff of the target's box: psi = 0
    double d2 = 0.0;
    const double pp[3] = {x, y, z};
    for (int d = 0; d < 3; d++) {
      const double e = fmax(0.0, fmax(T.blo[d] - pp[d], pp[d] - T.bhi[d]));
      d2 += e * e;
    }
    if (d2 < T.rfar2) {
      for (int i = ix0; i <= ix1; i++)
        for (int j = iy0; j <= iy1; j++) {
          const double lox = T.cx + i * T.cell, loy = T.cy + j * T.cell;
          const double dxm = fmax(0.0, fmax(lox - x, x - (lox + T.cell)));
          const double dym = fmax(0.0, fmax(loy - y, y - (loy + T.cell)));
          const double c2 = dxm * dxm + dym * dym;
          if (c2 >= R * R) continue;
          const double hh = sqrt(R * R - c2);
          const int k0 = clampi((int)floor((z - hh - T.cz) * T.inv_cell), 0, T.nz - 1);
          const int k1 = clampi((int)floor((z + hh - T.cz) * T.inv_cell), 0, T.nz - 1);
          const int b = T.nz * (j + T.ny * i);
          const int a0 = T.cstart[b + k0], a1 = T.cstart[b + k1 + 1];
          for (int a = a0 + lane; a < a1; a += 64) {
            const double4 at = T.atoms[a];
            const double dx = x - at.x, dy = y - at.y, dz = z - at.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < T.rc2cut) || r2 == 0.0) continue;
            const double r = sqrt(r2), rinv = 1.0 / r;
            double S, dS;
            switch_fn(T, r, S, dS);
            const double uc = rinv + T.cc1 + T.cc2 * r2;
            const double u1 = r * T.irc3 - rinv * rinv;
            // (1 - S) u and its radial derivative
            const double ph = (1.0 - S) * uc, ph1 = (1.0 - S) * u1 - dS * uc;
            v += at.w * ph;
            const double gr = at.w * ph1 * rinv;
            gx += gr * dx;
            gy += gr * dy;
            gz += gr * dz;
          }
        }
    }
    v = gsum<64>(v);
    gx = gsum<64>(gx);
    gy = gsum<64>(gy);
    gz = gsum<64>(gz);
    if (lane == 0) {
      val[pi] = v;
      if (grad) {
        grad[3 * pi] = gx;
        grad[3 * pi + 1] = gy;
        grad[3 * pi + 2] = gz;
      }
    }
  }
}

// Interpolating cubic B-spline coefficients from node values, one axis at a
// time (in place): the causal and anti-causal recursions of the pole
// z1 = sqrt(3) - 2 (Unser's direct B-spline filter) with the field taken as
// zero beyond the grid, where it vanishes.  One thread per grid line.
__global__ void bspline_prefilter_kernel(double *__restrict__ c, int n, uint64_t nlines,
                                         uint64_t stride, uint64_t inner) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  // line l: its first element; lines enumerate (outer, inner) index pairs
  const uint64_t base = (l / inner) * inner * (uint64_t)n + (l % inner);
  const double z1 = 1.7320508075688772 - 2.0;
  double *p = c + base;
  double cp = p[0];
  for (int k = 1; k < n; k++) {
    cp = p[(uint64_t)k * stride] + z1 * cp;
    p[(uint64_t)k * stride] = cp;
  }
  double cm = -z1 / (1.0 - z1 * z1) * p[(uint64_t)(n - 1) * stride];
  p[(uint64_t)(n - 1) * stride] = 6.0 * cm;
  for (int k = n - 2; k >= 0; k--) {
    cm = z1 * (cm - p[(uint64_t)k * stride]);
    p[(uint64_t)k * stride] = 6.0 * cm;
  }
}

// psi and grad psi of the tabulated spline at n points (validation).
__global__ void spline_eval_kernel(Tgt T, const double *__restrict__ xyz, uint64_t n,
                                   double *__restrict__ val, double *__restrict__ grad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  const double u[3] = {(p[0] - T.pox) * T.pih, (p[1] - T.poy) * T.pih, (p[2] - T.poz) * T.pih};
  long i0[3];
  double t[3];
  for (int d = 0; d < 3; d++) {
    const double f = floor(u[d]);
    i0[d] = (long)f - 1;
    t[d] = u[d] - f;
  }
  double v = 0, g[3] = {0, 0, 0};
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      for (int cc = 0; cc < 4; cc++) {
        const long ix = i0[0] + a, iy = i0[1] + b, iz = i0[2] + cc;
        if (ix < 0 || ix >= T.pnx || iy < 0 || iy >= T.pny || iz < 0 || iz >= T.pnz) continue;
        const double c = T.psi[((size_t)ix * T.pny + iy) * T.pnz + iz];
        double wx, dwx, ddx, wy, dwy, ddy, wz, dwz, ddz;
        bspline_piece(a, t[0], wx, dwx, ddx);
        bspline_piece(b, t[1], wy, dwy, ddy);
        bspline_piece(cc, t[2], wz, dwz, ddz);
        v += c * wx * wy * wz;
        g[0] += c * dwx * wy * wz * T.pih;
        g[1] += c * wx * dwy * wz * T.pih;
        g[2] += c * wx * wy * dwz * T.pih;
      }
  val[i] = v;
  for (int d = 0; d < 3; d++) grad[3 * i + d] = g[d];
}

// Distance from each point to the nearest target atom, capped at cap
// (cells within ceil(cap / cell) of the point's cell).
__global__ void mindist_kernel(Tgt T, const float *__restrict__ xyz, uint64_t n, double cap,
                               float *__restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
  const int R = (int)ceil(cap / T.cell);
  const int i = (int)floor((x - T.cx) / T.cell), j = (int)floor((y - T.cy) / T.cell),
            k = (int)floor((z - T.cz) / T.cell);
  double best = cap * cap;
  for (int a = max(0, i - R); a <= min(T.nx - 1, i + R); a++)
    for (int b = max(0, j - R); b <= min(T.ny - 1, j + R); b++)
      for (int c = max(0, k - R); c <= min(T.nz - 1, k + R); c++) {
        const int cell = c + T.nz * b + T.nz * T.ny * a;
        for (int q = T.cstart[cell]; q < T.cstart[cell + 1]; q++) {
          const double4 at = T.atoms[q];
          const double ex = at.x - x, ey = at.y - y, ez = at.z - z;
          best = fmin(best, ex * ex + ey * ey + ez * ez);
        }
      }
  out[p] = (float)sqrt(best);
}

// Positions of grid nodes [first, first + n) (x-major).
__global__ void grid_points_kernel(uint64_t first, uint64_t n, FarGrid g, float *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t q = first + i;
  const uint64_t ix = q / ((uint64_t)g.ny * g.nz), iy = (q / g.nz) % g.ny, iz = q % g.nz;
  out[3 * i] = (float)(g.ox + g.h * ix);
  out[3 * i + 1] = (float)(g.oy + g.h * iy);
  out[3 * i + 2] = (float)(g.oz + g.h * iz);
}

// Network output 0 at grid nodes, in psi units, tapered to zero between ts
// and te (distance to the nearest atom) where the exact field reaches zero.
__global__ void pack_network_kernel(const float *__restrict__ y, const float *__restrict__ d,
                                    uint64_t n, float ts, float te, double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float di = d[i];
  float t = 1.f;
  if (di >= te) t = 0.f;
  else if (di > ts) {
    const float x = (di - ts) / (te - ts);
    t = 1.f - x * x * (3.f - 2.f * x);
  }
  out[i] = (double)(y[2 * i] * t);
}

__global__ void f64_to_f32_kernel(const double *__restrict__ a, uint64_t n, float *__restrict__ b) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// Training targets: psi scaled, for both network outputs.
__global__ void targets_kernel(const double *__restrict__ v, uint64_t n, double s,
                               float *__restrict__ e) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  e[2 * i] = (float)(v[i] / s);
  e[2 * i + 1] = (float)(v[i] / s);
}

}  // namespace dev

namespace {

// Device buffer released on scope exit (error paths included).
template <class T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t n) { HIP_OK(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1))); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  operator T *() const { return p; }
};
double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
unsigned nblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

void far_eval(const Engine *E, const Tgt &T, const double *d_xyz, uint64_t n, double *d_val,
              double *d_grad, hipStream_t st) {
  FarGrid g{};
  const uint64_t blocks = std::min<uint64_t>((n + 3) / 4, (uint64_t)E->cus * 16);
  hipLaunchKernelGGL(far_eval_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, st, T, d_xyz, g, n,
                     d_val, d_grad);
  HIP_OK(hipGetLastError());
}

}  // namespace

bool split_capable(const Engine *E) { return (E->gk == 2 && E->ft == 6) || (E->gk == 1 && E->ft == 4); }

void engine_build_surrogate(Engine *E, const mccs_surrogate_opts &o, mccs_surrogate_report &rep,
                            float *params_out) {
  if (!split_capable(E))
    throw Error(MCCS_ERR_UNSUPPORTED,
                "the far-field surrogate is built for N2 force type 6 and He/Ar/co2 force type 4");
  HIP_OK(hipSetDevice(E->device));
  std::memset(&rep, 0, sizeof rep);
  const Tgt &T = E->T;
  const double sw1 = o.shell > 0 ? o.shell : 8.0, sw2 = sw1 + (o.blend > 0 ? o.blend : 4.0);
  const double h = o.grid > 0 ? o.grid : 0.5;
  if (sw2 > T.rc) throw Error(MCCS_ERR_INVALID, "shell + blend must not exceed the Coulomb cutoff");
  if (h < 0.05 || h > 2.0) throw Error(MCCS_ERR_INVALID, "grid spacing must lie in [0.05, 2] A");
  if (o.table_source < 0 || o.table_source > 2) throw Error(MCCS_ERR_INVALID, "table_source must be 0, 1 or 2");
  const bool network = o.table_source == 0;
  const uint64_t ntr = network ? (o.n_train ? o.n_train : 2000000ull) : 0, nval = 20000;
  hipStream_t st = E->stream;
  auto t0 = std::chrono::steady_clock::now();

  // the split target: the pair loop reaches sw2 (and the LJ cutoff)
  Tgt S = T;
  S.sw1 = sw1;
  S.sw2 = sw2;
  S.isw = 1.0 / (sw2 - sw1);
  S.sw2cut = sqrt_cut(sw2);
  S.rlong = std::max(T.rlj, sw2);
  // the lists of the exact kernel's capacity hold the shorter split lists
  E->shmem_split = E->shmem;

  // grid: the target's box plus the cutoff and six nodes of margin (the
  // spline prefilter takes the field as zero beyond the grid; it is, and its
  // coefficients decay by sqrt(3) - 2 per node from where it ends)
  FarGrid g{};
  const double marg = T.rc + 6.0 * h;
  g.h = h;
  g.ox = T.blo[0] - marg;
  g.oy = T.blo[1] - marg;
  g.oz = T.blo[2] - marg;
  g.nx = (int)std::ceil((T.bhi[0] - T.blo[0] + 2.0 * marg) / h) + 1;
  g.ny = (int)std::ceil((T.bhi[1] - T.blo[1] + 2.0 * marg) / h) + 1;
  g.nz = (int)std::ceil((T.bhi[2] - T.blo[2] + 2.0 * marg) / h) + 1;
  const uint64_t nodes = (uint64_t)g.nx * g.ny * g.nz;
  if (E->d_psi) HIP_OK(hipFree(E->d_psi));
  E->d_psi = nullptr;
  E->sur_built = false;
  HIP_OK(hipMalloc(&E->d_psi, sizeof(float) * nodes));
  DevBuf<double> node64(nodes);  // node values, then coefficients, in FP64

  // sample points where probes go: at least 2 A from every atom and within
  // the cutoff of some atom (psi vanishes beyond), uniform in the box;
  // nval of them are held out for validation
  const double dmin = 2.0;
  const uint64_t want = ntr + nval, chunk = 1u << 20;
  std::vector<double> pts;
  pts.reserve(3 * want);
  std::mt19937_64 rng(o.seed ? o.seed : 2104);
  std::uniform_real_distribution<double> U01(0.0, 1.0);
  {
    std::vector<float> hp(3 * chunk), hd(chunk);
    DevBuf<float> dp(3 * chunk), dd(chunk);
    for (int guard = 0; pts.size() < 3 * want; guard++) {
      if (guard > 1000) throw Error(MCCS_ERR_INVALID, "no room around the target to sample");
      for (uint64_t i = 0; i < chunk; i++)
        for (int d = 0; d < 3; d++)
          hp[3 * i + d] = (float)(T.blo[d] - T.rc + (T.bhi[d] - T.blo[d] + 2.0 * T.rc) * U01(rng));
      HIP_OK(hipMemcpyAsync(dp, hp.data(), sizeof(float) * 3 * chunk, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(mindist_kernel, dim3(nblk(chunk)), dim3(256), 0, st, T, (const float *)dp,
                         (uint64_t)chunk, T.rc + 1.0, dd);
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpyAsync(hd.data(), dd, sizeof(float) * chunk, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      for (uint64_t i = 0; i < chunk && pts.size() < 3 * want; i++)
        if (hd[i] >= dmin && hd[i] < T.rc)
          for (int d = 0; d < 3; d++) pts.push_back((double)hp[3 * i + d]);
    }
  }
  DevBuf<double> dx(3 * want), pv(want), pg(3 * want);
  HIP_OK(hipMemcpyAsync(dx, pts.data(), sizeof(double) * 3 * want, hipMemcpyHostToDevice, st));
  far_eval(E, S, dx, want, pv, pg, st);
  std::vector<double> hv(want), hg(3 * want);
  HIP_OK(hipMemcpyAsync(hv.data(), pv, sizeof(double) * want, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(hg.data(), pg, sizeof(double) * 3 * want, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  rep.seconds_sample = secs_since(t0);

  auto t2 = std::chrono::steady_clock::now();
  MlpSpec spec{};
  if (network) {
    // network: psi from its position; Fourier frequencies N(0, (L / flen)^2)
    // in units of the scaled input (1 / flen per A)
    auto t1 = std::chrono::steady_clock::now();
    double L = 0.0;
    for (int d = 0; d < 3; d++) {
      spec.centre[d] = 0.5 * (T.blo[d] + T.bhi[d]);
      L = std::max(L, 0.5 * (T.bhi[d] - T.blo[d]) + T.rc);
    }
    spec.in_scale = 1.0 / L;
    double ss = 0.0;
    for (uint64_t i = 0; i < ntr; i++) ss += hv[i] * hv[i];
    spec.out_scale[0] = spec.out_scale[1] = std::max(std::sqrt(ss / (double)ntr), 1e-9);
    std::normal_distribution<double> N01(0.0, 1.0);
    double ext = 0.0;
    for (int d = 0; d < 3; d++) ext = std::max(ext, T.bhi[d] - T.blo[d]);
    const double flen = o.feature_length > 0 ? o.feature_length : 0.23 * (ext + 2.0 * T.rc);
    for (int f = 0; f < MLP_NF; f++)
      for (int d = 0; d < 3; d++) spec.freq[f][d] = (float)(N01(rng) * L / flen);
    std::vector<float> params(MLP_NPARAM, 0.0f);
    auto init = [&](size_t off, int fin, int fout) {  // uniform +-sqrt(6 / (fin + fout))
      const double a = std::sqrt(6.0 / (fin + fout));
      for (size_t i = 0; i < (size_t)fin * fout; i++) params[off + i] = (float)(a * (2.0 * U01(rng) - 1.0));
    };
    init(MLP_W1, MLP_IN, MLP_H);
    init(MLP_W2, MLP_H, MLP_H);
    init(MLP_W3, MLP_H, MLP_H);
    init(MLP_W4, MLP_H, MLP_OUT);
    if (E->mlp) mlp_destroy(E->mlp);
    E->mlp = nullptr;
    E->mlp = mlp_create(E->device, spec, params.data(), 65536);
    // training data: positions, scaled psi, and the directional derivative of
    // psi along a random unit vector per point (matched by the network's
    // forward-mode derivative along it)
    DevBuf<float> txyz(3 * ntr), te(2 * ntr), tdir(3 * ntr), tder(2 * ntr);
    double wd = 1.0;
    {
      std::vector<float> dir(3 * ntr), der(2 * ntr);
      double sd = 0.0;
      for (uint64_t i = 0; i < ntr; i++) {
        double v[3], nn = 0.0;
        do {
          nn = 0.0;
          for (int d = 0; d < 3; d++) {
            v[d] = N01(rng);
            nn += v[d] * v[d];
          }
        } while (nn < 1e-12);
        nn = std::sqrt(nn);
        double dd = 0.0;
        for (int d = 0; d < 3; d++) {
          dir[3 * i + d] = (float)(v[d] / nn);
          dd += hg[3 * i + d] * (double)dir[3 * i + d];
        }
        der[2 * i] = der[2 * i + 1] = (float)(dd / spec.out_scale[0]);
        sd += 2.0 * (double)der[2 * i] * der[2 * i];
      }
      wd = 2.0 * ntr / std::max(sd, 1e-30);  // both terms of comparable size
      HIP_OK(hipMemcpy(tdir, dir.data(), sizeof(float) * 3 * ntr, hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(tder, der.data(), sizeof(float) * 2 * ntr, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(nblk(3 * ntr)), dim3(256), 0, st, (const double *)dx,
                       (uint64_t)(3 * ntr), txyz);
    hipLaunchKernelGGL(targets_kernel, dim3(nblk(ntr)), dim3(256), 0, st, (const double *)pv, ntr,
                       spec.out_scale[0], te);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    TrainOpts to;
    to.epochs = o.epochs > 0 ? o.epochs : 300;
    to.batch = o.batch > 0 ? o.batch : 8192;
    to.lr = o.lr > 0 ? o.lr : 5e-3;
    to.lr_final = to.lr / 20.0;
    to.seed = o.seed ? o.seed : 2104;
    to.d_dir = tdir;
    to.d_deriv = tder;
    to.w_deriv = wd;
    rep.train_loss = mlp_train(E->mlp, txyz, te, ntr, to, nullptr);
    rep.seconds_train = secs_since(t1);
    // held-out error of the network itself (output 0 and its gradient)
    std::vector<double> y(2 * nval), dy(6 * nval);
    mlp_eval(E->mlp, pts.data() + 3 * ntr, nval, y.data(), dy.data());
    double eu = 0, su = 0, ef = 0, sf = 0;
    for (uint64_t i = 0; i < nval; i++) {
      const double ue = hv[ntr + i];
      eu += (y[2 * i] - ue) * (y[2 * i] - ue);
      su += ue * ue;
      for (int d = 0; d < 3; d++) {
        const double ge = hg[3 * (ntr + i) + d], gm = dy[6 * i + d];
        ef += (gm - ge) * (gm - ge);
        sf += ge * ge;
      }
    }
    rep.val_rmse_u[1] = std::sqrt(eu / nval);
    rep.val_rms_u[1] = std::sqrt(su / nval);
    rep.val_rmse_f[1] = std::sqrt(ef / (3.0 * nval));
    rep.val_rms_f[1] = std::sqrt(sf / (3.0 * nval));
    t2 = std::chrono::steady_clock::now();
    // tabulate the network at the nodes (MFMA inference), tapered to zero
    // over the last 3 A before the cutoff distance from every atom
    const uint64_t gch = 1u << 20;
    DevBuf<float> gp(3 * gch), gd(gch), gy(2 * gch);
    for (uint64_t f0 = 0; f0 < nodes; f0 += gch) {
      const uint64_t c = std::min(gch, nodes - f0);
      hipLaunchKernelGGL(grid_points_kernel, dim3(nblk(c)), dim3(256), 0, st, f0, c, g, gp);
      hipLaunchKernelGGL(mindist_kernel, dim3(nblk(c)), dim3(256), 0, st, T, (const float *)gp, c,
                         T.rc + 1.0, gd);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(st));
      mlp_eval_device(E->mlp, gp, c, gy, nullptr);
      hipLaunchKernelGGL(pack_network_kernel, dim3(nblk(c)), dim3(256), 0, st, (const float *)gy,
                         (const float *)gd, c, (float)(T.rc - 3.0), (float)T.rc, node64 + f0);
      HIP_OK(hipGetLastError());
    }
  } else if (o.table_source == 2) {
    // no far field: the near terms alone (the ablation the network is
    // measured against)
    HIP_OK(hipMemsetAsync((double *)node64, 0, sizeof(double) * nodes, st));
  } else {
    // the exact field at every node (the linked-cell evaluator)
    const uint64_t blocks = std::min<uint64_t>((nodes + 3) / 4, (uint64_t)E->cus * 16);
    hipLaunchKernelGGL(far_eval_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, st, S,
                       (const double *)nullptr, g, nodes, (double *)node64, (double *)nullptr);
    HIP_OK(hipGetLastError());
  }
  // node values -> interpolating B-spline coefficients, axis by axis
  {
    const uint64_t nxy = (uint64_t)g.nx * g.ny, nyz = (uint64_t)g.ny * g.nz;
    hipLaunchKernelGGL(bspline_prefilter_kernel, dim3(nblk(nxy)), dim3(256), 0, st, (double *)node64, g.nz,
                       nxy, (uint64_t)1, (uint64_t)1);
    hipLaunchKernelGGL(bspline_prefilter_kernel, dim3(nblk((uint64_t)g.nx * g.nz)), dim3(256), 0, st,
                       (double *)node64, g.ny, (uint64_t)g.nx * g.nz, (uint64_t)g.nz, (uint64_t)g.nz);
    hipLaunchKernelGGL(bspline_prefilter_kernel, dim3(nblk(nyz)), dim3(256), 0, st, (double *)node64, g.nx,
                       nyz, nyz, nyz);
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(nblk(nodes)), dim3(256), 0, st, (const double *)node64,
                       nodes, E->d_psi);
    HIP_OK(hipGetLastError());
  }
  S.psi = E->d_psi;
  S.pox = g.ox;
  S.poy = g.oy;
  S.poz = g.oz;
  S.pih = 1.0 / h;
  S.pnx = g.nx;
  S.pny = g.ny;
  S.pnz = g.nz;
  HIP_OK(hipStreamSynchronize(st));
  rep.seconds_table = secs_since(t2);
  // held-out error of the table the trajectories use (against the exact field)
  {
    DevBuf<double> sv(nval), sg(3 * nval);
    hipLaunchKernelGGL(spline_eval_kernel, dim3(nblk(nval)), dim3(256), 0, st, S,
                       (const double *)(dx + 3 * ntr), nval, sv, sg);
    HIP_OK(hipGetLastError());
    std::vector<double> v(nval), gg(3 * nval);
    HIP_OK(hipMemcpyAsync(v.data(), sv, sizeof(double) * nval, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(gg.data(), sg, sizeof(double) * 3 * nval, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double eu = 0, su = 0, ef = 0, sf = 0;
    for (uint64_t i = 0; i < nval; i++) {
      const double ue = hv[ntr + i];
      eu += (v[i] - ue) * (v[i] - ue);
      su += ue * ue;
      for (int d = 0; d < 3; d++) {
        const double ge = hg[3 * (ntr + i) + d];
        ef += (gg[3 * i + d] - ge) * (gg[3 * i + d] - ge);
        sf += ge * ge;
      }
    }
    rep.val_rmse_u[0] = std::sqrt(eu / nval);
    rep.val_rms_u[0] = std::sqrt(su / nval);
    rep.val_rmse_f[0] = std::sqrt(ef / (3.0 * nval));
    rep.val_rms_f[0] = std::sqrt(sf / (3.0 * nval));
  }
  rep.n_train = ntr;
  rep.grid_nodes = nodes;
  if (network) {
    for (int f = 0; f < MLP_NF; f++)
      for (int d = 0; d < 3; d++) rep.spec.freq[f][d] = spec.freq[f][d];
    for (int d = 0; d < 3; d++) rep.spec.centre[d] = spec.centre[d];
    rep.spec.in_scale = spec.in_scale;
    rep.spec.out_scale[0] = spec.out_scale[0];
    rep.spec.out_scale[1] = spec.out_scale[1];
    if (params_out) mlp_get_params(E->mlp, params_out);
  }
  E->Ts = S;
  E->sur_built = true;
  if (E->split) {  // rebuilt while in use: the kernel sees the new table
    HIP_OK(hipMemcpy(E->d_tg, &E->Ts, sizeof(Tgt), hipMemcpyHostToDevice));
  }
}

void engine_use_surrogate(Engine *E, bool on) {
  HIP_OK(hipSetDevice(E->device));
  if (on && !E->sur_built) throw Error(MCCS_ERR_INVALID, "no surrogate has been built for this engine");
  E->split = on;
  HIP_OK(hipMemcpy(E->d_tg, on ? &E->Ts : &E->T, sizeof(Tgt), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------- batches
// Many targets in one run: targets are grouped by probe kind and force type
// (one kernel instantiation each); each group runs as ONE persistent launch
// whose device queue spans all its targets' trajectories, ordered by
// predicted cost (largest targets first) so that the long ones start early.
// The launches are captured once in a hipGraph and replayed.
namespace {

struct BatchClass {
  int gk = 1, ft = 2;
  bool split = false;        // members run with their range-split far fields
  std::vector<int> members;  // target indices, in queue order
  int lanes = 64, block = 256;
  double skin = 1.0;
  int tile = 0;              // LDS per block: 0 by the list-size rule, 1 half a CU, 2 a whole CU
  bool tuned = false;
  Tgt *d_tg = nullptr;
  unsigned long long *d_toff = nullptr;
  size_t shmem = 0;
  uint64_t blocks = 0, n = 0;
  TrajFn fn = nullptr;
  QueueWS q;
  double *d_u = nullptr;
  mccs_traj_record *d_rec = nullptr;
  uint64_t cap = 0;
  std::vector<unsigned long long> toff;
};

// Tuned launch shapes by class signature: probe kind, force type, cutoffs and
// the targets' sizes rounded to powers of two.
struct Shape {
  int lanes;
  double skin;
  int tile;
};
std::map<std::string, Shape> &tune_cache() {
  static std::map<std::string, Shape> c;
  return c;
}

}  // namespace

struct Batch {
  int device = 0, cus = 0;
  bool count = true, use_graph = true;
  bool split = false;  // far-field surrogates in use (where built)
  bool ordered = true;  // primary trajectories in the start order (likely-long first)
  int lanes_opt = 0;
  double skin_opt = 0.0;
  hipStream_t stream = nullptr;
  std::vector<Engine *> eng;  // per target: device-resident target data
  std::vector<std::string> sig;
  std::deque<BatchClass> cls;  // elements stay in place (they own device memory)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gkey;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;  // orders batch_run_device after / before a caller's stream
  ~Batch() {
    hipSetDevice(device);
    if (ev_in) hipEventDestroy(ev_in);
    if (ev_out) hipEventDestroy(ev_out);
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    for (auto &c : cls) {
      if (c.d_tg) hipFree(c.d_tg);
      if (c.d_toff) hipFree(c.d_toff);
      if (c.d_u) hipFree(c.d_u);
      if (c.d_rec) hipFree(c.d_rec);
    }
    for (Engine *e : eng) delete e;
    if (stream) hipStreamDestroy(stream);
  }
};

static std::string class_signature(const Batch &B, const BatchClass &c) {
  std::ostringstream o;
  o << c.gk << "/" << c.ft << (B.split ? "/far" : "");
  for (int t : c.members) {
    const Tgt &T = B.eng[t]->T;
    int lg = 0;
    while ((1 << lg) < T.natoms) lg++;
    o << "/" << lg << ":" << T.rlj << ":" << T.rc;
  }
  return o.str();
}

// (Re)builds a class's device target array for its launch shape: list
// capacity for the shape's LDS budget and the chosen skin.
static void configure_class(Batch &B, BatchClass &c) {
  c.split = B.split;
  for (int t : c.members) c.split = c.split && B.eng[t]->sur_built;
  c.fn = c.gk == 1 ? select_traj_multi1(c.ft, c.lanes, B.count, c.split)
                   : (c.gk == 2 ? select_traj_multi2(c.ft, c.lanes, B.count, c.split) : nullptr);
  if (!c.fn)
    throw Error(MCCS_ERR_UNSUPPORTED,
                "batched runs support He/Ar/co2 force types 2 and 4 and N2 force type 6");
  const int groups = c.block / c.lanes;
  const bool central = c.gk == 3;
  const size_t entry = sizeof(int);
  const size_t base = lds_bytes(0, 0, c.block, c.lanes, false, central);
  // half a CU per block when every member's estimated lists, scaled from a
  // 1 A skin to the class's (and to the split radius), fit in 90 % of it;
  // else the whole CU
  const size_t half = base < LDS_BLOCK_HALF ? (LDS_BLOCK_HALF - base) / (entry * groups) : 0;
  bool fit_half = true;
  for (int t : c.members) {
    const Engine *e = B.eng[t];
    const double r1 = e->T.rlong + 1.0 + LIST_OFF_MAX;
    const double rs = (c.split ? e->Ts.rlong : e->T.rlong) + c.skin + LIST_OFF_MAX;
    fit_half = fit_half && e->list_need * (rs / r1) * (rs / r1) * (rs / r1) <= 0.9 * (double)half;
  }
  const size_t budget = c.tile == 1 ? LDS_BLOCK_HALF
                        : (c.tile == 2 ? LDS_BLOCK_FULL : (fit_half ? LDS_BLOCK_HALF : LDS_BLOCK_FULL));
  const size_t room = base < budget ? budget - base : 0;
  const int cap = (int)(room / (entry * groups));
  std::vector<Tgt> tg;
  int maxcap = 0;
  for (int t : c.members) {
    Tgt T = c.split ? B.eng[t]->Ts : B.eng[t]->T;
    T.list_cap = std::min(T.natoms, cap);
    T.skin = c.skin;
    maxcap = std::max(maxcap, T.list_cap);
    tg.push_back(T);
  }
  c.shmem = lds_bytes(0, maxcap, c.block, c.lanes, false, central);
  if (!c.d_tg) HIP_OK(hipMalloc(&c.d_tg, sizeof(Tgt) * tg.size()));
  HIP_OK(hipMemcpy(c.d_tg, tg.data(), sizeof(Tgt) * tg.size(), hipMemcpyHostToDevice));
  set_shmem((const void *)c.fn, c.shmem);
  int per_cu = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)c.fn, c.block, c.shmem));
  if (per_cu < 1) throw Error(MCCS_ERR_DEVICE, "batched trajectory kernel does not fit on a CU");
  c.blocks = (uint64_t)per_cu * B.cus;
}

Batch *batch_create(const Prepared *const *targets, int ntargets, const mccs_batch_opts &o) {
  if (ntargets <= 0) throw Error(MCCS_ERR_INVALID, "a batch needs at least one target");
  std::unique_ptr<Batch> B(new Batch);
  B->device = o.device;
  B->count = o.count_work != 0;
  B->use_graph = o.use_graph != 0;
  B->lanes_opt = o.lanes;
  B->skin_opt = o.skin;
  if (o.order < 0 || o.order > 1) throw Error(MCCS_ERR_INVALID, "order must be 0 or 1");
  B->ordered = o.order == 0;
  if (o.lanes != 0 && o.lanes != 16 && o.lanes != 32 && o.lanes != 64)
    throw Error(MCCS_ERR_INVALID, "lanes must be 0, 16, 32 or 64");
  if (o.skin < 0.0) throw Error(MCCS_ERR_INVALID, "skin must be >= 0 (0 = tuned)");
  mccs_engine_opts eo{};
  eo.device = o.device;
  eo.lanes = 64;
  eo.use_lds = 0;
  eo.count_work = o.count_work;
  eo.spec_below = -1;
  for (int t = 0; t < ntargets; t++) {
    if (!targets[t]) throw Error(MCCS_ERR_INVALID, "null target");
    B->eng.push_back(engine_create(*targets[t], eo));
  }
  B->device = B->eng[0]->device;
  B->cus = B->eng[0]->cus;
  HIP_OK(hipSetDevice(B->device));
  HIP_OK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  // classes by (probe kind, force type); members by predicted cost: atoms
  // within the longest cutoff grow with the target, so larger targets first
  std::map<std::pair<int, int>, std::vector<int>> by;
  for (int t = 0; t < ntargets; t++) by[{B->eng[t]->gk, B->eng[t]->ft}].push_back(t);
  for (auto &kv : by) {
    B->cls.emplace_back();
    BatchClass &c = B->cls.back();
    c.gk = kv.first.first;
    c.ft = kv.first.second;
    c.members = kv.second;
    std::stable_sort(c.members.begin(), c.members.end(),
                     [&](int a, int b) { return B->eng[a]->T.natoms > B->eng[b]->T.natoms; });
    c.q.init(B->device);
    c.lanes = o.lanes ? o.lanes : 64;
    c.skin = o.skin > 0.0 ? o.skin : 1.0;
    auto it = tune_cache().find(class_signature(*B, c));
    if (it != tune_cache().end() && !o.lanes && !(o.skin > 0.0)) {
      c.lanes = it->second.lanes;
      c.skin = it->second.skin;
      c.tile = it->second.tile;
      c.tuned = true;
    }
    configure_class(*B, c);
  }
  return B.release();
}

void batch_destroy(Batch *b) { delete b; }

// Uploads the samples of n[t] trajectories per target into each class's
// concatenated buffers; returns whether any class changed size.
static bool stage_samples(Batch &B, const uint64_t *n, const double *const *samples) {
  bool changed = false;
  for (auto &c : B.cls) {
    std::vector<unsigned long long> off(1, 0);
    for (int t : c.members) off.push_back(off.back() + n[t]);
    const uint64_t tot = off.back();
    if (tot > c.cap) {
      if (c.d_u) HIP_OK(hipFree(c.d_u));
      if (c.d_rec) HIP_OK(hipFree(c.d_rec));
      c.d_u = nullptr;
      c.d_rec = nullptr;
      HIP_OK(hipMalloc(&c.d_u, sizeof(double) * 9 * std::max<uint64_t>(tot, 1)));
      HIP_OK(hipMalloc(&c.d_rec, sizeof(mccs_traj_record) * std::max<uint64_t>(tot, 1)));
      c.cap = tot;
      changed = true;
    }
    if (off != c.toff) changed = true;
    c.toff = off;
    c.n = tot;
    if (!c.d_toff) HIP_OK(hipMalloc(&c.d_toff, sizeof(unsigned long long) * (c.members.size() + 1)));
    HIP_OK(hipMemcpyAsync(c.d_toff, off.data(), sizeof(unsigned long long) * off.size(),
                          hipMemcpyHostToDevice, B.stream));
    for (size_t m = 0; m < c.members.size(); m++) {
      const int t = c.members[m];
      if (n[t] && !samples[t]) throw Error(MCCS_ERR_INVALID, "null samples");
      if (n[t])
        HIP_OK(hipMemcpyAsync(c.d_u + 9 * off[m], samples[t], sizeof(double) * 9 * n[t],
                              hipMemcpyDefault, B.stream));  // host or device samples
    }
  }
  return changed;
}

static void launch_class(Batch &B, BatchClass &c, hipStream_t s) {
  if (c.n == 0) return;
  const uint64_t gpb = c.block / c.lanes;
  const uint64_t blocks = std::min<uint64_t>(c.blocks, (c.n + gpb - 1) / gpb);
  Queue Q = c.q.make(c.n, (long long)(blocks * gpb), B.ordered);
  launch_queue_init(Q, c.n, s);
  const Tgt *tg = c.d_tg;
  const unsigned long long *toff = c.d_toff;
  int ntg = (int)c.members.size();
  const double *u = c.d_u;
  uint64_t n = c.n;
  launch_order(Q, tg, toff, ntg, u, n, s);
  mccs_traj_record *rec = c.d_rec;
  void *args[] = {&tg, &toff, &ntg, &u, &n, &rec, &Q};
  HIP_OK(hipLaunchKernel((const void *)c.fn, dim3((unsigned)blocks), dim3(c.block), args, c.shmem, s));
}

// Runs every class on the batch stream: replays the captured graph when the
// sizes and shapes are those it was captured with.
static void run_classes(Batch &B, bool sizes_changed) {
  if (!B.use_graph) {
    for (auto &c : B.cls) launch_class(B, c, B.stream);
    return;
  }
  std::vector<uint64_t> key;
  for (auto &c : B.cls) {
    key.push_back(c.n);
    key.push_back((uint64_t)c.lanes);
    key.push_back((uint64_t)(c.skin * 1000.0));
    key.push_back((uint64_t)c.split);
    key.push_back((uint64_t)c.tile);
  }
  if (!B.gexec || sizes_changed || key != B.gkey) {
    if (B.gexec) HIP_OK(hipGraphExecDestroy(B.gexec));
    if (B.graph) HIP_OK(hipGraphDestroy(B.graph));
    B.gexec = nullptr;
    B.graph = nullptr;
    // every queue workspace must exist before capture (no allocation inside)
    for (auto &c : B.cls) c.q.make(c.n, 0, B.ordered);
    HIP_OK(hipStreamSynchronize(B.stream));
    HIP_OK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    try {
      for (auto &c : B.cls) launch_class(B, c, B.stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      hipStreamEndCapture(B.stream, &g);
      if (g) hipGraphDestroy(g);
      throw;
    }
    HIP_OK(hipStreamEndCapture(B.stream, &B.graph));
    HIP_OK(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
    B.gkey = key;
  }
  HIP_OK(hipGraphLaunch(B.gexec, B.stream));
}

static void fetch_records(Batch &B, const uint64_t *n, mccs_traj_record *const *records) {
  for (auto &c : B.cls)
    for (size_t m = 0; m < c.members.size(); m++) {
      const int t = c.members[m];
      if (n[t] && !records[t]) throw Error(MCCS_ERR_INVALID, "null records");
      if (n[t])
        HIP_OK(hipMemcpyAsync(records[t], c.d_rec + c.toff[m], sizeof(mccs_traj_record) * n[t],
                              hipMemcpyDeviceToHost, B.stream));
    }
  HIP_OK(hipStreamSynchronize(B.stream));
  for (auto &c : B.cls) c.q.check_stall();
}

void batch_run(Batch *B, const uint64_t *n, const double *const *samples,
               mccs_traj_record *const *records) {
  HIP_OK(hipSetDevice(B->device));
  const bool changed = stage_samples(*B, n, samples);
  run_classes(*B, changed);
  fetch_records(*B, n, records);
}

// The same over device-resident samples and records, enqueued after the
// work already on `stream` and finished before later work on it (the batch
// runs on its own stream, which waits on and then releases the caller's).
void batch_run_device(Batch *B, const uint64_t *n, const double *const *samples_dev,
                      mccs_traj_record *const *records_dev, void *stream) {
  HIP_OK(hipSetDevice(B->device));
  hipStream_t s = stream ? (hipStream_t)stream : B->stream;
  if (!B->ev_in) {
    HIP_OK(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
  }
  if (s != B->stream) {
    HIP_OK(hipEventRecord(B->ev_in, s));
    HIP_OK(hipStreamWaitEvent(B->stream, B->ev_in, 0));
  }
  const bool changed = stage_samples(*B, n, samples_dev);
  run_classes(*B, changed);
  for (auto &c : B->cls)
    for (size_t m = 0; m < c.members.size(); m++) {
      const int t = c.members[m];
      if (n[t] && !records_dev[t]) throw Error(MCCS_ERR_INVALID, "null records");
      if (n[t])
        HIP_OK(hipMemcpyAsync(records_dev[t], c.d_rec + c.toff[m], sizeof(mccs_traj_record) * n[t],
                              hipMemcpyDeviceToDevice, B->stream));
    }
  if (s != B->stream) {
    HIP_OK(hipEventRecord(B->ev_out, B->stream));
    HIP_OK(hipStreamWaitEvent(s, B->ev_out, 0));
  }
}

void batch_last_work(Batch *B, uint64_t out[5]) {
  HIP_OK(hipSetDevice(B->device));
  HIP_OK(hipStreamSynchronize(B->stream));
  for (int k = 0; k < 5; k++) out[k] = 0;
  for (auto &c : B->cls) {
    uint64_t w[5];
    c.q.work(w);
    for (int k = 0; k < 5; k++) out[k] += w[k];
  }
}

// Autotuner: per class, times trial runs (the first n_trial[t] samples of
// every target) over lanes {16, 32, 64} x Verlet skin {0.5, 1, 2} A x LDS
// tile {half a CU, a whole CU per block} (the Verlet-list capacity: half a
// CU runs two blocks per CU, a whole CU longer lists that overflow less) and
// keeps the fastest; results are cached for this process by class signature.
std::string batch_autotune(Batch *B, const uint64_t *n_trial, const double *const *samples) {
  HIP_OK(hipSetDevice(B->device));
  std::ostringstream rep;
  rep << "[";
  const int lanes_c[3] = {16, 32, 64};
  const double skin_c[3] = {0.5, 1.0, 2.0};
  std::vector<uint64_t> nt(B->eng.size());
  for (size_t t = 0; t < nt.size(); t++) nt[t] = n_trial[t];
  bool first = true;
  for (auto &c : B->cls) {
    std::vector<Shape> cand;
    const bool split = B->split && [&] {
      for (int t : c.members)
        if (!B->eng[t]->sur_built) return false;
      return true;
    }();
    for (int L : lanes_c)
      for (double sk : skin_c)
        for (int tile : {1, 2}) {
          if (B->lanes_opt && L != B->lanes_opt) continue;
          if (B->skin_opt > 0.0 && sk != B->skin_opt) continue;
          cand.push_back({L, sk, tile});
        }
    if (B->skin_opt > 0.0 && cand.empty())
      for (int tile : {1, 2}) cand.push_back({B->lanes_opt ? B->lanes_opt : 64, B->skin_opt, tile});
    // trial samples of this class's members only
    std::vector<uint64_t> n_only(B->eng.size(), 0);
    for (int t : c.members) n_only[t] = nt[t];
    double best = 1e300;
    Shape pick{c.lanes, c.skin, c.tile};
    if (!first) rep << ",";
    first = false;
    rep << "{\"class\": \"" << class_signature(*B, c) << "\", \"trials\": [";
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    bool f2 = true;
    for (auto &cd : cand) {
      c.lanes = cd.lanes;
      c.skin = cd.skin;
      c.tile = cd.tile;
      float ms = 0.f;
      try {
        configure_class(*B, c);
        stage_samples(*B, n_only.data(), samples);
        HIP_OK(hipEventRecord(e0, B->stream));
        launch_class(*B, c, B->stream);
        HIP_OK(hipEventRecord(e1, B->stream));
        HIP_OK(hipEventSynchronize(e1));
        c.q.check_stall();
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      } catch (const Error &e) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        throw Error(e.code, "autotune trial (lanes " + std::to_string(cd.lanes) + ", skin " +
                                std::to_string(cd.skin) + ", tile " + std::to_string(cd.tile) +
                                "): " + e.what());
      }
      if (!f2) rep << ", ";
      f2 = false;
      rep << "{\"lanes\": " << cd.lanes << ", \"skin\": " << cd.skin << ", \"tile\": " << cd.tile
          << ", \"ms\": " << ms << "}";
      if (ms < best) {
        best = ms;
        pick = cd;
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c.lanes = pick.lanes;
    c.skin = pick.skin;
    c.tile = pick.tile;
    c.tuned = true;
    configure_class(*B, c);
    tune_cache()[class_signature(*B, c)] = pick;
    rep << "], \"lanes\": " << pick.lanes << ", \"skin\": " << pick.skin << ", \"tile\": " << pick.tile
        << "}";
  }
  rep << "]";
  // the graph (if any) was captured with the old shapes
  if (B->gexec) HIP_OK(hipGraphExecDestroy(B->gexec));
  if (B->graph) HIP_OK(hipGraphDestroy(B->graph));
  B->gexec = nullptr;
  B->graph = nullptr;
  B->gkey.clear();
  return rep.str();
}

// Far-field surrogates of every member with a dipole field (N2 type 6,
// He/Ar/co2 type 4), one build per target (the others run exact); reports
// (may be null) receive one report per target, zero for the others.
void batch_build_surrogate(Batch *B, const mccs_surrogate_opts &o, mccs_surrogate_report *reports) {
  for (size_t t = 0; t < B->eng.size(); t++) {
    mccs_surrogate_report r{};
    if (split_capable(B->eng[t])) engine_build_surrogate(B->eng[t], o, r, nullptr);
    if (reports) reports[t] = r;
  }
  if (B->split) {
    for (auto &c : B->cls) configure_class(*B, c);
    B->gkey.clear();
  }
}

void batch_use_surrogate(Batch *B, bool on) {
  HIP_OK(hipSetDevice(B->device));
  B->split = on;
  for (auto &c : B->cls) {
    // tuned shapes are cached per signature, which names the far field
    auto it = tune_cache().find(class_signature(*B, c));
    if (it != tune_cache().end() && !B->lanes_opt && !(B->skin_opt > 0.0)) {
      c.lanes = it->second.lanes;
      c.skin = it->second.skin;
      c.tile = it->second.tile;
    } else {
      if (!B->lanes_opt) c.lanes = 64;
      c.tile = 0;
    }
    configure_class(*B, c);
  }
  B->gkey.clear();
}

void batch_shapes(const Batch *B, int32_t *lanes, double *skin) {
  for (const auto &c : B->cls)
    for (int t : c.members) {
      if (lanes) lanes[t] = c.lanes;
      if (skin) skin[t] = c.skin;
    }
}

}  // namespace mccs
