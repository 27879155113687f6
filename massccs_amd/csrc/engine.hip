// MI355X (gfx950) trajectory engine: per-point site kernel, tally kernels and
// the host-side launcher behind the C ABI.
#include "kernels.hpp"

namespace mccs {
namespace dev {

// Per-point site forces: one group of L lanes per point.
template <class S, int L, bool LDS>
__global__ __launch_bounds__(512) void eval_kernel(Tgt T, const double *__restrict__ xyz, uint64_t n,
                                                   double qsite, double *__restrict__ fout,
                                                   double *__restrict__ uout) {
  extern __shared__ __align__(16) char smem[];
  Shm sh;
  stage<L, LDS, any_central<S>()>(T, sh, smem);
  const int lane = threadIdx.x % L;
  const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / L);
  Work w{0, 0, 0, 0};
  for (uint64_t g = (uint64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; g < n; g += groups) {
    double p[1][3] = {{xyz[3 * g], xyz[3 * g + 1], xyz[3 * g + 2]}};
    double qs[1] = {qsite};
    double f[1][3], U[1];
    evaluate<S, L, LDS, false>(T, sh, lane, p, qs, f, U, w, nullptr);
    if (lane == 0) {
      fout[3 * g] = f[0][0];
      fout[3 * g + 1] = f[0][1];
      fout[3 * g + 2] = f[0][2];
      uout[g] = U[0];
    }
  }
}

// Tally: one block per iteration, thread 0 sums its trajectories in index
// order exactly as src/System.cpp:287-307.
__global__ void tally_iter_kernel(const mccs_traj_record *__restrict__ rec, uint64_t n_traj,
                                  double *__restrict__ omega, int64_t *__restrict__ counts) {
  if (threadIdx.x != 0) return;
  const uint64_t i = blockIdx.x;
  const mccs_traj_record *r = rec + i * n_traj;
  double dO = 0.0;
  int nf = 0, ns = 0, nl = 0;
  for (uint64_t k = 0; k < n_traj; k++) {
    const mccs_traj_record x = r[k];
    dO += x.domega;
    if (!x.hit) nf++;
    else if (x.success) ns++;
    else nl++;
  }
  omega[i] = 1.0 / ((double)(float)(ns + nf)) * dO;
  counts[3 * i] = nf;
  counts[3 * i + 1] = ns;
  counts[3 * i + 2] = nl;
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

TrajFn select_traj(int gk, int ft, int L, bool lds, bool count) {
  switch (gk) {
    case 1: return select_traj_gk1(ft, L, lds, count);
    case 2: return select_traj_gk2(ft, L, lds, count);
    case 3: return select_traj_gk3(ft, L, lds, count);
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

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int gk = 1, ft = 2, lanes = 64, block = 256, cus = 0;
  bool lds = false, count = false;
  size_t shmem = 0;
  Tgt T{};
  double site_q[5] = {0, -0.4825, 0.965, -0.3256, 0.6512};
  std::vector<void *> allocs;
  // work queue: counters, each polled one on its own 256-byte line (ticket,
  // shead, stail, active, work[5], stalled), and a per-trajectory workspace
  // grown on demand
  unsigned long long *qctl = nullptr;
  char *qws = nullptr;
  uint64_t qcap = 0;
  unsigned long long wait_ticks = 0;
  int spec_below = -1;
  hipStream_t last_stream = nullptr;
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
    for (void *p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
};

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
  const double rlj = P.cfg.lj_cutoff, rc = P.cfg.coul_cutoff;
  T.rlj = rlj; T.rc = rc;
  T.rlj2lo = rlj * rlj * (1.0 - 1e-15); T.rlj2hi = rlj * rlj * (1.0 + 1e-15);
  T.rc2lo = rc * rc * (1.0 - 1e-15); T.rc2hi = rc * rc * (1.0 + 1e-15);
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
  E->qctl = reinterpret_cast<unsigned long long *>(E->dalloc(sizeof(unsigned long long) * QCTL_WORDS));
  HIP_OK(hipMemset(E->qctl, 0, sizeof(unsigned long long) * QCTL_WORDS));
  {  // a group waits for work at most 120 s (any single try takes far less)
    int khz = 0;
    HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, E->device));
    if (khz <= 0) khz = 100000;
    E->wait_ticks = 120ull * 1000ull * (unsigned long long)khz;
  }

  // launch shape: lanes per trajectory by target size unless given
  int L = o.lanes;
  if (L == 0) L = P.natoms <= 200 ? 16 : 64;
  if (L != 16 && L != 32 && L != 64) throw Error(MCCS_ERR_INVALID, "lanes must be 0, 16, 32 or 64");
  E->lanes = L;
  E->block = o.block ? o.block : 256;
  if (E->block != 128 && E->block != 256) throw Error(MCCS_ERR_INVALID, "block must be 0, 128 or 256");
  // Atoms (and their LJ pairs) are staged in LDS when the block's share fits
  // 80 KB (two blocks per CU); Verlet lists then hold 16-bit indices.
  const bool central = E->gk == 3;
  const size_t per_block = 80 * 1024;
  const bool fits = T.natoms <= 65535 &&
                    lds_bytes(T.natoms, T.natoms, E->block, L, true, central) <= per_block;
  E->lds = o.use_lds < 0 ? fits : (o.use_lds != 0);
  if (E->lds && T.natoms > 65535) throw Error(MCCS_ERR_INVALID, "target too large to stage in LDS");
  // list capacity: every atom for small targets, else what 120 KB of lists allow
  const int groups = E->block / L;
  const size_t lidx = E->lds ? 2 : 4;
  const size_t base = lds_bytes(T.natoms, 0, E->block, L, E->lds, central);
  const size_t room = base < 150 * 1024 ? 150 * 1024 - base : 0;
  T.list_cap = (int)std::min<size_t>((size_t)T.natoms, room / (lidx * groups));
  E->shmem = lds_bytes(T.natoms, T.list_cap, E->block, L, E->lds, central);
  if (E->shmem > 160 * 1024) throw Error(MCCS_ERR_INVALID, "target too large to stage in LDS");
  return E.release();
}

void engine_destroy(Engine *e) { delete e; }

static void set_shmem(const void *fn, size_t bytes) {
  if (bytes > 64 * 1024)
    HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void engine_run_device(Engine *E, const double *u_dev, uint64_t n, mccs_traj_record *rec_dev,
                       void *stream) {
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  hipStream_t s = stream ? (hipStream_t)stream : E->stream;
  TrajFn fn = select_traj(E->gk, E->ft, E->lanes, E->lds, E->count);
  if (!fn) {  // force types without tuned variants run 64 lanes with work counters
    E->lanes = 64;
    E->count = true;
    fn = select_traj(E->gk, E->ft, 64, E->lds, true);
  }
  if (!fn) throw Error(MCCS_ERR_UNSUPPORTED, "no trajectory kernel for this gas / force type");
  set_shmem((const void *)fn, E->shmem);
  int per_cu = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, E->block, E->shmem));
  if (per_cu < 1) throw Error(MCCS_ERR_DEVICE, "trajectory kernel does not fit on a CU");
  const uint64_t groups_per_block = E->block / E->lanes;
  uint64_t blocks = (uint64_t)per_cu * E->cus;
  blocks = std::min<uint64_t>(blocks, (n + groups_per_block - 1) / groups_per_block);
  // per trajectory: TRIES_MAX try outcomes, 8 spawn slots, claimed tries,
  // spawned and resolved flags
  constexpr size_t per_traj = TRIES_MAX * sizeof(TryOut) + 8 * sizeof(unsigned long long) + 3 * sizeof(int);
  if (n > E->qcap) {
    if (E->qws) {
      auto it = std::find(E->allocs.begin(), E->allocs.end(), (void *)E->qws);
      if (it != E->allocs.end()) E->allocs.erase(it);
      HIP_OK(hipFree(E->qws));
      E->qws = nullptr;
      E->qcap = 0;
    }
    const uint64_t cap = std::max<uint64_t>(n, 1024);
    E->qws = reinterpret_cast<char *>(E->dalloc(per_traj * cap + sizeof(unsigned long long) * SLOT_SLACK));
    E->qcap = cap;
  }
  Queue Q;
  Q.ticket = E->qctl + 0 * QLINE;
  Q.shead = E->qctl + 1 * QLINE;
  Q.stail = E->qctl + 2 * QLINE;
  Q.active = reinterpret_cast<long long *>(E->qctl + 3 * QLINE);
  Q.work = E->qctl + 4 * QLINE;
  Q.stalled = reinterpret_cast<int *>(E->qctl + 5 * QLINE);
  Q.tab = reinterpret_cast<TryOut *>(E->qws);
  Q.slots = reinterpret_cast<unsigned long long *>(E->qws + TRIES_MAX * sizeof(TryOut) * n);
  Q.nslots = 8 * n + SLOT_SLACK;
  Q.tries = reinterpret_cast<int *>(Q.slots + Q.nslots);
  Q.spawned = Q.tries + n;
  Q.resolved = Q.spawned + n;
  Q.wait_ticks = E->wait_ticks;
  // automatic: spawning starts once fewer trajectories than resident groups
  // are unresolved, i.e. when groups begin to fall idle
  Q.spec_below = E->spec_below >= 0 ? E->spec_below : (long long)(blocks * groups_per_block);
  hipLaunchKernelGGL(queue_init, dim3(256), dim3(256), 0, s, Q, n);
  HIP_OK(hipGetLastError());
  Tgt T = E->T;
  void *args[] = {&T, &u_dev, &n, &rec_dev, &Q};
  HIP_OK(hipLaunchKernel((const void *)fn, dim3((unsigned)blocks), dim3(E->block), args, E->shmem, s));
  E->last_stream = s;
}

// Throws when a group of the last run gave up waiting for work (its records
// are then incomplete).
static void check_stall(Engine *E) {
  unsigned long long c = 0;
  HIP_OK(hipMemcpy(&c, E->qctl + 5 * QLINE, sizeof c, hipMemcpyDeviceToHost));
  if ((int)(c & 0xffffffffu) != 0)
    throw Error(MCCS_ERR_DEVICE, "trajectory work queue stalled; records are incomplete");
}

void engine_last_work(Engine *E, uint64_t out[5]) {
  HIP_OK(hipSetDevice(E->device));
  HIP_OK(hipStreamSynchronize(E->last_stream ? E->last_stream : E->stream));
  check_stall(E);
  HIP_OK(hipMemcpy(out, E->qctl + 4 * QLINE, sizeof(uint64_t) * 5, hipMemcpyDeviceToHost));
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
    check_stall(E);
  } catch (...) {
    hipFree(du);
    hipFree(dr);
    throw;
  }
  hipFree(du);
  hipFree(dr);
}

void engine_eval_sites(Engine *E, int site, const double *xyz, uint64_t n, double *f, double *U) {
  if (n == 0) return;
  HIP_OK(hipSetDevice(E->device));
  const int ft = E->ft, gk = E->gk;
  bool lj = false, cen = false, coul = false, dip = false, cut = false;
  double q = 0.0;
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
  std::vector<double> hf(3 * n, 0.0), hu(n, 0.0);
  if (!(lj || coul || dip)) {  // N2 dummy of force types 1/2: no interaction
    std::memcpy(f, hf.data(), sizeof(double) * 3 * n);
    std::memcpy(U, hu.data(), sizeof(double) * n);
    return;
  }
  EvalFn fn = select_eval(lj, cen, coul, dip, cut, E->lds);
  if (!fn) throw Error(MCCS_ERR_UNSUPPORTED, "no site kernel for this role");
  const int block = 256;
  const size_t shmem = lds_bytes(E->T.natoms, E->T.list_cap, block, 64, E->lds, E->gk == 3);
  set_shmem((const void *)fn, shmem);
  double *dx = nullptr, *df = nullptr, *du = nullptr;
  HIP_OK(hipMalloc(&dx, sizeof(double) * 3 * n));
  HIP_OK(hipMalloc(&df, sizeof(double) * 3 * n));
  HIP_OK(hipMalloc(&du, sizeof(double) * n));
  HIP_OK(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, E->stream));
  const uint64_t gpb = block / 64;
  const uint64_t blocks = std::min<uint64_t>((n + gpb - 1) / gpb, (uint64_t)E->cus * 8);
  Tgt T = E->T;
  void *args[] = {&T, &dx, &n, &q, &df, &du};
  hipError_t err = hipLaunchKernel((const void *)fn, dim3((unsigned)blocks), dim3(block), args, shmem,
                                   E->stream);
  if (err == hipSuccess) {
    hipMemcpyAsync(f, df, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, E->stream);
    hipMemcpyAsync(U, du, sizeof(double) * n, hipMemcpyDeviceToHost, E->stream);
    err = hipStreamSynchronize(E->stream);
  }
  hipFree(dx);
  hipFree(df);
  hipFree(du);
  if (err != hipSuccess) throw Error(MCCS_ERR_DEVICE, std::string("site kernel: ") + hipGetErrorString(err));
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

}  // namespace mccs
