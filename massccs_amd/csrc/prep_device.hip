// Target preparation on the GPU for very large targets (10^6 atoms and
// more): the linked-cell binning and the stable reorder of the atoms by cell
// (src/LinkedCell.cpp:93-225), and the centre of mass and inertia tensor of
// the principal-axis step (src/MoleculeTarget.cpp:701-805).  The host
// preparation (host_target.cpp) stays authoritative: the cell order computed
// here is the same integer result; the FP64 sums of the inertia tensor are
// reduced in another order (agreement to rounding).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "host.hpp"
#include "massccs.h"

#define PREP_OK(x)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess)                                                                     \
      throw ::mccs::Error(MCCS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace mccs {
namespace {

// Flat cell index of every atom, LinkedCell::calculateIndex
// (src/LinkedCell.cpp:117-129): i = floor((x - corner) / cell), ...,
// index = k + nz j + nz ny i; atoms outside the grid are flagged.
__global__ void cell_index_kernel(const double *__restrict__ xyz, int64_t n, double cell, double cx,
                                  double cy, double cz, int nx, int ny, int nz, int *__restrict__ key,
                                  int *__restrict__ bad) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int i = (int)floor((xyz[3 * a] - cx) / cell);
  const int j = (int)floor((xyz[3 * a + 1] - cy) / cell);
  const int k = (int)floor((xyz[3 * a + 2] - cz) / cell);
  const long long c = (long long)k + (long long)nz * j + (long long)nz * ny * i;
  if (c < 0 || c >= (long long)nx * ny * nz) {
    atomicOr(bad, 1);
    key[a] = 0;
    return;
  }
  key[a] = (int)c;
}

__global__ void iota_kernel(int64_t n, int *__restrict__ v) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) v[a] = (int)a;
}

// cell_start[c] = first sorted position with key >= c, for c in [0, ncells]
// (the sorted keys are non-decreasing: a binary search per cell).
__global__ void cell_start_kernel(const int *__restrict__ sorted, int64_t n, int ncells,
                                  int *__restrict__ start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  start[c] = (int)lo;
}

// Per-block partial sums of the centre-of-mass numerators and total mass.
__global__ void com_kernel(const double *__restrict__ xyz, const double *__restrict__ m, int64_t n,
                           double *__restrict__ part) {
  __shared__ double s[4][256];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n;
       a += (int64_t)gridDim.x * blockDim.x) {
    v[0] += m[a] * xyz[3 * a];
    v[1] += m[a] * xyz[3 * a + 1];
    v[2] += m[a] * xyz[3 * a + 2];
    v[3] += m[a];
  }
  for (int q = 0; q < 4; q++) s[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; q++) s[q][threadIdx.x] += s[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) part[4 * blockIdx.x + q] = s[q][0];
}

// Per-block partial sums of the inertia tensor about c,
// I_ab = sum m (r^2 delta_ab - r_a r_b) (src/MoleculeTarget.cpp:701-745).
__global__ void inertia_kernel(const double *__restrict__ xyz, const double *__restrict__ m, int64_t n,
                               double cx, double cy, double cz, double *__restrict__ part) {
  __shared__ double s[6][256];
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // xx, yy, zz, xy, xz, yz
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n;
       a += (int64_t)gridDim.x * blockDim.x) {
    const double x = xyz[3 * a] - cx, y = xyz[3 * a + 1] - cy, z = xyz[3 * a + 2] - cz;
    const double mi = m[a], r2 = x * x + y * y + z * z;
    v[0] += mi * (r2 - x * x);
    v[1] += mi * (r2 - y * y);
    v[2] += mi * (r2 - z * z);
    v[3] -= mi * x * y;
    v[4] -= mi * x * z;
    v[5] -= mi * y * z;
  }
  for (int q = 0; q < 6; q++) s[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 6; q++) s[q][threadIdx.x] += s[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; q++) part[6 * blockIdx.x + q] = s[q][0];
}

template <class T>
struct Buf {
  T *p = nullptr;
  explicit Buf(size_t n) { PREP_OK(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1))); }
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

void device_cell_sort(int device, int64_t n, const double *xyz_dev, double cell, const double corner[3],
                      int nx, int ny, int nz, int *cell_start_dev, int *order_dev, void *stream) {
  PREP_OK(hipSetDevice(device));
  hipStream_t s = (hipStream_t)stream;
  const long long nc = (long long)nx * ny * nz;
  if (n < 0 || n > INT32_MAX || nc <= 0 || nc >= INT32_MAX)
    throw Error(MCCS_ERR_INVALID, "atom or cell count out of range");
  Buf<int> key(n), key_sorted(n), idx(n), bad(1);
  PREP_OK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  if (n) {
    hipLaunchKernelGGL(cell_index_kernel, dim3(blocks_for(n)), dim3(256), 0, s, xyz_dev, n, cell,
                       corner[0], corner[1], corner[2], nx, ny, nz, key.p, bad.p);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, idx.p);
    PREP_OK(hipGetLastError());
    // a stable (LSD radix) sort by cell index keeps the atoms of a cell in
    // their input order: the reference's reorder
    int end_bit = 1;
    while (end_bit < 31 && (1ll << end_bit) < nc) end_bit++;
    size_t tmp_bytes = 0;
    PREP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key_sorted.p, idx.p, order_dev,
                                               (int)n, 0, end_bit, s));
    Buf<char> tmp(tmp_bytes);
    PREP_OK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key_sorted.p, idx.p, order_dev,
                                               (int)n, 0, end_bit, s));
    hipLaunchKernelGGL(cell_start_kernel, dim3(blocks_for(nc + 1)), dim3(256), 0, s,
                       (const int *)key_sorted.p, n, (int)nc, cell_start_dev);
    PREP_OK(hipGetLastError());
    int hbad = 0;
    PREP_OK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PREP_OK(hipStreamSynchronize(s));
    if (hbad) throw Error(MCCS_ERR_INVALID, "target atom outside the linked-cell grid");
  } else {
    PREP_OK(hipMemsetAsync(cell_start_dev, 0, sizeof(int) * (size_t)(nc + 1), s));
    PREP_OK(hipStreamSynchronize(s));
  }
}

void device_inertia(int device, int64_t n, const double *xyz_dev, const double *m_dev, double com[3],
                    double *mass, double inertia[3][3], void *stream) {
  PREP_OK(hipSetDevice(device));
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
  Buf<double> part(6 * (size_t)nb);
  std::vector<double> h(6 * (size_t)nb);
  hipLaunchKernelGGL(com_kernel, dim3(nb), dim3(256), 0, s, xyz_dev, m_dev, n, part.p);
  PREP_OK(hipGetLastError());
  PREP_OK(hipMemcpyAsync(h.data(), part.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, s));
  PREP_OK(hipStreamSynchronize(s));
  double v[4] = {0, 0, 0, 0};
  for (unsigned b = 0; b < nb; b++)
    for (int q = 0; q < 4; q++) v[q] += h[4 * b + q];
  *mass = v[3];
  for (int d = 0; d < 3; d++) com[d] = v[3] != 0.0 ? v[d] / v[3] : 0.0;
  hipLaunchKernelGGL(inertia_kernel, dim3(nb), dim3(256), 0, s, xyz_dev, m_dev, n, com[0], com[1], com[2],
                     part.p);
  PREP_OK(hipGetLastError());
  PREP_OK(hipMemcpyAsync(h.data(), part.p, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, s));
  PREP_OK(hipStreamSynchronize(s));
  double w[6] = {0, 0, 0, 0, 0, 0};
  for (unsigned b = 0; b < nb; b++)
    for (int q = 0; q < 6; q++) w[q] += h[6 * b + q];
  inertia[0][0] = w[0];
  inertia[1][1] = w[1];
  inertia[2][2] = w[2];
  inertia[0][1] = inertia[1][0] = w[3];
  inertia[0][2] = inertia[2][0] = w[4];
  inertia[1][2] = inertia[2][1] = w[5];
}

}  // namespace mccs
