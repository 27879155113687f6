// Fourier-feature MLP of the far-field surrogate: BF16 MFMA forward,
// backward and Adam kernels with FP32 accumulation (gfx950).
//
// Matrix products use v_mfma_f32_16x16x16_bf16: one wavefront computes a
// 32 x 32 output tile as 2 x 2 MFMA tiles; a block of four wavefronts covers
// 64 x 64.  Operand fragments (lane l): A[m = l % 16][k = 4 (l / 16) + e],
// B[k = 4 (l / 16) + e][n = l % 16], e = 0..3; result D[m = 4 (l / 16) + e]
// [n = l % 16].  Products whose reduction runs over the batch (weight
// gradients) are split along k over blocks; each block writes its partial
// product and the partials are summed in k order, so that training gives the
// same weights on every run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "host.hpp"
#include "mlp.hpp"

#define MLP_OK(x)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess)                                                                     \
      throw ::mccs::Error(MCCS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));   \
  } while (0)

namespace mccs {
namespace mlpdev {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr float TWO_PI = 6.283185307179586f;

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct DSpec {
  float freq[MLP_NF][3];
  float c[3];
  float s;
};

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// 64 features per point: column f = sin(2 pi B_f . r'), column 32 + f = cos.
__global__ void features_kernel(const float *__restrict__ xyz, int n, DSpec sp,
                                uint16_t *__restrict__ X0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = i / MLP_NF, f = i % MLP_NF;
  if (p >= n) return;
  const float x = (xyz[3 * p] - sp.c[0]) * sp.s;
  const float y = (xyz[3 * p + 1] - sp.c[1]) * sp.s;
  const float z = (xyz[3 * p + 2] - sp.c[2]) * sp.s;
  const float a = TWO_PI * (sp.freq[f][0] * x + sp.freq[f][1] * y + sp.freq[f][2] * z);
  float sv, cv;
  sincosf(a, &sv, &cv);
  X0[(size_t)p * MLP_IN + f] = to_bf16(sv);
  X0[(size_t)p * MLP_IN + MLP_NF + f] = to_bf16(cv);
}

// C (M x N, FP32) = op(A) (M x K) * op(B) (K x N), BF16 operands.
// op(A)[m][k] = TA ? A[k lda + m] : A[m lda + k]; op(B)[k][n] = TB ? B[n ldb + k] : B[k ldb + n].
// blockIdx.z selects a k-chunk of kchunk (multiple of 16); SPLIT writes the
// chunk's product to slice z of C (slices of M ldc floats).
template <bool TA, bool TB, bool SPLIT>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const uint16_t *__restrict__ A,
                                                   int lda, const uint16_t *__restrict__ B, int ldb,
                                                   float *__restrict__ C, int ldc, int kchunk) {
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m0 = blockIdx.y * 64 + (w / 2) * 32;
  const int n0 = blockIdx.x * 64 + (w % 2) * 32;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int r16 = lane % 16, q = lane / 16;
  if (SPLIT) C += (size_t)blockIdx.z * M * ldc;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k = kb; k < ke; k += 16) {
    s4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int m = m0 + 16 * i + r16;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int kk = k + 4 * q + e;
        uint16_t v = 0;
        if (m < M && kk < ke) v = TA ? A[(size_t)kk * lda + m] : A[(size_t)m * lda + kk];
        a[i][e] = (short)v;
      }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int nn = n0 + 16 * j + r16;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int kk = k + 4 * q + e;
        uint16_t v = 0;
        if (nn < N && kk < ke) v = TB ? B[(size_t)nn * ldb + kk] : B[(size_t)kk * ldb + nn];
        b[j][e] = (short)v;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int m = m0 + 16 * i + 4 * q + e;
        const int nn = n0 + 16 * j + r16;
        if (m < M && nn < N) C[(size_t)m * ldc + nn] = acc[i][j][e];
      }
}

// out[i] += the sum of part[z][i] over z = 0 .. nz - 1, in that order.
__global__ void splitsum_kernel(const float *__restrict__ part, int nz, int n, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nz; z++) s += part[(size_t)z * n + i];
  out[i] += s;
}

// Z += bias; H = bf16(SiLU(Z)).
__global__ void bias_silu_kernel(float *__restrict__ Z, const float *__restrict__ bias, int n,
                                 uint16_t *__restrict__ H) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * MLP_H) return;
  const float z = Z[i] + bias[i % MLP_H];
  Z[i] = z;
  H[i] = to_bf16(z * sigm(z));
}

// Y = H3 W4 + b4 (BF16 operands, FP32 sums).
__global__ void out_layer_kernel(const uint16_t *__restrict__ H, const uint16_t *__restrict__ W4,
                                 const float *__restrict__ b4, int n, float *__restrict__ Y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float y0 = 0.f, y1 = 0.f;
  for (int j = 0; j < MLP_H; j++) {
    const float h = from_bf16(H[(size_t)p * MLP_H + j]);
    y0 += h * from_bf16(W4[2 * j]);
    y1 += h * from_bf16(W4[2 * j + 1]);
  }
  Y[2 * p] = y0 + b4[0];
  Y[2 * p + 1] = y1 + b4[1];
}

// dY = d(mean squared error)/dY; loss accumulated in double.
__global__ void loss_kernel(const float *__restrict__ Y, const float *__restrict__ E,
                            const int *__restrict__ idx, int n, float inv, float *__restrict__ dY,
                            double *__restrict__ loss) {
  __shared__ float red[256];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (p < n) {
    const int g = idx ? idx[p] : p;
    const float d0 = Y[2 * p] - E[2 * g], d1 = Y[2 * p + 1] - E[2 * g + 1];
    dY[2 * p] = 2.f * d0 * inv;
    dY[2 * p + 1] = 2.f * d1 * inv;
    l = (d0 * d0 + d1 * d1) * inv;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
}

// Seeds dY with output o (gradient of one output).
__global__ void seed_kernel(int n, int o, float *__restrict__ dY) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  dY[2 * p] = o == 0 ? 1.f : 0.f;
  dY[2 * p + 1] = o == 1 ? 1.f : 0.f;
}

// dZ3 = (dY W4^T) * SiLU'(Z3), FP32 and BF16 copies.
__global__ void out_back_kernel(const float *__restrict__ dY, const uint16_t *__restrict__ W4,
                                const float *__restrict__ Z, int n, float *__restrict__ dZf,
                                uint16_t *__restrict__ dZb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * MLP_H) return;
  const size_t p = i / MLP_H;
  const int j = (int)(i % MLP_H);
  const float dh = dY[2 * p] * from_bf16(W4[2 * j]) + dY[2 * p + 1] * from_bf16(W4[2 * j + 1]);
  const float z = Z[i], s = sigm(z);
  const float dz = dh * s * (1.f + z * (1.f - s));
  dZf[i] = dz;
  dZb[i] = to_bf16(dz);
}

// dZ = dH * SiLU'(Z).
__global__ void silu_back_kernel(const float *__restrict__ dH, const float *__restrict__ Z, int n,
                                 float *__restrict__ dZf, uint16_t *__restrict__ dZb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * MLP_H) return;
  const float z = Z[i], s = sigm(z);
  const float dz = dH[i] * s * (1.f + z * (1.f - s));
  dZf[i] = dz;
  dZb[i] = to_bf16(dz);
}

// out[c] += sum over rows of A[r][c] (block per column, rows strided).
__global__ void colsum_kernel(const float *__restrict__ A, int n, int ncol, float *__restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < n; r += blockDim.x) s += A[(size_t)r * ncol + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out + c, red[0]);
}

// gW4[j][o] += sum_p H3[p][j] dY[p][o]; gb4[o] += sum_p dY[p][o].
__global__ void w4grad_kernel(const uint16_t *__restrict__ H, const float *__restrict__ dY, int n,
                              float *__restrict__ gW4, float *__restrict__ gb4) {
  __shared__ float r0[256], r1[256];
  const int j = blockIdx.x;  // 0..MLP_H (MLP_H: bias)
  float s0 = 0.f, s1 = 0.f;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const float h = j < MLP_H ? from_bf16(H[(size_t)p * MLP_H + j]) : 1.f;
    s0 += h * dY[2 * p];
    s1 += h * dY[2 * p + 1];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      r0[threadIdx.x] += r0[threadIdx.x + k];
      r1[threadIdx.x] += r1[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (j < MLP_H) {
      gW4[2 * j] += r0[0];
      gW4[2 * j + 1] += r1[0];
    } else {
      gb4[0] += r0[0];
      gb4[1] += r1[0];
    }
  }
}

// dE/dr from dX0 (features) through the Fourier map.
__global__ void input_grad_kernel(const float *__restrict__ xyz, const float *__restrict__ dX, int n,
                                  DSpec sp, float scale, float *__restrict__ g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float x = (xyz[3 * p] - sp.c[0]) * sp.s;
  const float y = (xyz[3 * p + 1] - sp.c[1]) * sp.s;
  const float z = (xyz[3 * p + 2] - sp.c[2]) * sp.s;
  float gx = 0.f, gy = 0.f, gz = 0.f;
  for (int f = 0; f < MLP_NF; f++) {
    const float a = TWO_PI * (sp.freq[f][0] * x + sp.freq[f][1] * y + sp.freq[f][2] * z);
    float sv, cv;
    sincosf(a, &sv, &cv);
    const float d = (dX[(size_t)p * MLP_IN + f] * cv - dX[(size_t)p * MLP_IN + MLP_NF + f] * sv) *
                    TWO_PI * sp.s;
    gx += d * sp.freq[f][0];
    gy += d * sp.freq[f][1];
    gz += d * sp.freq[f][2];
  }
  g[3 * p] = gx * scale;
  g[3 * p + 1] = gy * scale;
  g[3 * p + 2] = gz * scale;
}

// Adam step on the flat parameters, then their BF16 copies.
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, uint16_t *__restrict__ wb, int np, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = to_bf16(wi);
  g[i] = 0.f;
}

__global__ void to_bf16_kernel(const float *__restrict__ w, int np, uint16_t *__restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) wb[i] = to_bf16(w[i]);
}

__global__ void gather_kernel(const float *__restrict__ xyz, const int *__restrict__ idx, int n,
                              float *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int g = idx[p];
  out[3 * p] = xyz[3 * g];
  out[3 * p + 1] = xyz[3 * g + 1];
  out[3 * p + 2] = xyz[3 * g + 2];
}

// Rows [0, B): the points; [B, 2B): point + h dir; [2B, 3B): point - h dir.
__global__ void sobolev_pts_kernel(const float *__restrict__ xyz, const float *__restrict__ dir,
                                   const int *__restrict__ idx, int B, float h, float *__restrict__ pts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= B) return;
  const int g = idx[p];
  for (int d = 0; d < 3; d++) {
    const float r = xyz[3 * g + d], u = dir[3 * g + d];
    pts[3 * p + d] = r;
    pts[3 * (B + p) + d] = r + h * u;
    pts[3 * (2 * B + p) + d] = r - h * u;
  }
}

// Energy error at the points plus w times the error of the central-difference
// directional derivative; seeds dY of all 3B rows.
__global__ void sobolev_loss_kernel(const float *__restrict__ Y, const float *__restrict__ E,
                                    const float *__restrict__ T, const int *__restrict__ idx, int B,
                                    float inv2h, float inv, float w, float *__restrict__ dY,
                                    double *__restrict__ loss) {
  __shared__ float red[256];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (p < B) {
    const int g = idx[p];
    for (int o = 0; o < 2; o++) {
      const float re = Y[2 * p + o] - E[2 * g + o];
      const float D = (Y[2 * (B + p) + o] - Y[2 * (2 * B + p) + o]) * inv2h;
      const float rd = D - T[2 * g + o];
      dY[2 * p + o] = 2.f * re * inv;
      dY[2 * (B + p) + o] = 2.f * w * rd * inv2h * inv;
      dY[2 * (2 * B + p) + o] = -2.f * w * rd * inv2h * inv;
      l += (re * re + w * rd * rd) * inv;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
}

__global__ void scale_out_kernel(const float *__restrict__ Y, int n, float s0, float s1,
                                 float *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  out[2 * p] = Y[2 * p] * s0;
  out[2 * p + 1] = Y[2 * p + 1] * s1;
}

}  // namespace mlpdev

using namespace mlpdev;

struct Mlp {
  int device = 0, maxb = 0;
  MlpSpec spec{};
  DSpec ds{};
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wb = nullptr;
  uint16_t *X0 = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dZb = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *Y = nullptr, *dY = nullptr, *dH = nullptr,
        *dZf = nullptr, *dX = nullptr, *pts = nullptr, *grd = nullptr;
  uint16_t *dZ1b = nullptr, *dZ2b = nullptr, *dZ3b = nullptr;
  int *idx = nullptr;
  double *loss = nullptr;
  float *part = nullptr;  // partial products of the split weight-gradient products
  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    MLP_OK(hipMalloc(&p, std::max<size_t>(sizeof(T) * count, 16)));
    allocs.push_back(p);
    return reinterpret_cast<T *>(p);
  }
  ~Mlp() {
    hipSetDevice(device);
    for (void *p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

unsigned blocks_for(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

constexpr int MAX_SPLIT = 64;  // k-chunks of a split product

// C = op(A) op(B).  Split (part != null): C += op(A) op(B), the k range cut
// into at most MAX_SPLIT chunks whose products go to part (MAX_SPLIT M N
// floats) and are summed in chunk order; C then has ldc == N.
template <bool TA, bool TB>
void gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
          float *C, int ldc, int splitk = 1, float *part = nullptr) {
  int kchunk = (K + splitk - 1) / splitk;
  kchunk = (kchunk + 15) / 16 * 16;
  splitk = (K + kchunk - 1) / kchunk;
  if (!part) {
    dim3 grid((N + 63) / 64, (M + 63) / 64, 1);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb,
                       C, ldc, K);
    return;
  }
  if (splitk > MAX_SPLIT || ldc != N) throw Error(MCCS_ERR_INVALID, "split product outside its workspace");
  dim3 grid((N + 63) / 64, (M + 63) / 64, splitk);
  hipLaunchKernelGGL((gemm_kernel<TA, TB, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb,
                     part, N, kchunk);
  hipLaunchKernelGGL(splitsum_kernel, dim3(blocks_for((size_t)M * N)), dim3(256), 0, s,
                     (const float *)part, splitk, M * N, C);
}

// Forward over n <= maxb points already in m->pts: leaves Z*, H*, Y.
void forward(Mlp *m, int n) {
  hipStream_t s = m->stream;
  hipLaunchKernelGGL(features_kernel, dim3(blocks_for((size_t)n * MLP_NF)), dim3(256), 0, s,
                     m->pts, n, m->ds, m->X0);
  gemm<false, false>(s, n, MLP_H, MLP_IN, m->X0, MLP_IN, m->wb + MLP_W1, MLP_H, m->Z1, MLP_H);
  hipLaunchKernelGGL(bias_silu_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->Z1,
                     m->w + MLP_B1, n, m->H1);
  gemm<false, false>(s, n, MLP_H, MLP_H, m->H1, MLP_H, m->wb + MLP_W2, MLP_H, m->Z2, MLP_H);
  hipLaunchKernelGGL(bias_silu_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->Z2,
                     m->w + MLP_B2, n, m->H2);
  gemm<false, false>(s, n, MLP_H, MLP_H, m->H2, MLP_H, m->wb + MLP_W3, MLP_H, m->Z3, MLP_H);
  hipLaunchKernelGGL(bias_silu_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->Z3,
                     m->w + MLP_B3, n, m->H3);
  hipLaunchKernelGGL(out_layer_kernel, dim3(blocks_for(n)), dim3(256), 0, s, m->H3, m->wb + MLP_W4,
                     m->w + MLP_B4, n, m->Y);
}

// Backward from m->dY; weight gradients accumulated into m->g when wgrad,
// input-feature gradient into m->dX when xgrad.
void backward(Mlp *m, int n, bool wgrad, bool xgrad) {
  hipStream_t s = m->stream;
  const int sk = std::max(1, std::min(64, n / 256));
  if (wgrad)
    hipLaunchKernelGGL(w4grad_kernel, dim3(MLP_H + 1), dim3(256), 0, s, m->H3, m->dY, n,
                       m->g + MLP_W4, m->g + MLP_B4);
  hipLaunchKernelGGL(out_back_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->dY,
                     m->wb + MLP_W4, m->Z3, n, m->dZf, m->dZ3b);
  if (wgrad) {
    hipLaunchKernelGGL(colsum_kernel, dim3(MLP_H), dim3(256), 0, s, m->dZf, n, MLP_H, m->g + MLP_B3);
    gemm<true, false>(s, MLP_H, MLP_H, n, m->H2, MLP_H, m->dZ3b, MLP_H, m->g + MLP_W3, MLP_H, sk, m->part);
  }
  gemm<false, true>(s, n, MLP_H, MLP_H, m->dZ3b, MLP_H, m->wb + MLP_W3, MLP_H, m->dH, MLP_H);
  hipLaunchKernelGGL(silu_back_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->dH,
                     m->Z2, n, m->dZf, m->dZ2b);
  if (wgrad) {
    hipLaunchKernelGGL(colsum_kernel, dim3(MLP_H), dim3(256), 0, s, m->dZf, n, MLP_H, m->g + MLP_B2);
    gemm<true, false>(s, MLP_H, MLP_H, n, m->H1, MLP_H, m->dZ2b, MLP_H, m->g + MLP_W2, MLP_H, sk, m->part);
  }
  gemm<false, true>(s, n, MLP_H, MLP_H, m->dZ2b, MLP_H, m->wb + MLP_W2, MLP_H, m->dH, MLP_H);
  hipLaunchKernelGGL(silu_back_kernel, dim3(blocks_for((size_t)n * MLP_H)), dim3(256), 0, s, m->dH,
                     m->Z1, n, m->dZf, m->dZ1b);
  if (wgrad) {
    hipLaunchKernelGGL(colsum_kernel, dim3(MLP_H), dim3(256), 0, s, m->dZf, n, MLP_H, m->g + MLP_B1);
    gemm<true, false>(s, MLP_IN, MLP_H, n, m->X0, MLP_IN, m->dZ1b, MLP_H, m->g + MLP_W1, MLP_H, sk, m->part);
  }
  if (xgrad)
    gemm<false, true>(s, n, MLP_IN, MLP_H, m->dZ1b, MLP_H, m->wb + MLP_W1, MLP_H, m->dX, MLP_IN);
}

DSpec device_spec(const MlpSpec &sp) {
  DSpec d{};
  for (int f = 0; f < MLP_NF; f++)
    for (int k = 0; k < 3; k++) d.freq[f][k] = sp.freq[f][k];
  for (int k = 0; k < 3; k++) d.c[k] = (float)sp.centre[k];
  d.s = (float)sp.in_scale;
  return d;
}

}  // namespace

Mlp *mlp_create(int device, const MlpSpec &spec, const float *params, int max_batch) {
  std::unique_ptr<Mlp> m(new Mlp);
  m->device = device;
  MLP_OK(hipSetDevice(device));
  MLP_OK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  m->spec = spec;
  m->ds = device_spec(spec);
  m->maxb = max_batch;
  const size_t B = (size_t)max_batch;
  m->w = m->alloc<float>(MLP_NPARAM);
  m->g = m->alloc<float>(MLP_NPARAM);
  m->m = m->alloc<float>(MLP_NPARAM);
  m->v = m->alloc<float>(MLP_NPARAM);
  m->wb = m->alloc<uint16_t>(MLP_NPARAM);
  m->X0 = m->alloc<uint16_t>(B * MLP_IN);
  m->H1 = m->alloc<uint16_t>(B * MLP_H);
  m->H2 = m->alloc<uint16_t>(B * MLP_H);
  m->H3 = m->alloc<uint16_t>(B * MLP_H);
  m->dZ1b = m->alloc<uint16_t>(B * MLP_H);
  m->dZ2b = m->alloc<uint16_t>(B * MLP_H);
  m->dZ3b = m->alloc<uint16_t>(B * MLP_H);
  m->Z1 = m->alloc<float>(B * MLP_H);
  m->Z2 = m->alloc<float>(B * MLP_H);
  m->Z3 = m->alloc<float>(B * MLP_H);
  m->dH = m->alloc<float>(B * MLP_H);
  m->dZf = m->alloc<float>(B * MLP_H);
  m->dX = m->alloc<float>(B * MLP_IN);
  m->Y = m->alloc<float>(B * MLP_OUT);
  m->dY = m->alloc<float>(B * MLP_OUT);
  m->pts = m->alloc<float>(B * 3);
  m->grd = m->alloc<float>(B * 3);
  m->idx = m->alloc<int>(B);
  m->loss = m->alloc<double>(1);
  m->part = m->alloc<float>((size_t)MAX_SPLIT * MLP_H * MLP_H);
  MLP_OK(hipMemcpy(m->w, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  MLP_OK(hipMemset(m->g, 0, sizeof(float) * MLP_NPARAM));
  MLP_OK(hipMemset(m->m, 0, sizeof(float) * MLP_NPARAM));
  MLP_OK(hipMemset(m->v, 0, sizeof(float) * MLP_NPARAM));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(MLP_NPARAM)), dim3(256), 0, m->stream, m->w,
                     (int)MLP_NPARAM, m->wb);
  MLP_OK(hipStreamSynchronize(m->stream));
  return m.release();
}

void mlp_destroy(Mlp *m) { delete m; }

void mlp_get_params(const Mlp *m, float *params) {
  MLP_OK(hipSetDevice(m->device));
  MLP_OK(hipMemcpy(params, m->w, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
}

void mlp_eval(Mlp *m, const double *xyz, uint64_t n, double *y, double *dy) {
  MLP_OK(hipSetDevice(m->device));
  std::vector<float> p(3 * (size_t)m->maxb), Y(2 * (size_t)m->maxb), G(3 * (size_t)m->maxb);
  for (uint64_t b0 = 0; b0 < n; b0 += m->maxb) {
    const int nb = (int)std::min<uint64_t>(m->maxb, n - b0);
    for (int i = 0; i < 3 * nb; i++) p[i] = (float)xyz[3 * b0 + i];
    MLP_OK(hipMemcpyAsync(m->pts, p.data(), sizeof(float) * 3 * nb, hipMemcpyHostToDevice, m->stream));
    forward(m, nb);
    MLP_OK(hipMemcpyAsync(Y.data(), m->Y, sizeof(float) * 2 * nb, hipMemcpyDeviceToHost, m->stream));
    MLP_OK(hipStreamSynchronize(m->stream));
    for (int i = 0; i < nb; i++)
      for (int o = 0; o < 2; o++) y[2 * (b0 + i) + o] = (double)Y[2 * i + o] * m->spec.out_scale[o];
    if (!dy) continue;
    for (int o = 0; o < 2; o++) {
      hipLaunchKernelGGL(seed_kernel, dim3(blocks_for(nb)), dim3(256), 0, m->stream, nb, o, m->dY);
      backward(m, nb, false, true);
      hipLaunchKernelGGL(input_grad_kernel, dim3(blocks_for(nb)), dim3(256), 0, m->stream, m->pts,
                         m->dX, nb, m->ds, (float)m->spec.out_scale[o], m->grd);
      MLP_OK(hipMemcpyAsync(G.data(), m->grd, sizeof(float) * 3 * nb, hipMemcpyDeviceToHost, m->stream));
      MLP_OK(hipStreamSynchronize(m->stream));
      for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++) dy[6 * (b0 + i) + 3 * o + d] = G[3 * i + d];
    }
  }
}

double mlp_loss_grad(Mlp *m, const double *xyz, const double *e, uint64_t n, float *grad) {
  MLP_OK(hipSetDevice(m->device));
  if (n == 0 || n > (uint64_t)m->maxb) throw Error(MCCS_ERR_INVALID, "batch larger than the network's workspace");
  std::vector<float> p(3 * n), E(2 * n);
  for (uint64_t i = 0; i < 3 * n; i++) p[i] = (float)xyz[i];
  for (uint64_t i = 0; i < n; i++)
    for (int o = 0; o < 2; o++) E[2 * i + o] = (float)(e[2 * i + o] / m->spec.out_scale[o]);
  float *dE = nullptr;
  MLP_OK(hipMalloc(&dE, sizeof(float) * 2 * n));
  double loss = 0.0;
  try {
    MLP_OK(hipMemcpyAsync(m->pts, p.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemcpyAsync(dE, E.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
    MLP_OK(hipMemsetAsync(m->g, 0, sizeof(float) * MLP_NPARAM, m->stream));
    forward(m, (int)n);
    hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(n)), dim3(256), 0, m->stream, m->Y, dE,
                       (const int *)nullptr, (int)n, 1.0f / (2.0f * n), m->dY, m->loss);
    backward(m, (int)n, true, false);
    MLP_OK(hipMemcpyAsync(grad, m->g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_OK(hipMemcpyAsync(&loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_OK(hipStreamSynchronize(m->stream));
    MLP_OK(hipMemsetAsync(m->g, 0, sizeof(float) * MLP_NPARAM, m->stream));
    MLP_OK(hipStreamSynchronize(m->stream));
  } catch (...) {
    hipFree(dE);
    throw;
  }
  hipFree(dE);
  return loss;
}

double mlp_train(Mlp *m, const float *d_xyz, const float *d_e, uint64_t n, const TrainOpts &o,
                 std::vector<double> *history) {
  MLP_OK(hipSetDevice(m->device));
  const bool sob = o.d_dir && o.d_deriv;
  const int B = std::min(o.batch, sob ? m->maxb / 3 : m->maxb);
  std::mt19937_64 rng(o.seed);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  int *d_perm = nullptr;
  MLP_OK(hipMalloc(&d_perm, sizeof(int) * n));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long step = 0;
  double last = 0.0;
  const uint64_t nb = (n + B - 1) / B;
  const double total = (double)o.epochs * nb;
  try {
    for (int ep = 0; ep < o.epochs; ep++) {
      std::shuffle(perm.begin(), perm.end(), rng);
      MLP_OK(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->stream));
      MLP_OK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
      for (uint64_t b = 0; b < nb; b++) {
        const int cnt = (int)std::min<uint64_t>(B, n - b * B);
        const int *ix = d_perm + b * B;
        if (sob) {
          hipLaunchKernelGGL(sobolev_pts_kernel, dim3(blocks_for(cnt)), dim3(256), 0, m->stream, d_xyz,
                             o.d_dir, ix, cnt, (float)o.fd_h, m->pts);
          forward(m, 3 * cnt);
          hipLaunchKernelGGL(sobolev_loss_kernel, dim3(blocks_for(cnt)), dim3(256), 0, m->stream, m->Y,
                             d_e, o.d_deriv, ix, cnt, (float)(0.5 / o.fd_h), 1.0f / (2.0f * cnt),
                             (float)o.w_deriv, m->dY, m->loss);
          backward(m, 3 * cnt, true, false);
        } else {
          hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(cnt)), dim3(256), 0, m->stream, d_xyz, ix,
                             cnt, m->pts);
          forward(m, cnt);
          hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(cnt)), dim3(256), 0, m->stream, m->Y, d_e, ix,
                             cnt, 1.0f / (2.0f * cnt), m->dY, m->loss);
          backward(m, cnt, true, false);
        }
        step++;
        // cosine decay from lr to lr_final
        const double t = step / total;
        const double lr = o.lr_final + 0.5 * (o.lr - o.lr_final) * (1.0 + std::cos(M_PI * t));
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)step)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)step)));
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(MLP_NPARAM)), dim3(256), 0, m->stream, m->w,
                           m->g, m->m, m->v, m->wb, (int)MLP_NPARAM, (float)lr, b1, b2, c1, c2, eps);
      }
      double l = 0.0;
      MLP_OK(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
      MLP_OK(hipStreamSynchronize(m->stream));
      last = l / (double)nb;
      if (history) history->push_back(last);
    }
  } catch (...) {
    hipFree(d_perm);
    throw;
  }
  hipFree(d_perm);
  return last;
}

double mlp_train_host(Mlp *m, const double *xyz, const double *e, uint64_t n, const TrainOpts &o,
                      std::vector<double> *history) {
  MLP_OK(hipSetDevice(m->device));
  std::vector<float> p(3 * n), t(2 * n);
  for (uint64_t i = 0; i < 3 * n; i++) p[i] = (float)xyz[i];
  for (uint64_t i = 0; i < n; i++)
    for (int k = 0; k < 2; k++) t[2 * i + k] = (float)(e[2 * i + k] / m->spec.out_scale[k]);
  float *dp = nullptr, *dt = nullptr;
  if (hipMalloc(&dp, sizeof(float) * 3 * n) != hipSuccess || hipMalloc(&dt, sizeof(float) * 2 * n) != hipSuccess) {
    hipFree(dp);
    throw Error(MCCS_ERR_NOMEM, "device allocation of training data failed");
  }
  double last = 0.0;
  try {
    MLP_OK(hipMemcpy(dp, p.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice));
    MLP_OK(hipMemcpy(dt, t.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    last = mlp_train(m, dp, dt, n, o, history);
  } catch (...) {
    hipFree(dp);
    hipFree(dt);
    throw;
  }
  hipFree(dp);
  hipFree(dt);
  return last;
}

void mlp_eval_device(Mlp *m, const float *d_xyz, uint64_t n, float *d_y, void *stream) {
  MLP_OK(hipSetDevice(m->device));
  for (uint64_t b0 = 0; b0 < n; b0 += m->maxb) {
    const int nb = (int)std::min<uint64_t>(m->maxb, n - b0);
    MLP_OK(hipMemcpyAsync(m->pts, d_xyz + 3 * b0, sizeof(float) * 3 * nb, hipMemcpyDeviceToDevice,
                          m->stream));
    forward(m, nb);
    hipLaunchKernelGGL(scale_out_kernel, dim3(blocks_for(nb)), dim3(256), 0, m->stream, m->Y, nb,
                       (float)m->spec.out_scale[0], (float)m->spec.out_scale[1], d_y + 2 * b0);
  }
  MLP_OK(hipStreamSynchronize(m->stream));
  (void)stream;
}

}  // namespace mccs
