// Fourier-feature MLP of the far-field surrogate: BF16 MFMA forward,
// backward and Adam kernels with FP32 accumulation (gfx950).
//
// Matrix products run on v_mfma_f32_16x16x32_bf16 over LDS-staged tiles
// (gemm_kernel), with the bias + SiLU of a forward layer and the SiLU'
// factor of a backward layer applied in the product's epilogue.  Products
// whose reduction runs over the batch (weight gradients) are split along k
// over blocks; each block writes its partial product and the partials are
// summed in k order, so that training gives the same weights on every run.
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
// SiLU'(z) and SiLU''(z) from sg = sigm(z)
__device__ __forceinline__ float silu1(float z, float sg) { return sg * (1.f + z * (1.f - sg)); }
__device__ __forceinline__ float silu2(float z, float sg) {
  return sg * (1.f - sg) * (2.f + z * (1.f - 2.f * sg));
}

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

// C (M x N, FP32) = op(A) (M x K) * op(B) (K x N), BF16 operands, FP32 sums.
// op(A)[m][k] = TA ? A[k lda + m] : A[m lda + k]; op(B)[k][n] = TB ? B[n ldb + k] : B[k ldb + n].
//
// A block of four wavefronts computes a GT x GT tile of C; each wavefront a
// 32 x 32 quarter as 2 x 2 tiles of v_mfma_f32_16x16x32_bf16.  The k range
// is staged through LDS GK at a time, both operands stored k-contiguous
// (As[m][k], Bs[n][k], rows padded to GP elements so that the 16 rows one
// fragment read touches start in distinct bank groups); every global load
// is one 16-byte vector of 8 BF16 along the operand's contiguous axis, and
// the next stage's loads are in flight while the current stage multiplies.
// The MFMA computes the transposed tile (B^T A^T) so that each lane ends
// with four consecutive columns of one row and stores them as one vector.
// Fragments (lane l, q = l / 16): first operand B[k = 8 q + e][n = l % 16],
// second A[m = l % 16][k = 8 q + e], e = 0..7; result C[m = l % 16][n = 4 q
// + e], e = 0..3.
//
// blockIdx.z selects a k-chunk of kchunk (a multiple of GK); SPLIT writes the
// chunk's product to slice z of C (slices of M ldc floats).  Epilogues:
// EPI_STORE C = product; EPI_BIAS_SILU (forward layer) C = Z = product +
// bias[n] and H = bf16(SiLU(Z)); EPI_SILU_BACK (backward layer) H = bf16(dZ),
// dZ = product * SiLU'(Zin), and C receives the block's column sums of dZ
// over its GT rows: C[blockIdx.y][n] (the partials of the bias gradient).
// Tangent rows (forward-mode derivative): EPI_TAN C = product and H =
// bf16(product * SiLU'(Zin)), Zin the value rows' pre-activations (forward:
// dZ -> dH; backward: the tangent's upstream gradient -> its dZ);
// EPI_SILU2_BACK as EPI_SILU_BACK with dZ = product * SiLU'(Zin) + Gt *
// SiLU''(Zin) * Zt (Zt the tangent pre-activations, Gt the tangent's
// upstream gradient), for the value rows of a tangent-carrying backward.
constexpr int GT = 64, GK = 64, GP = GK + 8;
enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_SILU_BACK = 2, EPI_TAN = 3, EPI_SILU2_BACK = 4 };

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// One 8-element chunk of an operand tile, read along the operand's
// contiguous axis: rows [row0, row0 + rows) x contiguous [c0, c0 + 8) of a
// matrix with leading dimension ld whose valid extent is (row_end, c_end).
__device__ __forceinline__ u16x8 load_chunk(const uint16_t *__restrict__ X, int ld, int row, int row_end,
                                            int c, int c_end) {
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (row >= row_end) return v;
  const uint16_t *p = X + (size_t)row * ld + c;
  if (c + 8 <= c_end) return *reinterpret_cast<const u16x8 *>(p);
#pragma unroll
  for (int e = 0; e < 8; e++)
    if (c + e < c_end) v[e] = p[e];
  return v;
}

// The GT x GK tile of an operand at (outer0, k0) into registers (two chunks
// per thread).  KC: the operand is k-contiguous in memory (rows = outer
// index); otherwise outer-contiguous (rows = k).
template <bool KC>
__device__ __forceinline__ void load_tile(const uint16_t *__restrict__ X, int ld, int outer0, int outer_end,
                                          int k0, int k_end, u16x8 r[2]) {
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = threadIdx.x + 256 * h;
    const int row = c >> 3, col = (c & 7) * 8;
    if (KC) r[h] = load_chunk(X, ld, outer0 + row, outer_end, k0 + col, k_end);
    else r[h] = load_chunk(X, ld, k0 + row, k_end, outer0 + col, outer_end);
  }
}

// Registers of load_tile into the k-contiguous LDS tile S[outer][k].
template <bool KC>
__device__ __forceinline__ void store_tile(uint16_t *S, const u16x8 r[2]) {
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = threadIdx.x + 256 * h;
    const int row = c >> 3, col = (c & 7) * 8;
    if (KC) {
      *reinterpret_cast<u16x8 *>(S + row * GP + col) = r[h];
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) S[(col + e) * GP + row] = r[h][e];
    }
  }
}

template <bool TA, bool TB, bool SPLIT, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const uint16_t *__restrict__ A,
                                                   int lda, const uint16_t *__restrict__ B, int ldb,
                                                   float *__restrict__ C, int ldc, int kchunk,
                                                   const float *__restrict__ bias,
                                                   const float *__restrict__ Zin,
                                                   uint16_t *__restrict__ H,
                                                   const float *__restrict__ Zt,
                                                   const float *__restrict__ Gt) {
  __shared__ __align__(16) uint16_t As[GT * GP];
  __shared__ __align__(16) uint16_t Bs[GT * GP];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int bm = blockIdx.y * GT, bn = blockIdx.x * GT;
  const int wm = (w / 2) * 32, wn = (w % 2) * 32;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int r16 = lane % 16, q = lane / 16;
  if (SPLIT) C += (size_t)blockIdx.z * M * ldc;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  u16x8 ra[2], rb[2];
  load_tile<!TA>(A, lda, bm, M, kb, ke, ra);
  load_tile<TB>(B, ldb, bn, N, kb, ke, rb);
  for (int k0 = kb; k0 < ke; k0 += GK) {
    __syncthreads();  // the previous stage's fragments are read
    store_tile<!TA>(As, ra);
    store_tile<TB>(Bs, rb);
    __syncthreads();
    if (k0 + GK < ke) {
      load_tile<!TA>(A, lda, bm, M, k0 + GK, ke, ra);
      load_tile<TB>(B, ldb, bn, N, k0 + GK, ke, rb);
    }
#pragma unroll
    for (int ks = 0; ks < GK; ks += 32) {
      bf8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++)
        a[i] = *reinterpret_cast<const bf8 *>(As + (wm + 16 * i + r16) * GP + ks + 8 * q);
#pragma unroll
      for (int j = 0; j < 2; j++)
        b[j] = *reinterpret_cast<const bf8 *>(Bs + (wn + 16 * j + r16) * GP + ks + 8 * q);
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  }
  if (EPI == EPI_SILU_BACK || EPI == EPI_SILU2_BACK) {
    // column sums over the wavefront's 32 rows: the two tiles of a lane,
    // then the 16 lanes of equal q (rows l % 16), read from lane q * 16
    float cs[2][4];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float v = 0.f;
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const int m = bm + wm + 16 * i + r16;
          const int nn = bn + wn + 16 * j + 4 * q + e;
          float dz = 0.f;
          if (m < M && nn < N) {
            const size_t o = (size_t)m * ldc + nn;
            const float z = Zin[o], sg = sigm(z);
            dz = acc[i][j][e] * silu1(z, sg);
            if (EPI == EPI_SILU2_BACK) dz += Gt[o] * silu2(z, sg) * Zt[o];
          }
          acc[i][j][e] = dz;
          v += dz;
        }
        cs[j][e] = v;
      }
#pragma unroll
    for (int sh = 1; sh < 16; sh <<= 1)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) cs[j][e] += __shfl_xor(cs[j][e], sh);
    __shared__ float colp[2][GT];
    if (r16 == 0)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) colp[w / 2][wn + 16 * j + 4 * q + e] = cs[j][e];
    __syncthreads();
    if (threadIdx.x < GT && bn + (int)threadIdx.x < N)
      C[(size_t)blockIdx.y * ldc + bn + threadIdx.x] = colp[0][threadIdx.x] + colp[1][threadIdx.x];
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int m = bm + wm + 16 * i + r16;
      const int n0 = bn + wn + 16 * j + 4 * q;  // N and ldc are multiples of 4
      if (m >= M || n0 >= N) continue;
      const size_t o = (size_t)m * ldc + n0;
      f4 v = acc[i][j];
      if (EPI == EPI_BIAS_SILU) {
        const f4 bb = *reinterpret_cast<const f4 *>(bias + n0);
        v += bb;
        *reinterpret_cast<f4 *>(C + o) = v;
      } else if (EPI == EPI_STORE || EPI == EPI_TAN) {
        *reinterpret_cast<f4 *>(C + o) = v;
      }
      if (EPI != EPI_STORE) {
        uint16_t hv[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          float h = v[e];
          if (EPI == EPI_BIAS_SILU) h = v[e] * sigm(v[e]);
          if (EPI == EPI_TAN) {
            const float z = Zin[o + e];
            h = v[e] * silu1(z, sigm(z));
          }
          hv[e] = to_bf16(h);
        }
        uint2 packed;
        packed.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
        packed.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
        *reinterpret_cast<uint2 *>(H + o) = packed;
      }
    }
}

// out[i] += the sum of part[z][i] over z = 0 .. nz - 1 (slices of width
// floats).  64 threads per output element: thread g sums z = g, g + 64, ...
// in order, and the 64 partial sums are combined by a fixed tree, so the
// result is the same on every run.
__global__ __launch_bounds__(256) void partsum_kernel(const float *__restrict__ part, int nz, int width,
                                                      float *__restrict__ out) {
  __shared__ float red[256];
  const int i = blockIdx.x * 4 + threadIdx.x / 64, g = threadIdx.x % 64;
  float s = 0.f;
  if (i < width)
    for (int z = g; z < nz; z += 64) s += part[(size_t)z * width + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (g < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (g == 0 && i < width) out[i] += red[threadIdx.x];
}

// Per-block partial gradients of the output layer over rows [b rpb, (b + 1)
// rpb): part[b][2 j + o] = sum_p H3[p][j] dY[p][o], part[b][2 MLP_H + o] =
// sum_p dY[p][o] (the bias).
constexpr int W4PART = 2 * MLP_H + 2;
__global__ __launch_bounds__(256) void w4part_kernel(const uint16_t *__restrict__ H, const float *__restrict__ dY,
                                                     int n, int nb, int rpb, float *__restrict__ part) {
  __shared__ float red[256][3];
  const int j = threadIdx.x % MLP_H, h = threadIdx.x / MLP_H;
  const int r0 = blockIdx.x * rpb, r1 = min(n, r0 + rpb);
  float s0 = 0.f, s1 = 0.f, sb0 = 0.f, sb1 = 0.f;
  for (int r = r0 + h; r < r1; r += 2) {
    const float hv = from_bf16(H[(size_t)r * MLP_H + j]);
    const float d0 = dY[2 * r], d1 = dY[2 * r + 1];
    s0 += hv * d0;
    s1 += hv * d1;
    if (r < nb) {  // tangent rows carry no bias
      sb0 += d0;
      sb1 += d1;
    }
  }
  red[threadIdx.x][0] = s0;
  red[threadIdx.x][1] = s1;
  red[threadIdx.x][2] = j == 0 ? sb0 : sb1;
  __syncthreads();
  if (h == 0) {
    float *o = part + (size_t)blockIdx.x * W4PART;
    o[2 * j] = red[j][0] + red[MLP_H + j][0];
    o[2 * j + 1] = red[j][1] + red[MLP_H + j][1];
    if (j < 2) {
      // thread j == 0 summed dY[.][0] in slot 2, thread 1 dY[.][1]
      o[2 * MLP_H + j] = red[j][2] + red[MLP_H + j][2];
    }
  }
}

// Y = H3 W4 + b4 (BF16 operands, FP32 sums); rows from nb on are tangent
// rows, which carry no bias.
__global__ void out_layer_kernel(const uint16_t *__restrict__ H, const uint16_t *__restrict__ W4,
                                 const float *__restrict__ b4, int n, int nb, float *__restrict__ Y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float y0 = 0.f, y1 = 0.f;
  for (int j = 0; j < MLP_H; j++) {
    const float h = from_bf16(H[(size_t)p * MLP_H + j]);
    y0 += h * from_bf16(W4[2 * j]);
    y1 += h * from_bf16(W4[2 * j + 1]);
  }
  Y[2 * p] = p < nb ? y0 + b4[0] : y0;
  Y[2 * p + 1] = p < nb ? y1 + b4[1] : y1;
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

// dZ3 = (dY W4^T) * SiLU'(Z3) as BF16 over rows [b rpb, (b + 1) rpb), and
// the block's column sums of dZ3 (the partials of the bias gradient):
// part[b][j].  Two rows per step, each handled by 128 consecutive threads.
__global__ __launch_bounds__(256) void out_back_kernel(const float *__restrict__ dY,
                                                       const uint16_t *__restrict__ W4,
                                                       const float *__restrict__ Z, int n, int rpb,
                                                       uint16_t *__restrict__ dZb, float *__restrict__ part) {
  __shared__ float red[256];
  const int j = threadIdx.x % MLP_H, h = threadIdx.x / MLP_H;
  const int r0 = blockIdx.x * rpb, r1 = min(n, r0 + rpb);
  const float w0 = from_bf16(W4[2 * j]), w1 = from_bf16(W4[2 * j + 1]);
  float cs = 0.f;
  for (int r = r0 + h; r < r1; r += 2) {
    const size_t i = (size_t)r * MLP_H + j;
    const float dh = dY[2 * r] * w0 + dY[2 * r + 1] * w1;
    const float z = Z[i], sg = sigm(z);
    const float dz = dh * sg * (1.f + z * (1.f - sg));
    dZb[i] = to_bf16(dz);
    cs += dz;
  }
  red[threadIdx.x] = cs;
  __syncthreads();
  if (h == 0) part[(size_t)blockIdx.x * MLP_H + j] = red[j] + red[MLP_H + j];
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

// Value rows [0, B) and tangent rows [B, 2B) of the features of points
// xyz[idx[p]] (idx may be null) along unit directions dir[idx[p]]: sin a,
// cos a and their derivatives cos a da, -sin a da with da = 2 pi s B . u.
__global__ void jvp_features_kernel(const float *__restrict__ xyz, const float *__restrict__ dir,
                                    const int *__restrict__ idx, int B, DSpec sp,
                                    uint16_t *__restrict__ X0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = i / MLP_NF, f = i % MLP_NF;
  if (p >= B) return;
  const int g = idx ? idx[p] : p;
  const float x = (xyz[3 * g] - sp.c[0]) * sp.s;
  const float y = (xyz[3 * g + 1] - sp.c[1]) * sp.s;
  const float z = (xyz[3 * g + 2] - sp.c[2]) * sp.s;
  const float a = TWO_PI * (sp.freq[f][0] * x + sp.freq[f][1] * y + sp.freq[f][2] * z);
  const float da = TWO_PI * sp.s *
                   (sp.freq[f][0] * dir[3 * g] + sp.freq[f][1] * dir[3 * g + 1] + sp.freq[f][2] * dir[3 * g + 2]);
  float sv, cv;
  sincosf(a, &sv, &cv);
  X0[(size_t)p * MLP_IN + f] = to_bf16(sv);
  X0[(size_t)p * MLP_IN + MLP_NF + f] = to_bf16(cv);
  X0[(size_t)(B + p) * MLP_IN + f] = to_bf16(cv * da);
  X0[(size_t)(B + p) * MLP_IN + MLP_NF + f] = to_bf16(-sv * da);
}

// Loss of value rows Y[0, B) against E and of tangent rows Y[B, 2B) against
// the directional derivatives T (weight w), per point g = idx[p]; seeds dY
// of all 2B rows.
__global__ void jvp_loss_kernel(const float *__restrict__ Y, const float *__restrict__ E,
                                const float *__restrict__ T, const int *__restrict__ idx, int B,
                                float inv, float w, float *__restrict__ dY, double *__restrict__ loss) {
  __shared__ float red[256];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (p < B) {
    const int g = idx ? idx[p] : p;
    for (int o = 0; o < 2; o++) {
      const float re = Y[2 * p + o] - E[2 * g + o];
      const float rd = Y[2 * (B + p) + o] - T[2 * g + o];
      dY[2 * p + o] = 2.f * re * inv;
      dY[2 * (B + p) + o] = 2.f * w * rd * inv;
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

// Output layer back through value and tangent rows: dH = dY W4^T on both,
// then dZ3 = dH SiLU'(Z3) + dH~ SiLU''(Z3) Z3~ (value rows) and dZ3~ = dH~
// SiLU'(Z3) (tangent rows, Z3 rows [B, 2B) hold Z3~) as BF16, over value
// rows [b rpb, (b + 1) rpb), with the block's column sums of the value
// rows' dZ3 (bias-gradient partials): part[b][j].
__global__ __launch_bounds__(256) void out_back_jvp_kernel(const float *__restrict__ dY,
                                                           const uint16_t *__restrict__ W4,
                                                           const float *__restrict__ Z, int B, int rpb,
                                                           uint16_t *__restrict__ dZb,
                                                           float *__restrict__ part) {
  __shared__ float red[256];
  const int j = threadIdx.x % MLP_H, h = threadIdx.x / MLP_H;
  const int r0 = blockIdx.x * rpb, r1 = min(B, r0 + rpb);
  const float w0 = from_bf16(W4[2 * j]), w1 = from_bf16(W4[2 * j + 1]);
  float cs = 0.f;
  for (int r = r0 + h; r < r1; r += 2) {
    const size_t i = (size_t)r * MLP_H + j, it = (size_t)(B + r) * MLP_H + j;
    const float dh = dY[2 * r] * w0 + dY[2 * r + 1] * w1;
    const float dht = dY[2 * (B + r)] * w0 + dY[2 * (B + r) + 1] * w1;
    const float z = Z[i], sg = sigm(z);
    const float dz = dh * silu1(z, sg) + dht * silu2(z, sg) * Z[it];
    dZb[i] = to_bf16(dz);
    dZb[it] = to_bf16(dht * silu1(z, sg));
    cs += dz;
  }
  red[threadIdx.x] = cs;
  __syncthreads();
  if (h == 0) part[(size_t)blockIdx.x * MLP_H + j] = red[j] + red[MLP_H + j];
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
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *Y = nullptr, *dY = nullptr,
         *dX = nullptr, *pts = nullptr, *grd = nullptr;
  uint16_t *dZ1b = nullptr, *dZ2b = nullptr, *dZ3b = nullptr;
  int *idx = nullptr;
  double *loss = nullptr;
  float *part = nullptr;  // partial products of the split weight-gradient products
  float *dHt = nullptr;   // tangent rows' upstream gradient of a layer (forward-mode training)
  float *dir = nullptr;   // directions of mlp_jvp_loss_grad's points
  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    MLP_OK(hipMalloc(&p, std::max<size_t>(sizeof(T) * count, 16)));
    allocs.push_back(p);
    return reinterpret_cast<T *>(p);
  }
  ~Mlp() {
    (void)hipSetDevice(device);
    for (void *p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

unsigned blocks_for(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

constexpr int MAX_SPLIT = 64;  // k-chunks of a split product

// C = op(A) op(B) with epilogue EPI (bias and Zin/H as gemm_kernel takes
// them).  Split (part != null, EPI_STORE only): C += op(A) op(B), the k range
// cut into at most MAX_SPLIT chunks whose products go to part (MAX_SPLIT M N
// floats) and are summed in chunk order; C then has ldc == N.  The operands'
// leading dimensions and base addresses must allow 16-byte loads.
template <bool TA, bool TB, int EPI = EPI_STORE>
void gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
          float *C, int ldc, int splitk = 1, float *part = nullptr, const float *bias = nullptr,
          const float *Zin = nullptr, uint16_t *H = nullptr, const float *Zt = nullptr,
          const float *Gt = nullptr) {
  if (lda % 8 || ldb % 8 || ((uintptr_t)A & 15) || ((uintptr_t)B & 15))
    throw Error(MCCS_ERR_INVALID, "matrix operand not aligned for 16-byte loads");
  if (N % 4 || ldc % 4 || ((uintptr_t)C & 15) || ((uintptr_t)H & 7))
    throw Error(MCCS_ERR_INVALID, "matrix result not aligned for vector stores");
  if ((EPI == EPI_SILU_BACK || EPI == EPI_SILU2_BACK) &&
      (size_t)((M + GT - 1) / GT) * ldc > (size_t)MAX_SPLIT * MLP_H * MLP_H)
    throw Error(MCCS_ERR_INVALID, "column partials outside their workspace");
  int kchunk = (K + splitk - 1) / splitk;
  kchunk = (kchunk + GK - 1) / GK * GK;
  splitk = (K + kchunk - 1) / kchunk;
  if (!part) {
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, 1);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, false, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb,
                       C, ldc, K, bias, Zin, H, Zt, Gt);
    return;
  }
  if (EPI != EPI_STORE || splitk > MAX_SPLIT || ldc != N)
    throw Error(MCCS_ERR_INVALID, "split product outside its workspace");
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, splitk);
  hipLaunchKernelGGL((gemm_kernel<TA, TB, true, EPI_STORE>), grid, dim3(256), 0, s, M, N, K, A, lda, B,
                     ldb, part, N, kchunk, (const float *)nullptr, (const float *)nullptr,
                     (uint16_t *)nullptr, (const float *)nullptr, (const float *)nullptr);
  hipLaunchKernelGGL(partsum_kernel, dim3((M * N + 3) / 4), dim3(256), 0, s, (const float *)part, splitk,
                     M * N, C);
}

// Row blocks of the column reductions: at least 64 rows each, at most
// MAX_ROWBLOCKS blocks (their partials fit the split-product workspace).
constexpr int MAX_ROWBLOCKS = 512;
int rows_per_block(int n) { return std::max(64, (n + MAX_ROWBLOCKS - 1) / MAX_ROWBLOCKS); }
static_assert((size_t)MAX_ROWBLOCKS * W4PART <= (size_t)MAX_SPLIT * MLP_H * MLP_H,
              "column partials exceed the workspace");

// out[c] += the sum over nb row blocks of part[b][c], c < MLP_H, in a fixed order.
void colsum(hipStream_t s, const float *part, int nb, float *out) {
  hipLaunchKernelGGL(partsum_kernel, dim3((MLP_H + 3) / 4), dim3(256), 0, s, part, nb, MLP_H, out);
}

// Forward over n <= maxb points already in m->pts: leaves Z*, H*, Y.
void forward(Mlp *m, int n) {
  hipStream_t s = m->stream;
  hipLaunchKernelGGL(features_kernel, dim3(blocks_for((size_t)n * MLP_NF)), dim3(256), 0, s,
                     m->pts, n, m->ds, m->X0);
  gemm<false, false, EPI_BIAS_SILU>(s, n, MLP_H, MLP_IN, m->X0, MLP_IN, m->wb + MLP_W1, MLP_H, m->Z1, MLP_H,
                                    1, nullptr, m->w + MLP_B1, nullptr, m->H1);
  gemm<false, false, EPI_BIAS_SILU>(s, n, MLP_H, MLP_H, m->H1, MLP_H, m->wb + MLP_W2, MLP_H, m->Z2, MLP_H,
                                    1, nullptr, m->w + MLP_B2, nullptr, m->H2);
  gemm<false, false, EPI_BIAS_SILU>(s, n, MLP_H, MLP_H, m->H2, MLP_H, m->wb + MLP_W3, MLP_H, m->Z3, MLP_H,
                                    1, nullptr, m->w + MLP_B3, nullptr, m->H3);
  hipLaunchKernelGGL(out_layer_kernel, dim3(blocks_for(n)), dim3(256), 0, s, m->H3, m->wb + MLP_W4,
                     m->w + MLP_B4, n, n, m->Y);
}

// Forward over B points xyz[idx[p]] (idx may be null) with the derivative
// along dir[idx[p]]: value rows [0, B) and tangent rows [B, 2B) of X0, Z*
// (rows [B, 2B): the tangent pre-activations), H* and Y.
void forward_jvp(Mlp *m, int B, const float *xyz, const float *dir, const int *idx) {
  hipStream_t s = m->stream;
  const size_t tb = (size_t)B * MLP_H;  // tangent-row offset of a 128-wide buffer
  hipLaunchKernelGGL(jvp_features_kernel, dim3(blocks_for((size_t)B * MLP_NF)), dim3(256), 0, s, xyz, dir,
                     idx, B, m->ds, m->X0);
  const uint16_t *in = m->X0;
  size_t tin = (size_t)B * MLP_IN;
  int K = MLP_IN;
  const size_t wo[3] = {MLP_W1, MLP_W2, MLP_W3}, bo[3] = {MLP_B1, MLP_B2, MLP_B3};
  float *Z[3] = {m->Z1, m->Z2, m->Z3};
  uint16_t *H[3] = {m->H1, m->H2, m->H3};
  for (int l = 0; l < 3; l++) {
    gemm<false, false, EPI_BIAS_SILU>(s, B, MLP_H, K, in, K, m->wb + wo[l], MLP_H, Z[l], MLP_H, 1, nullptr,
                                      m->w + bo[l], nullptr, H[l]);
    gemm<false, false, EPI_TAN>(s, B, MLP_H, K, in + tin, K, m->wb + wo[l], MLP_H, Z[l] + tb, MLP_H, 1,
                                nullptr, nullptr, Z[l], H[l] + tb);
    in = H[l];
    tin = tb;
    K = MLP_H;
  }
  hipLaunchKernelGGL(out_layer_kernel, dim3(blocks_for(2 * (size_t)B)), dim3(256), 0, s, m->H3,
                     m->wb + MLP_W4, m->w + MLP_B4, 2 * B, B, m->Y);
}

// Backward from m->dY; weight gradients accumulated into m->g when wgrad,
// input-feature gradient into m->dX when xgrad.
void backward(Mlp *m, int n, bool wgrad, bool xgrad) {
  hipStream_t s = m->stream;
  const int sk = std::max(1, std::min(MAX_SPLIT, n / 256));
  if (wgrad) {
    // gW4 and gb4 are adjacent in the flat layout: one fixed-order reduction
    static_assert(MLP_B4 == MLP_W4 + 2 * MLP_H, "output-layer gradient layout");
    const int rpb = rows_per_block(n), nb = (n + rpb - 1) / rpb;
    hipLaunchKernelGGL(w4part_kernel, dim3(nb), dim3(256), 0, s, m->H3, m->dY, n, n, rpb, m->part);
    hipLaunchKernelGGL(partsum_kernel, dim3((W4PART + 3) / 4), dim3(256), 0, s, (const float *)m->part, nb,
                       W4PART, m->g + MLP_W4);
  }
  {
    const int rpb = rows_per_block(n), nb = (n + rpb - 1) / rpb;
    hipLaunchKernelGGL(out_back_kernel, dim3(nb), dim3(256), 0, s, m->dY, m->wb + MLP_W4, m->Z3, n, rpb,
                       m->dZ3b, m->part);
    if (wgrad) colsum(s, m->part, nb, m->g + MLP_B3);
  }
  const int nbt = (n + GT - 1) / GT;  // row tiles of the backward products (column partials)
  if (wgrad)
    gemm<true, false>(s, MLP_H, MLP_H, n, m->H2, MLP_H, m->dZ3b, MLP_H, m->g + MLP_W3, MLP_H, sk, m->part);
  gemm<false, true, EPI_SILU_BACK>(s, n, MLP_H, MLP_H, m->dZ3b, MLP_H, m->wb + MLP_W3, MLP_H, m->part, MLP_H,
                                   1, nullptr, nullptr, m->Z2, m->dZ2b);
  if (wgrad) {
    colsum(s, m->part, nbt, m->g + MLP_B2);
    gemm<true, false>(s, MLP_H, MLP_H, n, m->H1, MLP_H, m->dZ2b, MLP_H, m->g + MLP_W2, MLP_H, sk, m->part);
  }
  gemm<false, true, EPI_SILU_BACK>(s, n, MLP_H, MLP_H, m->dZ2b, MLP_H, m->wb + MLP_W2, MLP_H, m->part, MLP_H,
                                   1, nullptr, nullptr, m->Z1, m->dZ1b);
  if (wgrad) {
    colsum(s, m->part, nbt, m->g + MLP_B1);
    gemm<true, false>(s, MLP_IN, MLP_H, n, m->X0, MLP_IN, m->dZ1b, MLP_H, m->g + MLP_W1, MLP_H, sk, m->part);
  }
  if (xgrad)
    gemm<false, true>(s, n, MLP_IN, MLP_H, m->dZ1b, MLP_H, m->wb + MLP_W1, MLP_H, m->dX, MLP_IN);
}

// Backward of forward_jvp from m->dY (value and tangent rows); parameter
// gradients accumulated into m->g.  Weight gradients sum over both row
// sets, bias gradients over the value rows.
void backward_jvp(Mlp *m, int B) {
  hipStream_t s = m->stream;
  const int n2 = 2 * B;
  const size_t tb = (size_t)B * MLP_H;
  const int sk = std::max(1, std::min(MAX_SPLIT, n2 / 256));
  {
    static_assert(MLP_B4 == MLP_W4 + 2 * MLP_H, "output-layer gradient layout");
    const int rpb = rows_per_block(n2), nb = (n2 + rpb - 1) / rpb;
    hipLaunchKernelGGL(w4part_kernel, dim3(nb), dim3(256), 0, s, m->H3, m->dY, n2, B, rpb, m->part);
    hipLaunchKernelGGL(partsum_kernel, dim3((W4PART + 3) / 4), dim3(256), 0, s, (const float *)m->part, nb,
                       W4PART, m->g + MLP_W4);
  }
  {
    const int rpb = rows_per_block(B), nb = (B + rpb - 1) / rpb;
    hipLaunchKernelGGL(out_back_jvp_kernel, dim3(nb), dim3(256), 0, s, m->dY, m->wb + MLP_W4, m->Z3, B, rpb,
                       m->dZ3b, m->part);
    colsum(s, m->part, nb, m->g + MLP_B3);
  }
  const int nbt = (B + GT - 1) / GT;
  // layers 3 -> 2 -> 1: weight gradient over both row sets, then the
  // tangent rows' gradient (EPI_TAN: also keeps the upstream tangent gradient
  // in dHt for the value rows), then the value rows' (EPI_SILU2_BACK)
  const uint16_t *Hin[3] = {m->X0, m->H1, m->H2};
  const int Kin[3] = {MLP_IN, MLP_H, MLP_H};
  const size_t wo[3] = {MLP_W1, MLP_W2, MLP_W3}, bo[3] = {MLP_B1, MLP_B2, MLP_B3};
  float *Z[3] = {m->Z1, m->Z2, m->Z3};
  uint16_t *dZb[3] = {m->dZ1b, m->dZ2b, m->dZ3b};
  for (int l = 2; l >= 0; l--) {
    gemm<true, false>(s, Kin[l], MLP_H, n2, Hin[l], Kin[l], dZb[l], MLP_H, m->g + wo[l], MLP_H, sk, m->part);
    if (l == 0) break;
    gemm<false, true, EPI_TAN>(s, B, MLP_H, MLP_H, dZb[l] + tb, MLP_H, m->wb + wo[l], MLP_H, m->dHt, MLP_H, 1,
                               nullptr, nullptr, Z[l - 1], dZb[l - 1] + tb);
    gemm<false, true, EPI_SILU2_BACK>(s, B, MLP_H, MLP_H, dZb[l], MLP_H, m->wb + wo[l], MLP_H, m->part, MLP_H,
                                      1, nullptr, nullptr, Z[l - 1], dZb[l - 1], Z[l - 1] + tb, m->dHt);
    colsum(s, m->part, nbt, m->g + bo[l - 1]);
  }
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
  if (max_batch <= 0 || max_batch > 524288)
    throw Error(MCCS_ERR_INVALID, "network batch must hold 1..524288 points");
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
  m->dX = m->alloc<float>(B * MLP_IN);
  m->Y = m->alloc<float>(B * MLP_OUT);
  m->dY = m->alloc<float>(B * MLP_OUT);
  m->pts = m->alloc<float>(B * 3);
  m->grd = m->alloc<float>(B * 3);
  m->idx = m->alloc<int>(B);
  m->loss = m->alloc<double>(1);
  m->part = m->alloc<float>((size_t)MAX_SPLIT * MLP_H * MLP_H);
  m->dHt = m->alloc<float>(B / 2 * MLP_H + MLP_H);
  m->dir = m->alloc<float>(B * 3);
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
    (void)hipFree(dE);
    throw;
  }
  (void)hipFree(dE);
  return loss;
}

double mlp_jvp_loss_grad(Mlp *m, const double *xyz, const double *dir, const double *e, const double *d,
                         double w, uint64_t n, float *grad, double *y, double *yd) {
  MLP_OK(hipSetDevice(m->device));
  if (n == 0 || n > (uint64_t)m->maxb / 2)
    throw Error(MCCS_ERR_INVALID, "batch larger than the network's workspace");
  std::vector<float> p(3 * n), u(3 * n), E(2 * n), D(2 * n);
  for (uint64_t i = 0; i < 3 * n; i++) {
    p[i] = (float)xyz[i];
    u[i] = (float)dir[i];
  }
  for (uint64_t i = 0; i < n; i++)
    for (int o = 0; o < 2; o++) {
      E[2 * i + o] = (float)(e[2 * i + o] / m->spec.out_scale[o]);
      D[2 * i + o] = (float)(d[2 * i + o] / m->spec.out_scale[o]);
    }
  float *dE = nullptr;
  MLP_OK(hipMalloc(&dE, sizeof(float) * 4 * n));
  double loss = 0.0;
  std::vector<float> Y(4 * n);
  try {
    MLP_OK(hipMemcpyAsync(m->pts, p.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemcpyAsync(m->dir, u.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemcpyAsync(dE, E.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemcpyAsync(dE + 2 * n, D.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
    MLP_OK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
    MLP_OK(hipMemsetAsync(m->g, 0, sizeof(float) * MLP_NPARAM, m->stream));
    forward_jvp(m, (int)n, m->pts, m->dir, nullptr);
    MLP_OK(hipMemcpyAsync(Y.data(), m->Y, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, m->stream));
    hipLaunchKernelGGL(jvp_loss_kernel, dim3(blocks_for(n)), dim3(256), 0, m->stream, m->Y, dE, dE + 2 * n,
                       (const int *)nullptr, (int)n, 1.0f / (2.0f * n), (float)w, m->dY, m->loss);
    backward_jvp(m, (int)n);
    MLP_OK(hipMemcpyAsync(grad, m->g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_OK(hipMemcpyAsync(&loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_OK(hipStreamSynchronize(m->stream));
    MLP_OK(hipMemsetAsync(m->g, 0, sizeof(float) * MLP_NPARAM, m->stream));
    MLP_OK(hipStreamSynchronize(m->stream));
  } catch (...) {
    (void)hipFree(dE);
    throw;
  }
  (void)hipFree(dE);
  for (uint64_t i = 0; i < 2 * n; i++) {
    if (y) y[i] = Y[i];
    if (yd) yd[i] = Y[2 * n + i];
  }
  return loss;
}

double mlp_train(Mlp *m, const float *d_xyz, const float *d_e, uint64_t n, const TrainOpts &o,
                 std::vector<double> *history) {
  MLP_OK(hipSetDevice(m->device));
  const bool sob = o.d_dir && o.d_deriv;
  const int B = std::min(o.batch, sob ? m->maxb / 2 : m->maxb);
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
          forward_jvp(m, cnt, d_xyz, o.d_dir, ix);
          hipLaunchKernelGGL(jvp_loss_kernel, dim3(blocks_for(cnt)), dim3(256), 0, m->stream, m->Y, d_e,
                             o.d_deriv, ix, cnt, 1.0f / (2.0f * cnt), (float)o.w_deriv, m->dY, m->loss);
          backward_jvp(m, cnt);
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
    (void)hipFree(d_perm);
    throw;
  }
  (void)hipFree(d_perm);
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
    (void)hipFree(dp);
    throw Error(MCCS_ERR_NOMEM, "device allocation of training data failed");
  }
  double last = 0.0;
  try {
    MLP_OK(hipMemcpy(dp, p.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice));
    MLP_OK(hipMemcpy(dt, t.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    last = mlp_train(m, dp, dt, n, o, history);
  } catch (...) {
    (void)hipFree(dp);
    (void)hipFree(dt);
    throw;
  }
  (void)hipFree(dp);
  (void)hipFree(dt);
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
