// Fourier-feature MLP of the far-field surrogate (internal to libmassccs_amd).
//
// Network: r (A) -> r' = (r - centre) * in_scale -> 64 Fourier features
// [sin(2 pi B r'), cos(2 pi B r')] (B: 32 x 3, fixed) -> 128 SiLU -> 128 SiLU
// -> 128 SiLU -> 2 outputs (energy of the N2 nitrogen site and of the N2
// charge/dipole site, in units of out_scale kcal/mol).  Weights are FP32
// masters with BF16 copies; every matrix product runs on BF16 MFMA with FP32
// accumulation.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "massccs.h"

namespace mccs {

constexpr int MLP_NF = 32;    // Fourier frequencies
constexpr int MLP_IN = 64;    // features
constexpr int MLP_H = 128;    // hidden width
constexpr int MLP_OUT = 2;    // outputs

// flat FP32 parameter layout (row-major W[in][out])
constexpr size_t MLP_W1 = 0;
constexpr size_t MLP_B1 = MLP_W1 + (size_t)MLP_IN * MLP_H;
constexpr size_t MLP_W2 = MLP_B1 + MLP_H;
constexpr size_t MLP_B2 = MLP_W2 + (size_t)MLP_H * MLP_H;
constexpr size_t MLP_W3 = MLP_B2 + MLP_H;
constexpr size_t MLP_B3 = MLP_W3 + (size_t)MLP_H * MLP_H;
constexpr size_t MLP_W4 = MLP_B3 + MLP_H;
constexpr size_t MLP_B4 = MLP_W4 + (size_t)MLP_H * MLP_OUT;
constexpr size_t MLP_NPARAM = MLP_B4 + MLP_OUT;

struct MlpSpec {
  float freq[MLP_NF][3];   // B
  double centre[3];
  double in_scale;         // 1 / length scale
  double out_scale[MLP_OUT];
};

// Device workspace and parameters of one network.
struct Mlp;

Mlp *mlp_create(int device, const MlpSpec &spec, const float *params, int max_batch);
void mlp_destroy(Mlp *m);
void mlp_get_params(const Mlp *m, float *params);
// Outputs (n x 2, kcal/mol) and optionally their gradients (n x 2 x 3,
// kcal/mol/A) at host points (n x 3).
void mlp_eval(Mlp *m, const double *xyz, uint64_t n, double *y, double *dy);
// Loss (mean over points and outputs of the squared error of the scaled
// outputs) and its gradient with respect to the flat parameters.
double mlp_loss_grad(Mlp *m, const double *xyz, const double *e, uint64_t n, float *grad);
// Adam training on device-resident data: n points (float3, scaled energies
// float2); returns the last epoch's mean loss.
struct TrainOpts {
  int epochs = 30, batch = 8192;
  double lr = 2e-3, lr_final = 1e-4;
  uint64_t seed = 2104;
  // optional derivative matching: per point a unit direction (float3) and
  // the scaled directional derivatives of both outputs (float2), matched by
  // the network's forward-mode derivative along the direction (a tangent
  // row beside each value row, back-propagated with it), weight w_deriv
  const float *d_dir = nullptr, *d_deriv = nullptr;
  double w_deriv = 1.0;
};
double mlp_train(Mlp *m, const float *d_xyz, const float *d_e, uint64_t n, const TrainOpts &o,
                 std::vector<double> *history);
// Loss of the values and the directional derivatives along unit vectors
// dir (n x 3) against e and d (n x 2, kcal/mol and kcal/mol/A; scaled
// here): sum of (y - e)^2 + w (y' - d)^2 over points and outputs / 2n, with
// the gradient with respect to the flat parameters; y and yd (may be null)
// receive the scaled outputs and derivatives (n x 2).  n <= max_batch / 2.
double mlp_jvp_loss_grad(Mlp *m, const double *xyz, const double *dir, const double *e, const double *d,
                         double w, uint64_t n, float *grad, double *y, double *yd);
// The same from host points and energies (kcal/mol; scaled here).
double mlp_train_host(Mlp *m, const double *xyz, const double *e, uint64_t n, const TrainOpts &o,
                      std::vector<double> *history);
// Energies at device points (float3) into device FP32 (n x 2), kcal/mol.
void mlp_eval_device(Mlp *m, const float *d_xyz, uint64_t n, float *d_y, void *stream);

}  // namespace mccs
