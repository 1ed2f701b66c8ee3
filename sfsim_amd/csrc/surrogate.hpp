// MLP surrogate for scattering orders >= 2 (no counterpart in the reference).
//
// Input: 4 LUT coordinates of an entry of the scatter grid (index / (size - 1)) and 6 atmosphere
// parameters mapped to [0, 1]; 8-band Fourier features (x, sin 2^b pi x, cos 2^b pi x), 170 values
// padded to 192; three hidden layers of 128 with SiLU; 16 outputs of which the first 15 are used:
// output 3 (k - 1) + c is log(P_k + eps) + 10 for channel c of P_k = dS_1 + ... + dS_k, the sum of the
// first k scattering tables of the plain iteration (k = 1..5).  Weights are kept in FP32 (Adam state too); every matrix product runs as bf16 x bf16
// with FP32 accumulation on the matrix cores (v_mfma_f32_32x32x8bf16_1k).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_params.hpp"

namespace sfsim {

constexpr int kMlpIn = 10;
constexpr int kMlpBands = 17;  // features per input: x, sin and cos of 2^b pi x for b = 0..7
constexpr int kMlpFeat = 192;
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 16;
constexpr int kMlpOrders = 5;
constexpr int kMlpUsed = 3 * kMlpOrders;
constexpr int kMaxSeededOrders = 32;  // scattering orders of a seeded build (the network predicts the first 5)
constexpr float kMlpEps = 1e-8f;     // floor inside the log
constexpr float kMlpShift = 10.0f;   // targets are log(P + eps) + shift
constexpr int kMaxSplits = 32;       // K splits of the weight-gradient products

// Offsets of the parameter blocks in the flat FP32 vector (W1 b1 W2 b2 W3 b3 W4 b4, row-major W[in][out]).
struct MlpLayout {
  static constexpr size_t W1 = 0;
  static constexpr size_t b1 = W1 + (size_t)kMlpFeat * kMlpHidden;
  static constexpr size_t W2 = b1 + kMlpHidden;
  static constexpr size_t b2 = W2 + (size_t)kMlpHidden * kMlpHidden;
  static constexpr size_t W3 = b2 + kMlpHidden;
  static constexpr size_t b3 = W3 + (size_t)kMlpHidden * kMlpHidden;
  static constexpr size_t W4 = b3 + kMlpHidden;
  static constexpr size_t b4 = W4 + (size_t)kMlpHidden * kMlpOut;
  static constexpr size_t total = b4 + kMlpOut;
};

// The six conditioning parameters of an atmosphere, mapped to [0, 1] over the training ranges:
// Rayleigh scale height 6..10 km, Rayleigh coefficient scale 0.5..2 (relative to 5.8e-6, 13.5e-6,
// 33.1e-6 /m), Mie scale height 0.8..2 km, Mie scattering coefficient 5e-6..5e-5 /m, Mie g 0.6..0.85,
// ground albedo 0.1..0.5.
struct AtmosphereCode {
  float v[6];
};

struct Surrogate {
  int batch_capacity = 0;
  bool trained = false;
  // device buffers
  float* theta = nullptr;      // FP32 master weights (MlpLayout::total)
  uint16_t* theta_bf = nullptr;  // bf16 copies of W1..W4 (same offsets)
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;
  // activations for one batch
  uint16_t* X = nullptr;       // [B][192] bf16 features
  float* Z[3] = {nullptr, nullptr, nullptr};      // [B][128] FP32 pre-activations
  uint16_t* A[3] = {nullptr, nullptr, nullptr};   // [B][128] bf16 activations
  float* Y = nullptr;          // [B][16] outputs
  float* T = nullptr;          // [B][16] targets
  uint16_t* dY = nullptr;      // [B][16] bf16
  float* dA = nullptr;         // [B][128] FP32 upstream gradient
  float* dZf = nullptr;        // [B][128] FP32 gradient before the activation (bias gradients)
  uint16_t* dZ = nullptr;      // [B][128] bf16
  float* loss = nullptr;       // [1 + B / 256] partial sums
  float* partial = nullptr;    // [kMaxSplits][192][128] split-K partial products
  float* ztab = nullptr;       // layer-1 tables of the fused prediction
  size_t ztab_capacity = 0;
  int adam_step = 0;
};

// Allocates / frees the surrogate's device state for batches up to `batch`.
hipError_t surrogate_alloc(Surrogate& s, int batch);
void surrogate_free(Surrogate& s);
// Deterministic initialisation (splitmix64 stream from `seed`), Adam state reset.
hipError_t surrogate_init(Surrogate& s, uint64_t seed, hipStream_t stream);
// Refresh the bf16 weight copies from the FP32 master weights.
hipError_t surrogate_sync_bf16(Surrogate& s, hipStream_t stream);

// Features of `rows` explicit input rows (host layout [rows][10], values in [0, 1]) -> s.X.
hipError_t surrogate_encode_rows(Surrogate& s, const float* inputs_dev, int rows, hipStream_t stream);
// Forward pass over s.X (rows <= capacity) -> s.Y.
hipError_t surrogate_forward(Surrogate& s, int rows, hipStream_t stream);
// Loss and gradient for targets s.T with output mask (first kMlpUsed outputs) -> s.grad, s.loss.
hipError_t surrogate_backward(Surrogate& s, int rows, hipStream_t stream);
// One Adam step on s.grad.
hipError_t surrogate_adam(Surrogate& s, float lr, hipStream_t stream);

// Training batch: `rows` entries drawn with a counter-based generator from `natm` atmospheres of
// `entries` scatter-grid entries each; targets[atm][entry][9] hold log(P + eps) + shift.
hipError_t surrogate_gather_batch(Surrogate& s, const DevParams& P, const AtmosphereCode* codes_dev, int natm,
                                  const float* targets_dev, uint64_t seed, int step, int rows, hipStream_t stream);
// Output biases b4[0 .. kMlpUsed) = mean of each target column over `rows` target rows.
hipError_t surrogate_init_output_bias(Surrogate& s, const float* targets_dev, long long rows, hipStream_t stream);
// Targets of one atmosphere from its order sums P_1..P_5 (float4 tables).
hipError_t surrogate_make_targets(const float4* const* sums, long long entries, float* targets_dev, hipStream_t stream);
// Allocates the layer-1 tables of a prediction over P's scatter grid (before a graph capture).
hipError_t surrogate_reserve_tables(Surrogate& s, const DevParams& P);
// Predicts the order sums P_k for every entry of the scatter grid of one atmosphere into out[k - 1]
// (entries of `out` may be null: not written): layer-1 tables, then one fused kernel for layers 2-4
// and the decoding.
hipError_t surrogate_predict_tables(Surrogate& s, const DevParams& P, const AtmosphereCode& code, float4* const* out,
                                    hipStream_t stream);

// Matrix product on the matrix cores: C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n]), A(m, k) at
// A[m * sam + k * sak], B(k, n) at B[k * sbk + n * sbn] (bf16), C row-major FP32 with leading dim ldc.
// act: 0 none, 1 SiLU; out_bf16 (nullable) receives act(C) as bf16 with leading dimension ldc.
hipError_t mfma_gemm(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                     long long sbn, float* C, int ldc, const float* bias, int act, uint16_t* out_bf16, int M, int N,
                     int K, hipStream_t stream);
// The same product with K split over up to kMaxSplits workgroups per tile (C row-major, ldc = N);
// partial holds kMaxSplits x M x N floats.
hipError_t mfma_gemm_split_k(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                             long long sbn, float* C, int M, int N, int K, float* partial, hipStream_t stream);

AtmosphereCode atmosphere_code(const double* rayleigh_base, double rayleigh_scale, double mie_base, double mie_scale,
                               double mie_g, double albedo);

}  // namespace sfsim
