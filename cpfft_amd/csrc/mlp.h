// mlp.h -- buffers and launchers of the warm-start network (mlp.hip) and of
// the per-voxel feature / warm-start kernels (cpfft_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpfft {

constexpr int kMlpIn = 42, kMlpHid = 128, kMlpOut = 12;
// flat FP32 parameter vector [W1 (128x42) | b1 | W2 (128x128) | b2 | W3 (12x128) | b3];
// W rows are output units
constexpr long kOffW1 = 0;
constexpr long kOffB1 = kOffW1 + (long)kMlpHid * kMlpIn;
constexpr long kOffW2 = kOffB1 + kMlpHid;
constexpr long kOffB2 = kOffW2 + (long)kMlpHid * kMlpHid;
constexpr long kOffW3 = kOffB2 + kMlpHid;
constexpr long kOffB3 = kOffW3 + (long)kMlpOut * kMlpHid;
constexpr long kMlpParams = kOffB3 + kMlpOut;
// normalisation [mean_x (42) | sd_x (42) | mean_t (12) | sd_t (12)]
constexpr int kMlpNorm = 2 * kMlpIn + 2 * kMlpOut;

struct MlpBuffers {
  float* params = nullptr;  // FP32 master weights
  uint16_t* pbf = nullptr;  // BF16 copy used by the GEMMs
  float* grads = nullptr;
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* norm = nullptr;
  float* logX = nullptr;    // logged raw features [cap][42]
  float* logT = nullptr;    // logged converged slip increments [cap][12]
  long cap = 0;
  int rows = 0;             // batch / chunk capacity of the work buffers
  uint16_t* Xb = nullptr;   // [rows][42] normalised BF16 inputs
  float* Xf = nullptr;      // [rows][42] raw FP32 features (inference chunks)
  float* Tb = nullptr;      // [rows][12] normalised targets
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dZ1 = nullptr, *dZ2 = nullptr;
  uint16_t *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ1b = nullptr, *dZ2b = nullptr;
  double* part = nullptr;   // loss partial sums
  float* ws = nullptr;      // predicted slip increments for the warm start [12][V]
};

void mlp_to_bf16(long n, const float* x, uint16_t* y, hipStream_t s);
void mlp_weights_to_bf16(const MlpBuffers& b, hipStream_t s);
// Y[rows][12] = network(Xb) (normalised units); H1, H2 kept for mlp_backward
void mlp_forward(const MlpBuffers& b, int rows, const uint16_t* Xb, float* Y, hipStream_t s);
// gradient of mean((Y - T)^2) over rows x 12 into grads (b.Y from mlp_forward);
// part[0..mlp_loss_parts(rows)) receives partial sums of (Y - T)^2
void mlp_backward(const MlpBuffers& b, int rows, const uint16_t* Xb, const float* T, float* grads,
                  double* part, hipStream_t s);
int mlp_loss_parts(int rows);
void mlp_adam(const MlpBuffers& b, float lr, long t, hipStream_t s);
void mlp_gather_batch(const MlpBuffers& b, int rows, long count, uint32_t seed, uint32_t step,
                      hipStream_t s);
void mlp_norm_inputs(const MlpBuffers& b, long rows, const float* X, uint16_t* Xb, hipStream_t s);
void mlp_stats(const MlpBuffers& b, long count, hipStream_t s);

}  // namespace cpfft
