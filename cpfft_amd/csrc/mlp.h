// mlp.h -- buffers and launchers of the warm-start network (mlp.hip) and of
// the per-voxel feature / warm-start kernels (cpfft_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpfft {

constexpr int kMlpHid = 128;
constexpr int kMlpMaxOut = 24;  // slip systems of the largest set the warm start serves

// Sizes of the network for a crystal of ns slip systems: nin = 18 + 2 ns
// inputs (see voxel_features), nout = ns outputs.  Flat FP32 parameter
// vector [W1 (128 x nin) | b1 | W2 (128 x 128) | b2 | W3 (nout x 128) | b3],
// W rows the output units; normalisation [mean_x | sd_x | mean_t | sd_t].
struct MlpDims {
  int nin = 0, nout = 0;
  long offW1 = 0, offB1 = 0, offW2 = 0, offB2 = 0, offW3 = 0, offB3 = 0, nparam = 0;
  int nnorm = 0;
};
inline MlpDims mlp_dims(int ns) {
  MlpDims d;
  d.nin = 18 + 2 * ns;
  d.nout = ns;
  d.offW1 = 0;
  d.offB1 = d.offW1 + (long)kMlpHid * d.nin;
  d.offW2 = d.offB1 + kMlpHid;
  d.offB2 = d.offW2 + (long)kMlpHid * kMlpHid;
  d.offW3 = d.offB2 + kMlpHid;
  d.offB3 = d.offW3 + (long)d.nout * kMlpHid;
  d.nparam = d.offB3 + d.nout;
  d.nnorm = 2 * d.nin + 2 * d.nout;
  return d;
}
// inputs padded for the inference kernel's K steps of 32
inline int mlp_nin_pad(int nin) { return (nin + 31) / 32 * 32; }

struct MlpBuffers {
  MlpDims d;
  float* params = nullptr;  // FP32 master weights
  uint16_t* pbf = nullptr;  // BF16 copy used by the GEMMs
  float* grads = nullptr;
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* norm = nullptr;
  float* logX = nullptr;    // logged raw features [cap][nin]
  float* logT = nullptr;    // logged converged slip increments [cap][nout]
  long cap = 0;
  int rows = 0;             // batch capacity of the training buffers
  uint16_t* Xb = nullptr;   // [rows][nin] normalised BF16 inputs
  float* Xf = nullptr;      // [rows][nin] FP32 inputs staged from the host
  float* Tb = nullptr;      // [rows][nout] normalised targets
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dZ1 = nullptr, *dZ2 = nullptr;
  uint16_t *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ1b = nullptr, *dZ2b = nullptr;
  double* part = nullptr;   // loss partial sums
  uint16_t* Xv = nullptr;   // normalised BF16 inputs of every voxel [V][mlp_nin_pad(nin)]
  float* ws = nullptr;      // predicted slip increments for the warm start [nout][V]
};

void mlp_to_bf16(long n, const float* x, uint16_t* y, hipStream_t s);
void mlp_weights_to_bf16(const MlpBuffers& b, hipStream_t s);
// Y[rows][nout] = network(Xb) (normalised units); H1, H2 kept for mlp_backward
void mlp_forward(const MlpBuffers& b, int rows, const uint16_t* Xb, float* Y, hipStream_t s);
// gradient of mean((Y - T)^2) over rows x nout into grads (b.Y from
// mlp_forward); part[0..mlp_loss_parts(b, rows)) receives partial sums of
// (Y - T)^2
void mlp_backward(const MlpBuffers& b, int rows, const uint16_t* Xb, const float* T, float* grads,
                  double* part, hipStream_t s);
int mlp_loss_parts(const MlpBuffers& b, int rows);
void mlp_adam(const MlpBuffers& b, float lr, long t, hipStream_t s);
void mlp_gather_batch(const MlpBuffers& b, int rows, long count, uint32_t seed, uint32_t step,
                      hipStream_t s);
void mlp_stats(const MlpBuffers& b, long count, hipStream_t s);
// Inference over all voxels in one launch: ws[k][e] = the network's output
// k for the normalised inputs Xv[e] (row stride mlp_nin_pad(nin)),
// denormalised (the weights must be in pbf).  BF16 operands, FP32
// accumulation, the same rounding points as mlp_forward.
void mlp_infer(const MlpBuffers& b, long V, hipStream_t s);

}  // namespace cpfft
