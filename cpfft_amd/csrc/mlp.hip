// mlp.hip -- the learned warm start of the per-voxel Newton solve: a
// 42 -> 128 -> 128 -> 12 tanh network trained and evaluated with BF16
// matrix-core (MFMA) GEMMs and FP32 accumulation.  It has no counterpart in
// the reference; it only supplies the starting point of mm10_solve
// (mm10_a.f:2945-3023), whose FP64 Newton iteration and tolerances are
// unchanged.
//
// Matrices are row-major, one sample per row.  A GEMM computes
// C[m][n] = sum_k A(m,k) B(k,n) with A(m,k) = A[m*sam + k*sak] and
// B(k,n) = B[k*sbk + n*sbn] (BF16), so transposed operands are read in
// place.  Each wavefront owns one 16x16 output tile and steps K by 16 with
// v_mfma_f32_16x16x16_bf16: lane l supplies A rows l%16 and B columns l%16
// at k = 4*(l/16)..+3, and holds C rows 4*(l/16)..+3 of column l%16.
#include "mlp.h"

namespace cpfft {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (inputs are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

static constexpr int kGT = 256;  // threads per GEMM block: 4 wavefronts, a 64x16 tile

// EPI 0: C = AB.  1: C = tanh(AB + bias), with a BF16 copy.  2: C = AB + bias.
// 3: C = AB (1 - H^2) with H an FP32 matrix of C's shape (tanh derivative),
// with a BF16 copy.
template <int EPI>
__global__ __launch_bounds__(kGT) void k_gemm(int M, int N, int K, const uint16_t* __restrict__ A,
                                             long sam, long sak, const uint16_t* __restrict__ B,
                                             long sbk, long sbn, float* __restrict__ C, long ldc,
                                             const float* __restrict__ bias,
                                             const float* __restrict__ H,
                                             uint16_t* __restrict__ Cb) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 64 + w * 16, n0 = blockIdx.y * 16;
  const int li = lane & 15, kq = 4 * (lane >> 4);
  const int am = m0 + li, bn = n0 + li;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + kq + j;
      a[j] = (am < M && k < K) ? (short)A[am * sam + k * sak] : (short)0;
      b[j] = (bn < N && k < K) ? (short)B[k * sbk + bn * sbn] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int n = n0 + li;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = m0 + kq + j;
    if (r < M && n < N) {
      float v = acc[j];
      if (EPI == 1 || EPI == 2) v += bias[n];
      if (EPI == 1) v = tanhf(v);
      if (EPI == 3) {
        const float h = H[(long)r * ldc + n];
        v *= 1.0f - h * h;
      }
      C[(long)r * ldc + n] = v;
      if (EPI == 1 || EPI == 3) Cb[(long)r * ldc + n] = f2bf(v);
    }
  }
}

template <int EPI>
static void gemm(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B,
                 long sbk, long sbn, float* C, long ldc, const float* bias, const float* H,
                 uint16_t* Cb, hipStream_t s) {
  dim3 g((M + 63) / 64, (N + 15) / 16);
  k_gemm<EPI><<<g, kGT, 0, s>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, H, Cb);
}

__global__ void k_to_bf16(long n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// dY = 2 (Y - T) / (rows * cols) (FP32 and BF16) and per-block partial sums of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(long n, const float* __restrict__ Y,
                                             const float* __restrict__ T, float* __restrict__ dY,
                                             uint16_t* __restrict__ dYb, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  const float sc = 2.0f / (float)n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    acc += (double)d * d;
    const float g = sc * d;
    dY[i] = g;
    dYb[i] = f2bf(g);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// out[c] = sum_r X[r][c], one block per column, fixed summation order
__global__ __launch_bounds__(256) void k_colsum(int rows, int cols, const float* __restrict__ X,
                                                float* __restrict__ out) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float acc = 0.f;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) acc += X[(long)r * cols + c];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

// mean and standard deviation of each column of a [rows][cols] FP32 matrix
// (two passes, FP64 accumulation); a zero deviation is replaced by 1
__global__ __launch_bounds__(256) void k_colstats(long rows, int cols, const float* __restrict__ X,
                                                  float* __restrict__ mean, float* __restrict__ sd) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (long r = threadIdx.x; r < rows; r += blockDim.x) acc += X[r * cols + c];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double mu = sh[0] / (double)rows;
  __syncthreads();
  acc = 0.0;
  for (long r = threadIdx.x; r < rows; r += blockDim.x) {
    const double d = X[r * cols + c] - mu;
    acc += d * d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = sqrt(sh[0] / (double)rows);
    mean[c] = (float)mu;
    sd[c] = v > 0.0 ? (float)v : 1.0f;
  }
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// one training batch: rows drawn with replacement from the log by a hash of
// (seed, step, row), normalised; inputs as BF16, targets as FP32
__global__ void k_gather_batch(int Bn, long count, uint32_t seed, uint32_t step,
                               const float* __restrict__ logX, const float* __restrict__ logT,
                               const float* __restrict__ nrm, uint16_t* __restrict__ Xb,
                               float* __restrict__ Tb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Bn) return;
  const uint32_t h = mix32(seed * 0x9e3779b9u ^ mix32(step * 0x85ebca6bu ^ (uint32_t)r));
  const long src = (long)(h % (uint32_t)count);
  const float* mx = nrm;
  const float* sx = nrm + kMlpIn;
  const float* mt = nrm + 2 * kMlpIn;
  const float* st = nrm + 2 * kMlpIn + kMlpOut;
  for (int c = 0; c < kMlpIn; ++c)
    Xb[(long)r * kMlpIn + c] = f2bf((logX[src * kMlpIn + c] - mx[c]) / sx[c]);
  for (int c = 0; c < kMlpOut; ++c) Tb[(long)r * kMlpOut + c] = (logT[src * kMlpOut + c] - mt[c]) / st[c];
}

// raw features [rows][42] -> normalised BF16
__global__ void k_norm_inputs(long rows, const float* __restrict__ X, const float* __restrict__ nrm,
                              uint16_t* __restrict__ Xb) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < rows * kMlpIn;
       i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % kMlpIn);
    Xb[i] = f2bf((X[i] - nrm[c]) / nrm[kMlpIn + c]);
  }
}

// Adam (Kingma & Ba 2015) on FP32 master weights
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

static unsigned grid1(long n) {
  long g = (n + 255) / 256;
  return (unsigned)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
}

// ------------------------------------------------------------- host side
void mlp_to_bf16(long n, const float* x, uint16_t* y, hipStream_t s) {
  k_to_bf16<<<grid1(n), 256, 0, s>>>(n, x, y);
}

void mlp_weights_to_bf16(const MlpBuffers& b, hipStream_t s) {
  k_to_bf16<<<grid1(kMlpParams), 256, 0, s>>>(kMlpParams, b.params, b.pbf);
}

void mlp_forward(const MlpBuffers& b, int rows, const uint16_t* Xb, float* Y, hipStream_t s) {
  const float* P = b.params;
  const uint16_t* Pb = b.pbf;
  // Z1 = X W1^T + b1, H1 = tanh(Z1)
  gemm<1>(rows, kMlpHid, kMlpIn, Xb, kMlpIn, 1, Pb + kOffW1, 1, kMlpIn, b.H1, kMlpHid, P + kOffB1,
          nullptr, b.H1b, s);
  gemm<1>(rows, kMlpHid, kMlpHid, b.H1b, kMlpHid, 1, Pb + kOffW2, 1, kMlpHid, b.H2, kMlpHid,
          P + kOffB2, nullptr, b.H2b, s);
  gemm<2>(rows, kMlpOut, kMlpHid, b.H2b, kMlpHid, 1, Pb + kOffW3, 1, kMlpHid, Y, kMlpOut,
          P + kOffB3, nullptr, nullptr, s);
}

void mlp_backward(const MlpBuffers& b, int rows, const uint16_t* Xb, const float* T, float* grads,
                  double* part, hipStream_t s) {
  const uint16_t* Pb = b.pbf;
  const long ny = (long)rows * kMlpOut;
  const unsigned gm = grid1(ny);
  k_mse<<<gm, 256, 0, s>>>(ny, b.Y, T, b.dY, b.dYb, part);
  // output layer
  gemm<0>(kMlpOut, kMlpHid, rows, b.dYb, 1, kMlpOut, b.H2b, kMlpHid, 1, grads + kOffW3, kMlpHid,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<kMlpOut, 256, 0, s>>>(rows, kMlpOut, b.dY, grads + kOffB3);
  // dZ2 = (dY W3) (1 - H2^2)
  gemm<3>(rows, kMlpHid, kMlpOut, b.dYb, kMlpOut, 1, Pb + kOffW3, kMlpHid, 1, b.dZ2, kMlpHid,
          nullptr, b.H2, b.dZ2b, s);
  gemm<0>(kMlpHid, kMlpHid, rows, b.dZ2b, 1, kMlpHid, b.H1b, kMlpHid, 1, grads + kOffW2, kMlpHid,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<kMlpHid, 256, 0, s>>>(rows, kMlpHid, b.dZ2, grads + kOffB2);
  // dZ1 = (dZ2 W2) (1 - H1^2)
  gemm<3>(rows, kMlpHid, kMlpHid, b.dZ2b, kMlpHid, 1, Pb + kOffW2, kMlpHid, 1, b.dZ1, kMlpHid,
          nullptr, b.H1, b.dZ1b, s);
  gemm<0>(kMlpHid, kMlpIn, rows, b.dZ1b, 1, kMlpHid, Xb, kMlpIn, 1, grads + kOffW1, kMlpIn,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<kMlpHid, 256, 0, s>>>(rows, kMlpHid, b.dZ1, grads + kOffB1);
}

int mlp_loss_parts(int rows) { return (int)grid1((long)rows * kMlpOut); }

void mlp_adam(const MlpBuffers& b, float lr, long t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  k_adam<<<grid1(kMlpParams), 256, 0, s>>>(kMlpParams, b.params, b.grads, b.m, b.v, lr, b1, b2, eps,
                                           c1, c2);
}

void mlp_gather_batch(const MlpBuffers& b, int rows, long count, uint32_t seed, uint32_t step,
                      hipStream_t s) {
  k_gather_batch<<<(rows + 255) / 256, 256, 0, s>>>(rows, count, seed, step, b.logX, b.logT, b.norm,
                                                    b.Xb, b.Tb);
}

void mlp_norm_inputs(const MlpBuffers& b, long rows, const float* X, uint16_t* Xb, hipStream_t s) {
  k_norm_inputs<<<grid1(rows * kMlpIn), 256, 0, s>>>(rows, X, b.norm, Xb);
}

void mlp_stats(const MlpBuffers& b, long count, hipStream_t s) {
  k_colstats<<<kMlpIn, 256, 0, s>>>(count, kMlpIn, b.logX, b.norm, b.norm + kMlpIn);
  k_colstats<<<kMlpOut, 256, 0, s>>>(count, kMlpOut, b.logT, b.norm + 2 * kMlpIn,
                                     b.norm + 2 * kMlpIn + kMlpOut);
}

}  // namespace cpfft
