// mlp.hip -- the learned warm start of the per-voxel Newton solve: a
// nin -> 128 -> 128 -> nout tanh network (nin = 18 + 2 ns, nout = ns for a
// crystal of ns slip systems) trained and evaluated with BF16 matrix-core
// (MFMA) GEMMs and FP32 accumulation.  It has no counterpart in
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
__global__ void k_gather_batch(int Bn, long count, uint32_t seed, uint32_t step, int nin,
                               int nout, const float* __restrict__ logX,
                               const float* __restrict__ logT, const float* __restrict__ nrm,
                               uint16_t* __restrict__ Xb, float* __restrict__ Tb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Bn) return;
  const uint32_t h = mix32(seed * 0x9e3779b9u ^ mix32(step * 0x85ebca6bu ^ (uint32_t)r));
  const long src = (long)(h % (uint32_t)count);
  const float* mx = nrm;
  const float* sx = nrm + nin;
  const float* mt = nrm + 2 * nin;
  const float* st = nrm + 2 * nin + nout;
  for (int c = 0; c < nin; ++c)
    Xb[(long)r * nin + c] = f2bf((logX[src * nin + c] - mx[c]) / sx[c]);
  for (int c = 0; c < nout; ++c) Tb[(long)r * nout + c] = (logT[src * nout + c] - mt[c]) / st[c];
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
  k_to_bf16<<<grid1(b.d.nparam), 256, 0, s>>>(b.d.nparam, b.params, b.pbf);
}

void mlp_forward(const MlpBuffers& b, int rows, const uint16_t* Xb, float* Y, hipStream_t s) {
  const float* P = b.params;
  const uint16_t* Pb = b.pbf;
  const MlpDims& d = b.d;
  // Z1 = X W1^T + b1, H1 = tanh(Z1)
  gemm<1>(rows, kMlpHid, d.nin, Xb, d.nin, 1, Pb + d.offW1, 1, d.nin, b.H1, kMlpHid, P + d.offB1,
          nullptr, b.H1b, s);
  gemm<1>(rows, kMlpHid, kMlpHid, b.H1b, kMlpHid, 1, Pb + d.offW2, 1, kMlpHid, b.H2, kMlpHid,
          P + d.offB2, nullptr, b.H2b, s);
  gemm<2>(rows, d.nout, kMlpHid, b.H2b, kMlpHid, 1, Pb + d.offW3, 1, kMlpHid, Y, d.nout,
          P + d.offB3, nullptr, nullptr, s);
}

void mlp_backward(const MlpBuffers& b, int rows, const uint16_t* Xb, const float* T, float* grads,
                  double* part, hipStream_t s) {
  const uint16_t* Pb = b.pbf;
  const MlpDims& d = b.d;
  const long ny = (long)rows * d.nout;
  const unsigned gm = grid1(ny);
  k_mse<<<gm, 256, 0, s>>>(ny, b.Y, T, b.dY, b.dYb, part);
  // output layer
  gemm<0>(d.nout, kMlpHid, rows, b.dYb, 1, d.nout, b.H2b, kMlpHid, 1, grads + d.offW3, kMlpHid,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<d.nout, 256, 0, s>>>(rows, d.nout, b.dY, grads + d.offB3);
  // dZ2 = (dY W3) (1 - H2^2)
  gemm<3>(rows, kMlpHid, d.nout, b.dYb, d.nout, 1, Pb + d.offW3, kMlpHid, 1, b.dZ2, kMlpHid,
          nullptr, b.H2, b.dZ2b, s);
  gemm<0>(kMlpHid, kMlpHid, rows, b.dZ2b, 1, kMlpHid, b.H1b, kMlpHid, 1, grads + d.offW2, kMlpHid,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<kMlpHid, 256, 0, s>>>(rows, kMlpHid, b.dZ2, grads + d.offB2);
  // dZ1 = (dZ2 W2) (1 - H1^2)
  gemm<3>(rows, kMlpHid, kMlpHid, b.dZ2b, kMlpHid, 1, Pb + d.offW2, kMlpHid, 1, b.dZ1, kMlpHid,
          nullptr, b.H1, b.dZ1b, s);
  gemm<0>(kMlpHid, d.nin, rows, b.dZ1b, 1, kMlpHid, Xb, d.nin, 1, grads + d.offW1, d.nin,
          nullptr, nullptr, nullptr, s);
  k_colsum<<<kMlpHid, 256, 0, s>>>(rows, kMlpHid, b.dZ1, grads + d.offB1);
}

int mlp_loss_parts(const MlpBuffers& b, int rows) { return (int)grid1((long)rows * b.d.nout); }

void mlp_adam(const MlpBuffers& b, float lr, long t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  k_adam<<<grid1(b.d.nparam), 256, 0, s>>>(b.d.nparam, b.params, b.grads, b.m, b.v, lr, b1, b2,
                                           eps, c1, c2);
}

void mlp_gather_batch(const MlpBuffers& b, int rows, long count, uint32_t seed, uint32_t step,
                      hipStream_t s) {
  k_gather_batch<<<(rows + 255) / 256, 256, 0, s>>>(rows, count, seed, step, b.d.nin, b.d.nout,
                                                    b.logX, b.logT, b.norm, b.Xb, b.Tb);
}

void mlp_stats(const MlpBuffers& b, long count, hipStream_t s) {
  const MlpDims& d = b.d;
  k_colstats<<<d.nin, 256, 0, s>>>(count, d.nin, b.logX, b.norm, b.norm + d.nin);
  k_colstats<<<d.nout, 256, 0, s>>>(count, d.nout, b.logT, b.norm + 2 * d.nin,
                                    b.norm + 2 * d.nin + d.nout);
}

// ------------------------------------------------------------ inference
// One launch over all voxels.  The block stages the BF16 weights (rows
// padded by eight elements against LDS bank conflicts) and the FP32 biases
// and output normalisation in LDS; each wavefront then takes 16 voxels at a
// time through the three layers with v_mfma_f32_16x16x32_bf16, the weights
// as the A operand (rows = output units) and the activations as the B
// operand (columns = the 16 voxels).  Lane l supplies A row l%16 and B
// column l%16 at k = 8*(l/16)..+7 of each K step of 32, and holds C rows
// 4*(l/16)..+3 of column l%16.  So after a layer, lane l holds units
// 16t + 4g + j (g = l/16, j < 4) of tile t for voxel l%16, and the eight
// values of tiles 2s and 2s+1 are that lane's B operand of the next layer's
// K step s, once the K index of the next weights is permuted the same way:
// position 8g + 4h + j of K step s is unit 32s + 16h + 4g + j.  The
// weights are stored permuted in LDS, the activations never leave the
// registers.  The rounding points are those of mlp_forward: FP32 bias and
// tanh, BF16 hidden activations, FP32 output.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
static constexpr int kIW = 8;               // wavefronts per inference block
static constexpr int kHS = kMlpHid + 8;     // padded row of W2, W3

// the unit at K position p (< 128) of the permuted order
__device__ __forceinline__ int perm_k(int p) {
  const int s = p >> 5, q = p & 31, g = q >> 3, h = (q >> 2) & 1, j = q & 3;
  return 32 * s + 16 * h + 4 * g + j;
}

template <int NINP, int NOUTP>
__global__ __launch_bounds__(64 * kIW) void k_mlp_infer(long V, int nin, int nout,
                                                        const uint16_t* __restrict__ Xv,
                                                        const uint16_t* __restrict__ pbf,
                                                        const float* __restrict__ params,
                                                        MlpDims d, const float* __restrict__ nrm,
                                                        float* __restrict__ ws) {
  constexpr int S1 = NINP + 8;
  __shared__ __align__(16) uint16_t W1[kMlpHid * S1];
  __shared__ __align__(16) uint16_t W2[kMlpHid * kHS];   // K permuted
  __shared__ __align__(16) uint16_t W3[NOUTP * kHS];     // K permuted
  __shared__ float b1[kMlpHid], b2[kMlpHid], b3[NOUTP], mt[NOUTP], st[NOUTP];
  const int T = 64 * kIW;
  for (int i = threadIdx.x; i < kMlpHid * S1; i += T) {
    const int n = i / S1, k = i - n * S1;
    W1[i] = k < nin ? pbf[d.offW1 + (long)n * nin + k] : 0;
  }
  for (int i = threadIdx.x; i < kMlpHid * kHS; i += T) {
    const int n = i / kHS, k = i - n * kHS;
    W2[i] = k < kMlpHid ? pbf[d.offW2 + (long)n * kMlpHid + perm_k(k)] : 0;
  }
  for (int i = threadIdx.x; i < NOUTP * kHS; i += T) {
    const int n = i / kHS, k = i - n * kHS;
    W3[i] = (n < nout && k < kMlpHid) ? pbf[d.offW3 + (long)n * kMlpHid + perm_k(k)] : 0;
  }
  for (int i = threadIdx.x; i < kMlpHid; i += T) {
    b1[i] = params[d.offB1 + i];
    b2[i] = params[d.offB2 + i];
  }
  for (int i = threadIdx.x; i < NOUTP; i += T) {
    b3[i] = i < nout ? params[d.offB3 + i] : 0.f;
    mt[i] = i < nout ? nrm[2 * nin + i] : 0.f;
    st[i] = i < nout ? nrm[2 * nin + nout + i] : 0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const long nrb = (V + 15) / 16;
  for (long rb = (long)blockIdx.x * kIW + w; rb < nrb; rb += (long)gridDim.x * kIW) {
    const long r = rb * 16 + li;  // this lane's voxel (B column)
    // layer 1: B = the voxel's inputs straight from HBM
    bf16x8 x1[NINP / 32];
#pragma unroll
    for (int ks = 0; ks < NINP / 32; ++ks) {
      if (r < V) {
        x1[ks] = *reinterpret_cast<const bf16x8*>(Xv + r * NINP + ks * 32 + kg * 8);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) x1[ks][j] = (__bf16)0.0f;
      }
    }
    // a layer's output as the next layer's B operand: K step s packs tiles
    // 2s (elements 0-3) and 2s+1 (elements 4-7), two BF16 per 32-bit word
    u32x4 h[kMlpHid / 32];
    // bias, tanh and BF16 rounding of tile t's four values, packed in pairs
    auto act = [&](f32x4 acc, const float* bias, int t, uint32_t* p) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float v0 = tanhf(acc[2 * q] + bias[16 * t + 4 * kg + 2 * q]);
        const float v1 = tanhf(acc[2 * q + 1] + bias[16 * t + 4 * kg + 2 * q + 1]);
        p[q] = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
      }
    };
#pragma unroll
    for (int s2 = 0; s2 < kMlpHid / 32; ++s2) {
      f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < NINP / 32; ++ks) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(W1 + (32 * s2 + li) * S1 + ks * 32 + kg * 8);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(W1 + (32 * s2 + 16 + li) * S1 + ks * 32 + kg * 8);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, x1[ks], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, x1[ks], acc1, 0, 0, 0);
      }
      uint32_t p[4];
      act(acc0, b1, 2 * s2, p);
      act(acc1, b1, 2 * s2 + 1, p + 2);
      h[s2] = u32x4{p[0], p[1], p[2], p[3]};
      // keep the scheduler from hoisting every tile's weight loads (the
      // registers they would hold cost occupancy)
      __builtin_amdgcn_sched_barrier(0);
    }
    // layer 2
    u32x4 h2[kMlpHid / 32];
#pragma unroll
    for (int s2 = 0; s2 < kMlpHid / 32; ++s2) {
      f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kMlpHid / 32; ++ks) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(W2 + (32 * s2 + li) * kHS + ks * 32 + kg * 8);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(W2 + (32 * s2 + 16 + li) * kHS + ks * 32 + kg * 8);
        const bf16x8 b = __builtin_bit_cast(bf16x8, h[ks]);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc1, 0, 0, 0);
      }
      uint32_t p[4];
      act(acc0, b2, 2 * s2, p);
      act(acc1, b2, 2 * s2 + 1, p + 2);
      h2[s2] = u32x4{p[0], p[1], p[2], p[3]};
      __builtin_amdgcn_sched_barrier(0);
    }
    // layer 3 and the output normalisation
#pragma unroll
    for (int jt = 0; jt < NOUTP / 16; ++jt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kMlpHid / 32; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(W3 + (16 * jt + li) * kHS + ks * 32 + kg * 8);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, h2[ks]), acc, 0, 0, 0);
      }
      if (r < V) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = 16 * jt + 4 * kg + j;
          if (n < nout) ws[n * V + r] = (acc[j] + b3[n]) * st[n] + mt[n];
        }
      }
    }
  }
}

void mlp_infer(const MlpBuffers& b, long V, hipStream_t s) {
  const unsigned g = 512;  // two blocks per CU (their weights fit the LDS); each loops over row blocks
  const MlpDims& d = b.d;
  if (mlp_nin_pad(d.nin) == 64 && d.nout <= 16)
    k_mlp_infer<64, 16><<<g, 64 * kIW, 0, s>>>(V, d.nin, d.nout, b.Xv, b.pbf, b.params, d, b.norm,
                                                b.ws);
  else
    k_mlp_infer<96, 32><<<g, 64 * kIW, 0, s>>>(V, d.nin, d.nout, b.Xv, b.pbf, b.params, d, b.norm,
                                                b.ws);
}

}  // namespace cpfft
