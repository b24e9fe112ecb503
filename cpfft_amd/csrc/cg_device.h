// Device helpers of the CG solver (fftPcg, FFT_nr3.f:214-360): the CG
// scalars, the fixed-order sum of per-block partial sums and the
// convergence test.  The sums are finished by single-block kernels: a
// "last block of the launch" hand-off would need a device-scope release
// fence per block, which on a chip with one L2 per XCD writes back that L2
// each time and cost more than the launch it saves.
#pragma once
#include <hip/hip_runtime.h>

namespace cpfft {

// CG state kept on the device, one struct per system being solved.
struct CGScalars {
  double rr;       // r.r of the current residual
  double alpha, beta;
  double tolb, tol;
  double resnorm;  // |r| after the last update
  int k;           // iterations completed
  int maxit;
  int done;        // 0 running, 1 converged, 2 iteration limit reached
  int pad;
};

// Sum of v over a block of B threads (B a power of two, at least 64), in
// the order of the halving tree "x[t] += x[t + s] for t < s, s = B/2, ...,
// 1" and so with its bits, but with one barrier instead of one per level:
// after the values pass through LDS, each lane of wave 0 folds its column
// x[t], x[t+64], ... in the tree's order for s >= 64, and the last six
// levels run as shuffles inside the wave (lane t < s adds lane t + s, as
// the tree does).  Every thread must call it; the sum is valid in thread 0
// only.  A caller that uses sh again must pass a barrier first.
template <int B>
__device__ inline double tree_sum0(double v, double* sh) {
  static_assert(B >= 64 && (B & (B - 1)) == 0, "block size must be a power of two >= 64");
  const int t = threadIdx.x;
  if constexpr (B > 64) {
    sh[t] = v;
    __syncthreads();
    if (t < 64) {
      double y[B / 64];
#pragma unroll
      for (int j = 0; j < B / 64; ++j) y[j] = sh[t + 64 * j];
#pragma unroll
      for (int s = B / 128; s > 0; s >>= 1)
#pragma unroll
        for (int j = 0; j < s; ++j) y[j] += y[j + s];
      v = y[0];
    }
  }
  if (t < 64) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  }
  return v;
}

// Sum of part[0..np) in a fixed order for a block of B threads: thread t
// adds part[t], part[t+B], part[t+2B], ... in turn, then the halving tree
// over the block (tree_sum0).  The loads are issued eight at a time ahead
// of the additions, which leaves the order (and so the bits) of the sum
// unchanged.  Every thread of the block must call it; the sum is valid in
// thread 0.
template <int B>
__device__ inline double ordered_sum(const double* part, int np, double* sh) {
  const int t = threadIdx.x;
  double acc = 0.0;
  int i = t;
  for (; i + 7 * B < np; i += 8 * B) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[i + u * B];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; i < np; i += B) acc += part[i];
  return tree_sum0<B>(acc, sh);
}

// Convergence test and direction coefficient after r has been updated
// (FFT_nr3.f:304-306): converged at |r| <= tol |b| or |r| <= tol.  A solve
// that needed maxit iterations fails even when the last one passed the
// test, as fftPcg aborts on itercount >= maxIter (FFT_nr3.f:335-339).
__device__ inline void cg_check(CGScalars* cs, double rr1) {
  cs->k += 1;
  const double res = sqrt(rr1);
  cs->resnorm = res;
  if (res <= cs->tolb || res <= cs->tol) {
    cs->done = cs->k >= cs->maxit ? 2 : 1;
  } else {
    cs->beta = rr1 / cs->rr;
    cs->rr = rr1;
    if (cs->k >= cs->maxit) cs->done = 2;
  }
}

}  // namespace cpfft
