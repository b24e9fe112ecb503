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

// Sum of part[0..np) in a fixed order for a block of B threads: thread t
// adds part[t], part[t+B], part[t+2B], ... in turn, then a tree over the
// block.  The loads are issued eight at a time ahead of the additions,
// which leaves the order (and so the bits) of the sum unchanged.  Every
// thread of the block must call it; all get the sum.
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
  sh[t] = acc;
  __syncthreads();
#pragma unroll
  for (int s = B / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Convergence test and direction coefficient after r has been updated
// (FFT_nr3.f:304-306): converged at |r| <= tol |b| or |r| <= tol.
__device__ inline void cg_check(CGScalars* cs, double rr1) {
  cs->k += 1;
  const double res = sqrt(rr1);
  cs->resnorm = res;
  if (res <= cs->tolb || res <= cs->tol) {
    cs->done = 1;
  } else {
    cs->beta = rr1 / cs->rr;
    cs->rr = rr1;
    if (cs->k >= cs->maxit) cs->done = 2;
  }
}

}  // namespace cpfft
