// green.h -- the Green projection of G_K_dF at one frequency, shared by the
// generic and the compile-time-planned FFT paths.
//
// Frequency (x, y, kz) of an N^3 grid (x, y in [0, N), kz in [0, N/2]) is
// f = (x or x-N, y or y-N, kz).  Textbook operator (FFT_init.f:272-385 on the
// DFT frequencies): out_ij = q_j (A q)_i / |q|^2 with q = f, and for even N
// the Nyquist planes removed.  Reference operator for even N
// (G_K_dF.f:20-87, the phase shift by Nhalf = N/2+1 bins with the real part
// taken): M(f) = (Gamma(w(f-1)) + Gamma(w(-f-1)))/2, w wrapping to the
// symmetric range.
#pragma once

#include "cpfft_common.h"

namespace cpfft {

// One row of ddot42n (G_K_dF.f:241-268): sum_c K[c] F[c] with the
// reference's pairwise order t1 + (((t2+t6)+(t4+t8)) + ((t3+t7)+(t5+t9)));
// products are rounded before the sum (no contraction) wherever it is used,
// so every kernel that forms K4:F gets the same bits.
__device__ inline double ddot42n_row(const double* K, const double* F) {
#pragma clang fp contract(off)
  double t[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) t[d] = K[d] * F[d];
  return t[0] + (((t[1] + t[5]) + (t[3] + t[7])) + ((t[2] + t[6]) + (t[4] + t[8])));
}

// out += w * q_j (A q)_i / |q|^2
__device__ inline void gamma_acc(const double q[3], const cplx A[9], cplx out[9], double w) {
  double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  if (qq <= 1e-10) return;
  double s = w / qq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    cplx aq{A[3 * i].re * q[0] + A[3 * i + 1].re * q[1] + A[3 * i + 2].re * q[2],
            A[3 * i].im * q[0] + A[3 * i + 1].im * q[1] + A[3 * i + 2].im * q[2]};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      out[3 * i + j].re += aq.re * q[j] * s;
      out[3 * i + j].im += aq.im * q[j] * s;
    }
  }
}

__device__ inline int wrapf(int v, int N) {
  // representative in [-N/2, N/2) for even N, (-N/2, N/2] for odd N
  int h = N / 2;
  int m = ((v + h) % N + N) % N;
  return m - h;
}

__device__ inline void project9(int N, int gamma_mode, int x, int y, int kz, const cplx A[9],
                                cplx out[9]) {
#pragma unroll
  for (int c = 0; c < 9; ++c) out[c] = cplx{0.0, 0.0};
  const bool even = (N % 2) == 0;
  const int f[3] = {x <= N / 2 ? x : x - N, y <= N / 2 ? y : y - N, kz};
  if (gamma_mode == 0 && even) {
    double q1[3], q2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      q1[a] = (double)wrapf(f[a] - 1, N);
      q2[a] = (double)wrapf(-f[a] - 1, N);
    }
    gamma_acc(q1, A, out, 0.5);
    gamma_acc(q2, A, out, 0.5);
  } else {
    const bool nyq = even && (abs(f[0]) == N / 2 || abs(f[1]) == N / 2 || f[2] == N / 2);
    if (!nyq) {
      double q[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) q[a] = (double)wrapf(f[a], N);
      gamma_acc(q, A, out, 1.0);
    }
  }
}

}  // namespace cpfft
