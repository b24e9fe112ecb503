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

// (A q) for one row (A3 = A_i1, A_i2, A_i3).  Here and in gamma_term
// products are rounded before the sums, so every kernel that projects gets
// the same bits.
__device__ __forceinline__ cplx gamma_aq(const double q[3], const cplx A3[3]) {
#pragma clang fp contract(off)
  return cplx{A3[0].re * q[0] + A3[1].re * q[1] + A3[2].re * q[2],
              A3[0].im * q[0] + A3[1].im * q[1] + A3[2].im * q[2]};
}
// o += (A q)_i q_j s
__device__ __forceinline__ void gamma_term(cplx aq, double qj, double s, cplx& o) {
#pragma clang fp contract(off)
  o.re += aq.re * qj * s;
  o.im += aq.im * qj * s;
}

// out += w * q_j (A q)_i / |q|^2
__device__ inline void gamma_acc(const double q[3], const cplx A[9], cplx out[9], double w) {
  double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  if (qq <= 1e-10) return;
  double s = w / qq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const cplx aq = gamma_aq(q, A + 3 * i);
#pragma unroll
    for (int j = 0; j < 3; ++j) gamma_term(aq, q[j], s, out[3 * i + j]);
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

// project9 one component at a time, for kernels that hold one tensor row.
// proj_vectors: the wave vectors of project9 at one frequency (q2 for the
// reference's even-N operator only); proj_weights: their weights
// s = w / |q|^2, 0 where project9 adds nothing (|q|^2 <= 1e-10, the
// textbook operator's Nyquist planes); project1w: component (i, j) of
// project9's result from row i of A (A3) and those weights.  Where a term is
// added the operations and operands are project9's; a zero weight adds a
// signed zero, which leaves the sum's bits as project9 leaves them.
__device__ inline bool proj_vectors(int N, int gamma_mode, int x, int y, int kz, double q1[3],
                                    double q2[3]) {
  const bool even = (N % 2) == 0;
  const int f[3] = {x <= N / 2 ? x : x - N, y <= N / 2 ? y : y - N, kz};
  const bool two = gamma_mode == 0 && even;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q1[a] = (double)wrapf(two ? f[a] - 1 : f[a], N);
    q2[a] = (double)wrapf(-f[a] - 1, N);
  }
  return two;
}
__device__ inline void proj_weights(int N, int gamma_mode, int x, int y, int kz, double& s1,
                                    double& s2) {
  double q1[3], q2[3];
  const bool two = proj_vectors(N, gamma_mode, x, y, kz, q1, q2);
  const bool even = (N % 2) == 0;
  const int f[3] = {x <= N / 2 ? x : x - N, y <= N / 2 ? y : y - N, kz};
  const bool nyq = !two && even && (abs(f[0]) == N / 2 || abs(f[1]) == N / 2 || f[2] == N / 2);
  const double w = two ? 0.5 : 1.0;
  const double qq1 = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
  const double qq2 = q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2];
  s1 = (nyq || qq1 <= 1e-10) ? 0.0 : w / qq1;
  s2 = (!two || qq2 <= 1e-10) ? 0.0 : w / qq2;
}
__device__ inline cplx project1w(int N, int gamma_mode, int x, int y, int kz, const cplx A3[3],
                                 int j, double s1, double s2) {
  cplx o{0.0, 0.0};
  double q1[3], q2[3];
  const bool two = proj_vectors(N, gamma_mode, x, y, kz, q1, q2);
  gamma_term(gamma_aq(q1, A3), j == 0 ? q1[0] : j == 1 ? q1[1] : q1[2], s1, o);
  if (two) gamma_term(gamma_aq(q2, A3), j == 0 ? q2[0] : j == 1 ? q2[1] : q2[2], s2, o);
  return o;
}

}  // namespace cpfft
