// mm10_core.h -- per-voxel finite-strain crystal-plasticity stress update
// (kinematics + mm10 Voce model + consistent tangent + exact dP/dF), the
// body of the drive_eps_sig kernel.  One thread updates one voxel; all
// per-voxel temporaries live in registers.
//
// Reference routines (src/): do_nleps_block drive_eps_sig.f:57-339, polar
// decomposition polar.f:18-307, getrm1 polar.f:680-880, qmply1.f:15-49,
// mm10 mm10_a.f:29-355, mm10_setup mm10_a.f:830-962, mm10_solve_strup
// mm10_a.f:2628-2845, mm10_solve mm10_a.f:2860-3295, mm10_tangent
// mm10_a.f:658-815, mm10_update_rotation mm10_a.f:3310-3415, mm10_output
// mm10_a.f:3429-3689, residual/Jacobian/Voce law mm10_b.f:63-2071,
// cep2A cep2A.f:86-284, cs2p drive_eps_sig.f:1182-1224.
//
// Differences in arithmetic organisation from the reference (results agree
// to rounding): every sum over slip systems is accumulated with the
// orientation's lattice-frame Schmid tensors m_s and spin vectors q_s and
// rotated once (sum_s g_s (E m_s) = E sum_s g_s m_s, with E the rotation by
// Rp_n^T), so no per-voxel copy of the rotated tensors is kept; the power
// law is evaluated once per slip system and evaluation point, by repeated
// squaring when the rate exponent is an integer; the Jacobian used by the
// tangent is re-formed from the saved iterate instead of being kept across
// the line search; dP/dF is assembled one (k,l) column at a time from 3x3
// blocks.  The polar decomposition is evaluated with the reference's
// operation order and without FMA contraction because its Cardano
// eigenvalue step amplifies rounding.
#pragma once

#include "cpfft_common.h"

namespace cpfft {

// ---------------------------------------------------------------- params
struct MatParams {
  double rate_n, tau_y, tau_v, theta_0, voche_m, eps_dot_0_y, dt;
  double atol, atol1, rtol, rtol1;
  int miter, alter_mode, jac_exact, int_rate;  // int_rate: rate_n integral (>=2)
  int m_is_one;                                // voche_m == 1
};

// per-orientation invariants (setup_mm10_rknstr, drive_eps_sig.f:537-1002):
// rotated stiffness C (row-major 6x6), ms (Voigt 6, eng. shear), qs (skew 3)
template <int NS>
struct Orient {
  double C[36];
  double ms[NS][6];
  double qs[NS][3];
};

struct Counters {
  int pred_it, upd_it, cuts, failed, ls_evals;
  int ws_fallback;  // warm-started solves that failed and were redone from the plain start
};

// ---------------------------------------------------------------- helpers
CPFFT_HD inline double sgn1(double x) { return copysign(1.0, x); }

CPFFT_HD inline double ipow(double x, int n) {
  double r = 1.0, b = x;
  while (n > 0) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

// |u|^(n-2) for the power law (n = rate exponent)
CPFFT_HD inline double pw_nm2(const MatParams& p, double au) {
  return p.int_rate ? ipow(au, (int)p.rate_n - 2) : pow(au, p.rate_n - 2.0);
}

// Gaussian elimination with partial pivoting on [A | B] followed by back
// substitution, N <= 7, row-major A (destroyed).  Every index is static
// once the loops are unrolled (the row exchange of each step is done with
// selects), so A and B stay in registers.  The arithmetic is that of an LU
// factorisation followed by forward and back substitution.
template <int N, int NR>
CPFFT_HD inline void lu_solve(double (&A)[N][N], double (&B)[N][NR]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    int p = j;
    double amax = fabs(A[j][j]);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      const double v = fabs(A[i][j]);
      if (v > amax) { amax = v; p = i; }
    }
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      const bool sw = (p == i);
#pragma unroll
      for (int k = j; k < N; ++k) {
        const double a = A[j][k], b = A[i][k];
        A[j][k] = sw ? b : a;
        A[i][k] = sw ? a : b;
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double a = B[j][r], b = B[i][r];
        B[j][r] = sw ? b : a;
        B[i][r] = sw ? a : b;
      }
    }
    const double rp = 1.0 / A[j][j];
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      const double l = A[i][j] * rp;
#pragma unroll
      for (int k = j + 1; k < N; ++k) A[i][k] -= l * A[j][k];
#pragma unroll
      for (int r = 0; r < NR; ++r) B[i][r] -= l * B[j][r];
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      double s = B[i][r];
#pragma unroll
      for (int k = i + 1; k < N; ++k) s -= A[i][k] * B[k][r];
      B[i][r] = s / A[i][i];
    }
  }
}

CPFFT_HD inline void rt2rve(const double rt[3][3], double rv[6][6]) {
  for (int r = 0; r < 3; ++r) {
    rv[r][0] = rt[r][0] * rt[r][0];
    rv[r][1] = rt[r][1] * rt[r][1];
    rv[r][2] = rt[r][2] * rt[r][2];
    rv[r][3] = 2.0 * rt[r][0] * rt[r][1];
    rv[r][4] = 2.0 * rt[r][2] * rt[r][1];
    rv[r][5] = 2.0 * rt[r][0] * rt[r][2];
  }
  const int a[3] = {0, 1, 0}, b[3] = {1, 2, 2};
  for (int q = 0; q < 3; ++q) {
    int i = a[q], j = b[q];
    rv[3 + q][0] = rt[i][0] * rt[j][0];
    rv[3 + q][1] = rt[i][1] * rt[j][1];
    rv[3 + q][2] = rt[i][2] * rt[j][2];
    rv[3 + q][3] = rt[i][0] * rt[j][1] + rt[j][0] * rt[i][1];
    rv[3 + q][4] = rt[i][1] * rt[j][2] + rt[i][2] * rt[j][1];
    rv[3 + q][5] = rt[i][0] * rt[j][2] + rt[i][2] * rt[j][0];
  }
}

CPFFT_HD inline void rt2rvw(const double rt[3][3], double rv[3][3]) {
  rv[0][0] = rt[1][1] * rt[2][2] - rt[1][2] * rt[2][1];
  rv[0][1] = rt[1][0] * rt[2][2] - rt[1][2] * rt[2][0];
  rv[0][2] = rt[1][0] * rt[2][1] - rt[1][1] * rt[2][0];
  rv[1][0] = rt[0][1] * rt[2][2] - rt[0][2] * rt[2][1];
  rv[1][1] = rt[0][0] * rt[2][2] - rt[0][2] * rt[2][0];
  rv[1][2] = rt[0][0] * rt[2][1] - rt[0][1] * rt[2][0];
  rv[2][0] = rt[0][1] * rt[1][2] - rt[0][2] * rt[1][1];
  rv[2][1] = rt[0][0] * rt[1][2] - rt[0][2] * rt[1][0];
  rv[2][2] = rt[0][0] * rt[1][1] - rt[0][1] * rt[1][0];
}

// mm10_symsw (mm10_b.f:1505-1533)
CPFFT_HD inline void symsw(const double s[6], const double w[3], double sw[6]) {
  sw[0] = s[3] * w[2] - s[5] * w[1];
  sw[1] = s[3] * w[2] - s[4] * w[0];
  sw[2] = s[5] * w[1] + s[4] * w[0];
  sw[3] = 0.5 * (w[2] * (s[0] - s[1]) + w[0] * s[5] - w[1] * s[4]);
  sw[4] = 0.5 * (w[0] * (s[1] - s[2]) + w[1] * s[3] + w[2] * s[5]);
  sw[5] = 0.5 * (w[1] * (s[0] - s[2]) + w[0] * s[3] - w[2] * s[4]);
}

// ---------------------------------------------------------- kinematics
// Polar decomposition R = F U^-1 with Cardano eigenvalues (polar.f:18-307).
// Operation order follows the reference; no FMA contraction.
CPFFT_HD inline void polar_R(const double f[3][3], double r[3][3]) {
#pragma clang fp contract(off)
  double c[6], cc[6];
  c[0] = f[0][0] * f[0][0] + f[1][0] * f[1][0] + f[2][0] * f[2][0];
  c[1] = f[0][0] * f[0][1] + f[1][0] * f[1][1] + f[2][0] * f[2][1];
  c[2] = f[0][1] * f[0][1] + f[1][1] * f[1][1] + f[2][1] * f[2][1];
  c[3] = f[0][0] * f[0][2] + f[1][0] * f[1][2] + f[2][0] * f[2][2];
  c[4] = f[0][1] * f[0][2] + f[1][1] * f[1][2] + f[2][1] * f[2][2];
  c[5] = f[0][2] * f[0][2] + f[1][2] * f[1][2] + f[2][2] * f[2][2];
  cc[0] = c[0] * c[0] + c[1] * c[1] + c[3] * c[3];
  cc[1] = c[0] * c[1] + c[1] * c[2] + c[3] * c[4];
  cc[2] = c[1] * c[1] + c[2] * c[2] + c[4] * c[4];
  cc[3] = c[0] * c[3] + c[1] * c[4] + c[3] * c[5];
  cc[4] = c[1] * c[3] + c[2] * c[4] + c[4] * c[5];
  cc[5] = c[3] * c[3] + c[4] * c[4] + c[5] * c[5];
  const double third = 0.3333333333333333333, oneroot3 = 0.5773502691896258;
  double m11 = c[0], m12 = c[1], m13 = c[3], m22 = c[2], m23 = c[4], m33 = c[5];
  double de = m12 * m23, dd = m12 * m12, ee = m23 * m23, ff = m13 * m13;
  double m = m11 + m22 + m33;
  double c1 = (m11 * m22 + m11 * m33 + m22 * m33) - (dd + ee + ff);
  double c0 = m33 * dd + m11 * ee + m22 * ff - m11 * m22 * m33 - 2.0 * m13 * de;
  double p = m * m - 3.0 * c1;
  double q = m * (p - 1.5 * c1) - 13.5 * c0;
  double sqrtp = sqrt(fabs(p));
  double phi = 27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0));
  phi = third * atan2(sqrt(fabs(phi)), q);
  double cphi = sqrtp * cos(phi);
  double sphi = oneroot3 * sqrtp * sin(phi);
  double e2 = third * (m - cphi);
  double e3 = e2 + sphi;
  double e1 = e2 + cphi;
  e2 = e2 - sphi;
  double t;
  if (e2 < e1) { t = e1; e1 = e2; e2 = t; }
  if (e3 < e1) { t = e1; e1 = e3; e3 = t; }
  if (e3 < e2) { t = e2; e2 = e3; e3 = t; }
  e1 = sqrt(e1);
  e2 = sqrt(e2);
  e3 = sqrt(e3);
  double iu = e1 + e2 + e3;
  double iiu = e1 * e2 + e2 * e3 + e1 * e3;
  double iiiu = e1 * e2 * e3;
  double a2 = 1.0 / (iiiu * (iu * iiu - iiiu));
  double b2 = iu * iiu * iiu - iiiu * (iu * iu + iiu);
  double c2 = -iiiu - iu * (iu * iu - 2.0 * iiu);
  double d2 = iu;
  double ui[6];
  ui[0] = a2 * (b2 + c2 * c[0] + d2 * cc[0]);
  ui[1] = a2 * (c2 * c[1] + d2 * cc[1]);
  ui[2] = a2 * (b2 + c2 * c[2] + d2 * cc[2]);
  ui[3] = a2 * (c2 * c[3] + d2 * cc[3]);
  ui[4] = a2 * (c2 * c[4] + d2 * cc[4]);
  ui[5] = a2 * (b2 + c2 * c[5] + d2 * cc[5]);
  for (int i = 0; i < 3; ++i) {
    r[i][0] = f[i][0] * ui[0] + f[i][1] * ui[1] + f[i][2] * ui[3];
    r[i][1] = f[i][0] * ui[1] + f[i][1] * ui[2] + f[i][2] * ui[4];
    r[i][2] = f[i][0] * ui[3] + f[i][1] * ui[4] + f[i][2] * ui[5];
  }
}

// inv33 (drive_eps_sig.f:1017-1107)
CPFFT_HD inline double inv33(const double a[3][3], double g[3][3]) {
#pragma clang fp contract(off)
  double j1 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  double j2 = a[1][0] * a[2][2] - a[1][2] * a[2][0];
  double j3 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  double d = a[0][0] * j1 - a[0][1] * j2 + a[0][2] * j3;
  g[0][0] = j1 / d;
  g[1][0] = -j2 / d;
  g[2][0] = j3 / d;
  g[0][1] = (a[2][1] * a[0][2] - a[0][1] * a[2][2]) / d;
  g[1][1] = (a[0][0] * a[2][2] - a[2][0] * a[0][2]) / d;
  g[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / d;
  g[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  g[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / d;
  g[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
  return d;
}

// ---------------------------------------------------------- material
// Constants of the local problem at one (sub)step of one voxel (the
// crystal_state of mm10_setup, mm10_a.f:830-962).  The current Schmid
// tensor and spin vector of slip system s are E m_s and W q_s, with m_s and
// q_s the orientation's values (Orient::ms, Orient::qs).
struct Point {
  double D[6];      // unrotated strain increment
  double dg, tinc;  // slip-rate pre-factor, time increment
  double tn[3][3];  // Rp_n^T; E = mm10_rt2rve(tn) is applied through rotE/rotET
  double W[3][3];   // spin rotation by R_{n+1} Rp_n^T (mm10_rt2rvw)
};

// E v with E = rt2rve(t): the Voigt components (11,22,33,12,23,13) of
// t V t^T, V the symmetric tensor with those components
CPFFT_HD inline void rotE(const double t[3][3], const double v[6], double o[6]) {
  const double V[3][3] = {{v[0], v[3], v[5]}, {v[3], v[1], v[4]}, {v[5], v[4], v[2]}};
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = t[i][0] * V[0][j] + t[i][1] * V[1][j] + t[i][2] * V[2][j];
  auto tv = [&](int i, int j) { return A[i][0] * t[j][0] + A[i][1] * t[j][1] + A[i][2] * t[j][2]; };
  o[0] = tv(0, 0);
  o[1] = tv(1, 1);
  o[2] = tv(2, 2);
  o[3] = tv(0, 1);
  o[4] = tv(1, 2);
  o[5] = tv(0, 2);
}
// E^T s with E = rt2rve(t): X = t^T S t with S the symmetric tensor whose
// off-diagonal entries are half the shear components of s; the result is
// (X11, X22, X33, 2 X12, 2 X23, 2 X13)
CPFFT_HD inline void rotET(const double t[3][3], const double s[6], double o[6]) {
  const double S[3][3] = {{s[0], 0.5 * s[3], 0.5 * s[5]},
                          {0.5 * s[3], s[1], 0.5 * s[4]},
                          {0.5 * s[5], 0.5 * s[4], s[2]}};
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = S[i][0] * t[0][j] + S[i][1] * t[1][j] + S[i][2] * t[2][j];
  auto tv = [&](int i, int j) { return t[0][i] * A[0][j] + t[1][i] * A[1][j] + t[2][i] * A[2][j]; };
  o[0] = tv(0, 0);
  o[1] = tv(1, 1);
  o[2] = tv(2, 2);
  o[3] = 2.0 * tv(0, 1);
  o[4] = 2.0 * tv(1, 2);
  o[5] = 2.0 * tv(0, 2);
}

struct StateN {
  double stress[6];
  double Rp[3][3];
  double tt, tt_rate, D[6];
};

CPFFT_HD inline void mv6(const double A[6][6], const double v[6], double o[6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a)
    o[a] = A[a][0] * v[0] + A[a][1] * v[1] + A[a][2] * v[2] + A[a][3] * v[3] + A[a][4] * v[4] +
           A[a][5] * v[5];
}
CPFFT_HD inline void mtv6(const double A[6][6], const double v[6], double o[6]) {
#pragma unroll
  for (int b = 0; b < 6; ++b)
    o[b] = A[0][b] * v[0] + A[1][b] * v[1] + A[2][b] * v[2] + A[3][b] * v[3] + A[4][b] * v[4] +
           A[5][b] * v[5];
}
CPFFT_HD inline void mv3(const double A[3][3], const double v[3], double o[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) o[a] = A[a][0] * v[0] + A[a][1] * v[1] + A[a][2] * v[2];
}
// C (row-major 6x6 in memory) times v
CPFFT_HD inline void Cv6(const double* __restrict__ C, const double v[6], double o[6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a)
    o[a] = C[6 * a] * v[0] + C[6 * a + 1] * v[1] + C[6 * a + 2] * v[2] + C[6 * a + 3] * v[3] +
           C[6 * a + 4] * v[4] + C[6 * a + 5] * v[5];
}

// slip increment of one system: u = rs/tt, p = |u|^(n-2),
// slip = dg |u|^(n-1) u (mm10_slipinc, mm10_b.f:1676-1723, Voce)
struct Slip {
  double u, au, p, g;
};
CPFFT_HD inline Slip slip_of(const MatParams& P, double dg, double rs, double tt) {
  Slip r;
  r.u = rs / tt;
  r.au = fabs(r.u);
  r.p = pw_nm2(P, r.au);
  r.g = dg * (r.p * r.au) * r.u;
  return r;
}

// Sums over slip systems at stress s and resistance tt, in the lattice
// frame: mb = sum_s slip_s m_s, qb = sum_s slip_s q_s, sabs = sum |slip_s|.
// The resolved shear is rs_s = s . (E m_s) = (E^T s) . m_s.
template <int NS>
CPFFT_HD inline void slip_sums(const MatParams& P, const Orient<NS>& o, const Point& pt,
                               const double s[6], double tt, double mb[6], double qb[3],
                               double* sabs) {
  double sE[6];
  rotET(pt.tn, s, sE);
  double sa = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) mb[k] = 0.0;
  qb[0] = qb[1] = qb[2] = 0.0;
#pragma unroll 2
  for (int i = 0; i < NS; ++i) {
    const double* __restrict__ m = o.ms[i];
    const double* __restrict__ q = o.qs[i];
    const double rs = sE[0] * m[0] + sE[1] * m[1] + sE[2] * m[2] + sE[3] * m[3] + sE[4] * m[4] +
                      sE[5] * m[5];
    const Slip sl = slip_of(P, pt.dg, rs, tt);
#pragma unroll
    for (int k = 0; k < 6; ++k) mb[k] += sl.g * m[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) qb[k] += sl.g * q[k];
    sa += fabs(sl.g);
  }
  *sabs = sa;
}

// R1 (mm10_formR1, mm10_b.f:1065-1113): s - s_n - C (D - dbar) + 2 sym(s W)
// with dbar = E mb, wp = W qb
template <int NS>
CPFFT_HD inline void form_R1(const MatParams& P, const Orient<NS>& o, const Point& pt,
                             const StateN& n, const double s[6], double tt, double R1[6],
                             double* sabs) {
  double mb[6], qb[3];
  slip_sums<NS>(P, o, pt, s, tt, mb, qb, sabs);
  double dbar[6], wp[3], sw[6], w1[6], c[6];
  rotE(pt.tn, mb, dbar);
  mv3(pt.W, qb, wp);
  symsw(s, wp, sw);
#pragma unroll
  for (int k = 0; k < 6; ++k) w1[k] = pt.D[k] - dbar[k];
  Cv6(o.C, w1, c);
#pragma unroll
  for (int a = 0; a < 6; ++a) R1[a] = s[a] - n.stress[a] - c[a] + 2.0 * sw[a];
}

// tt from the Voce law (mm10_h_voche, mm10_b.f:1888-1916), tau_l = 0
CPFFT_HD inline double voce_h(const MatParams& P, const StateN& n, double tt, double sabs) {
  double h_term = 1.0 - (tt - P.tau_y) / P.tau_v;
  double f = P.m_is_one ? h_term : pow(fabs(h_term), P.voche_m) * sgn1(h_term);
  return n.tt + P.theta_0 * (f * sabs);
}

// full residual R (mm10_formR, mm10_b.f:1029-1063); returns |R|^2 and tt_rate
template <int NS>
CPFFT_HD inline double form_R(const MatParams& P, const Orient<NS>& o, const Point& pt,
                              const StateN& n, const double x[7], double R[7], double* tt_rate) {
  double sabs;
  form_R1<NS>(P, o, pt, n, x, x[6], R, &sabs);
  double h = voce_h(P, n, x[6], sabs);
  R[6] = x[6] - h;
  *tt_rate = (h - n.tt) / pt.tinc;
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) d += R[k] * R[k];
  return d;
}

// Jacobian (mm10_formJ11/J12/J21/J22, mm10_b.f:177-481, 1918-1983).
// NJ = 6: stress block only (predictor); NJ = 7: full system.
// J11 = I + IW(wp) + sum_s dgdt_s (C E m_s + 2 sym(s, W q_s)) (E m_s)^T
//     = I + IW(wp) + (C E M + 2 [sym(s, (W Q)_c)]_c) E^T
// with M = sum_s dgdt_s m_s m_s^T (upper triangle kept) and
// Q = sum_s dgdt_s q_s m_s^T.
template <int NS, int NJ>
CPFFT_HD inline void form_J(const MatParams& P, const Orient<NS>& o, const Point& pt,
                            const double s[6], double tt, double (&J)[NJ][NJ]) {
  const double n_ = P.rate_n;
  const double h_term = 1.0 - (tt - P.tau_y) / P.tau_v;
  const double fh = P.m_is_one ? h_term : pow(fabs(h_term), P.voche_m) * sgn1(h_term);
  const double absh_m = P.m_is_one ? fabs(h_term) : pow(fabs(h_term), P.voche_m);
  double sE[6];
  rotET(pt.tn, s, sE);
  double M[6][6], Q[3][6], mb[6], qb[3], mf[6];
  double etau = 0.0;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
#pragma unroll
    for (int b = 0; b < 6; ++b) M[a][b] = 0.0;
    mb[a] = 0.0;
    mf[a] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    qb[k] = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) Q[k][b] = 0.0;
  }
#pragma unroll 2
  for (int i = 0; i < NS; ++i) {
    const double* __restrict__ m = o.ms[i];
    const double* __restrict__ q = o.qs[i];
    const double rs = sE[0] * m[0] + sE[1] * m[1] + sE[2] * m[2] + sE[3] * m[3] + sE[4] * m[4] +
                      sE[5] * m[5];
    const Slip sl = slip_of(P, pt.dg, rs, tt);
    const double dgdt = pt.dg * n_ / tt * (sl.p * sl.au);  // d slip / d rs
    double dm[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) dm[a] = dgdt * m[a];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) M[a][b] += dm[a] * m[b];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int b = 0; b < 6; ++b) Q[k][b] += q[k] * dm[b];
#pragma unroll
    for (int k = 0; k < 6; ++k) mb[k] += sl.g * m[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) qb[k] += sl.g * q[k];
    if (NJ == 7) {
      const double f21 = fh * sl.p * sl.u;  // |h|^m sgn(h) |u|^(n-2) u
#pragma unroll
      for (int k = 0; k < 6; ++k) mf[k] += f21 * m[k];
      const double as = fabs(sl.g);
      const double sgs = P.jac_exact ? 1.0 : sgn1(sl.g);
      etau += (P.voche_m * (-1.0 / P.tau_v) * as / fabs(h_term) -
               as * n_ / tt * sgn1(h_term) * sgs) *
              absh_m;
    }
  }
  // T[:, c] = C E M[:, c] + 2 sym(s, W Q[:, c]); row a of J11 is E T[a, :]^T
  {
    double T[6][6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double mc[6], em[6], tc[6], wq[3], sw[6];
#pragma unroll
      for (int a = 0; a < 6; ++a) mc[a] = a <= c ? M[a][c] : M[c][a];
      rotE(pt.tn, mc, em);
      Cv6(o.C, em, tc);
#pragma unroll
      for (int k = 0; k < 3; ++k) wq[k] = pt.W[k][0] * Q[0][c] + pt.W[k][1] * Q[1][c] + pt.W[k][2] * Q[2][c];
      symsw(s, wq, sw);
#pragma unroll
      for (int a = 0; a < 6; ++a) T[a][c] = tc[a] + 2.0 * sw[a];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      double ja[6];
      rotE(pt.tn, T[a], ja);
#pragma unroll
      for (int b = 0; b < 6; ++b) J[a][b] = ja[b];
    }
  }
  double wp[3];
  mv3(pt.W, qb, wp);
  if (NJ == 7) {
    // J12 = -n/tt sum_s slip_s (C E m_s + 2 sym(s, W q_s)) = -n/tt (C E mb + 2 sym(s, wp))
    double emb[6], cm[6], sw[6];
    rotE(pt.tn, mb, emb);
    Cv6(o.C, emb, cm);
    symsw(s, wp, sw);
#pragma unroll
    for (int a = 0; a < 6; ++a) J[a][NJ - 1] = (cm[a] + 2.0 * sw[a]) * (-n_ / tt);
    // J21 = -theta_0 dg n/tt E mf, J22 = 1 - theta_0 etau
    double emf[6];
    rotE(pt.tn, mf, emf);
    const double fac = P.theta_0 * pt.dg * n_ / tt;
#pragma unroll
    for (int b = 0; b < 6; ++b) J[NJ - 1][b] = -emf[b] * fac;
    J[NJ - 1][NJ - 1] = 1.0 - P.theta_0 * etau;
  }
  // IW(wp) (mm10_iw, mm10_b.f:1605-1634) and identity
  J[0][3] += 2.0 * wp[2];
  J[0][5] -= 2.0 * wp[1];
  J[1][3] += 2.0 * wp[2];
  J[1][4] -= 2.0 * wp[0];
  J[2][4] += 2.0 * wp[0];
  J[2][5] += 2.0 * wp[1];
  J[3][0] += wp[2];
  J[3][1] -= wp[2];
  J[3][4] -= wp[1];
  J[3][5] += wp[0];
  J[4][1] += wp[0];
  J[4][2] -= wp[0];
  J[4][3] += wp[1];
  J[4][5] += wp[2];
  J[5][0] += wp[1];
  J[5][2] -= wp[1];
  J[5][3] += wp[0];
  J[5][4] -= wp[2];
#pragma unroll
  for (int a = 0; a < 6; ++a) J[a][a] += 1.0;
}

// mm10_setup (mm10_a.f:830-962) for strain increment D over time tinc
CPFFT_HD inline void setup_point(const MatParams& P, const double R[3][3], const StateN& n,
                                 const double D[6], double tinc, Point& pt) {
#pragma unroll
  for (int k = 0; k < 6; ++k) pt.D[k] = D[k];
  pt.tinc = tinc;
  double t1 = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  double t2 = D[3] * D[3] + D[4] * D[4] + D[5] * D[5];
  pt.dg = P.alter_mode ? P.eps_dot_0_y * tinc : sqrt((2.0 / 3.0) * (t1 + 0.5 * t2));
  double w[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) pt.tn[i][j] = n.Rp[j][i];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      w[i][j] = R[i][0] * pt.tn[0][j] + R[i][1] * pt.tn[1][j] + R[i][2] * pt.tn[2][j];
  rt2rvw(w, pt.W);
}

// mm10_solve (mm10_a.f:2860-3295). x in/out; xprev receives the iterate at
// which the last full Jacobian was formed. Returns fail.
// x0w: optional starting point [stress, tt] (the warm start); without it
// the predictor starts from x with tt extrapolated (mm10_a.f:2993-3023).
template <int NS>
CPFFT_HD inline int local_solve(const MatParams& P, const Orient<NS>& o, const Point& pt,
                                const StateN& n, double dtinc, double x[7], double xprev[7],
                                double* tt_rate, Counters& cnt, const double* x0w = nullptr,
                                double x0w_res = 0.0) {
  const double cls = 1.0e-4, red = 0.5;
  const int mls = 10;
  int fail = 0;
  bool warm = false;
  double inR1;
  // ---- predictor: 6x6 Newton on R1 at extrapolated tt (mm10_a.f:2975-3156)
  {
    double d1[6], d2[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { d1[k] = pt.D[k]; d2[k] = n.D[k]; }
    double dtr = (d1[0] + d1[1] + d1[2]) / 3.0;
    d1[0] -= dtr; d1[1] -= dtr; d1[2] -= dtr;
    double t1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double t2 = d1[3] * d1[3] + d1[4] * d1[4] + d1[5] * d1[5];
    double nn = (t1 + t2 == 0.0) ? 0.0 : 1.0 / sqrt(t1 + t2);
#pragma unroll
    for (int k = 0; k < 6; ++k) d1[k] = (t1 + t2 == 0.0) ? 0.0 : d1[k] * nn;
    dtr = (d2[0] + d2[1] + d2[2]) / 3.0;
    d2[0] -= dtr; d2[1] -= dtr; d2[2] -= dtr;
    t1 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    t2 = d2[3] * d2[3] + d2[4] * d2[4] + d2[5] * d2[5];
    nn = (t1 + t2 == 0.0) ? 0.0 : 1.0 / sqrt(t1 + t2);
#pragma unroll
    for (int k = 0; k < 6; ++k) d2[k] = (t1 + t2 == 0.0) ? 0.0 : d2[k] * nn;
    double cosang = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) cosang += d1[k] * d2[k];
    cosang = fmax(cosang, 0.0);
    double x2 = x[6] + cosang * n.tt_rate * dtinc;
    double x1[6], R1[6], sabs;
#pragma unroll
    for (int k = 0; k < 6; ++k) x1[k] = x[k];
    form_R1<NS>(P, o, pt, n, x1, x2, R1, &sabs);
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) dot += R1[k] * R1[k];
    double nR1 = sqrt(dot);
    inR1 = nR1;
    if (x0w && x0w_res < nR1) {
      // warm start: the full state x0w (stress and slip resistance) has a
      // full residual below the plain start's predictor residual, so the
      // predictor is skipped and the full Newton starts there
#pragma unroll
      for (int k = 0; k < 7; ++k) x[k] = x0w[k];
      warm = true;
    }
    int it = 0, pfail = 0;
    while (!warm && nR1 > P.atol1 && nR1 / inR1 > P.rtol1) {
      double A[6][6], dx[6][1], gJ[6];
      form_J<NS, 6>(P, o, pt, x1, x2, A);
      // line-search slope J^T R before A is factorised
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        double wv = 0.0;
#pragma unroll
        for (int a = 0; a < 6; ++a) wv += A[a][b] * R1[a];
        gJ[b] = wv;
      }
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        dx[a][0] = R1[a];
#pragma unroll
        for (int b = 0; b < 6; ++b) A[a][b] = -A[a][b];
      }
      lu_solve<6, 1>(A, dx);
      double ls1 = 0.5 * dot, ls2 = 0.0;
#pragma unroll
      for (int b = 0; b < 6; ++b) ls2 += dx[b][0] * gJ[b];
      ls2 *= cls;
      double alpha = 1.0;
      for (int ls = 0;; ++ls) {
        double nlsx = ls1 + ls2 * alpha, xn[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) xn[k] = x1[k] + alpha * dx[k][0];
        form_R1<NS>(P, o, pt, n, xn, x2, R1, &sabs);
        cnt.ls_evals++;
        dot = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) dot += R1[k] * R1[k];
        nR1 = sqrt(dot);
        if (0.5 * dot <= nlsx || ls > mls) {
#pragma unroll
          for (int k = 0; k < 6; ++k) x1[k] = xn[k];
          break;
        }
        alpha *= red;
      }
      ++it;
      bool nan = false;
#pragma unroll
      for (int k = 0; k < 6; ++k) nan |= (x1[k] != x1[k]);
      if (it > P.miter || nan) { pfail = 1; break; }
    }
    cnt.pred_it += it;
    if (pfail) {
      fail = 1;
    } else if (!warm) {
#pragma unroll
      for (int k = 0; k < 6; ++k) x[k] = x1[k];
      x[6] = x2;
    }
  }
  // ---- full Newton on [R1; R2] (mm10_a.f:3162-3293)
  {
    double R[7];
    double dot = form_R<NS>(P, o, pt, n, x, R, tt_rate);
    double nR = sqrt(dot), inR = nR;
    if (inR == 0.0) inR = inR1;
    int it = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) xprev[k] = x[k];
    while ((nR > P.atol && nR / inR > P.rtol) || it < 1) {
      double A[7][7], dx[7][1], gJ[7];
      form_J<NS, 7>(P, o, pt, x, x[6], A);
#pragma unroll
      for (int k = 0; k < 7; ++k) xprev[k] = x[k];
#pragma unroll
      for (int b = 0; b < 7; ++b) {
        double wv = 0.0;
#pragma unroll
        for (int a = 0; a < 7; ++a) wv += A[a][b] * R[a];
        gJ[b] = wv;
      }
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        dx[a][0] = R[a];
#pragma unroll
        for (int b = 0; b < 7; ++b) A[a][b] = -A[a][b];
      }
      lu_solve<7, 1>(A, dx);
      double ls1 = 0.5 * dot, ls2 = 0.0;
#pragma unroll
      for (int b = 0; b < 7; ++b) ls2 += dx[b][0] * gJ[b];
      ls2 *= cls;
      double alpha = 1.0;
      for (int ls = 0;; ++ls) {
        double nlsx = ls1 + ls2 * alpha, xn[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) xn[k] = x[k] + alpha * dx[k][0];
        dot = form_R<NS>(P, o, pt, n, xn, R, tt_rate);
        cnt.ls_evals++;
        nR = sqrt(dot);
        if (0.5 * dot <= nlsx || ls > mls) {
#pragma unroll
          for (int k = 0; k < 7; ++k) x[k] = xn[k];
          break;
        }
        alpha *= red;
      }
      ++it;
      bool nan = false;
#pragma unroll
      for (int k = 0; k < 7; ++k) nan |= (x[k] != x[k]);
      if (it > P.miter || nan) { fail = 1; break; }
    }
    cnt.upd_it += it;
  }
  return fail;
}

// ------------------------------------------------------- exact dP/dF
// cep2A_a (cep2A.f:86-284) for one voxel; C = consistent tangent (row-major
// 6x6), t = unrotated Cauchy stress (Voigt), K written column (k,l) at a time.
template <class Store>
CPFFT_HD inline void dPdF(const double Fn[3][3], const double Fn1[3][3], const double t6[6],
                          const double C[6][6], const double Rh[3][3], const double fnhinv[3][3],
                          const double R[3][3], const double finv[3][3], double J,
                          Store&& store) {
  double t[3][3] = {{t6[0], t6[3], t6[5]}, {t6[3], t6[1], t6[4]}, {t6[5], t6[4], t6[2]}};
  auto mm = [](const double A[3][3], const double B[3][3], double Cc[3][3]) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Cc[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
  };
  auto det = [](const double A[3][3]) {
    return A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
           A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
  };
  double Rt[3][3], U[3][3], Y[3][3], sigma[3][3], Uinv[3][3], tUi[3][3], RT_[3][3];
  double RY[3][3], RYR[3][3], RtRF[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i][j] = R[j][i];
  mm(Rt, Fn1, U);
  double trU = U[0][0] + U[1][1] + U[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Y[i][j] = (i == j ? trU : 0.0) - U[i][j];
  mm(R, t, RT_);
  mm(RT_, Rt, sigma);
  mm(finv, R, Uinv);
  mm(t, Uinv, tUi);
  mm(R, Y, RY);
  mm(RY, Rt, RYR);
  double dY1 = 1.0 / det(Y);
  mm(RT_, Uinv, RtRF);
  double Rht[3][3], Fnh[3][3], dFn[3][3], Uh[3][3], Yh[3][3], RYRh[3][3], RYh[3][3], FpF[3][3];
  double L[3][3], D2[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Rht[i][j] = Rh[j][i];
      Fnh[i][j] = 0.5 * (Fn1[i][j] + Fn[i][j]);
      dFn[i][j] = Fn1[i][j] - Fn[i][j];
      FpF[i][j] = 0.5 * fnhinv[i][j];
    }
  mm(Rht, Fnh, Uh);
  double trUh = Uh[0][0] + Uh[1][1] + Uh[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Yh[i][j] = (i == j ? trUh : 0.0) - Uh[i][j];
  mm(Rh, Yh, RYh);
  mm(RYh, Rht, RYRh);
  double dYh1 = 0.5 / det(Yh);
  mm(dFn, FpF, L);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D2[i][j] = L[i][j] + L[j][i];
  const int vg[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
  // row l / column k of a 3x3 matrix for a run-time k, l (selects keep the
  // matrices in registers)
  auto colk = [](const double A[3][3], int k, double o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = k == 0 ? A[i][0] : (k == 1 ? A[i][1] : A[i][2]);
  };
  auto rowl = [](const double A[3][3], int l, double o[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = l == 0 ? A[0][i] : (l == 1 ? A[1][i] : A[2][i]);
  };
  // D2 Rh does not depend on (k, l)
  double A1[3][3];
  mm(D2, Rh, A1);
#pragma unroll 1
  for (int kl = 0; kl < 9; ++kl) {
    const int k = kl / 3, l = kl - 3 * k;
    double RYRk[3], Yl[3], RYl[3], RYk[3], RYRhk[3], Yhl[3], RYhl[3], RYhk[3], FpFl[3], Lk[3];
    double fk[3], fl[3];
    colk(RYR, k, RYRk);
    rowl(Y, l, Yl);
    colk(RY, l, RYl);
    rowl(RY, k, RYk);
    colk(RYRh, k, RYRhk);
    rowl(Yh, l, Yhl);
    colk(RYh, l, RYhl);
    rowl(RYh, k, RYhk);
    rowl(FpF, l, FpFl);
    colk(L, k, Lk);
    colk(finv, k, fk);  // finv[.][k]
    rowl(finv, l, fl);  // finv[l][.]
    const double flk = fk[0] * (l == 0) + fk[1] * (l == 1) + fk[2] * (l == 2);
    // dR(i,m)/dF(k,l), dRh(i,m)/dF(k,l), dL(i,m)/dF(k,l)
    double dR[3][3], dRh[3][3], dL[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        dR[i][m] = dY1 * (RYRk[i] * Yl[m] - RYl[i] * RYk[m]);
        dRh[i][m] = dYh1 * (RYRhk[i] * Yhl[m] - RYhl[i] * RYhk[m]);
        dL[i][m] = ((i == k) ? FpFl[m] : 0.0) - Lk[i] * FpFl[m];
      }
    // dt(i,j)/dF(k,l): dRh^T D2 Rh + Rh^T (dL + dL^T) Rh + Rh^T D2 dRh
    double S[3][3], A2[3][3], dd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) S[i][j] = dL[i][j] + dL[j][i];
    mm(S, Rh, A2);  // S Rh
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) v += dRh[m][i] * A1[m][j] + Rh[m][i] * A2[m][j];
        double w = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          double dm = 0.0;
#pragma unroll
          for (int q = 0; q < 3; ++q) dm += D2[m][q] * dRh[q][j];
          w += Rh[m][i] * dm;
        }
        dd[i][j] = v + w;
      }
    // tangent contraction dt'(i,j) = sum_ab cep99(ij,ab) dd(b,a)
    double dt[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) v += C[vg[a][b]][vg[i][j]] * dd[b][a];
        dt[i][j] = v;
      }
    // dP(i,j)/dF(k,l)
    double Rdt[3][3], RdtU[3][3];
    mm(R, dt, Rdt);
    mm(Rdt, Uinv, RdtU);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v = RtRF[i][j] * (J * flk) + J * RdtU[i][j];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          v -= J * sigma[i][m] * fk[j] * fl[m];
          v += J * dR[i][m] * tUi[m][j];
          double s4 = 0.0;
#pragma unroll
          for (int q = 0; q < 3; ++q) s4 += dR[q][m] * finv[j][q];
          v += J * RT_[i][m] * s4;
        }
        store(i * 27 + j * 9 + kl, v);
      }
  }
}

// ------------------------------------------------------- voxel update
// Kinematics of the step (drive_eps_sig.f:207-245): R_{n+1/2} and R_{n+1}
// (polar decompositions), F_{n+1/2}^-1, and the unrotated strain increment
// uddt = Q(R_{n+1/2}) sym(dF F_{n+1/2}^-1) (getrm1 opt 1 + qmply1).
CPFFT_HD inline void step_kinematics(const double Fn[3][3], const double Fn1[3][3],
                                     double rnh[3][3], double rn1[3][3], double fnhinv[3][3],
                                     double uddt[6]) {
  double fnh[3][3], dfn[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      fnh[i][j] = 0.5 * (Fn[i][j] + Fn1[i][j]);
      dfn[i][j] = Fn1[i][j] - Fn[i][j];
    }
  polar_R(fnh, rnh);
  polar_R(Fn1, rn1);
  inv33(fnh, fnhinv);
  double ddt[6];
  {
#pragma clang fp contract(off)
  const double(*A)[3] = dfn;
  const double(*B)[3] = fnhinv;
  ddt[0] = A[0][0] * B[0][0] + A[0][1] * B[1][0] + A[0][2] * B[2][0];
  ddt[3] = A[1][0] * B[0][0] + A[1][1] * B[1][0] + A[1][2] * B[2][0] + A[0][0] * B[0][1] +
           A[0][1] * B[1][1] + A[0][2] * B[2][1];
  ddt[5] = A[2][0] * B[0][0] + A[2][1] * B[1][0] + A[2][2] * B[2][0] + A[0][0] * B[0][2] +
           A[0][1] * B[1][2] + A[0][2] * B[2][2];
  ddt[1] = A[1][0] * B[0][1] + A[1][1] * B[1][1] + A[1][2] * B[2][1];
  ddt[4] = A[2][0] * B[0][1] + A[2][1] * B[1][1] + A[2][2] * B[2][1] + A[1][0] * B[0][2] +
           A[1][1] * B[1][2] + A[1][2] * B[2][2];
  ddt[2] = A[2][0] * B[0][2] + A[2][1] * B[1][2] + A[2][2] * B[2][2];
  const double(*r)[3] = rnh;
  double q[6][6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q[a][0] = r[0][a] * r[0][a];
    q[a][1] = r[1][a] * r[1][a];
    q[a][2] = r[2][a] * r[2][a];
    q[a][3] = r[0][a] * r[1][a];
    q[a][4] = r[2][a] * r[1][a];
    q[a][5] = r[0][a] * r[2][a];
  }
  const int ia[3] = {0, 1, 0}, ib[3] = {1, 2, 2};
#pragma unroll
  for (int z = 0; z < 3; ++z) {
    const int a = ia[z], b = ib[z];
    q[3 + z][0] = 2.0 * r[0][a] * r[0][b];
    q[3 + z][1] = 2.0 * r[1][a] * r[1][b];
    q[3 + z][2] = 2.0 * r[2][a] * r[2][b];
    q[3 + z][3] = r[0][a] * r[1][b] + r[0][b] * r[1][a];
    q[3 + z][4] = r[1][a] * r[2][b] + r[2][a] * r[1][b];
    q[3 + z][5] = r[0][a] * r[2][b] + r[2][a] * r[0][b];
  }
  // reference index pattern for rows 4..6 (polar.f:735-758)
  q[4][1] = 2.0 * r[1][2] * r[1][1];
  q[4][2] = 2.0 * r[2][1] * r[2][2];
  q[4][3] = r[0][1] * r[1][2] + r[1][1] * r[0][2];
  q[4][5] = r[0][1] * r[2][2] + r[2][1] * r[0][2];
  q[5][3] = r[0][0] * r[1][2] + r[1][0] * r[0][2];
  q[3][4] = r[1][0] * r[2][1] + r[2][0] * r[1][1];
#pragma unroll
  for (int j = 0; j < 6; ++j)
    uddt[j] = q[j][0] * ddt[0] + q[j][1] * ddt[1] + q[j][2] * ddt[2] + q[j][3] * ddt[3] +
              q[j][4] * ddt[4] + q[j][5] * ddt[5];
  }
}

// ------------------------------------------------------- warm start
// The network predicts each system's slip increment through its normalised
// resolved shear u_s = rs_s / tt = sgn(g_s) (|g_s| / dg)^(1/n), a one-to-one
// map of the slip increment g_s = dg |u_s|^n sgn(u_s) (Voce power law with
// alter_mode) that stays of order one where g_s spans many decades.
CPFFT_HD inline double slip_to_u(const MatParams& P, double dg, double g) {
  return g == 0.0 ? 0.0 : sgn1(g) * pow(fabs(g) / dg, 1.0 / P.rate_n);
}

// Starting point of the local solve from predicted u_s:
// tt0 from the discrete Voce law at S = sum_s dg |u_s|^n (closed form for
// m = 1: tt = (tt_n + theta_0 S (1 + tau_y/tau_v)) / (1 + theta_0 S / tau_v));
// the stress s0 minimises sum_s w_s (s . E m_s - tt0 u_s)^2 + lam |s - s_tr|^2
// with weights w_s = |u_s|^n (the slipping systems), s_tr = s_n + C:D the
// elastic trial stress and lam small, so that directions the slipping
// systems do not constrain (and elastic voxels) take the trial stress.
template <int NS>
CPFFT_HD inline void warm_x0(const MatParams& P, const Orient<NS>& o, const Point& pt,
                             const StateN& n, const float* ws, long wstride, double x0[7]) {
  double S = 0.0;
  for (int i = 0; i < NS; ++i) S += pt.dg * pw_nm2(P, fabs((double)ws[i * wstride])) *
                                    ((double)ws[i * wstride] * (double)ws[i * wstride]);
  double tt0;
  if (P.m_is_one) {
    tt0 = (n.tt + P.theta_0 * S * (1.0 + P.tau_y / P.tau_v)) / (1.0 + P.theta_0 * S / P.tau_v);
  } else {
    const double h = 1.0 - (n.tt - P.tau_y) / P.tau_v;
    tt0 = n.tt + P.theta_0 * pow(fabs(h), P.voche_m) * sgn1(h) * S;
  }
  double A[6][6], b[6][1];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    b[a][0] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) A[a][c] = 0.0;
  }
  double wsum = 0.0;
  for (int i = 0; i < NS; ++i) {
    const double u = (double)ws[i * wstride];
    const double w = pw_nm2(P, fabs(u)) * u * u;  // |u|^n
    if (w == 0.0) continue;
    double m[6], mc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) m[k] = o.ms[i][k];
    rotE(pt.tn, m, mc);
    const double rs = tt0 * u;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      b[a][0] += w * rs * mc[a];
#pragma unroll
      for (int c = 0; c < 6; ++c) A[a][c] += w * mc[a] * mc[c];
      wsum += w * mc[a] * mc[a];
    }
  }
  double c[6], st[6];
  Cv6(o.C, pt.D, c);
#pragma unroll
  for (int k = 0; k < 6; ++k) st[k] = n.stress[k] + c[k];
  const double lam = 1.0e-6 * wsum + 1.0e-300;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    A[a][a] += lam;
    b[a][0] += lam * st[a];
  }
  lu_solve<6, 1>(A, b);
#pragma unroll
  for (int k = 0; k < 6; ++k) x0[k] = b[k][0];
  x0[6] = tt0;
}

// ---- the network's inputs, in FP32 (the network reads them in BF16)

// inverse of a 3x3 FP32 matrix by cofactors
CPFFT_HD inline void inv33f(const float a[3][3], float g[3][3]) {
  g[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  g[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
  g[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
  g[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  g[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
  g[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
  g[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  g[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
  g[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
  const float id = 1.0f / (a[0][0] * g[0][0] + a[0][1] * g[1][0] + a[0][2] * g[2][0]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) g[i][j] *= id;
}

// The unrotated strain increment uddt of step_kinematics to FP32 accuracy:
// F_{n+1/2} and dF formed in FP64, the rotation of F_{n+1/2} by Higham's
// Newton iteration R <- (R + R^-T)/2 from R = F (quadratic convergence:
// five steps reach FP32 accuracy for stretches up to about 40%), then the
// same D = sym(dF F_{n+1/2}^-1) and Voigt rotation as step_kinematics.
CPFFT_HD inline void uddt_f32(const double Fn[3][3], const double Fn1[3][3], float uddt[6]) {
  float fnh[3][3], dfn[3][3], r[3][3], ri[3][3], B[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      fnh[i][j] = (float)(0.5 * (Fn[i][j] + Fn1[i][j]));
      dfn[i][j] = (float)(Fn1[i][j] - Fn[i][j]);
      r[i][j] = fnh[i][j];
    }
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    inv33f(r, ri);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) r[i][j] = 0.5f * (r[i][j] + ri[j][i]);
  }
  inv33f(fnh, B);
  const float(*A)[3] = dfn;
  float ddt[6];
  ddt[0] = A[0][0] * B[0][0] + A[0][1] * B[1][0] + A[0][2] * B[2][0];
  ddt[3] = A[1][0] * B[0][0] + A[1][1] * B[1][0] + A[1][2] * B[2][0] + A[0][0] * B[0][1] +
           A[0][1] * B[1][1] + A[0][2] * B[2][1];
  ddt[5] = A[2][0] * B[0][0] + A[2][1] * B[1][0] + A[2][2] * B[2][0] + A[0][0] * B[0][2] +
           A[0][1] * B[1][2] + A[0][2] * B[2][2];
  ddt[1] = A[1][0] * B[0][1] + A[1][1] * B[1][1] + A[1][2] * B[2][1];
  ddt[4] = A[2][0] * B[0][1] + A[2][1] * B[1][1] + A[2][2] * B[2][1] + A[1][0] * B[0][2] +
           A[1][1] * B[1][2] + A[1][2] * B[2][2];
  ddt[2] = A[2][0] * B[0][2] + A[2][1] * B[1][2] + A[2][2] * B[2][2];
  float q[6][6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q[a][0] = r[0][a] * r[0][a];
    q[a][1] = r[1][a] * r[1][a];
    q[a][2] = r[2][a] * r[2][a];
    q[a][3] = r[0][a] * r[1][a];
    q[a][4] = r[2][a] * r[1][a];
    q[a][5] = r[0][a] * r[2][a];
  }
  const int ia[3] = {0, 1, 0}, ib[3] = {1, 2, 2};
#pragma unroll
  for (int z = 0; z < 3; ++z) {
    const int a = ia[z], b = ib[z];
    q[3 + z][0] = 2.0f * r[0][a] * r[0][b];
    q[3 + z][1] = 2.0f * r[1][a] * r[1][b];
    q[3 + z][2] = 2.0f * r[2][a] * r[2][b];
    q[3 + z][3] = r[0][a] * r[1][b] + r[0][b] * r[1][a];
    q[3 + z][4] = r[1][a] * r[2][b] + r[2][a] * r[1][b];
    q[3 + z][5] = r[0][a] * r[2][b] + r[2][a] * r[0][b];
  }
  // the reference's index pattern for rows 4..6 (polar.f:735-758)
  q[4][1] = 2.0f * r[1][2] * r[1][1];
  q[4][2] = 2.0f * r[2][1] * r[2][2];
  q[4][3] = r[0][1] * r[1][2] + r[1][1] * r[0][2];
  q[4][5] = r[0][1] * r[2][2] + r[2][1] * r[0][2];
  q[5][3] = r[0][0] * r[1][2] + r[1][0] * r[0][2];
  q[3][4] = r[1][0] * r[2][1] + r[2][0] * r[1][1];
#pragma unroll
  for (int j = 0; j < 6; ++j)
    uddt[j] = q[j][0] * ddt[0] + q[j][1] * ddt[1] + q[j][2] * ddt[2] + q[j][3] * ddt[3] +
              q[j][4] * ddt[4] + q[j][5] * ddt[5];
}

// slip_to_u in FP32 after an exact binary split |g|/dg = m 2^e (FP64 slip
// increments reach 1e-50 and below, out of FP32's range): 2^((e + log2 m)/n)
CPFFT_HD inline float slip_to_u_f32(const MatParams& P, double dg, double g) {
  if (g == 0.0) return 0.0f;
  int e;
  const double m = frexp(fabs(g) / dg, &e);  // m in [0.5, 1)
  const float u = exp2f(((float)e + log2f((float)m)) / (float)P.rate_n);
  return g < 0.0 ? -u : u;
}

// Inputs of the warm-start network for one voxel (18 + 2 NS values: 42 for
// fcc, 66 for bcc24), in FP32: the trial stress s_n + C:D rotated to the
// lattice frame (Rp_n^T S Rp_n as a full 3x3, row-major), the slip
// resistance tt_n once per slip system (the Voce law keeps a single one),
// the strain increment D in the lattice frame (3x3 tensor, row-major), and
// the slip increments of the previous step (as u_s, see slip_to_u).
template <int NS, class LoadF, class LoadS>
CPFFT_HD inline void voxel_features(const MatParams& P, const Orient<NS>& o, int step, LoadF&& ldF,
                                    LoadS&& ldn, float* feat) {
  float uddt[6];
  {
    double Fn[3][3], Fn1[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Fn[i][j] = ldF(0, 3 * i + j);
        Fn1[i][j] = ldF(1, 3 * i + j);
      }
    uddt_f32(Fn, Fn1, uddt);
  }
  float sn[6], Rp[3][3], tt;
  if (step == 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) sn[k] = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Rp[i][j] = (i == j) ? 1.0f : 0.0f;
    tt = (float)(P.tau_y + 1.0e-5);
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) sn[k] = (float)ldn(k);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) Rp[i][j] = (float)ldn(6 + 3 * j + i);
    tt = (float)ldn(15);
  }
  float st[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    float c = 0.0f;
#pragma unroll
    for (int b = 0; b < 6; ++b) c += (float)o.C[a * 6 + b] * uddt[b];
    st[a] = sn[a] + c;
  }
  const float S[3][3] = {{st[0], st[3], st[5]}, {st[3], st[1], st[4]}, {st[5], st[4], st[2]}};
  const float D[3][3] = {{uddt[0], 0.5f * uddt[3], 0.5f * uddt[5]},
                         {0.5f * uddt[3], uddt[1], 0.5f * uddt[4]},
                         {0.5f * uddt[5], 0.5f * uddt[4], uddt[2]}};
  // lattice frame: Rp^T X Rp (tn = Rp^T maps lattice to unrotated)
  float SR[3][3], DR[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      SR[k][j] = S[k][0] * Rp[0][j] + S[k][1] * Rp[1][j] + S[k][2] * Rp[2][j];
      DR[k][j] = D[k][0] * Rp[0][j] + D[k][1] * Rp[1][j] + D[k][2] * Rp[2][j];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      feat[3 * i + j] = Rp[0][i] * SR[0][j] + Rp[1][i] * SR[1][j] + Rp[2][i] * SR[2][j];
      feat[9 + NS + 3 * i + j] = Rp[0][i] * DR[0][j] + Rp[1][i] * DR[1][j] + Rp[2][i] * DR[2][j];
    }
#pragma unroll
  for (int s = 0; s < NS; ++s) feat[9 + s] = tt;
  const double dg = P.alter_mode ? P.eps_dot_0_y * P.dt : 1.0;
  for (int s = 0; s < NS; ++s)
    feat[18 + NS + s] = step == 1 ? 0.0f : slip_to_u_f32(P, dg, ldn(23 + s));
}

// Per-voxel stress update.  ldF(0|1, 3i+j) loads F_n | F_{n+1}; ldn(w)
// loads state-n word w; stn1(w, v) stores state n+1; stP(3i+j, v) stores
// P; stT(6a+b, v) stores the symmetrised material tangent (row-major 6x6)
// that voxel_dPdF pulls back to dP/dF.
// Warm start (off unless ws or xprev_it is given): ws holds predicted slip
// increments ws[s * wstride] of this voxel (warm_x0 turns them into a full
// state), xprev_it the stress and slip resistance (7 values) of this step's
// previous global iteration.  The candidate of smaller full residual is
// handed to local_solve, which starts its full Newton there when that
// residual is below the plain start's predictor residual; a warm solve that
// fails is redone from the plain start.
template <int NS, class LoadF, class LoadS, class StoreS, class StoreP, class StoreT>
CPFFT_HD inline void voxel_update(const MatParams& P, const Orient<NS>& o, int step, int iter,
                                  LoadF&& ldF, LoadS&& ldn, StoreS&& stn1, StoreP&& stP,
                                  StoreT&& stT, Counters& cnt, const float* ws = nullptr,
                                  long wstride = 0, const double* xprev_it = nullptr) {
  double uddt[6], rn1[3][3];
  {
    double Fn[3][3], Fn1[3][3], rnh[3][3], fnhinv[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Fn[i][j] = ldF(0, 3 * i + j);
        Fn1[i][j] = ldF(1, 3 * i + j);
      }
    step_kinematics(Fn, Fn1, rnh, rn1, fnhinv, uddt);
  }

  // ---- state n (mm10_copy_cc_hist; re-initialised at step 1, mm10_a.f:89-95)
  StateN n;
  if (step == 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) { n.stress[k] = 0.0; n.D[k] = 0.0; }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) n.Rp[i][j] = (i == j) ? 1.0 : 0.0;
    n.tt = P.tau_y + 1.0e-5;
    n.tt_rate = ldn(16);
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) n.stress[k] = ldn(k);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) n.Rp[i][j] = ldn(6 + 3 * j + i);
    n.tt = ldn(15);
    n.tt_rate = ldn(16);
#pragma unroll
    for (int k = 0; k < 6; ++k) n.D[k] = ldn(17 + k);
  }

  const double* __restrict__ C = o.C;
  Point pt;
  setup_point(P, rn1, n, uddt, P.dt, pt);

  double s1[6], tt1, ttrate1 = 0.0;
  double tangent[6][6];
  double Rp1[3][3];
  double winc = 0.0, pwinc = 0.0, psinc = 0.0;
  int failed = 0;
  bool plastic = false;  // slip increments come from a converged plastic update
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Rp1[i][j] = 0.0;

  double temp1 = 0.0, temp2 = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) { temp1 += uddt[k] * uddt[k]; temp2 += n.stress[k] * n.stress[k]; }
  const bool no_load = (temp2 == 0.0) && (temp1 == 0.0);
  if (iter == 0 || no_load) {
    // explicit update with the elastic tangent (mm10_a.f:2735-2753)
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) tangent[a][b] = C[a * 6 + b];
#pragma unroll
    for (int k = 0; k < 6; ++k) s1[k] = n.stress[k];
    if (!no_load) {
      double R1[6], sabs;
      form_R1<NS>(P, o, pt, n, n.stress, n.tt, R1, &sabs);
#pragma unroll
      for (int k = 0; k < 6; ++k) s1[k] = n.stress[k] - R1[k];
    }
    tt1 = n.tt;
  } else {
    // adaptive sub-stepping (mm10_a.f:2759-2843).  A successful pass ends
    // with frac + stp == 1 exactly (dyadic fractions), so pt then holds the
    // full-step point again.
    double x[7], xprev[7], ost[6], ott = n.tt;
#pragma unroll
    for (int k = 0; k < 6; ++k) { x[k] = n.stress[k]; ost[k] = x[k]; }
    double frac = 0.0, stp = 1.0;
    int cuts = 0, fail = 0;
    if (ws || xprev_it) {
      double x0[7], xc[7], R[7], tr, best = 1.0e300;
      auto consider = [&](const double* c) {
        const double d = sqrt(form_R<NS>(P, o, pt, n, c, R, &tr));
        if (d < best) {  // false for NaN
          best = d;
#pragma unroll
          for (int k = 0; k < 7; ++k) x0[k] = c[k];
        }
      };
      if (ws) {
        warm_x0<NS>(P, o, pt, n, ws, wstride, xc);
        consider(xc);
      }
      if (xprev_it) consider(xprev_it);
      x[6] = n.tt;
      fail = local_solve<NS>(P, o, pt, n, P.dt, x, xprev, &ttrate1, cnt,
                             best < 1.0e300 ? x0 : nullptr, best);
      bool ok = !fail && (x[6] == x[6]);
#pragma unroll
      for (int k = 0; k < 6; ++k) ok &= (x[k] == x[k]);
      if (ok) {
        frac = 1.0;
      } else {
        cnt.ws_fallback++;
        fail = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = ost[k];
      }
    }
    while (frac < 1.0) {
      if (frac != 0.0 || stp != 1.0) {
        double Dw[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) Dw[k] = uddt[k] * (stp + frac);
        setup_point(P, rn1, n, Dw, P.dt * (stp + frac), pt);
      }
      x[6] = n.tt;
      fail = local_solve<NS>(P, o, pt, n, P.dt * stp, x, xprev, &ttrate1, cnt);
      if (fail) {
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = ost[k];
        x[6] = ott;
        stp *= 0.5;
        ++cuts;
        cnt.cuts++;
        if (cuts > 4) break;
        fail = 0;
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) ost[k] = x[k];
        ott = x[6];
        frac += stp;
      }
    }
    bool bad = fail || (x[6] != x[6]);
#pragma unroll
    for (int k = 0; k < 6; ++k) bad |= (x[k] != x[k]);
    if (bad) {
      // The reference abandons the rest of its element block here
      // (mm10_a.f:113-116, 255-258); this implementation reverts the voxel
      // to state n with the elastic tangent and counts it.
      failed = 1;
#pragma unroll
      for (int k = 0; k < 6; ++k) s1[k] = n.stress[k];
      tt1 = n.tt;
      ttrate1 = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Rp1[i][j] = n.Rp[i][j];
#pragma unroll
      for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b) tangent[a][b] = C[a * 6 + b];
    } else {
      plastic = true;
#pragma unroll
      for (int k = 0; k < 6; ++k) s1[k] = x[k];
      tt1 = x[6];
      // consistent tangent (mm10_tangent) from the Jacobian at xprev,
      // symmetrised (mm10_a.f:1134-1142)
      {
        double J[7][7];
        form_J<NS, 7>(P, o, pt, xprev, xprev[6], J);
        double JJ[6][6], B[6][6];
        const double j22 = J[6][6];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int b = 0; b < 6; ++b) JJ[a][b] = J[a][b] - J[a][6] * (J[6][b] / j22);
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int b = 0; b < 6; ++b) B[a][b] = C[a * 6 + b];
        lu_solve<6, 6>(JJ, B);
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int b = 0; b < 6; ++b) tangent[a][b] = 0.5 * (B[a][b] + B[b][a]);
      }
      // slip sums at the converged state
      double mb[6], qb[3], sabs;
      slip_sums<NS>(P, o, pt, s1, tt1, mb, qb, &sabs);
      // plastic rotation (mm10_update_rotation): Rp = expm(W(wbarp)) Rp_n
      {
        double RW[3][3], tn[3][3], wb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) tn[i][j] = n.Rp[j][i];
        rt2rvw(tn, RW);
        mv3(RW, qb, wb);
        double W[3][3] = {{0.0, wb[2], wb[1]}, {-wb[2], 0.0, wb[0]}, {-wb[1], -wb[0], 0.0}};
        double al = sqrt(W[1][2] * W[1][2] + W[0][2] * W[0][2] + W[0][1] * W[0][1]);
        double E[3][3];
        if (al < 1.0e-16) {
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) E[i][j] = 0.0;
        } else {
          double ca = (1.0 - cos(al)) / (al * al), sa = sin(al) / al;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
              E[i][j] = sa * W[i][j] + ca * (W[i][0] * W[0][j] + W[i][1] * W[1][j] + W[i][2] * W[2][j]);
        }
        E[0][0] += 1.0; E[1][1] += 1.0; E[2][2] += 1.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            Rp1[i][j] = E[i][0] * n.Rp[0][j] + E[i][1] * n.Rp[1][j] + E[i][2] * n.Rp[2][j];
      }
      // work terms kept in history (mm10_output)
      winc = s1[0] * uddt[0] + s1[1] * uddt[1] + s1[2] * uddt[2] + s1[3] * uddt[3] +
             s1[4] * uddt[4] + s1[5] * uddt[5];
      double ee[6];
      {
        double S[6][6], b[6][1];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
          b[a][0] = s1[a];
#pragma unroll
          for (int c = 0; c < 6; ++c) S[a][c] = C[a * 6 + c];
        }
        lu_solve<6, 1>(S, b);  // C^-1 sigma (SPD solve, mm10_a.f:3537)
        double er[6][6], bv[6];
        rt2rve(rn1, er);
#pragma unroll
        for (int c = 0; c < 6; ++c) bv[c] = b[c][0];
        mv6(er, bv, ee);
      }
      double dbar[6], wp[3], ewwe[6], ep[6];
      rotE(pt.tn, mb, dbar);
      mv3(pt.W, qb, wp);
      symsw(ee, wp, ewwe);
#pragma unroll
      for (int k = 0; k < 6; ++k) ep[k] = dbar[k] + ewwe[k];
      double a1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
      double a2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
      psinc = sqrt(2.0 / 3.0 * (a1 + 0.5 * a2));
      pwinc = s1[0] * ep[0] + s1[1] * ep[1] + s1[2] * ep[2] + s1[3] * ep[3] + s1[4] * ep[4] +
              s1[5] * ep[5];
    }
  }
  cnt.failed += failed;

  // ---- store state n+1 (mm10_store_cryhist, mm10_a_store_crystal)
#pragma unroll
  for (int k = 0; k < 6; ++k) stn1(k, s1[k]);
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) stn1(6 + 3 * j + i, Rp1[i][j]);
  stn1(15, tt1);
  stn1(16, ttrate1);
#pragma unroll
  for (int k = 0; k < 6; ++k) stn1(17 + k, uddt[k]);
  {
    // slip increments and cumulative slip (zero outside a converged plastic update)
    double sE[6];
    rotET(pt.tn, s1, sE);
#pragma unroll 2
    for (int i = 0; i < NS; ++i) {
      double g = 0.0;
      if (plastic) {
        const double* __restrict__ m = o.ms[i];
        const double rs = sE[0] * m[0] + sE[1] * m[1] + sE[2] * m[2] + sE[3] * m[3] +
                          sE[4] * m[4] + sE[5] * m[5];
        g = slip_of(P, pt.dg, rs, tt1).g;
      }
      stn1(23 + i, g);
      stn1(23 + NS + i, (step == 1 ? 0.0 : ldn(23 + NS + i)) + g);
    }
  }
  const double w0 = step == 1 ? 0.0 : ldn(23 + 2 * NS + 0);
  const double w1 = step == 1 ? 0.0 : ldn(23 + 2 * NS + 1);
  const double w2 = step == 1 ? 0.0 : ldn(23 + 2 * NS + 2);
  stn1(23 + 2 * NS + 0, w0 + winc);
  stn1(23 + 2 * NS + 1, w1 + pwinc);
  stn1(23 + 2 * NS + 2, w2 + psinc);

#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) stT(6 * a + b, tangent[a][b]);

  // ---- Cauchy stress sigma = R t R^T and P = J sigma F^-T (drive_eps_sig.f:262-273, cs2p)
  double Fn1[3][3], fn1inv[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Fn1[i][j] = ldF(1, 3 * i + j);
  const double detF = inv33(Fn1, fn1inv);
  double cs[3][3], tmp[3][3];
  const double ts[3][3] = {{s1[0], s1[3], s1[5]}, {s1[3], s1[1], s1[4]}, {s1[5], s1[4], s1[2]}};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      tmp[i][j] = rn1[i][0] * ts[0][j] + rn1[i][1] * ts[1][j] + rn1[i][2] * ts[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      cs[i][j] = tmp[i][0] * rn1[j][0] + tmp[i][1] * rn1[j][1] + tmp[i][2] * rn1[j][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      stP(3 * i + j,
          detF * (cs[i][0] * fn1inv[j][0] + cs[i][1] * fn1inv[j][1] + cs[i][2] * fn1inv[j][2]));
}

// K4 = dP/dF (cep2A, cep2A.f:86-284) of one voxel from F_n, F_{n+1}, the
// unrotated Cauchy stress t (state n+1 words 0..5) and the material tangent
// written by voxel_update; the kinematics are re-evaluated from F.
template <class LoadF, class LoadT, class StoreK>
CPFFT_HD inline void voxel_dPdF(LoadF&& ldF, const double t6[6], LoadT&& ldT, StoreK&& stK) {
  double Fn[3][3], Fn1[3][3], rnh[3][3], rn1[3][3], fnhinv[3][3], fn1inv[3][3], fnh[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Fn[i][j] = ldF(0, 3 * i + j);
      Fn1[i][j] = ldF(1, 3 * i + j);
      fnh[i][j] = 0.5 * (Fn[i][j] + Fn1[i][j]);
    }
  polar_R(fnh, rnh);
  polar_R(Fn1, rn1);
  inv33(fnh, fnhinv);
  const double detF = inv33(Fn1, fn1inv);
  double tangent[6][6];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) tangent[a][b] = ldT(6 * a + b);
  dPdF(Fn, Fn1, t6, tangent, rnh, fnhinv, rn1, fn1inv, detF, stK);
}

}  // namespace cpfft
