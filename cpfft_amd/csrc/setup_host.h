// setup_host.h -- host-side set-up of the per-voxel update: slip systems,
// elastic stiffness, per-orientation invariants and material parameters.
// Plain C++ (no HIP): shared by the library and its host-side test build.
#pragma once

#include <math.h>

#include <algorithm>

#include "../../include/cpfft.h"
#include "mm10_core.h"

namespace cpfft {

// ------------------------------------------------------------------ setup
// crystal-frame slip systems: mod_crystals.f:438-515 (fcc) and the first 24
// systems of bcc48, mod_crystals.f:812-1207 ({110} then {112})
inline void slip_systems(int set, int& ns, double b[24][3], double n[24][3]) {
  if (set == 1) {
    static const int bb[12][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0}, {0, 1, 1},  {1, 0, 1},  {1, -1, 0},
                                  {0, -1, 1}, {-1, 0, -1}, {1, 1, 0}, {0, 1, 1},  {1, 0, -1}, {-1, -1, 0}};
    static const int nn[12][3] = {{1, 1, 1},   {1, 1, 1},   {1, 1, 1},   {-1, -1, 1}, {-1, -1, 1}, {-1, -1, 1},
                                  {-1, 1, 1},  {-1, 1, 1},  {-1, 1, 1},  {1, -1, 1},  {1, -1, 1},  {1, -1, 1}};
    ns = 12;
    const double f2 = 1 / sqrt(2.0), f3 = 1 / sqrt(3.0);
    for (int s = 0; s < 12; ++s)
      for (int k = 0; k < 3; ++k) { b[s][k] = bb[s][k] * f2; n[s][k] = nn[s][k] * f3; }
  } else {
    static const int bb[24][3] = {{-1, 1, 1}, {1, -1, 1}, {1, 1, 1},  {1, 1, -1}, {1, 1, -1}, {-1, 1, 1},
                                  {1, 1, 1},  {1, -1, 1}, {1, 1, -1}, {1, -1, 1}, {1, 1, 1},  {-1, 1, 1},
                                  {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1}, {1, 1, -1},
                                  {1, 1, 1},  {-1, 1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1}};
    static const int nn[24][3] = {{1, 1, 0},  {1, 1, 0},  {1, -1, 0}, {1, -1, 0}, {1, 0, 1},  {1, 0, 1},
                                  {1, 0, -1}, {1, 0, -1}, {0, 1, 1},  {0, 1, 1},  {0, 1, -1}, {0, 1, -1},
                                  {1, 1, 2},  {-1, 1, 2}, {1, -1, 2}, {1, 1, -2}, {1, 2, 1},  {-1, 2, 1},
                                  {1, -2, 1}, {1, 2, -1}, {2, 1, 1},  {-2, 1, 1}, {2, -1, 1}, {2, 1, -1}};
    ns = 24;
    const double f2 = 1.0 / sqrt(2.0), f3 = 1.0 / sqrt(3.0), f112 = 1.0 / sqrt(6.0), f211 = 2.0 / sqrt(6.0);
    for (int s = 0; s < 24; ++s)
      for (int k = 0; k < 3; ++k) {
        b[s][k] = bb[s][k] * f3;
        int v = nn[s][k];
        n[s][k] = s < 12 ? v * f2 : (v == 2 ? f211 : v == -2 ? -f211 : v * f112);
      }
  }
}

// 6x6 inverse by Gauss-Jordan with partial pivoting (compliance -> stiffness)
inline bool inv6(double A[6][6]) {
  double M[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 12; ++j) M[i][j] = j < 6 ? A[i][j] : (j - 6 == i ? 1.0 : 0.0);
  for (int c = 0; c < 6; ++c) {
    int p = c;
    for (int i = c + 1; i < 6; ++i)
      if (fabs(M[i][c]) > fabs(M[p][c])) p = i;
    if (M[p][c] == 0.0) return false;
    for (int j = 0; j < 12; ++j) std::swap(M[c][j], M[p][j]);
    double d = M[c][c];
    for (int j = 0; j < 12; ++j) M[c][j] /= d;
    for (int i = 0; i < 6; ++i)
      if (i != c) {
        double f = M[i][c];
        for (int j = 0; j < 12; ++j) M[i][j] -= f * M[c][j];
      }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) A[i][j] = M[i][j + 6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < i; ++j) A[i][j] = A[j][i];
  return true;
}

// Kocks angles (degrees) -> rotation g (mm10_rotation_matrix, mm10_a.f:1287-1345)
inline void kocks_g(const double a[3], double g[3][3]) {
  const double pi = 3.141592653589793;
  double psi = a[0] * pi / 180.0, th = a[1] * pi / 180.0, phi = a[2] * pi / 180.0;
  g[0][0] = -sin(psi) * sin(phi) - cos(psi) * cos(phi) * cos(th);
  g[0][1] = cos(psi) * sin(phi) - sin(psi) * cos(phi) * cos(th);
  g[0][2] = cos(phi) * sin(th);
  g[1][0] = sin(psi) * cos(phi) - cos(psi) * sin(phi) * cos(th);
  g[1][1] = -cos(psi) * cos(phi) - sin(psi) * sin(phi) * cos(th);
  g[1][2] = sin(phi) * sin(th);
  g[2][0] = cos(psi) * sin(th);
  g[2][1] = sin(psi) * sin(th);
  g[2][2] = cos(th);
}

template <int NS>
inline void build_orient(const double stiff[6][6], const double b[24][3], const double n[24][3],
                         const double ang[3], Orient<NS>& o) {
  double g[3][3], t[3][3];
  kocks_g(ang, g);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = g[j][i];
  for (int s = 0; s < NS; ++s) {
    double bs[3], nv[3];
    for (int i = 0; i < 3; ++i) {
      bs[i] = t[i][0] * b[s][0] + t[i][1] * b[s][1] + t[i][2] * b[s][2];
      nv[i] = t[i][0] * n[s][0] + t[i][1] * n[s][1] + t[i][2] * n[s][2];
    }
    double A[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = bs[i] * nv[j];
    auto sy = [&](int i, int j) { return 0.5 * (A[i][j] + A[j][i]); };
    auto sk = [&](int i, int j) { return 0.5 * (A[i][j] - A[j][i]); };
    o.ms[s][0] = sy(0, 0);
    o.ms[s][1] = sy(1, 1);
    o.ms[s][2] = sy(2, 2);
    o.ms[s][3] = 2.0 * sy(0, 1);
    o.ms[s][4] = 2.0 * sy(1, 2);
    o.ms[s][5] = 2.0 * sy(0, 2);
    o.qs[s][0] = sk(1, 2);
    o.qs[s][1] = sk(0, 2);
    o.qs[s][2] = sk(0, 1);
  }
  double Rs[6][6], tmp[6][6];
  rt2rve(t, Rs);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0.0;
      for (int k = 0; k < 6; ++k) v += stiff[i][k] * Rs[j][k];
      tmp[i][j] = v;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0.0;
      for (int k = 0; k < 6; ++k) v += Rs[i][k] * tmp[k][j];
      o.C[i * 6 + j] = v;
    }
}

// cubic elasticity through the compliance (mod_crystals.f:1833-1879)
inline bool cubic_stiffness(const cpfft_material& m, double st[6][6]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) st[i][j] = 0.0;
  st[0][0] = st[1][1] = st[2][2] = 1 / m.e;
  st[0][1] = st[0][2] = st[1][0] = st[1][2] = st[2][0] = st[2][1] = -m.nu / m.e;
  st[3][3] = st[4][4] = st[5][5] = 1 / m.mu;
  return inv6(st);
}

inline MatParams make_params(const cpfft_material& m, double dt) {
  MatParams P{};
  P.rate_n = m.rate_n;
  P.tau_y = m.tau_y;
  P.tau_v = m.tau_v;
  P.theta_0 = m.theta_0;
  P.voche_m = m.voche_m;
  P.eps_dot_0_y = m.gamma_bar;
  P.dt = dt;
  P.atol = m.atol;
  P.atol1 = m.atol1;
  P.rtol = m.rtol;
  P.rtol1 = m.rtol1;
  P.miter = m.miter;
  P.alter_mode = m.alter_mode;
  P.jac_exact = m.jac_exact;
  P.int_rate = (m.rate_n == floor(m.rate_n) && m.rate_n >= 2.0 && m.rate_n <= 64.0) ? 1 : 0;
  P.m_is_one = (m.voche_m == 1.0) ? 1 : 0;
  return P;
}

}  // namespace cpfft
