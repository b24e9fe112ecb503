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
// Crystal-frame slip systems as integer Miller triples, numbered as the
// reference's slip_type: fcc (1, mod_crystals.f:438-515), bcc (2, :516-593:
// {110}<111>, the fcc table with directions and normals exchanged), roters
// (6, :604-704: fcc {111}<110> in another order and sign), bcc12 (7,
// :705-811: the first 12 systems of bcc48) and bcc48 (8, :812-1207:
// {110}<111>, {112}<111>, {123}<111>, in that order); 80 is the first 24
// systems of bcc48.  A unit vector has components sign(v) |v|/sqrt(|v|^2),
// the reference's named factors (f3 = 1/sqrt(3), f211 = 2/sqrt(6),
// f312 = 3/sqrt(14), ...) to the bit.
constexpr int kMaxSlip = 48;
enum {
  SLIP_FCC = 1, SLIP_BCC = 2, SLIP_ROTERS = 6, SLIP_BCC12 = 7, SLIP_BCC48 = 8, SLIP_BCC24 = 80
};

inline int slip_count(int set) {
  switch (set) {
    case SLIP_FCC: case SLIP_BCC: case SLIP_ROTERS: case SLIP_BCC12: return 12;
    case SLIP_BCC48: return 48;
    case SLIP_BCC24: return 24;
    default: return 0;
  }
}

inline void unit_miller(const int v[3], double out[3]) {
  const double m = (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) {
    const double a = (double)(v[k] < 0 ? -v[k] : v[k]) / sqrt(m);
    out[k] = v[k] < 0 ? -a : (v[k] > 0 ? a : 0.0);
  }
}

// set: one of the SLIP_* sets above; ns = 0 for another set
inline void slip_systems(int set, int& ns, double b[kMaxSlip][3], double n[kMaxSlip][3]) {
  static const int fb[12][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0}, {0, 1, 1},  {1, 0, 1},  {1, -1, 0},
                                {0, -1, 1}, {-1, 0, -1}, {1, 1, 0}, {0, 1, 1},  {1, 0, -1}, {-1, -1, 0}};
  static const int fn[12][3] = {{1, 1, 1},  {1, 1, 1},  {1, 1, 1},  {-1, -1, 1}, {-1, -1, 1}, {-1, -1, 1},
                                {-1, 1, 1}, {-1, 1, 1}, {-1, 1, 1}, {1, -1, 1},  {1, -1, 1},  {1, -1, 1}};
  static const int bb[48][3] = {
      {-1, 1, 1}, {1, -1, 1}, {1, 1, 1},  {1, 1, -1}, {1, 1, -1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1},
      {1, 1, -1}, {1, -1, 1}, {1, 1, 1},  {-1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},
      {1, -1, 1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1},
      {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1},
      {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1},
      {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1}};
  static const int bn[48][3] = {
      {1, 1, 0},  {1, 1, 0},  {1, -1, 0}, {1, -1, 0}, {1, 0, 1},  {1, 0, 1},  {1, 0, -1}, {1, 0, -1},
      {0, 1, 1},  {0, 1, 1},  {0, 1, -1}, {0, 1, -1}, {1, 1, 2},  {-1, 1, 2}, {1, -1, 2}, {1, 1, -2},
      {1, 2, 1},  {-1, 2, 1}, {1, -2, 1}, {1, 2, -1}, {2, 1, 1},  {-2, 1, 1}, {2, -1, 1}, {2, 1, -1},
      {1, 2, 3},  {-1, 2, 3}, {1, -2, 3}, {1, 2, -3}, {1, 3, 2},  {-1, 3, 2}, {1, -3, 2}, {1, 3, -2},
      {2, 1, 3},  {-2, 1, 3}, {2, -1, 3}, {2, 1, -3}, {2, 3, 1},  {-2, 3, 1}, {2, -3, 1}, {2, 3, -1},
      {3, 1, 2},  {-3, 1, 2}, {3, -1, 2}, {3, 1, -2}, {3, 2, 1},  {-3, 2, 1}, {3, -2, 1}, {3, 2, -1}};
  static const int rb[12][3] = {{1, -1, 0}, {1, 0, -1}, {0, 1, -1}, {1, 1, 0}, {1, 0, 1},  {0, 1, -1},
                                {1, 1, 0},  {1, 0, -1}, {0, 1, 1},  {1, -1, 0}, {1, 0, 1}, {0, 1, 1}};
  static const int rn[12][3] = {{1, 1, 1},   {-1, -1, -1}, {1, 1, 1},   {1, -1, -1},
                                {-1, 1, 1},  {1, -1, -1},  {1, -1, 1},  {1, -1, 1},
                                {-1, 1, -1}, {-1, -1, 1},  {-1, -1, 1}, {1, 1, -1}};
  ns = slip_count(set);
  const int(*tb)[3] = bb;
  const int(*tn)[3] = bn;
  if (set == SLIP_FCC) { tb = fb; tn = fn; }
  else if (set == SLIP_BCC) { tb = fn; tn = fb; }
  else if (set == SLIP_ROTERS) { tb = rb; tn = rn; }
  for (int s = 0; s < ns; ++s) {
    unit_miller(tb[s], b[s]);
    unit_miller(tn[s], n[s]);
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
inline void build_orient(const double stiff[6][6], const double b[kMaxSlip][3], const double n[kMaxSlip][3],
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
