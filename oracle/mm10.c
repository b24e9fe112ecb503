/*
 * mm10.c -- oracle restatement of the per-voxel stress update of CPFFT:
 * kinematics (drive_eps_sig.f, polar.f, qmply1.f), the WARP3D mm10 crystal
 * plasticity model with Voce hardening (mm10_a.f, mm10_b.f), the consistent
 * tangent and its exact pull-back to dP/dF (gptns1.f, cep2A.f).
 *
 * TEST INFRASTRUCTURE ONLY (see cpfft_oracle.h).
 *
 * Matrices are C row-major m[i][j] == Fortran m(i+1,j+1).  Expressions keep
 * the reference's evaluation order (Fortran left-to-right within a
 * precedence level); this file must be compiled with -ffp-contract=off.
 */
#include "cpfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define NSMAX ORC_MAX_SLIP

static const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, half = 0.5;
static const double pi = 3.141592653589793, one_eighty = 180.0;

/* ------------------------------------------------------------------------ */
/* small dense helpers                                                       */
/* ------------------------------------------------------------------------ */

/* LU with partial pivoting + solve, nrhs right-hand sides (DGESV semantics,
 * used at mm10_a.f:791,810,3065,3212 and FFT_nr3.f:417).  A is n x n
 * row-major (destroyed), B is n x nrhs row-major (overwritten).  Returns
 * 0 on success, k>0 if U(k,k) is exactly zero (LAPACK info convention). */
static int lu_solve(int n, double* A, int nrhs, double* B) {
  int ipiv[16];
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double amax = fabs(A[j * n + j]);
    for (int i = j + 1; i < n; ++i) {
      double v = fabs(A[i * n + j]);
      if (v > amax) { amax = v; p = i; }
    }
    ipiv[j] = p;
    if (A[p * n + j] != zero) {
      if (p != j)
        for (int k = 0; k < n; ++k) {
          double t = A[j * n + k]; A[j * n + k] = A[p * n + k]; A[p * n + k] = t;
        }
      double rp = one / A[j * n + j];
      for (int i = j + 1; i < n; ++i) A[i * n + j] *= rp;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int i = j + 1; i < n; ++i) {
      double l = A[i * n + j];
      for (int k = j + 1; k < n; ++k) A[i * n + k] -= l * A[j * n + k];
    }
  }
  if (info) return info;
  for (int j = 0; j < n; ++j) {
    int p = ipiv[j];
    if (p != j)
      for (int r = 0; r < nrhs; ++r) {
        double t = B[j * nrhs + r]; B[j * nrhs + r] = B[p * nrhs + r]; B[p * nrhs + r] = t;
      }
  }
  for (int r = 0; r < nrhs; ++r) {
    for (int i = 1; i < n; ++i) {
      double s = B[i * nrhs + r];
      for (int k = 0; k < i; ++k) s -= A[i * n + k] * B[k * nrhs + r];
      B[i * nrhs + r] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = B[i * nrhs + r];
      for (int k = i + 1; k < n; ++k) s -= A[i * n + k] * B[k * nrhs + r];
      B[i * nrhs + r] = s / A[i * n + i];
    }
  }
  return 0;
}

/* Cholesky solve of an SPD 6x6 system (DPOSV, mm10_a.f:3537). */
static void chol_solve6(const double S[6][6], double b[6]) {
  double L[6][6];
  memset(L, 0, sizeof L);
  for (int j = 0; j < 6; ++j) {
    double d = S[j][j];
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
    L[j][j] = sqrt(d);
    for (int i = j + 1; i < 6; ++i) {
      double s = S[i][j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      L[i][j] = s / L[j][j];
    }
  }
  for (int i = 0; i < 6; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i][k] * b[k];
    b[i] = s / L[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < 6; ++k) s -= L[k][i] * b[k];
    b[i] = s / L[i][i];
  }
}

/* mm10_a_mult_type_1 (mm10_a.f:3702): a = b * c (3x3) */
static void mult33(double a[3][3], const double b[3][3], const double c[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = b[i][0] * c[0][j] + b[i][1] * c[1][j] + b[i][2] * c[2][j];
}
/* mm10_a_mult_type_2 (mm10_a.f:3750): a = b * c (6x6 * 6) */
static void mv6(double a[6], const double b[6][6], const double c[6]) {
  for (int i = 0; i < 6; ++i) a[i] = b[i][0] * c[0];
  for (int j = 1; j < 6; ++j)
    for (int i = 0; i < 6; ++i) a[i] = a[i] + b[i][j] * c[j];
}
/* mm10_a_mult_type_4 (mm10_a.f:3777): a = b * c + const * d */
static void mv6_plus(double a[6], const double b[6][6], const double c[6], const double d[6],
                     double cst) {
  for (int i = 0; i < 6; ++i) a[i] = cst * d[i];
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) a[i] = a[i] + b[i][j] * c[j];
}
/* mm10_a_mult_type_3 (mm10_a.f:3877): a = b * c (3x3 * 3) */
static void mv3(double a[3], const double b[3][3], const double c[3]) {
  for (int i = 0; i < 3; ++i) a[i] = b[i][0] * c[0] + b[i][1] * c[1] + b[i][2] * c[2];
}

/* mm10_rt2rve (mm10_a.f:1400-1447) */
static void rt2rve(const double rt[3][3], double rv[6][6]) {
  rv[0][0] = rt[0][0] * rt[0][0];
  rv[0][1] = rt[0][1] * rt[0][1];
  rv[0][2] = rt[0][2] * rt[0][2];
  rv[0][3] = two * rt[0][0] * rt[0][1];
  rv[0][4] = two * rt[0][2] * rt[0][1];
  rv[0][5] = two * rt[0][0] * rt[0][2];
  rv[1][0] = rt[1][0] * rt[1][0];
  rv[1][1] = rt[1][1] * rt[1][1];
  rv[1][2] = rt[1][2] * rt[1][2];
  rv[1][3] = two * rt[1][0] * rt[1][1];
  rv[1][4] = two * rt[1][2] * rt[1][1];
  rv[1][5] = two * rt[1][0] * rt[1][2];
  rv[2][0] = rt[2][0] * rt[2][0];
  rv[2][1] = rt[2][1] * rt[2][1];
  rv[2][2] = rt[2][2] * rt[2][2];
  rv[2][3] = two * rt[2][0] * rt[2][1];
  rv[2][4] = two * rt[2][2] * rt[2][1];
  rv[2][5] = two * rt[2][0] * rt[2][2];
  rv[3][0] = rt[0][0] * rt[1][0];
  rv[3][1] = rt[0][1] * rt[1][1];
  rv[3][2] = rt[0][2] * rt[1][2];
  rv[3][3] = rt[0][0] * rt[1][1] + rt[1][0] * rt[0][1];
  rv[3][4] = rt[0][1] * rt[1][2] + rt[0][2] * rt[1][1];
  rv[3][5] = rt[0][0] * rt[1][2] + rt[0][2] * rt[1][0];
  rv[4][0] = rt[1][0] * rt[2][0];
  rv[4][1] = rt[2][1] * rt[1][1];
  rv[4][2] = rt[1][2] * rt[2][2];
  rv[4][3] = rt[1][0] * rt[2][1] + rt[1][1] * rt[2][0];
  rv[4][4] = rt[1][1] * rt[2][2] + rt[2][1] * rt[1][2];
  rv[4][5] = rt[1][0] * rt[2][2] + rt[1][2] * rt[2][0];
  rv[5][0] = rt[0][0] * rt[2][0];
  rv[5][1] = rt[0][1] * rt[2][1];
  rv[5][2] = rt[0][2] * rt[2][2];
  rv[5][3] = rt[0][0] * rt[2][1] + rt[0][1] * rt[2][0];
  rv[5][4] = rt[0][1] * rt[2][2] + rt[0][2] * rt[2][1];
  rv[5][5] = rt[0][0] * rt[2][2] + rt[2][0] * rt[0][2];
}

/* mm10_rt2rvw (mm10_a.f:1461-1479) */
static void rt2rvw(const double rt[3][3], double rv[3][3]) {
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

/* mm10_symsw (mm10_b.f:1505-1533) */
static void symsw(const double s[6], const double w[3], double sw[6]) {
  sw[0] = s[3] * w[2] - s[5] * w[1];
  sw[1] = s[3] * w[2] - s[4] * w[0];
  sw[2] = s[5] * w[1] + s[4] * w[0];
  sw[3] = half * (w[2] * (s[0] - s[1]) + w[0] * s[5] - w[1] * s[4]);
  sw[4] = half * (w[0] * (s[1] - s[2]) + w[1] * s[3] + w[2] * s[5]);
  sw[5] = half * (w[1] * (s[0] - s[2]) + w[0] * s[3] - w[2] * s[4]);
}

/* mm10_iw (mm10_b.f:1605-1634) */
static void iw_mat(const double w[3], double iw[6][6]) {
  memset(iw, 0, 36 * sizeof(double));
  iw[0][3] = two * w[2];
  iw[0][5] = -two * w[1];
  iw[1][3] = two * w[2];
  iw[1][4] = -two * w[0];
  iw[2][4] = two * w[0];
  iw[2][5] = two * w[1];
  iw[3][0] = w[2];
  iw[3][1] = -w[2];
  iw[3][4] = -w[1];
  iw[3][5] = w[0];
  iw[4][1] = w[0];
  iw[4][2] = -w[0];
  iw[4][3] = w[1];
  iw[4][5] = w[2];
  iw[5][0] = w[1];
  iw[5][2] = -w[1];
  iw[5][3] = w[0];
  iw[5][4] = -w[2];
}

static double fsign(double x) { return copysign(one, x); } /* Fortran sign(one,x) */

/* ------------------------------------------------------------------------ */
/* crystal setup                                                             */
/* ------------------------------------------------------------------------ */

/* symmetric 6x6 inverse (mm10_invsym, mm10_a.f:1358) by Gauss-Jordan */
static void invert6(double A[6][6]) {
  double M[36], B[36];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { M[i * 6 + j] = A[i][j]; B[i * 6 + j] = (i == j) ? one : zero; }
  lu_solve(6, M, 6, B);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) A[i][j] = B[i * 6 + j];
  /* symmetrise the way mm10_invsym copies the upper triangle down */
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < i; ++j) A[i][j] = A[j][i];
}

int orc_crystal_init(orc_crystal* c, int slip_set, double e, double nu, double mu,
                     double rate_n, double tau_y, double tau_v, double theta_0,
                     double voche_m, int alter_mode, double gamma_bar) {
  memset(c, 0, sizeof *c);
  /* slip systems: mod_crystals.f:438-515 (slip_type 1, fcc), 516-593 (2,
   * bcc: the fcc directions as normals and normals as directions), 604-704
   * (6, roters), 705-811 (7, bcc12: the first 12 systems of bcc48) and
   * 812-1207 (8, bcc48: {110}<111>, {112}<111>, {123}<111>); slip set 80
   * is the first 24 systems of bcc48.  The reference assigns each entry a
   * named factor (f = 1/sqrt(2), f3 = 1/sqrt(3), f211 = 2/sqrt(6),
   * f312 = 3/sqrt(14), ...): component v of a Miller triple of squared length
   * m becomes |v|/sqrt(m) with v's sign, which is that factor to the bit. */
  static const int fb[12][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0}, {0, 1, 1},
                                {1, 0, 1},  {1, -1, 0}, {0, -1, 1}, {-1, 0, -1},
                                {1, 1, 0},  {0, 1, 1},  {1, 0, -1}, {-1, -1, 0}};
  static const int fn[12][3] = {{1, 1, 1},  {1, 1, 1},  {1, 1, 1},  {-1, -1, 1},
                                {-1, -1, 1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, 1},
                                {-1, 1, 1}, {1, -1, 1},  {1, -1, 1}, {1, -1, 1}};
  /* bcc48 directions <111> and normals: 1..12 {110}, 13..24 {112},
   * 25..48 {123} */
  static const int bb[48][3] = {
      {-1, 1, 1}, {1, -1, 1}, {1, 1, 1},  {1, 1, -1}, {1, 1, -1}, {-1, 1, 1},
      {1, 1, 1},  {1, -1, 1}, {1, 1, -1}, {1, -1, 1}, {1, 1, 1},  {-1, 1, 1},
      {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1}, {1, 1, -1},
      {1, 1, 1},  {-1, 1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1},
      {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, -1, 1}, {1, 1, -1},
      {1, 1, 1},  {-1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},
      {1, -1, 1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1}, {-1, 1, 1}, {1, 1, 1},
      {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},  {1, 1, -1}, {1, -1, 1}};
  static const int bn[48][3] = {
      {1, 1, 0},  {1, 1, 0},  {1, -1, 0}, {1, -1, 0}, {1, 0, 1},  {1, 0, 1},
      {1, 0, -1}, {1, 0, -1}, {0, 1, 1},  {0, 1, 1},  {0, 1, -1}, {0, 1, -1},
      {1, 1, 2},  {-1, 1, 2}, {1, -1, 2}, {1, 1, -2}, {1, 2, 1},  {-1, 2, 1},
      {1, -2, 1}, {1, 2, -1}, {2, 1, 1},  {-2, 1, 1}, {2, -1, 1}, {2, 1, -1},
      {1, 2, 3},  {-1, 2, 3}, {1, -2, 3}, {1, 2, -3}, {1, 3, 2},  {-1, 3, 2},
      {1, -3, 2}, {1, 3, -2}, {2, 1, 3},  {-2, 1, 3}, {2, -1, 3}, {2, 1, -3},
      {2, 3, 1},  {-2, 3, 1}, {2, -3, 1}, {2, 3, -1}, {3, 1, 2},  {-3, 1, 2},
      {3, -1, 2}, {3, 1, -2}, {3, 2, 1},  {-3, 2, 1}, {3, -2, 1}, {3, 2, -1}};
  /* roters (slip_type 6): directions <110>, normals {111} */
  static const int rb[12][3] = {{1, -1, 0}, {1, 0, -1}, {0, 1, -1}, {1, 1, 0},
                                {1, 0, 1},  {0, 1, -1}, {1, 1, 0},  {1, 0, -1},
                                {0, 1, 1},  {1, -1, 0}, {1, 0, 1},  {0, 1, 1}};
  static const int rn[12][3] = {{1, 1, 1},   {-1, -1, -1}, {1, 1, 1},   {1, -1, -1},
                                {-1, 1, 1},  {1, -1, -1},  {1, -1, 1},  {1, -1, 1},
                                {-1, 1, -1}, {-1, -1, 1},  {-1, -1, 1}, {1, 1, -1}};
  const int(*tb)[3], (*tn)[3];
  if (slip_set == 1) { c->nslip = 12; tb = fb; tn = fn; }
  else if (slip_set == 2) { c->nslip = 12; tb = fn; tn = fb; }
  else if (slip_set == 6) { c->nslip = 12; tb = rb; tn = rn; }
  else if (slip_set == 7) { c->nslip = 12; tb = bb; tn = bn; }
  else if (slip_set == 8) { c->nslip = 48; tb = bb; tn = bn; }
  else if (slip_set == 80) { c->nslip = 24; tb = bb; tn = bn; }
  else return -1;
  for (int s = 0; s < c->nslip; ++s) {
    const int* v[2] = {tb[s], tn[s]};
    double* o[2] = {c->bi[s], c->ni[s]};
    for (int a = 0; a < 2; ++a) {
      double m = (double)(v[a][0] * v[a][0] + v[a][1] * v[a][1] + v[a][2] * v[a][2]);
      for (int k = 0; k < 3; ++k) {
        double mag = (double)abs(v[a][k]) / sqrt(m);
        o[a][k] = v[a][k] < 0 ? -mag : (v[a][k] > 0 ? mag : zero);
      }
    }
  }
  /* cubic elasticity via the compliance (mod_crystals.f:1833-1879) */
  double fl[6][6];
  memset(fl, 0, sizeof fl);
  fl[0][0] = 1 / e; fl[0][1] = -nu / e; fl[0][2] = -nu / e;
  fl[1][0] = -nu / e; fl[1][1] = 1 / e; fl[1][2] = -nu / e;
  fl[2][0] = -nu / e; fl[2][1] = -nu / e; fl[2][2] = 1 / e;
  fl[3][3] = 1 / mu; fl[4][4] = 1 / mu; fl[5][5] = 1 / mu;
  invert6(fl);
  memcpy(c->stiff, fl, sizeof fl);
  c->rate_n = rate_n;
  c->tau_y = tau_y;
  c->tau_v = tau_v;
  c->theta_0 = theta_0;
  c->voche_m = voche_m;
  c->eps_dot_0_y = gamma_bar;
  c->iD_v = 0.0;
  c->alter_mode = alter_mode;
  /* local solver defaults, mod_crystals.f:392-403 */
  c->atol = 1.0e-5;
  c->atol1 = 1.0e-5;
  c->rtol = 5.0e-5;
  c->rtol1 = 1.0e-5;
  c->miter = 30;
  return 0;
}

void orc_crystal_stiffness(const orc_crystal* c, double out36[36]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) out36[i * 6 + j] = c->stiff[i][j];
}

/* mm10_rotation_matrix, Kocks convention, degrees (mm10_a.f:1287-1345) */
void orc_rotation_matrix(const double angles[3], double g[9]) {
  double psi = angles[0] * pi / one_eighty;
  double theta = angles[1] * pi / one_eighty;
  double phi = angles[2] * pi / one_eighty;
  g[0] = -sin(psi) * sin(phi) - cos(psi) * cos(phi) * cos(theta);
  g[1] = cos(psi) * sin(phi) - sin(psi) * cos(phi) * cos(theta);
  g[2] = cos(phi) * sin(theta);
  g[3] = sin(psi) * cos(phi) - cos(psi) * sin(phi) * cos(theta);
  g[4] = -cos(psi) * cos(phi) - sin(psi) * sin(phi) * cos(theta);
  g[5] = sin(phi) * sin(theta);
  g[6] = cos(psi) * sin(theta);
  g[7] = sin(psi) * sin(theta);
  g[8] = cos(theta);
}

/* per-voxel invariants: setup_mm10_rknstr (drive_eps_sig.f:537-1002) */
typedef struct {
  double g[3][3];
  double C[6][6];
  double ms[NSMAX][6], qs[NSMAX][3];
} orient_t;

static void setup_orient(const orc_crystal* cp, const double angles[3], orient_t* o) {
  double gv[9];
  orc_rotation_matrix(angles, gv);
  double tr[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { o->g[i][j] = gv[i * 3 + j]; tr[j][i] = gv[i * 3 + j]; }
  for (int s = 0; s < cp->nslip; ++s) {
    double bs[3], ns[3], A[3][3], At[3][3];
    mv3(bs, tr, cp->bi[s]);
    mv3(ns, tr, cp->ni[s]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) { A[i][j] = bs[i] * ns[j]; At[j][i] = A[i][j]; }
    double As[3][3], Aa[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        As[i][j] = half * (A[i][j] + At[i][j]);
        Aa[i][j] = half * (A[i][j] - At[i][j]);
      }
    /* mm10_et2ev / mm10_wt2wv (mm10_a.f:1492-1534) */
    o->ms[s][0] = As[0][0];
    o->ms[s][1] = As[1][1];
    o->ms[s][2] = As[2][2];
    o->ms[s][3] = two * As[0][1];
    o->ms[s][4] = two * As[1][2];
    o->ms[s][5] = two * As[0][2];
    o->qs[s][0] = Aa[1][2];
    o->qs[s][1] = Aa[0][2];
    o->qs[s][2] = Aa[0][1];
  }
  double Rs[6][6], tmp[6][6];
  rt2rve(tr, Rs);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = zero;
      for (int k = 0; k < 6; ++k) s = s + cp->stiff[i][k] * Rs[j][k];
      tmp[i][j] = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = zero;
      for (int k = 0; k < 6; ++k) s = s + Rs[i][k] * tmp[k][j];
      o->C[i][j] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* kinematics                                                                */
/* ------------------------------------------------------------------------ */

/* evcmp1_new (polar.f:224-307): Cardano eigenvalues of symmetric 3x3 c */
static void evcmp1_new(const double c[6], double lam[3]) {
  const double third = 0.3333333333333333333, oneptfive = 1.5, thirteenptfive = 13.5,
               twentyseven = 27.0, quarter = 0.25, sixpt75 = 6.75,
               oneroot3 = 0.5773502691896258;
  double m11 = c[0], m12 = c[1], m13 = c[3], m22 = c[2], m23 = c[4], m33 = c[5];
  double de = m12 * m23;
  double dd = m12 * m12;
  double ee = m23 * m23;
  double ff = m13 * m13;
  double m = m11 + m22 + m33;
  double c1 = (m11 * m22 + m11 * m33 + m22 * m33) - (dd + ee + ff);
  double c0 = m33 * dd + m11 * ee + m22 * ff - m11 * m22 * m33 - two * m13 * de;
  double p = m * m - three * c1;
  double q = m * (p - oneptfive * c1) - thirteenptfive * c0;
  double sqrtp = sqrt(fabs(p));
  double phi = twentyseven * (quarter * c1 * c1 * (p - c1) + c0 * (q + sixpt75 * c0));
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
  lam[0] = e1;
  lam[1] = e2;
  lam[2] = e3;
}

/* rtcmp1 / irscp1 / ivcmp1 (polar.f:18-211): R = F U^-1 */
static void rtcmp1(const double f[3][3], double r[3][3]) {
  double c[6], cc[6], ev[3];
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
  evcmp1_new(c, ev);
  ev[0] = sqrt(ev[0]);
  ev[1] = sqrt(ev[1]);
  ev[2] = sqrt(ev[2]);
  double iu = ev[0] + ev[1] + ev[2];
  double iiu = ev[0] * ev[1] + ev[1] * ev[2] + ev[0] * ev[2];
  double iiiu = ev[0] * ev[1] * ev[2];
  double a2 = one / (iiiu * (iu * iiu - iiiu));
  double b2 = iu * iiu * iiu - iiiu * (iu * iu + iiu);
  double c2 = -iiiu - iu * (iu * iu - two * iiu);
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

void orc_rtcmp1(const double F[9], double R[9]) {
  double f[3][3], r[3][3];
  for (int i = 0; i < 9; ++i) f[i / 3][i % 3] = F[i];
  rtcmp1(f, r);
  for (int i = 0; i < 9; ++i) R[i] = r[i / 3][i % 3];
}

/* inv33 (drive_eps_sig.f:1017-1107) */
static void inv33(const double a[3][3], double g[3][3], double* dj) {
  double j1 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  double j2 = a[1][0] * a[2][2] - a[1][2] * a[2][0];
  double j3 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  double d = a[0][0] * j1 - a[0][1] * j2 + a[0][2] * j3;
  *dj = d;
  g[0][0] = j1 / d;
  g[1][0] = -j2 / d;
  g[2][0] = j3 / d;
  g[0][1] = (a[2][1] * a[0][2] - a[0][1] * a[2][2]) / d;
  g[1][1] = (a[0][0] * a[2][2] - a[2][0] * a[0][2]) / d;
  g[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / d;
  g[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  g[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / d;
  g[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
}

/* mul33 (drive_eps_sig.f:1110-1166): Voigt (engineering shear) of AA*BB */
static void mul33(const double A[3][3], const double B[3][3], double C[6]) {
  C[0] = A[0][0] * B[0][0] + A[0][1] * B[1][0] + A[0][2] * B[2][0];
  C[3] = A[1][0] * B[0][0] + A[1][1] * B[1][0] + A[1][2] * B[2][0] + A[0][0] * B[0][1] +
         A[0][1] * B[1][1] + A[0][2] * B[2][1];
  C[5] = A[2][0] * B[0][0] + A[2][1] * B[1][0] + A[2][2] * B[2][0] + A[0][0] * B[0][2] +
         A[0][1] * B[1][2] + A[0][2] * B[2][2];
  C[1] = A[1][0] * B[0][1] + A[1][1] * B[1][1] + A[1][2] * B[2][1];
  C[4] = A[2][0] * B[0][1] + A[2][1] * B[1][1] + A[2][2] * B[2][1] + A[1][0] * B[0][2] +
         A[1][1] * B[1][2] + A[1][2] * B[2][2];
  C[2] = A[2][0] * B[0][2] + A[2][1] * B[1][2] + A[2][2] * B[2][2];
}

/* getrm1 opt 1 (polar.f:680-760): unrotation of engineering-shear strain */
static void getrm1_opt1(const double r[3][3], double q[6][6]) {
  q[0][0] = r[0][0] * r[0][0];
  q[0][1] = r[1][0] * r[1][0];
  q[0][2] = r[2][0] * r[2][0];
  q[0][3] = r[0][0] * r[1][0];
  q[0][4] = r[2][0] * r[1][0];
  q[0][5] = r[0][0] * r[2][0];
  q[1][0] = r[0][1] * r[0][1];
  q[1][1] = r[1][1] * r[1][1];
  q[1][2] = r[2][1] * r[2][1];
  q[1][3] = r[0][1] * r[1][1];
  q[1][4] = r[2][1] * r[1][1];
  q[1][5] = r[0][1] * r[2][1];
  q[2][0] = r[0][2] * r[0][2];
  q[2][1] = r[1][2] * r[1][2];
  q[2][2] = r[2][2] * r[2][2];
  q[2][3] = r[0][2] * r[1][2];
  q[2][4] = r[2][2] * r[1][2];
  q[2][5] = r[0][2] * r[2][2];
  q[3][0] = two * r[0][0] * r[0][1];
  q[3][1] = two * r[1][0] * r[1][1];
  q[3][2] = two * r[2][0] * r[2][1];
  q[3][3] = r[0][0] * r[1][1] + r[0][1] * r[1][0];
  q[3][4] = r[1][0] * r[2][1] + r[2][0] * r[1][1];
  q[3][5] = r[0][0] * r[2][1] + r[2][0] * r[0][1];
  q[4][0] = two * r[0][1] * r[0][2];
  q[4][1] = two * r[1][2] * r[1][1];
  q[4][2] = two * r[2][1] * r[2][2];
  q[4][3] = r[0][1] * r[1][2] + r[1][1] * r[0][2];
  q[4][4] = r[1][1] * r[2][2] + r[1][2] * r[2][1];
  q[4][5] = r[0][1] * r[2][2] + r[2][1] * r[0][2];
  q[5][0] = two * r[0][0] * r[0][2];
  q[5][1] = two * r[1][0] * r[1][2];
  q[5][2] = two * r[2][0] * r[2][2];
  q[5][3] = r[0][0] * r[1][2] + r[1][0] * r[0][2];
  q[5][4] = r[1][0] * r[2][2] + r[2][0] * r[1][2];
  q[5][5] = r[0][0] * r[2][2] + r[0][2] * r[2][0];
}

/* getrm1 opt 2 (polar.f:762-805): rotation of Cauchy stress T = R t R^T */
static void getrm1_opt2(const double r[3][3], double q[6][6]) {
  q[0][0] = r[0][0] * r[0][0];
  q[0][1] = r[0][1] * r[0][1];
  q[0][2] = r[0][2] * r[0][2];
  q[0][3] = two * r[0][0] * r[0][1];
  q[0][4] = two * r[0][2] * r[0][1];
  q[0][5] = two * r[0][0] * r[0][2];
  q[1][0] = r[1][0] * r[1][0];
  q[1][1] = r[1][1] * r[1][1];
  q[1][2] = r[1][2] * r[1][2];
  q[1][3] = two * r[1][0] * r[1][1];
  q[1][4] = two * r[1][2] * r[1][1];
  q[1][5] = two * r[1][0] * r[1][2];
  q[2][0] = r[2][0] * r[2][0];
  q[2][1] = r[2][1] * r[2][1];
  q[2][2] = r[2][2] * r[2][2];
  q[2][3] = two * r[2][0] * r[2][1];
  q[2][4] = two * r[2][2] * r[2][1];
  q[2][5] = two * r[2][0] * r[2][2];
  q[3][0] = r[0][0] * r[1][0];
  q[3][1] = r[0][1] * r[1][1];
  q[3][2] = r[0][2] * r[1][2];
  q[3][3] = r[0][0] * r[1][1] + r[1][0] * r[0][1];
  q[3][4] = r[0][1] * r[1][2] + r[0][2] * r[1][1];
  q[3][5] = r[0][0] * r[1][2] + r[0][2] * r[1][0];
  q[4][0] = r[1][0] * r[2][0];
  q[4][1] = r[2][1] * r[1][1];
  q[4][2] = r[1][2] * r[2][2];
  q[4][3] = r[1][0] * r[2][1] + r[1][1] * r[2][0];
  q[4][4] = r[1][1] * r[2][2] + r[2][1] * r[1][2];
  q[4][5] = r[1][0] * r[2][2] + r[1][2] * r[2][0];
  q[5][0] = r[0][0] * r[2][0];
  q[5][1] = r[0][1] * r[2][1];
  q[5][2] = r[0][2] * r[2][2];
  q[5][3] = r[0][0] * r[2][1] + r[0][1] * r[2][0];
  q[5][4] = r[0][1] * r[2][2] + r[0][2] * r[2][1];
  q[5][5] = r[0][0] * r[2][2] + r[2][0] * r[0][2];
}

/* qmply1 (qmply1.f:15-49) */
static void qmply1(const double q[6][6], const double m1[6], double m2[6]) {
  for (int j = 0; j < 6; ++j)
    m2[j] = q[j][0] * m1[0] + q[j][1] * m1[1] + q[j][2] * m1[2] + q[j][3] * m1[3] +
            q[j][4] * m1[4] + q[j][5] * m1[5];
}

/* cs2p (drive_eps_sig.f:1182-1224): P = J sigma F^-T */
static void cs2p(const double cs[6], const double fi[3][3], double dF, double P[9]) {
  P[0] = dF * (cs[0] * fi[0][0] + cs[3] * fi[0][1] + cs[5] * fi[0][2]);
  P[1] = dF * (cs[0] * fi[1][0] + cs[3] * fi[1][1] + cs[5] * fi[1][2]);
  P[2] = dF * (cs[0] * fi[2][0] + cs[3] * fi[2][1] + cs[5] * fi[2][2]);
  P[3] = dF * (cs[3] * fi[0][0] + cs[1] * fi[0][1] + cs[4] * fi[0][2]);
  P[4] = dF * (cs[3] * fi[1][0] + cs[1] * fi[1][1] + cs[4] * fi[1][2]);
  P[5] = dF * (cs[3] * fi[2][0] + cs[1] * fi[2][1] + cs[4] * fi[2][2]);
  P[6] = dF * (cs[5] * fi[0][0] + cs[4] * fi[0][1] + cs[2] * fi[0][2]);
  P[7] = dF * (cs[5] * fi[1][0] + cs[4] * fi[1][1] + cs[2] * fi[1][2]);
  P[8] = dF * (cs[5] * fi[2][0] + cs[4] * fi[2][1] + cs[2] * fi[2][2]);
}

/* ------------------------------------------------------------------------ */
/* mm10 Voce model                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  double R[3][3], Rp[3][3];
  double stress[6], D[6];
  double slip_incs[NSMAX];
  double tau_tilde, tt_rate;
  double tinc, dg, mu_harden;
  double work_inc, p_work_inc, p_strain_inc;
  double ms[NSMAX][6], qs[NSMAX][3], qc[NSMAX][3];
  double tangent[6][6];
} cstate;

typedef struct {
  const orc_crystal* cp;
  const orient_t* o;  /* props%ms, props%qs, props%stiffness (rotated) */
} props_t;

#define NS (P->cp->nslip)

/* mm10_setup_np1 (mm10_a.f:1659-1732) */
static void setup_np1(const double R[3][3], const double D[6], double dt, cstate* s) {
  memset(s, 0, sizeof *s);
  s->tinc = dt;
  memcpy(s->D, D, 6 * sizeof(double));
  memcpy(s->R, R, 9 * sizeof(double));
}

/* mm10_setup (mm10_a.f:830-962) + mm10_setup_voche (mm10_a.f:2057-2082).
 * tau_l is identically zero on this path (gradFeinv = 0, mm10_a.f:924-950). */
static void mm10_setup(const props_t* P, cstate* np1, const cstate* n) {
  const double twothirds = 2.0 / 3.0;
  double t1 = np1->D[0] * np1->D[0] + np1->D[1] * np1->D[1] + np1->D[2] * np1->D[2];
  double t2 = np1->D[3] * np1->D[3] + np1->D[4] * np1->D[4] + np1->D[5] * np1->D[5];
  np1->dg = sqrt(twothirds * (t1 + half * t2));
  double tn[3][3], RE[6][6], RW[3][3], work[3][3], RWC[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tn[i][j] = n->Rp[j][i];
  rt2rve(tn, RE);
  rt2rvw(tn, RW);
  mult33(work, np1->R, tn);
  rt2rvw(work, RWC);
  for (int i = 0; i < NS; ++i) {
    mv6(np1->ms[i], RE, P->o->ms[i]);
    mv3(np1->qs[i], RW, P->o->qs[i]);
    mv3(np1->qc[i], RWC, P->o->qs[i]);
  }
  np1->mu_harden = P->o->C[5][5];
  if (P->cp->alter_mode) np1->dg = P->cp->eps_dot_0_y * np1->tinc;
}

/* mm10_rs (mm10_b.f:1837-1858) */
static double rs_of(const cstate* np1, const double s[6], int i) {
  return s[0] * np1->ms[i][0] + s[1] * np1->ms[i][1] + s[2] * np1->ms[i][2] +
         s[3] * np1->ms[i][3] + s[4] * np1->ms[i][4] + s[5] * np1->ms[i][5];
}
/* mm10_slipinc (mm10_b.f:1805-1835) */
static double slipinc_of(const props_t* P, const cstate* np1, const double s[6], double tt,
                         int i) {
  double rs = rs_of(np1, s, i);
  return np1->dg / tt * pow(fabs(rs / tt), P->cp->rate_n - one) * rs;
}

/* mm10_form_dbarp (mm10_b.f:1161-1229), Voce branch */
static void form_dbarp(const props_t* P, const cstate* np1, const double s[6], double tt,
                       double dbar[6]) {
  for (int k = 0; k < 6; ++k) dbar[k] = zero;
  for (int i = 0; i < NS; ++i) {
    double si = slipinc_of(P, np1, s, tt, i);
    double rs = rs_of(np1, s, i);
    double f = rs * np1->tinc * P->cp->iD_v + si;
    for (int k = 0; k < 6; ++k) dbar[k] = dbar[k] + f * np1->ms[i][k];
  }
}

/* mm10_form_wp (mm10_b.f:1304-1391) / mm10_form_wbarp (mm10_b.f:1231-1302),
 * Voce branch; use_qc selects qc (wp) or qs (wbarp) */
static void form_w(const props_t* P, const cstate* np1, const double s[6], double tt,
                   int use_qc, double w[3]) {
  w[0] = w[1] = w[2] = zero;
  for (int i = 0; i < NS; ++i) {
    const double* q = use_qc ? np1->qc[i] : np1->qs[i];
    double si = slipinc_of(P, np1, s, tt, i);
    for (int k = 0; k < 3; ++k) w[k] = w[k] + si * q[k];
    double rs = rs_of(np1, s, i);
    for (int k = 0; k < 3; ++k) w[k] = w[k] + rs * np1->tinc * P->cp->iD_v * q[k];
  }
}

/* mm10_formR1 (mm10_b.f:1065-1113) */
static void formR1(const props_t* P, const cstate* np1, const cstate* n, const double s[6],
                   double tt, double R1[6]) {
  double dbarp[6], wp[3], symTW[6], w1[6], w2[6];
  form_dbarp(P, np1, s, tt, dbarp);
  form_w(P, np1, s, tt, 1, wp);
  symsw(s, wp, symTW);
  for (int k = 0; k < 6; ++k) w1[k] = np1->D[k] - dbarp[k];
  mv6(w2, P->o->C, w1);
  for (int k = 0; k < 6; ++k) R1[k] = s[k] - n->stress[k] - w2[k] + two * symTW[k];
}

/* mm10_h_voche (mm10_b.f:1888-1916) */
static double h_voche(const props_t* P, const cstate* np1, const cstate* n, const double s[6],
                      double tt) {
  const orc_crystal* c = P->cp;
  double h = zero;
  for (int i = 0; i < NS; ++i) {
    double si = slipinc_of(P, np1, s, tt, i);
    double h_term = one - (tt - c->tau_y) / c->tau_v + zero / (tt - c->tau_y);
    h = h + pow(fabs(h_term), c->voche_m) * fsign(h_term) * fabs(si);
  }
  return n->tau_tilde + c->theta_0 * h;
}

/* mm10_formR2 (mm10_b.f:63-113); also sets np1%tt_rate */
static double formR2(const props_t* P, cstate* np1, const cstate* n, const double s[6],
                     double tt) {
  double h = h_voche(P, np1, n, s, tt);
  np1->tt_rate = (h - n->tau_tilde) / np1->tinc;
  return tt - h;
}

/* mm10_formR (mm10_b.f:1029-1063) */
static void formR(const props_t* P, cstate* np1, const cstate* n, const double x[7],
                  double R[7]) {
  formR1(P, np1, n, x, x[6], R);
  R[6] = formR2(P, np1, n, x, x[6]);
}

/* mm10_symswmat column i */
static void symtq_of(const double s[6], const double w[3], double sw[6]) { symsw(s, w, sw); }

/* mm10_dgdt_voche (mm10_b.f:2005-2031) */
static void dgdt_voche(const props_t* P, const cstate* np1, const double s[6], double tt,
                       double dgdt[NSMAX]) {
  const orc_crystal* c = P->cp;
  for (int i = 0; i < NS; ++i) {
    double rs = rs_of(np1, s, i);
    double v = pow(fabs(rs), c->rate_n - one);
    v = np1->dg * c->rate_n / pow(tt, c->rate_n) * v;
    dgdt[i] = v + np1->tinc * c->iD_v;
  }
}

/* mm10_formJ11 (mm10_b.f:177-269) */
static void formJ11(const props_t* P, const cstate* np1, const double s[6], double tt,
                    double J11[6][6]) {
  double dgdt[NSMAX];
  memset(J11, 0, 36 * sizeof(double));
  dgdt_voche(P, np1, s, tt, dgdt);
  for (int i = 0; i < NS; ++i) {
    double symtq[6], wv[6];
    symtq_of(s, np1->qc[i], symtq);
    mv6_plus(wv, P->o->C, np1->ms[i], symtq, two);
    /* DGER: J11 += dgdt(i) * wv * ms_i^T */
    for (int b = 0; b < 6; ++b) {
      double t = dgdt[i] * np1->ms[i][b];
      for (int a = 0; a < 6; ++a) J11[a][b] = J11[a][b] + wv[a] * t;
    }
  }
  double wp[3], Iw[6][6];
  form_w(P, np1, s, tt, 1, wp);
  iw_mat(wp, Iw);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) J11[a][b] = J11[a][b] + Iw[a][b];
  for (int a = 0; a < 6; ++a) J11[a][a] = J11[a][a] + one;
}

/* mm10_formJ (mm10_b.f:901-967): 7x7 Jacobian, J12 (mm10_b.f:271-367),
 * J21 = -estress (mm10_b.f:1918-1949), J22 = ehard (mm10_b.f:1951-1983) */
static void formJ(const props_t* P, const cstate* np1, const double x[7], double J[7][7]) {
  const orc_crystal* c = P->cp;
  const double* s = x;
  double tt = x[6];
  double J11[6][6];
  formJ11(P, np1, s, tt, J11);
  /* J12 */
  double J12[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NS; ++i) {
    double symtq[6], tm[6];
    symtq_of(s, np1->qc[i], symtq);
    mv6_plus(tm, P->o->C, np1->ms[i], symtq, two);
    double dgam = slipinc_of(P, np1, s, tt, i);
    double dgdtt = -c->rate_n / tt * dgam;
    double t = one * dgdtt;
    for (int a = 0; a < 6; ++a) J12[a] = J12[a] + tm[a] * t;
  }
  /* J21 = -estr */
  double et[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NS; ++i) {
    double rs = rs_of(np1, s, i);
    double h_term = one - (tt - c->tau_y) / c->tau_v + zero / (tt - c->tau_y);
    double f = pow(fabs(h_term), c->voche_m) * fsign(h_term) * pow(fabs(rs), c->rate_n - two) * rs;
    for (int a = 0; a < 6; ++a) et[a] = et[a] + f * np1->ms[i][a];
  }
  double fac = c->theta_0 * np1->dg * c->rate_n / pow(tt, c->rate_n);
  for (int a = 0; a < 6; ++a) et[a] = fac * et[a];
  /* J22 = ehard */
  double etau = zero;
  for (int i = 0; i < NS; ++i) {
    double si = slipinc_of(P, np1, s, tt, i);
    double h_term = one - (tt - c->tau_y) / c->tau_v + zero / (tt - c->tau_y);
    double dd = (tt - c->tau_y) * (tt - c->tau_y);
    double sgs = c->jac_exact ? one : fsign(si);
    etau = etau + (c->voche_m * (-one / c->tau_v - zero / dd) * fabs(si) / fabs(h_term) -
                   fabs(si) * c->rate_n / tt * fsign(h_term) * sgs) *
                      pow(fabs(h_term), c->voche_m);
  }
  etau = c->theta_0 * etau;
  etau = one - etau;
  for (int a = 0; a < 6; ++a) {
    for (int b = 0; b < 6; ++b) J[a][b] = J11[a][b];
    J[a][6] = J12[a];
    J[6][a] = -et[a];
  }
  J[6][6] = etau;
}

/* mm10_solve (mm10_a.f:2860-3295): predictor + full Newton. Returns fail. */
static int mm10_solve(const props_t* P, cstate* np1, const cstate* n, double stress[6],
                      double* tt, double dtinc, double Jmat[7][7], int stats[5]) {
  const orc_crystal* c = P->cp;
  const double cls = 1.0e-4, red = 0.5;
  const int mls = 10, mmin = 1;
  int fail = 0;
  double x[7];
  for (int k = 0; k < 6; ++k) x[k] = stress[k];
  x[6] = *tt;
  double inR1;
  /* ---- mm10_solve_predict (mm10_a.f:2975-3156) ---- */
  {
    int iter = 0;
    double x1[6], x2, d1[6], d2[6], t1, t2, dtrace, R1[6], nR1;
    for (int k = 0; k < 6; ++k) x1[k] = x[k];
    double dt = dtinc;
    for (int k = 0; k < 6; ++k) d1[k] = np1->D[k];
    dtrace = (d1[0] + d1[1] + d1[2]) / three;
    d1[0] = d1[0] - dtrace; d1[1] = d1[1] - dtrace; d1[2] = d1[2] - dtrace;
    t1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    t2 = d1[3] * d1[3] + d1[4] * d1[4] + d1[5] * d1[5];
    if (t1 + t2 == zero) for (int k = 0; k < 6; ++k) d1[k] = zero;
    else { double nn = sqrt(t1 + t2); for (int k = 0; k < 6; ++k) d1[k] = d1[k] / nn; }
    for (int k = 0; k < 6; ++k) d2[k] = n->D[k];
    dtrace = (d2[0] + d2[1] + d2[2]) / three;
    d2[0] = d2[0] - dtrace; d2[1] = d2[1] - dtrace; d2[2] = d2[2] - dtrace;
    t1 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    t2 = d2[3] * d2[3] + d2[4] * d2[4] + d2[5] * d2[5];
    if (t1 + t2 == zero) for (int k = 0; k < 6; ++k) d2[k] = zero;
    else { double nn = sqrt(t1 + t2); for (int k = 0; k < 6; ++k) d2[k] = d2[k] / nn; }
    t1 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    t2 = d1[3] * d2[3] + d1[4] * d2[4] + d1[5] * d2[5];
    double cos_ang = fmax(t1 + t2, zero);
    x2 = x[6] + cos_ang * n->tt_rate * dt;
    formR1(P, np1, n, x1, x2, R1);
    nR1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] + R1[3] * R1[3] + R1[4] * R1[4] +
               R1[5] * R1[5]);
    inR1 = nR1;
    int pfail = 0;
    while ((nR1 > c->atol1) && (nR1 / inR1 > c->rtol1)) {
      if (getenv("ORC_DEBUG")) fprintf(stderr, "  pred it %d nR1 %g x %g %g %g\n", iter, nR1, x1[0], x1[1], x1[2]);
      double J11[6][6], A[36], dx1[6];
      formJ11(P, np1, x1, x2, J11);
      for (int a = 0; a < 6; ++a) {
        dx1[a] = R1[a];
        for (int b = 0; b < 6; ++b) A[a * 6 + b] = -J11[a][b];
      }
      lu_solve(6, A, 1, dx1);
      double alpha = one;
      double dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] + R1[3] * R1[3] +
                     R1[4] * R1[4] + R1[5] * R1[5];
      double ls1 = half * dotR1;
      double wv[6];
      for (int b = 0; b < 6; ++b)
        wv[b] = J11[0][b] * R1[0] + J11[1][b] * R1[1] + J11[2][b] * R1[2] + J11[3][b] * R1[3] +
                J11[4][b] * R1[4] + J11[5][b] * R1[5];
      double ls2 = cls * (dx1[0] * wv[0] + dx1[1] * wv[1] + dx1[2] * wv[2] + dx1[3] * wv[3] +
                          dx1[4] * wv[4] + dx1[5] * wv[5]);
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[6];
        for (int k = 0; k < 6; ++k) xn[k] = x1[k] + alpha * dx1[k];
        formR1(P, np1, n, xn, x2, R1);
        stats[4]++;
        dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] + R1[3] * R1[3] + R1[4] * R1[4] +
                R1[5] * R1[5];
        nR1 = sqrt(dotR1);
        double nRs = half * dotR1;
        if ((nRs <= nlsx) || (ls > mls)) {
          for (int k = 0; k < 6; ++k) x1[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      int nan = 0;
      for (int k = 0; k < 6; ++k) nan |= isnan(x1[k]);
      if ((iter > c->miter) || nan) { pfail = 1; break; }
    }
    stats[0] += iter;
    if (pfail) {
      fail = 1; /* predictor failure leaves x unchanged; the update still runs */
    } else {
      for (int k = 0; k < 6; ++k) x[k] = x1[k];
      x[6] = x2;
    }
  }
  /* ---- mm10_solve_update (mm10_a.f:3162-3293) ---- */
  {
    int iter = 0;
    double R[7], nR, inR;
    formR(P, np1, n, x, R);
    double dot = zero;
    for (int k = 0; k < 7; ++k) dot = dot + R[k] * R[k];
    nR = sqrt(dot);
    inR = nR;
    if (inR == zero) inR = inR1;
    while (((nR > c->atol) && (nR / inR > c->rtol)) || (iter < mmin)) {
      if (getenv("ORC_DEBUG")) fprintf(stderr, "  upd it %d nR %g x %g %g %g %g\n", iter, nR, x[0], x[1], x[2], x[6]);
      double J[7][7], A[49], dx[7];
      formJ(P, np1, x, J);
      memcpy(Jmat, J, sizeof J);
      for (int a = 0; a < 7; ++a) {
        dx[a] = R[a];
        for (int b = 0; b < 7; ++b) A[a * 7 + b] = -J[a][b];
      }
      lu_solve(7, A, 1, dx);
      double alpha = one;
      double dotR = zero;
      for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
      double ls1 = half * dotR;
      double wv[7];
      for (int b = 0; b < 7; ++b) {
        double sacc = zero;
        for (int a = 0; a < 7; ++a) sacc = sacc + J[a][b] * R[a];
        wv[b] = sacc;
      }
      double ls2 = zero;
      for (int k = 0; k < 7; ++k) ls2 = ls2 + dx[k] * wv[k];
      ls2 = cls * ls2;
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[7];
        for (int k = 0; k < 7; ++k) xn[k] = x[k] + alpha * dx[k];
        formR(P, np1, n, xn, R);
        stats[4]++;
        dotR = zero;
        for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
        nR = sqrt(dotR);
        double nRs = half * dotR;
        if ((nRs <= nlsx) || (ls > mls)) {
          for (int k = 0; k < 7; ++k) x[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      int nan = 0;
      for (int k = 0; k < 7; ++k) nan |= isnan(x[k]);
      if ((iter > c->miter) || nan) { fail = 1; break; }
    }
    stats[1] += iter;
  }
  for (int k = 0; k < 6; ++k) stress[k] = x[k];
  *tt = x[6];
  return fail;
}

/* test hook: residual and Jacobian of the local problem at x for the
 * np1/n states built from (Fn, Fn1, sn) exactly as orc_voxel_update does */
static void setup_orient(const orc_crystal* cp, const double angles[3], orient_t* o);
static void rtcmp1(const double f[3][3], double r[3][3]);
static void inv33(const double a[3][3], double g[3][3], double* dj);
static void mul33(const double A[3][3], const double B[3][3], double C[6]);
static void getrm1_opt1(const double r[3][3], double q[6][6]);
static void qmply1(const double q[6][6], const double m1[6], double m2[6]);
int orc_local_RJ(const orc_crystal* cp, const double angles[3], double dt, const double Fnv[9],
                 const double Fn1v[9], const double* sn, const double x[7], double R[7],
                 double J[49]) {
  orient_t o;
  setup_orient(cp, angles, &o);
  props_t P = {cp, &o};
  double fn[3][3], fn1[3][3], fnh[3][3], dfn[3][3];
  for (int i = 0; i < 9; ++i) { fn[i / 3][i % 3] = Fnv[i]; fn1[i / 3][i % 3] = Fn1v[i]; }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { fnh[i][j] = half * (fn[i][j] + fn1[i][j]); dfn[i][j] = fn1[i][j] - fn[i][j]; }
  double rnh[3][3], rn1[3][3], fnhinv[3][3], detFh, ddt[6], qnh[6][6], uddt[6];
  rtcmp1(fnh, rnh); rtcmp1(fn1, rn1); inv33(fnh, fnhinv, &detFh); mul33(dfn, fnhinv, ddt);
  getrm1_opt1(rnh, qnh); qmply1(qnh, ddt, uddt);
  cstate n, np1;
  memset(&n, 0, sizeof n);
  for (int k = 0; k < 6; ++k) n.stress[k] = sn[ORC_W_SIG + k];
  for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) n.Rp[i][j] = sn[ORC_W_RP + 3 * j + i];
  n.tau_tilde = sn[ORC_W_TT]; n.tt_rate = sn[ORC_W_TTRATE];
  for (int k = 0; k < 6; ++k) n.D[k] = sn[ORC_W_D + k];
  setup_np1(rn1, uddt, dt, &np1);
  mm10_setup(&P, &np1, &n);
  double JJ[7][7];
  formR(&P, &np1, &n, x, R);
  formJ(&P, &np1, x, JJ);
  for (int a = 0; a < 49; ++a) J[a] = JJ[a / 7][a % 7];
  return 0;
}

/* mm10_tangent (mm10_a.f:658-815), Voce: ed = 0, dgammadd = 0 */
static void mm10_tangent(const props_t* P, cstate* np1, double J[7][7]) {
  double JJ[6][6], J12[6], beta[6], alpha = J[6][6];
  for (int a = 0; a < 6; ++a) {
    for (int b = 0; b < 6; ++b) JJ[a][b] = J[a][b];
    J12[a] = J[a][6];
    beta[a] = J[6][a];
  }
  /* DGESV of the 1x1 system J22 * beta = [J21 | ed] */
  for (int b = 0; b < 6; ++b) beta[b] = beta[b] / alpha;
  /* JJ = JJ - J12 * beta (DGEMM alpha=-1, beta=1) */
  for (int b = 0; b < 6; ++b) {
    double t = -one * beta[b];
    for (int a = 0; a < 6; ++a) JJ[a][b] = JJ[a][b] + t * J12[a];
  }
  /* JR = C - JA - JB with JA = JB = 0 exactly */
  double A[36], B[36];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      A[a * 6 + b] = JJ[a][b];
      B[a * 6 + b] = P->o->C[a][b] - zero - zero;
    }
  lu_solve(6, A, 6, B);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) np1->tangent[a][b] = B[a * 6 + b];
}

/* mm10_expmw3x3 (mm10_a.f:3382-3415) */
static void expmw3x3(const double w[3][3], double a[3][3]) {
  double alpha = sqrt(w[1][2] * w[1][2] + w[0][2] * w[0][2] + w[0][1] * w[0][1]);
  if (alpha < 1.0e-16) {
    memset(a, 0, 9 * sizeof(double));
  } else {
    double ca = (one - cos(alpha)) / (alpha * alpha);
    double sa = sin(alpha) / alpha;
    for (int j = 0; j < 3; ++j) {
      for (int i = 0; i < 3; ++i) a[i][j] = sa * w[i][j];
      for (int l = 0; l < 3; ++l) {
        double t = ca * w[l][j];
        for (int i = 0; i < 3; ++i) a[i][j] = a[i][j] + t * w[i][l];
      }
    }
  }
  a[0][0] = a[0][0] + one;
  a[1][1] = a[1][1] + one;
  a[2][2] = a[2][2] + one;
}

/* mm10_solve_strup (mm10_a.f:2628-2845) + remainder of mm10_solve_crystal
 * (mm10_a.f:1080-1157).  Returns fail flag (after 4 cuts). */
static int mm10_solve_crystal(const props_t* P, cstate* np1, const cstate* n, int iter0,
                              int stats[5]) {
  const orc_crystal* c = P->cp;
  double stress[6], ostress[6], tt = n->tau_tilde, ott = tt;
  for (int k = 0; k < 6; ++k) stress[k] = ostress[k] = n->stress[k];
  mm10_setup(P, np1, n);
  double temp1 = np1->D[0] * np1->D[0] + np1->D[1] * np1->D[1] + np1->D[2] * np1->D[2] +
                 np1->D[3] * np1->D[3] + np1->D[4] * np1->D[4] + np1->D[5] * np1->D[5];
  double temp2 = stress[0] * stress[0] + stress[1] * stress[1] + stress[2] * stress[2] +
                 stress[3] * stress[3] + stress[4] * stress[4] + stress[5] * stress[5];
  int no_load = (temp2 == zero) && (temp1 == zero);
  double curr_tt_rate = zero;
  double Jmat[7][7];
  int fail = 0;
  memset(Jmat, 0, sizeof Jmat);
  if (iter0 || no_load) {
    /* mm10_solve_strup_a (mm10_a.f:2735-2753) */
    memcpy(np1->tangent, P->o->C, 36 * sizeof(double));
    if (!no_load) {
      double R1[6];
      formR1(P, np1, n, stress, tt, R1);
      for (int k = 0; k < 6; ++k) stress[k] = stress[k] - R1[k];
    }
    curr_tt_rate = zero;
  } else {
    /* mm10_solve_strup_iterate (mm10_a.f:2759-2843) */
    const double mult = 0.5;
    const int mcuts = 4;
    double frac = zero, step = one;
    int cuts = 0;
    cstate curr;
    while (frac < one) {
      double Dw[6];
      for (int k = 0; k < 6; ++k) Dw[k] = np1->D[k] * (step + frac);
      setup_np1(np1->R, Dw, np1->tinc * (step + frac), &curr);
      mm10_setup(P, &curr, n);
      tt = n->tau_tilde;
      fail = mm10_solve(P, &curr, n, stress, &tt, np1->tinc * step, Jmat, stats);
      if (getenv("ORC_DEBUG"))
        fprintf(stderr, "substep frac=%g step=%g fail=%d s=%g %g %g tt=%g\n", frac, step, fail,
                stress[0], stress[1], stress[2], tt);
      if (fail) {
        for (int k = 0; k < 6; ++k) stress[k] = ostress[k];
        tt = ott;
        step = step * mult;
        cuts = cuts + 1;
        stats[2]++;
        if (cuts > mcuts) break;
        fail = 0;
      } else {
        for (int k = 0; k < 6; ++k) ostress[k] = stress[k];
        ott = tt;
        frac = frac + step;
      }
    }
    int bad = fail || isnan(tt);
    for (int k = 0; k < 6; ++k) bad |= isnan(stress[k]);
    if (bad) {
      stats[3] = 1;
      return 1;
    }
    curr_tt_rate = curr.tt_rate;
  }
  for (int k = 0; k < 6; ++k) np1->stress[k] = stress[k];
  np1->tau_tilde = tt;
  np1->tt_rate = curr_tt_rate;
  if (iter0 || no_load) return 0;

  /* consistent tangent, symmetrised (mm10_a.f:1134-1142) */
  mm10_tangent(P, np1, Jmat);
  {
    double w[6][6];
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < 6; ++i) w[i][j] = (np1->tangent[i][j] + np1->tangent[j][i]) * half;
    memcpy(np1->tangent, w, sizeof w);
  }
  /* mm10_update_rotation (mm10_a.f:3310-3343) */
  {
    double wbarp[3], W[3][3], ex[3][3];
    form_w(P, np1, np1->stress, np1->tau_tilde, 0, wbarp);
    W[0][0] = zero; W[0][1] = wbarp[2]; W[0][2] = wbarp[1];
    W[1][0] = -wbarp[2]; W[1][1] = zero; W[1][2] = wbarp[0];
    W[2][0] = -wbarp[1]; W[2][1] = -wbarp[0]; W[2][2] = zero;
    expmw3x3(W, ex);
    mult33(np1->Rp, ex, n->Rp);
  }
  /* mm10_output (mm10_a.f:3429-3689): the quantities that enter the
   * history words kept here (slip increments, work terms) */
  {
    double dif[NSMAX], ed[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < NS; ++i)
      np1->slip_incs[i] = slipinc_of(P, np1, np1->stress, np1->tau_tilde, i);
    for (int i = 0; i < NS; ++i) dif[i] = rs_of(np1, np1->stress, i) * np1->tinc * c->iD_v;
    for (int i = 0; i < NS; ++i)
      for (int k = 0; k < 6; ++k) ed[k] = ed[k] + dif[i] * np1->ms[i][k];
    np1->work_inc = np1->stress[0] * np1->D[0] + np1->stress[1] * np1->D[1] +
                    np1->stress[2] * np1->D[2] + np1->stress[3] * np1->D[3] +
                    np1->stress[4] * np1->D[4] + np1->stress[5] * np1->D[5];
    double eun[6], erot[6][6], ee[6], dbarp[6] = {0, 0, 0, 0, 0, 0}, wp[3], ewwe[6], ep[6];
    for (int k = 0; k < 6; ++k) eun[k] = np1->stress[k];
    chol_solve6(P->o->C, eun);
    rt2rve(np1->R, erot);
    mv6(ee, erot, eun);
    for (int i = 0; i < NS; ++i)
      for (int k = 0; k < 6; ++k) dbarp[k] = dbarp[k] + np1->slip_incs[i] * np1->ms[i][k];
    form_w(P, np1, np1->stress, np1->tau_tilde, 1, wp);
    symsw(ee, wp, ewwe);
    for (int k = 0; k < 6; ++k) ep[k] = dbarp[k] + ewwe[k];
    for (int k = 0; k < 6; ++k) ep[k] = ep[k] + ed[k];
    double t1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
    double t2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
    np1->p_strain_inc = sqrt(two / three * (t1 + half * t2));
    np1->p_work_inc = np1->stress[0] * ep[0] + np1->stress[1] * ep[1] + np1->stress[2] * ep[2] +
                      np1->stress[3] * ep[3] + np1->stress[4] * ep[4] + np1->stress[5] * ep[5];
    for (int i = 0; i < NS; ++i) np1->slip_incs[i] = np1->slip_incs[i] + dif[i];
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* cep2A (cep2A.f:14-284): exact pull-back dt/dd -> dP/dF                     */
/* ------------------------------------------------------------------------ */

static void m33mul(const double A[3][3], const double B[3][3], double C[3][3]) {
  /* cep2A_multiply33 (cep2A.f:300-317) */
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}
static void m33t(const double A[3][3], double B[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[j][i] = A[i][j];
}
static double det33c(const double A[3][3]) {
  /* cep2A_det33 (cep2A.f:348-362) */
  double d = zero;
  d = A[0][0] * A[1][1] * A[2][2];
  d = d + A[0][1] * A[1][2] * A[2][0];
  d = d + A[0][2] * A[1][0] * A[2][1];
  d = d - A[0][0] * A[1][2] * A[2][1];
  d = d - A[0][1] * A[1][0] * A[2][2];
  d = d - A[0][2] * A[1][1] * A[2][0];
  return d;
}

#define IDX4(i, j, k, l) ((i) * 27 + (j) * 9 + (k) * 3 + (l))

static void cep2A_a(const double Fn[3][3], const double t[3][3], const double C[6][6],
                    const double Rh[3][3], const double fnhinv[3][3], const double R[3][3],
                    const double Fn1[3][3], const double finv[3][3], double detFn1,
                    double dPdF[81]) {
  double transR[3][3], U[3][3], Y[3][3], temp[3][3], sigma[3][3], Uinv[3][3], t_Uinv[3][3];
  double R_t[3][3], RY[3][3], RYR[3][3], RtRF[3][3], dJdF[3][3], dFn[3][3], Fnh[3][3];
  double transRh[3][3], Uh[3][3], Yh[3][3], RYRh[3][3], RYh[3][3], FpFinv[3][3], L33[3][3],
      D2[3][3];
  m33t(R, transR);
  m33mul(transR, Fn1, U);
  double traceU = U[0][0] + U[1][1] + U[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Y[i][j] = ((i == j) ? traceU : zero) - U[i][j];
  m33mul(R, t, temp);
  m33mul(temp, transR, sigma);
  m33mul(finv, R, Uinv);
  m33mul(t, Uinv, t_Uinv);
  m33mul(R, t, R_t);
  m33mul(R, Y, RY);
  m33mul(RY, transR, RYR);
  double detY_1 = one / det33c(Y);
  m33mul(R_t, Uinv, RtRF);
  m33t(finv, dJdF);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dJdF[i][j] = dJdF[i][j] * detFn1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      dFn[i][j] = Fn1[i][j] - Fn[i][j];
      Fnh[i][j] = half * (Fn1[i][j] + Fn[i][j]);
    }
  m33t(Rh, transRh);
  m33mul(transRh, Fnh, Uh);
  double traceUh = Uh[0][0] + Uh[1][1] + Uh[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Yh[i][j] = ((i == j) ? traceUh : zero) - Uh[i][j];
  m33mul(Rh, Yh, temp);
  m33mul(temp, transRh, RYRh);
  double detYh_1 = one / det33c(Yh);
  m33mul(Rh, Yh, RYh);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FpFinv[i][j] = half * fnhinv[i][j];
  m33mul(dFn, FpFinv, L33);
  m33t(L33, temp);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D2[i][j] = L33[i][j] + temp[i][j];

  /* cep99 = C(index_voigt, index_voigt), index_voigt = [1,4,6,4,2,5,6,5,3] */
  static const int iv[9] = {0, 3, 5, 3, 1, 4, 5, 4, 2};
  double cep99[9][9];
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b) cep99[a][b] = C[iv[a]][iv[b]];

  double dRdF[81], dRhdF[81], dLdF[81];
  for (int m = 0; m < 81; ++m) {
    int i = m / 27, j = (m / 9) % 3, k = (m / 3) % 3, l = m % 3;
    dRdF[m] = detY_1 * (RYR[i][k] * Y[l][j] - RY[i][l] * RY[k][j]);
    dRhdF[m] = detYh_1 * (RYRh[i][k] * Yh[l][j] - RYh[i][l] * RYh[k][j]);
  }
  for (int m = 0; m < 81; ++m) dRhdF[m] = dRhdF[m] * half;
  for (int m = 0; m < 81; ++m) {
    int i = m / 27, j = (m / 9) % 3, k = (m / 3) % 3, l = m % 3;
    dLdF[m] = -L33[i][k] * FpFinv[l][j];
    if (i == k) dLdF[m] = dLdF[m] + FpFinv[l][j];
  }
  double d1[81], d2[81], d3[81], dtdF[81], dt1[81];
  for (int a = 0; a < 81; ++a) {
    int i = a / 27, j = (a / 9) % 3, p = (a / 3) % 3, q = a % 3;
    double s1 = zero, s2 = zero, s3 = zero;
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) {
        s1 = s1 + dRhdF[IDX4(m, i, p, q)] * D2[m][n] * Rh[n][j];
        s2 = s2 + Rh[m][i] * dLdF[IDX4(m, n, p, q)] * Rh[n][j];
        s2 = s2 + Rh[m][i] * dLdF[IDX4(n, m, p, q)] * Rh[n][j];
        s3 = s3 + Rh[m][i] * D2[m][n] * dRhdF[IDX4(n, j, p, q)];
      }
    d1[a] = s1; d2[a] = s2; d3[a] = s3;
  }
  for (int a = 0; a < 81; ++a) dtdF[a] = d1[a] + d2[a] + d3[a];
  memcpy(dt1, dtdF, sizeof dt1);
  /* cep2A_ddot44 (cep2A.f:331-346): C4_ijmn = sum_kl A4_ijkl B4_lkmn with
   * A4 = cep99 read in Fortran column-major order, i.e. A4(9(ii-1)+t) =
   * cep99(t,ii) */
  static const int tmp2[9] = {0, 27, 54, 9, 36, 63, 18, 45, 72};
  for (int ii = 0; ii < 9; ++ii)
    for (int jj = 0; jj < 9; ++jj) {
      double s = zero;
      for (int tt = 0; tt < 9; ++tt) s = s + cep99[tt][ii] * dt1[tmp2[tt] + jj];
      dtdF[ii * 9 + jj] = s;
    }
  double p0[81], p1[81], p2[81], p3[81], p4[81];
  for (int p = 0; p < 81; ++p) {
    int i = p / 27, j = (p / 9) % 3, k = (p / 3) % 3, l = p % 3;
    p0[p] = RtRF[i][j] * dJdF[k][l];
    double s1 = zero, s2 = zero, s3 = zero, s4 = zero;
    for (int m = 0; m < 3; ++m) {
      s1 = s1 - detFn1 * sigma[i][m] * finv[j][k] * finv[l][m];
      s2 = s2 + detFn1 * dRdF[IDX4(i, m, k, l)] * t_Uinv[m][j];
      for (int n = 0; n < 3; ++n) {
        s3 = s3 + detFn1 * R[i][m] * dtdF[IDX4(m, n, k, l)] * Uinv[n][j];
        s4 = s4 + detFn1 * R_t[i][m] * dRdF[IDX4(n, m, k, l)] * finv[j][n];
      }
    }
    p1[p] = s1; p2[p] = s2; p3[p] = s3; p4[p] = s4;
  }
  for (int p = 0; p < 81; ++p) dPdF[p] = p0[p] + p1[p] + p2[p] + p3[p] + p4[p];
}

void orc_cep2A(const double Fn[9], const double t6[6], const double C36[36], const double Rh[9],
               double detFh, const double fnhinv[9], const double R[9], const double Fn1[9],
               const double finv[9], double detF, double dPdF[81]) {
  (void)detFh;
  double a[3][3], t[3][3], C[6][6], rh[3][3], fh[3][3], r[3][3], f1[3][3], fi[3][3];
  for (int i = 0; i < 9; ++i) {
    a[i / 3][i % 3] = Fn[i];
    rh[i / 3][i % 3] = Rh[i];
    fh[i / 3][i % 3] = fnhinv[i];
    r[i / 3][i % 3] = R[i];
    f1[i / 3][i % 3] = Fn1[i];
    fi[i / 3][i % 3] = finv[i];
  }
  for (int i = 0; i < 36; ++i) C[i / 6][i % 6] = C36[i];
  t[0][0] = t6[0]; t[1][0] = t6[3]; t[2][0] = t6[5];
  t[0][1] = t6[3]; t[1][1] = t6[1]; t[2][1] = t6[4];
  t[0][2] = t6[5]; t[1][2] = t6[4]; t[2][2] = t6[2];
  cep2A_a(a, t, C, rh, fh, r, f1, fi, detF, dPdF);
}

/* ------------------------------------------------------------------------ */
/* per-voxel driver: do_nleps_block (drive_eps_sig.f:57-339) for one voxel   */
/* ------------------------------------------------------------------------ */

void orc_state_init(const orc_crystal* cp, long V, double* s) {
  int W = ORC_NWORDS(cp->nslip);
  for (int w = 0; w < W; ++w)
    for (long e = 0; e < V; ++e) s[(long)w * V + e] = 0.0;
  for (long e = 0; e < V; ++e) {
    s[(long)(ORC_W_RP + 0) * V + e] = 1.0;
    s[(long)(ORC_W_RP + 4) * V + e] = 1.0;
    s[(long)(ORC_W_RP + 8) * V + e] = 1.0;
    s[(long)ORC_W_TT * V + e] = cp->tau_y + 1.0e-5; /* mm10_init_voche (mm10_a.f:2042) */
  }
}

int orc_voxel_update(const orc_crystal* cp, const double angles[3], int step, int iter,
                     double dt, const double Fnv[9], const double Fn1v[9], const double* sn,
                     double* snp1, double Pout[9], double K4[81], int stats[5]) {
  const int ns = cp->nslip;
  orient_t o;
  setup_orient(cp, angles, &o);
  props_t P = {cp, &o};

  double fn[3][3], fn1[3][3], fnh[3][3], dfn[3][3];
  for (int i = 0; i < 9; ++i) { fn[i / 3][i % 3] = Fnv[i]; fn1[i / 3][i % 3] = Fn1v[i]; }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      fnh[i][j] = half * (fn[i][j] + fn1[i][j]);
      dfn[i][j] = fn1[i][j] - fn[i][j];
    }
  double rnh[3][3], rn1[3][3], fnhinv[3][3], detFh, ddt[6], qnh[6][6], uddt[6];
  rtcmp1(fnh, rnh);
  rtcmp1(fn1, rn1);
  inv33(fnh, fnhinv, &detFh);
  mul33(dfn, fnhinv, ddt);
  getrm1_opt1(rnh, qnh);
  qmply1(qnh, ddt, uddt);

  /* state n (mm10_copy_cc_hist, mm10_a.f:2464-2561); at step 1 the history
   * is re-initialised on every call (mm10_a.f:89-95, 222-229) */
  cstate n;
  memset(&n, 0, sizeof n);
  double slipsum_n[NSMAX], work_n[3];
  if (step == 1) {
    for (int k = 0; k < 6; ++k) n.stress[k] = zero;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) n.Rp[i][j] = (i == j) ? one : zero;
    n.tau_tilde = cp->tau_y + 1.0e-5;
    n.tt_rate = sn[ORC_W_TTRATE];
    for (int k = 0; k < 6; ++k) n.D[k] = zero;
    for (int i = 0; i < ns; ++i) slipsum_n[i] = zero;
    work_n[0] = work_n[1] = work_n[2] = zero;
  } else {
    for (int k = 0; k < 6; ++k) n.stress[k] = sn[ORC_W_SIG + k];
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) n.Rp[i][j] = sn[ORC_W_RP + 3 * j + i];
    n.tau_tilde = sn[ORC_W_TT];
    n.tt_rate = sn[ORC_W_TTRATE];
    for (int k = 0; k < 6; ++k) n.D[k] = sn[ORC_W_D + k];
    for (int i = 0; i < ns; ++i) slipsum_n[i] = sn[ORC_W_SLIPSUM(ns) + i];
    for (int k = 0; k < 3; ++k) work_n[k] = sn[ORC_W_WORK(ns) + k];
  }

  cstate np1;
  setup_np1(rn1, uddt, dt, &np1);
  int fail = mm10_solve_crystal(&P, &np1, &n, iter == 0, stats);
  if (fail) {
    /* The reference abandons the rest of the element block here
     * (mm10_a.f:113-116, 255-258) and leaves its history unset.  This
     * restatement reverts the voxel to state n with the elastic tangent. */
    memcpy(np1.stress, n.stress, sizeof np1.stress);
    np1.tau_tilde = n.tau_tilde;
    np1.tt_rate = zero;
    memcpy(np1.Rp, n.Rp, sizeof np1.Rp);
    memcpy(np1.tangent, o.C, sizeof np1.tangent);
    memset(np1.slip_incs, 0, sizeof np1.slip_incs);
    np1.work_inc = np1.p_work_inc = np1.p_strain_inc = zero;
  }

  /* mm10_store_cryhist + mm10_a_store_crystal (mm10_a.f:285-318, 976-1055) */
  for (int k = 0; k < 6; ++k) snp1[ORC_W_SIG + k] = np1.stress[k];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) snp1[ORC_W_RP + 3 * j + i] = np1.Rp[i][j];
  snp1[ORC_W_TT] = np1.tau_tilde;
  snp1[ORC_W_TTRATE] = np1.tt_rate;
  for (int k = 0; k < 6; ++k) snp1[ORC_W_D + k] = np1.D[k];
  for (int i = 0; i < ns; ++i) snp1[ORC_W_SLIP + i] = np1.slip_incs[i];
  for (int i = 0; i < ns; ++i) snp1[ORC_W_SLIPSUM(ns) + i] = slipsum_n[i] + np1.slip_incs[i] / 1.0;
  snp1[ORC_W_WORK(ns) + 0] = work_n[0] + np1.work_inc / 1.0;
  snp1[ORC_W_WORK(ns) + 1] = work_n[1] + np1.p_work_inc / 1.0;
  snp1[ORC_W_WORK(ns) + 2] = work_n[2] + np1.p_strain_inc / 1.0;

  /* Cauchy stress and P (drive_eps_sig.f:262-273) */
  double qtn1[6][6], cs[6], fn1inv[3][3], detF;
  getrm1_opt2(rn1, qtn1);
  qmply1(qtn1, np1.stress, cs);
  inv33(fn1, fn1inv, &detF);
  cs2p(cs, fn1inv, detF, Pout);

  /* K4 = dP/dF (drive_eps_sig.f:285-305) */
  double t[3][3];
  t[0][0] = np1.stress[0]; t[1][0] = np1.stress[3]; t[2][0] = np1.stress[5];
  t[0][1] = np1.stress[3]; t[1][1] = np1.stress[1]; t[2][1] = np1.stress[4];
  t[0][2] = np1.stress[5]; t[1][2] = np1.stress[4]; t[2][2] = np1.stress[2];
  cep2A_a(fn, t, np1.tangent, rnh, fnhinv, rn1, fn1, fn1inv, detF, K4);
  return fail;
}

int orc_drive_eps_sig(const orc_crystal* cp, long V, const double* angles, int step, int iter,
                      double dt, const double* Fn, const double* Fn1, const double* sn,
                      double* snp1, double* Pn1, double* K4, long* stats_out) {
  const int W = ORC_NWORDS(cp->nslip);
  long s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : s0, s1, s2, s3, s4)
  for (long e = 0; e < V; ++e) {
    double fa[9], fb[9], a[3], P[9], K[81];
    double sa[ORC_NWORDS(ORC_MAX_SLIP)], sb[ORC_NWORDS(ORC_MAX_SLIP)];
    int st[5] = {0, 0, 0, 0, 0};
    for (int c = 0; c < 9; ++c) { fa[c] = Fn[c * V + e]; fb[c] = Fn1[c * V + e]; }
    for (int k = 0; k < 3; ++k) a[k] = angles[k * V + e];
    for (int w = 0; w < W; ++w) sa[w] = sn[w * V + e];
    orc_voxel_update(cp, a, step, iter, dt, fa, fb, sa, sb, P, K, st);
    for (int w = 0; w < W; ++w) snp1[w * V + e] = sb[w];
    for (int c = 0; c < 9; ++c) Pn1[c * V + e] = P[c];
    for (int c = 0; c < 81; ++c) K4[c * V + e] = K[c];
    s0 += st[0]; s1 += st[1]; s2 += st[2]; s3 += st[3]; s4 += st[4];
  }
  if (stats_out) {
    stats_out[0] = s0; stats_out[1] = s1; stats_out[2] = s2; stats_out[3] = s3; stats_out[4] = s4;
  }
  return (int)s3;
}
