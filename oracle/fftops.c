/*
 * fftops.c -- oracle restatement of CPFFT's projection operator G_K_dF and
 * the global solver around it: fftPcg, tangent_homo, NBC_update, FFT_nr3.
 *
 * TEST INFRASTRUCTURE ONLY (see cpfft_oracle.h).
 *
 * G_K_dF is restated literally: nine complex-to-complex 3-D DFTs of the
 * phase-shifted fields (G_K_dF.f:101-227, phase tables FFT_init.f:354-385),
 * contraction with the Ghat4 table built on q = idx - Nhalf (FFT_init.f:
 * 272-340) separately on the real and imaginary parts, nine backward DFTs
 * scaled by 1/N^3 and the real part of the inverse phase shift.  The DFTs
 * are textbook: iterative radix-2 for powers of two, direct summation
 * otherwise.
 */
#include "cpfft_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct orc_fft {
  int N;
  long V;
  int Nhalf;
  int mode;
  double* c1;     /* coeffs1 = cos(theta)  */
  double* c2;     /* coeffs2 = -sin(theta) */
  double* twr;    /* exp(-2 pi i k / N), k = 0..N-1 */
  double* twi;
  double* re;     /* work: 9 components, split complex */
  double* im;
  double* line_r;
  double* line_i;
};

orc_fft* orc_fft_create(int N) { return orc_fft_create2(N, 0); }

/* mode 0: the reference operator, phase shift of Nhalf bins (FFT_init.f:
 * 146-147, 370-379), which for even N pairs frequency f with Gamma(f-1).
 * mode 1: the textbook Green operator: phase shift of N/2 bins for even N
 * and the Nyquist planes (any |q_i| = N/2) removed, so that the operator is
 * an orthogonal projection.  Identical to mode 0 for odd N. */
orc_fft* orc_fft_create2(int N, int mode) {
  orc_fft* f = (orc_fft*)calloc(1, sizeof *f);
  long V = (long)N * N * N;
  f->N = N;
  f->V = V;
  f->mode = mode;
  /* FFT_init.f:146-147 */
  f->Nhalf = (N % 2 == 1) ? (N + 1) / 2 : N / 2 + 1;
  f->c1 = (double*)malloc(V * sizeof(double));
  f->c2 = (double*)malloc(V * sizeof(double));
  /* formfftshift (FFT_init.f:354-385) */
  const double pi = 4.0 * atan(1.0);
  int shift = (mode == 1 && N % 2 == 0) ? N / 2 : f->Nhalf;
  double NhN = (double)shift * 2.0 * pi / (double)N;
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double t = NhN * (double)(i + j + k);
        long e = (long)k * N * N + (long)j * N + i;
        f->c1[e] = cos(t);
        f->c2[e] = -sin(t);
      }
  f->twr = (double*)malloc(N * sizeof(double));
  f->twi = (double*)malloc(N * sizeof(double));
  for (int k = 0; k < N; ++k) {
    f->twr[k] = cos(2.0 * pi * (double)k / (double)N);
    f->twi[k] = -sin(2.0 * pi * (double)k / (double)N);
  }
  f->re = (double*)malloc(9 * V * sizeof(double));
  f->im = (double*)malloc(9 * V * sizeof(double));
  f->line_r = (double*)malloc(2 * N * sizeof(double));
  f->line_i = (double*)malloc(2 * N * sizeof(double));
  return f;
}

void orc_fft_destroy(orc_fft* f) {
  if (!f) return;
  free(f->c1); free(f->c2); free(f->twr); free(f->twi);
  free(f->re); free(f->im); free(f->line_r); free(f->line_i);
  free(f);
}

/* in-place 1-D DFT of a strided line; sign = -1 forward, +1 backward */
static void dft1(const orc_fft* f, double* xr, double* xi, long stride, int sign, double* wr,
                 double* wi) {
  int N = f->N;
  for (int n = 0; n < N; ++n) { wr[n] = xr[n * stride]; wi[n] = xi[n * stride]; }
  if ((N & (N - 1)) == 0) {
    /* bit reversal */
    for (int i = 1, j = 0; i < N; ++i) {
      int bit = N >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) {
        double t = wr[i]; wr[i] = wr[j]; wr[j] = t;
        t = wi[i]; wi[i] = wi[j]; wi[j] = t;
      }
    }
    for (int len = 2; len <= N; len <<= 1) {
      int half = len / 2, st = N / len;
      for (int i = 0; i < N; i += len)
        for (int k = 0; k < half; ++k) {
          double cr = f->twr[k * st], ci = sign < 0 ? f->twi[k * st] : -f->twi[k * st];
          double ar = wr[i + k + half], ai = wi[i + k + half];
          double tr = ar * cr - ai * ci, ti = ar * ci + ai * cr;
          wr[i + k + half] = wr[i + k] - tr;
          wi[i + k + half] = wi[i + k] - ti;
          wr[i + k] = wr[i + k] + tr;
          wi[i + k] = wi[i + k] + ti;
        }
    }
    for (int n = 0; n < N; ++n) { xr[n * stride] = wr[n]; xi[n * stride] = wi[n]; }
  } else {
    for (int k = 0; k < N; ++k) {
      double sr = 0.0, si = 0.0;
      for (int n = 0; n < N; ++n) {
        int m = (int)(((long)k * n) % N);
        double cr = f->twr[m], ci = sign < 0 ? f->twi[m] : -f->twi[m];
        sr += wr[n] * cr - wi[n] * ci;
        si += wr[n] * ci + wi[n] * cr;
      }
      xr[k * stride] = sr;
      xi[k * stride] = si;
    }
  }
}

/* 3-D DFT as three passes of 1-D transforms (lines in parallel) */
static void dft3(const orc_fft* f, double* re, double* im, int sign) {
  int N = f->N;
  long NN = (long)N * N;
#pragma omp parallel
  {
    double wr[256], wi[256];
#pragma omp for schedule(static)
    for (long a = 0; a < NN; ++a) dft1(f, re + a * N, im + a * N, 1, sign, wr, wi);
#pragma omp for schedule(static)
    for (long a = 0; a < NN; ++a) {
      long k = a / N, i = a % N;
      dft1(f, re + k * NN + i, im + k * NN + i, N, sign, wr, wi);
    }
#pragma omp for schedule(static)
    for (long a = 0; a < NN; ++a) dft1(f, re + a, im + a, NN, sign, wr, wi);
  }
}

/* ddot42n (G_K_dF.f:241-268) for one voxel with stride */
static inline double ddot_row(const double* A, const double* B, long sa, long sb) {
  double t[9];
  for (int c = 0; c < 9; ++c) t[c] = A[c * sa] * B[c * sb];
  return t[0] + (((t[1] + t[5]) + (t[3] + t[7])) + ((t[2] + t[6]) + (t[4] + t[8])));
}

void orc_ddot42n(long V, const double* A4, const double* B2, double* C2) {
  for (int r = 0; r < 9; ++r)
    for (long e = 0; e < V; ++e) C2[r * V + e] = ddot_row(A4 + (long)9 * r * V + e, B2 + e, V, V);
}

/* Ghat4 row of voxel e (formG, FFT_init.f:272-340): q = idx - Nhalf with
 * q1 on the slowest axis; entries with i != k are zero */
static void ghat_voxel(const orc_fft* f, long e, double g[81]) {
  int N = f->N;
  int ii = (int)(e / ((long)N * N)), jj = (int)((e / N) % N), kk = (int)(e % N);
  double q[3] = {(double)(ii + 1 - f->Nhalf), (double)(jj + 1 - f->Nhalf),
                 (double)(kk + 1 - f->Nhalf)};
  double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  int nyq = 0;
  if (f->mode == 1 && N % 2 == 0)
    nyq = (ii == 0) || (jj == 0) || (kk == 0); /* q_i = -N/2 */
  if (nyq) qq = 0.0;
  for (int m = 0; m < 81; ++m) {
    int i = m / 27, j = (m / 9) % 3, k = (m / 3) % 3, l = m % 3;
    double v = (i == k) ? q[j] * q[l] : 0.0;
    g[m] = (fabs(qq) <= 1e-10) ? 0.0 : v / qq;
  }
}

void orc_gkdf(orc_fft* f, const double* K4, const double* F, double* GKF, int flgK) {
  long V = f->V;
  double* re = f->re;
  double* im = f->im;
  /* real1 = K4:F or F */
  if (flgK) orc_ddot42n(V, K4, F, GKF);
  else memcpy(GKF, F, 9 * V * sizeof(double));
  /* fftfem3d: phase pre-multiply and forward transform, per component */
  for (int c = 0; c < 9; ++c) {
    double* r = re + c * V;
    double* i = im + c * V;
    const double* x = GKF + c * V;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < V; ++e) { r[e] = x[e] * f->c1[e]; i[e] = x[e] * f->c2[e]; }
    dft3(f, r, i, -1);
  }
  /* Ghat contraction on re and im */
#pragma omp parallel for schedule(static)
  for (long e = 0; e < V; ++e) {
    double g[81];
    ghat_voxel(f, e, g);
    double rr[9], ri[9];
    for (int row = 0; row < 9; ++row) {
      rr[row] = ddot_row(g + 9 * row, re + e, 1, V);
      ri[row] = ddot_row(g + 9 * row, im + e, 1, V);
    }
    for (int row = 0; row < 9; ++row) { re[row * V + e] = rr[row]; im[row * V + e] = ri[row]; }
  }
  /* ifftfem3d: backward transform (scale 1/N^3) and phase post-multiply */
  double sc = 1.0 / (double)V;
  for (int c = 0; c < 9; ++c) {
    double* r = re + c * V;
    double* i = im + c * V;
    dft3(f, r, i, +1);
    double* y = GKF + c * V;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < V; ++e) {
      double a = (r[e] * sc) * f->c1[e];
      double b = (i[e] * sc) * f->c2[e];
      y[e] = a + b;
    }
  }
}

static double nrm2(long n, const double* x) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += x[i] * x[i];
  return sqrt(s);
}
static double dot(long n, const double* x, const double* y) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* fftPcg (FFT_nr3.f:214-360): unpreconditioned CG on G(K4:x) = b, x0 = 0,
 * stop when ||r|| <= tol*||b|| or ||r|| <= tol, at most 1000 iterations */
int orc_pcg(orc_fft* f, const double* K4, const double* b, double* x, double tol, int* iters,
            double* relres) {
  return orc_pcg_limit(f, K4, b, x, tol, 1000, iters, relres);
}

/* fftPcg with an iteration limit other than the reference's 1000 (a timing
 * sample runs a fixed number of iterations) */
int orc_pcg_limit(orc_fft* f, const double* K4, const double* b, double* x, double tol,
                  int maxIter, int* iters, double* relres) {
  long n = 9 * f->V;
  for (long i = 0; i < n; ++i) x[i] = 0.0;
  *iters = 0;
  *relres = 0.0;
  double n2b = nrm2(n, b);
  double tolb = tol * n2b;
  if (tol <= 2.220446049250313e-16 || tol >= 1.0) return -1;
  if (n2b <= 2.220446049250313e-16) return 1;
  double* r = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* q = (double*)malloc(n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  memcpy(p, b, n * sizeof(double));
  double rr = dot(n, r, r), resnorm = n2b;
  int k, status = -2;
  for (k = 1; k <= maxIter; ++k) {
    orc_gkdf(f, K4, p, q, 1);
    double alpha = rr / dot(n, p, q);
    for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    resnorm = nrm2(n, r);
    if (resnorm <= tolb || resnorm <= tol) { status = 0; break; }
    double rr1 = dot(n, r, r);
    double beta = rr1 / rr;
    rr = rr1;
    for (long i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
  }
  *iters = (k > maxIter) ? maxIter : k;
  *relres = resnorm / n2b;
  free(r); free(p); free(q);
  if (*iters >= maxIter) status = -2;
  return status;
}

/* tangent_homo (tangent_homo.f:11-73).  Returns 0, or the first negative
 * status of its nine fftPcg solves (fftPcg aborts there, FFT_nr3.f:248-251,
 * 335-339); C_homo is filled either way.  *iters (if not NULL) receives the
 * CG iterations of the nine solves. */
int orc_tangent_homo_it(orc_fft* f, const double* K4, double tol, double C_homo[81],
                        int* iters) {
  int status = 0;
  if (iters) *iters = 0;
  long V = f->V, n = 9 * V;
  double* C = (double*)malloc(n * sizeof(double));
  double* b = (double*)malloc(n * sizeof(double));
  for (int i = 0; i < 9; ++i) {
    for (int j = 0; j < 9; ++j) memcpy(C + j * V, K4 + (long)(9 * j + i) * V, V * sizeof(double));
    orc_gkdf(f, K4, C, b, 0);
    for (long a = 0; a < n; ++a) b[a] = -b[a];
    int it; double rel;
    int st = orc_pcg(f, K4, b, C, tol, &it, &rel);
    if (st < 0 && status == 0) status = st;
    if (iters) *iters += it;
    if (getenv("ORC_PROGRESS")) {
      fprintf(stderr, "oracle: tangent_homo column %d (%d CG)\n", i + 1, it);
      fflush(stderr);
    }
    for (long e = 0; e < V; ++e) C[i * V + e] += 1.0;
    for (int j = 0; j < 9; ++j) {
      double s = dot(n, K4 + (long)9 * j * V, C);
      C_homo[9 * j + i] = s / (double)V;
    }
  }
  free(C); free(b);
  return status;
}

int orc_tangent_homo(orc_fft* f, const double* K4, double tol, double C_homo[81]) {
  return orc_tangent_homo_it(f, K4, tol, C_homo, NULL);
}

/* NBC_update (FFT_nr3.f:375-433) */
int orc_nbc_update(const double C_homo[81], double DbarF[9], const double P_bar[9],
                   const double PBC[9], const int isNBC[9]) {
  double A[81], bb[9];
  int piv[9];
  for (int i = 0; i < 9; ++i) {
    if (isNBC[i]) {
      for (int m = 0; m < 9; ++m) A[i * 9 + m] = C_homo[9 * i + m];
      bb[i] = PBC[i] - P_bar[i];
    } else {
      for (int m = 0; m < 9; ++m) A[i * 9 + m] = 0.0;
      A[i * 9 + i] = 1.0;
      bb[i] = DbarF[i];
    }
  }
  /* dgesv */
  for (int j = 0; j < 9; ++j) {
    int p = j;
    for (int i = j + 1; i < 9; ++i)
      if (fabs(A[i * 9 + j]) > fabs(A[p * 9 + j])) p = i;
    piv[j] = p;
    if (A[p * 9 + j] == 0.0) return j + 1;
    if (p != j) {
      for (int k = 0; k < 9; ++k) { double t = A[j * 9 + k]; A[j * 9 + k] = A[p * 9 + k]; A[p * 9 + k] = t; }
      double t = bb[j]; bb[j] = bb[p]; bb[p] = t;
    }
    for (int i = j + 1; i < 9; ++i) {
      double l = A[i * 9 + j] / A[j * 9 + j];
      for (int k = j + 1; k < 9; ++k) A[i * 9 + k] -= l * A[j * 9 + k];
      bb[i] -= l * bb[j];
    }
  }
  (void)piv;
  for (int i = 8; i >= 0; --i) {
    double s = bb[i];
    for (int k = i + 1; k < 9; ++k) s -= A[i * 9 + k] * bb[k];
    bb[i] = s / A[i * 9 + i];
  }
  for (int i = 0; i < 9; ++i) DbarF[i] = bb[i];
  return 0;
}

static void axpy(long n, double a, const double* x, double* y) {
  for (long i = 0; i < n; ++i) y[i] += a * x[i];
}

/* FFT_nr3 (FFT_nr3.f:14-200), preceded by the elastic call of `compute`
 * (FFT_finite_3d.f:145: drive_eps_sig(1,0)).  With the environment variable
 * ORC_PROGRESS set, each Newton iteration prints a line to stderr (the
 * reference's verbose residual line, FFT_nr3.f:114). */
int orc_fft_nr3(orc_problem* pb, int nstep, const double* BC_all, const int isNBC[9],
                double* Pbar_out, int* counters_out) {
  const int progress = getenv("ORC_PROGRESS") != NULL;
  const orc_crystal* cp = pb->cp;
  long V = pb->V, n = 9 * V;
  int W = ORC_NWORDS(cp->nslip);
  orc_fft* f = orc_fft_create2(pb->N, pb->gamma_mode);
  double* b = (double*)malloc(n * sizeof(double));
  double* dFm = (double*)malloc(n * sizeof(double));
  double barF_t[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, barF[9], DbarF[9], P_bar[9] = {0};
  double C_homo[81];
  int existNBC = 0, rc = 0;
  for (int i = 0; i < 9; ++i) existNBC |= isNBC[i];

  if (pb->tolPCG <= 2.220446049250313e-16 || pb->tolPCG >= 1.0) { rc = -1; goto done; }
  orc_drive_eps_sig(cp, V, pb->angles, 1, 0, pb->dt, pb->Fn, pb->Fn1, pb->sn, pb->snp1, pb->Pn1,
                    pb->K4, NULL);
  if ((rc = orc_tangent_homo(f, pb->K4, pb->tolPCG, C_homo)) < 0) goto done;

  for (int step = 1; step <= nstep; ++step) {
    int newton = 0, cgits = 0, nbc_loops = 0, calls = 0;
    double PBC[9] = {0}, FBC[9] = {0};
    for (int i = 0; i < 9; ++i) {
      DbarF[i] = 0.0;
      double v = BC_all[(step - 1) * 9 + i];
      if (isNBC[i]) PBC[i] = v;
      else { FBC[i] = v; DbarF[i] = FBC[i] - barF_t[i]; }
    }
    if (existNBC && orc_nbc_update(C_homo, DbarF, P_bar, PBC, isNBC)) { rc = -3; goto done; }
    for (int i = 0; i < 9; ++i) barF[i] = barF_t[i] + DbarF[i];
    int iiter_NBC = 0;
    for (;;) {
      for (int c = 0; c < 9; ++c)
        for (long e = 0; e < V; ++e) dFm[c * V + e] = DbarF[c];
      double Fnorm = nrm2(n, pb->Fn1);
      axpy(n, 1.0, dFm, pb->Fn1);
      orc_gkdf(f, pb->K4, dFm, b, 1);
      for (long a = 0; a < n; ++a) b[a] = -b[a];
      int it; double rel;
      if ((rc = orc_pcg(f, pb->K4, b, dFm, pb->tolPCG, &it, &rel)) < 0) goto done;
      rc = 0;
      cgits += it;
      axpy(n, 1.0, dFm, pb->Fn1);
      int iiter_EBC = 0;
      double resfft = 1.0;
      while (resfft > pb->tolNR) {
        orc_drive_eps_sig(cp, V, pb->angles, step, iiter_EBC, pb->dt, pb->Fn, pb->Fn1, pb->sn,
                          pb->snp1, pb->Pn1, pb->K4, NULL);
        calls++;
        orc_gkdf(f, pb->K4, pb->Pn1, b, 0);
        for (long a = 0; a < n; ++a) b[a] = -b[a];
        if ((rc = orc_pcg(f, pb->K4, b, dFm, pb->tolPCG, &it, &rel)) < 0) goto done;
        rc = 0;
        cgits += it;
        axpy(n, 1.0, dFm, pb->Fn1);
        resfft = nrm2(n, dFm) / Fnorm;
        if (progress) {
          fprintf(stderr, "oracle: step %d mixed-BC loop %d Newton %d |dF|/|F| = %.4e (%d CG)\n",
                  step, iiter_NBC, iiter_EBC, resfft, it);
          fflush(stderr);
        }
        if (iiter_EBC == pb->maxIter) { rc = -4; goto done; }
        iiter_EBC++;
        newton++;
      }
      orc_drive_eps_sig(cp, V, pb->angles, step, iiter_EBC, pb->dt, pb->Fn, pb->Fn1, pb->sn,
                        pb->snp1, pb->Pn1, pb->K4, NULL);
      calls++;
      double res1 = 0.0, res2 = 0.0;
      for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (long e = 0; e < V; ++e) s += pb->Pn1[i * V + e];
        P_bar[i] = s / (double)V;
        res2 += P_bar[i] * P_bar[i];
        if (!isNBC[i]) continue;
        res1 += (P_bar[i] - PBC[i]) * (P_bar[i] - PBC[i]);
      }
      double res3 = (res2 < 1.0e-8) ? sqrt(res1) : sqrt(res1 / res2);
      nbc_loops++;
      if (res3 <= pb->tolNR) break;
      if (iiter_NBC > pb->maxIter) { rc = -5; goto done; }
      /* the step's CG iterations include tangent_homo's (as the library's
       * iteration counter does) */
      int hit = 0;
      if ((rc = orc_tangent_homo_it(f, pb->K4, pb->tolPCG, C_homo, &hit)) < 0) goto done;
      cgits += hit;
      for (int i = 0; i < 9; ++i) DbarF[i] = 0.0;
      if (orc_nbc_update(C_homo, DbarF, P_bar, PBC, isNBC)) { rc = -3; goto done; }
      for (int i = 0; i < 9; ++i) barF[i] = barF[i] + DbarF[i];
      iiter_NBC++;
    }
    for (int i = 0; i < 9; ++i) barF_t[i] = barF[i];
    memcpy(pb->Fn, pb->Fn1, n * sizeof(double));
    memcpy(pb->Pn, pb->Pn1, n * sizeof(double));
    memcpy(pb->sn, pb->snp1, (long)W * V * sizeof(double)); /* update (update.f:16-193) */
    for (int i = 0; i < 9; ++i) Pbar_out[(step - 1) * 9 + i] = P_bar[i];
    if (counters_out) {
      counters_out[(step - 1) * 4 + 0] = newton;
      counters_out[(step - 1) * 4 + 1] = cgits;
      counters_out[(step - 1) * 4 + 2] = nbc_loops;
      counters_out[(step - 1) * 4 + 3] = calls;
    }
  }
done:
  free(b); free(dFm);
  orc_fft_destroy(f);
  return rc;
}
