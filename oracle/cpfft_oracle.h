/*
 * cpfft_oracle.h -- CPU restatement of the CPFFT finite-strain FFT solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X implementation in cpfft_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links,
 * loads or calls it.
 *
 * Every routine restates a reference routine and cites it (file:line,
 * relative to the reference's src/ directory).  Arithmetic is IEEE FP64
 * throughout, operation order follows the reference where it matters for
 * rounding (ddot42n pairwise order, Cardano polar decomposition).
 *
 * Layouts (identical to the product's SoA layout):
 *   second-order fields  : [9][V]   component c = 3(i-1)+j (row-major 11,12,..)
 *   fourth-order fields  : [81][V]  index 27(i-1)+9(j-1)+3(k-1)+l
 *   voxel index          : e = ix*N*N + iy*N + iz   (x slowest)
 *   per-voxel state words: see ORC_W_* below, [W][V]
 */
#ifndef CPFFT_ORACLE_H
#define CPFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SLIP 48

/* crystal properties (mod_crystals.f:24-61 crystal_props subset used by the
 * Voce path; defaults per mod_crystals.f:231-410) */
typedef struct {
  int    nslip;
  double bi[ORC_MAX_SLIP][3];   /* slip directions, crystal frame */
  double ni[ORC_MAX_SLIP][3];   /* slip plane normals, crystal frame */
  double stiff[6][6];           /* elast_stiff (crystal frame, Voigt, eng. shear) */
  double rate_n, tau_y, tau_v, theta_0, voche_m, eps_dot_0_y, iD_v;
  int    alter_mode;
  double atol, atol1, rtol, rtol1;
  int    miter;
  int    jac_exact;   /* 0: J22 as mm10_ehard_voche (mm10_b.f:1951-1983) writes it,
                         1: the exact derivative of R2 (drops its sign(slipinc) factor) */
} orc_crystal;

/* state word offsets for a crystal with nslip systems */
#define ORC_W_SIG      0   /* unrotated Cauchy stress, Voigt 6 (hist crystal term 1) */
#define ORC_W_RP       6   /* plastic rotation Rp, column-major 9 (term 3)          */
#define ORC_W_TT      15   /* tau_tilde (term 7)                                     */
#define ORC_W_TTRATE  16   /* tt_rate (term 9)                                       */
#define ORC_W_D       17   /* unrotated strain increment uddt, Voigt 6 (term 4)      */
#define ORC_W_SLIP    23   /* slip increments [nslip] (term 6)                       */
#define ORC_W_SLIPSUM(ns) (23 + (ns))          /* cumulative slip [nslip] (common 5)  */
#define ORC_W_WORK(ns)    (23 + 2 * (ns))      /* work, plastic work, plastic strain  */
#define ORC_NWORDS(ns)    (26 + 2 * (ns))

/* ---- crystal setup ---------------------------------------------------- */
/* slip_set: 1 = fcc (12), 8 = bcc48 (48), 80 = first 24 systems of bcc48 ({110}+{112}) */
int  orc_crystal_init(orc_crystal* c, int slip_set, double e, double nu, double mu,
                      double rate_n, double tau_y, double tau_v, double theta_0,
                      double voche_m, int alter_mode, double gamma_bar);
void orc_crystal_stiffness(const orc_crystal* c, double out36[36]);

/* ---- per-voxel stress update (drive_eps_sig.f:57-339 for one voxel) ---- */
/* Fn, Fn1 row-major (F11,F12,...,F33); angles = Kocks Euler angles in degrees.
 * sn / snp1: state words (stride 1). stats[0..4] = predictor iters, update
 * iters, sub-step cuts, failed flag, line-search residual evaluations.     */
int orc_voxel_update(const orc_crystal* cp, const double angles[3], int step, int iter,
                     double dt, const double Fn[9], const double Fn1[9],
                     const double* sn, double* snp1, double P[9], double K4[81],
                     int stats[5]);

/* field version: SoA arrays, OpenMP over voxels. stats_out may be NULL,
 * else receives 5 summed counters. angles: [3][V]. */
int orc_drive_eps_sig(const orc_crystal* cp, long V, const double* angles, int step,
                      int iter, double dt, const double* Fn, const double* Fn1,
                      const double* sn, double* snp1, double* Pn1, double* K4,
                      long* stats_out);

/* initial state words (mm10_init_cc_hist0, mm10_init_general_hist) */
void orc_state_init(const orc_crystal* cp, long V, double* s);

/* pieces exposed for unit tests */
void orc_rotation_matrix(const double angles_deg[3], double g[9]);  /* row-major */
void orc_rtcmp1(const double F[9], double R[9]);                     /* row-major */
void orc_cep2A(const double Fn[9], const double t6[6], const double C36[36],
               const double Rh[9], double detFh, const double fnhinv[9],
               const double R[9], const double Fn1[9], const double finv[9],
               double detF, double dPdF[81]);

/* ---- projection operator G_K_dF (G_K_dF.f:11-87) ----------------------- */
typedef struct orc_fft orc_fft;
orc_fft* orc_fft_create(int N);
/* mode 0 = reference operator (even-N shift quirk kept), 1 = textbook Gamma */
orc_fft* orc_fft_create2(int N, int mode);
void     orc_fft_destroy(orc_fft* f);
/* GKF = G(K4:F) (flgK != 0) or G(F); literal C2C + phase-shift restatement */
void orc_gkdf(orc_fft* f, const double* K4, const double* F, double* GKF, int flgK);
/* ddot42n (G_K_dF.f:241-268): C[r] = sum_c A[9r+c] B[c] in the reference order */
void orc_ddot42n(long V, const double* A4, const double* B2, double* C2);

/* ---- solver pieces (FFT_nr3.f) ---------------------------------------- */
/* fftPcg (FFT_nr3.f:214-360); returns 0 ok, 1 zero RHS, -1 bad tol, -2 maxit */
int  orc_pcg_limit(orc_fft* f, const double* K4, const double* b, double* x, double tol,
                    int maxIter, int* iters, double* relres);
int  orc_pcg(orc_fft* f, const double* K4, const double* b, double* x, double tol,
             int* iters, double* relres);
int  orc_tangent_homo(orc_fft* f, const double* K4, double tol, double C_homo[81]);
int  orc_tangent_homo_it(orc_fft* f, const double* K4, double tol, double C_homo[81],
                         int* iters);
int  orc_nbc_update(const double C_homo[81], double DbarF[9], const double P_bar[9],
                    const double PBC[9], const int isNBC[9]);

/* full FFT_nr3 loop (FFT_nr3.f:14-200) over nstep load steps.
 * BC_all: [nstep][9]; Pbar_out: [nstep][9]; counters_out: [nstep][4] =
 * (Newton iterations, CG iterations of every solve in the step, tangent_homo's
 * included, mixed-BC loops, drive_eps_sig calls). */
typedef struct {
  int N; long V;
  const orc_crystal* cp;
  const double* angles;   /* [3][V] */
  double dt, tolPCG, tolNR;
  int maxIter;
  int gamma_mode;         /* 0 = reference operator, 1 = textbook Gamma */
  /* state owned by caller: */
  double *Fn, *Fn1, *Pn, *Pn1, *K4, *sn, *snp1;
} orc_problem;
int orc_fft_nr3(orc_problem* pb, int nstep, const double* BC_all, const int isNBC[9],
                double* Pbar_out, int* counters_out);

#ifdef __cplusplus
}
#endif
#endif
