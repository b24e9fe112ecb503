/*
 * cpfft.h -- C-ABI of the MI355X implementation of CPFFT's finite-strain
 * FFT solver step (libcpfft_amd.so).
 *
 * The reference is a Fortran program whose hot path is the load-step loop
 * FFT_nr3 (src/FFT_nr3.f:14-200) around three operators:
 *   G_K_dF         src/G_K_dF.f:11-87        projection Gamma:(K4:F)
 *   fftPcg         src/FFT_nr3.f:214-360     conjugate gradients on it
 *   drive_eps_sig  src/drive_eps_sig.f:16-45 per-voxel stress update
 * plus tangent_homo (src/tangent_homo.f:11-73), update (src/update.f:16)
 * and the BLAS-1 calls FFT_nr3 makes (dnrm2/daxpy/dscal/dcopy).
 *
 * Each entry point below replaces one of those routines.  All state lives
 * in device memory owned by an opaque context; fields are exchanged as
 * device pointers (plain addresses, obtained from cpfft_field()).  The
 * reference's module globals (module fft, src/mod_fft.f:1-79) become
 * context fields with the same names.  A Fortran caller binds these with
 * ISO_C_BINDING interfaces (see cpfft_amd/fortran/cpfft_bindings.f90).
 *
 * Layout (identical to the reference's Fortran arrays F(N3,9), K4(N3,81)):
 *   second-order field  [9][V]  component c = 3(i-1)+j  (11,12,13,21,...)
 *   K4                  [81][V] index 27(i-1)+9(j-1)+3(k-1)+l
 *   voxel               e = ix*N*N + iy*N + iz (x slowest)
 *   material state      [W][V]  W = 26 + 2*nslip words, see CPFFT_W_*
 *
 * Errors: every call returns 0 on success, a positive number for a
 * non-fatal condition (CG not converged, local update failures) and a
 * negative number for a fatal error; cpfft_last_error() describes it.
 */
#ifndef CPFFT_H
#define CPFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cpfft_ctx cpfft_ctx;

/* Crystal plasticity material (one crystal per voxel, Voce hardening).
 * Mirrors the crystal properties the deck command `crystal` sets
 * (src/incrystal.f, defaults src/mod_crystals.f:231-410). */
typedef struct {
  int32_t slip_set;   /* the reference's slip_type: 1 fcc {111}<110> (12),
                         2 bcc {110}<111> (12), 6 roters (fcc, 12), 7 bcc12
                         (12), 8 bcc48 {110}+{112}+{123}<111> (48); and 80,
                         the first 24 systems of bcc48 ({110} and {112})    */
  double e, nu, mu;   /* cubic elasticity through the compliance (mod_crystals.f:1833-1879) */
  double rate_n;      /* harden_n: power-law rate exponent                   */
  double tau_y, tau_v, theta_0, voche_m;   /* Voce hardening                 */
  double gamma_bar;   /* eps_dot_0_y, reference slip rate (alter_mode)        */
  int32_t alter_mode; /* 1: dg = gamma_bar * dt (mm10_a.f:2073)               */
  double atol, atol1, rtol, rtol1;         /* local Newton tolerances          */
  int32_t miter;      /* local Newton iteration limit                         */
  int32_t jac_exact;  /* 0: local Jacobian J22 exactly as mm10_ehard_voche
                         (mm10_b.f:1951-1983) writes it; 1: its exact
                         derivative (no spurious sign(slipinc) factor)        */
} cpfft_material;

/* Green operator variants (G_K_dF.f:11-87 with FFT_init.f:272-385) */
enum {
  CPFFT_GAMMA_REFERENCE = 0, /* as the reference: for even N the phase shift of
                                Nhalf = N/2+1 bins pairs frequency f with Gamma(f-1) */
  CPFFT_GAMMA_TEXTBOOK = 1   /* orthogonal projection: Gamma(f), Nyquist planes
                                removed; identical to the reference for odd N        */
};

/* device fields owned by the context (module fft globals, mod_fft.f:51-53);
 * CPFFT_EPS [6][V] is the element strain of the reference's element blocks
 * (eps_n, mod_eleblocks.f:68): the sum over the rolled-over load steps of
 * the unrotated strain increment of each step's last update (state words
 * CPFFT_W_D), accumulated by cpfft_update (rstgp1.f:1378-1397, update.f:85)
 * and written by the result files (ouresult.f:230-243)                    */
enum {
  CPFFT_FN = 0, CPFFT_FN1 = 1, CPFFT_PN = 2, CPFFT_PN1 = 3, CPFFT_DFM = 4, CPFFT_B = 5,
  CPFFT_K4 = 6, CPFFT_STATE_N = 7, CPFFT_STATE_NP1 = 8, CPFFT_WORK = 9, CPFFT_EPS = 10,
  CPFFT_NFIELDS = 11
};

/* material state words (history words of mm10, mm10_d.f:174-254) */
#define CPFFT_W_SIG 0            /* unrotated Cauchy stress, Voigt 6           */
#define CPFFT_W_RP 6             /* plastic rotation, column-major 9           */
#define CPFFT_W_TT 15            /* slip resistance tau_tilde                  */
#define CPFFT_W_TTRATE 16        /* its rate, used by the next step's predictor */
#define CPFFT_W_D 17             /* unrotated strain increment, Voigt 6        */
#define CPFFT_W_SLIP 23          /* slip increments [nslip]                    */
#define CPFFT_W_SLIPSUM(ns) (23 + (ns))  /* cumulative slip [nslip]            */
#define CPFFT_W_WORK(ns) (23 + 2 * (ns)) /* work, plastic work, plastic strain */
#define CPFFT_NWORDS(ns) (26 + 2 * (ns))

/* ---- context ---------------------------------------------------------- */
/* FFT_init (FFT_init.f:16-177) + fftAllocate (:191-258) + the per-voxel
 * orientation setup of setup_mm10_rknstr (drive_eps_sig.f:537-1002).
 * angles_deg: host array [3][V] of Kocks Euler angles in degrees (the angle
 * file of inmat_cp); dt: time step (`time step`, indypm.f:56-58).         */
int cpfft_create(int N, const cpfft_material* mat, const double* angles_deg, double dt,
                 int gamma_mode, cpfft_ctx** out);
int cpfft_destroy(cpfft_ctx* ctx);
const char* cpfft_last_error(void);
int cpfft_nslip(const cpfft_ctx* ctx);
/* device address of a context field */
double* cpfft_field(cpfft_ctx* ctx, int field);
/* host <-> device copies of whole fields */
int cpfft_upload(cpfft_ctx* ctx, int field, const double* host);
int cpfft_download(cpfft_ctx* ctx, int field, double* host);
/* voxels [e0, e0+n) of every component of a field, host[k][i] =
 * field[k][e0+i] (a slab of a large field without copying all of it)     */
int cpfft_download_voxels(cpfft_ctx* ctx, int field, long e0, long n, double* host);
/* reset Fn = Fn1 = I, Pn = Pn1 = 0, state = initial (FFT_init.f:153-160) */
int cpfft_reset(cpfft_ctx* ctx);
int cpfft_sync(cpfft_ctx* ctx);
/* Host state a load-step driver carries from one step to the next, the
 * saved locals of FFT_nr3 (FFT_nr3.f:23-34): barF_t(9), P_bar(9),
 * C_homo(81), in that order (99 doubles owned by the context, set to
 * (I, 0, 0) by cpfft_create and cpfft_reset).  Keeping it in the context
 * lets one driver run several RVEs, one after another or at once.        */
double* cpfft_driver_state(cpfft_ctx* ctx);

/* ---- operators -------------------------------------------------------- */
/* G_K_dF (G_K_dF.f:11-19): GKF = Gamma:(K4:F) if flgK else Gamma:F.  F and
 * GKF are device fields [9][V]; may not alias.                            */
int cpfft_gkdf(cpfft_ctx* ctx, const double* F, double* GKF, int flgK);

/* fftPcg (FFT_nr3.f:214-360): x = CG solution of Gamma:(K4:x) = b, x0 = 0,
 * stop at ||r|| <= tol*||b|| or ||r|| <= tol, 1000 iterations at most.
 * b must lie in the range of Gamma (b = Gamma:(.)), as every right-hand
 * side FFT_nr3 and tangent_homo form does: with the textbook operator (a
 * symmetric projection) p.q is then formed as p.(K4:p).
 * Returns 0, 1 (zero right-hand side, x = 0), -1 (tol outside (eps,1)),
 * -2 (no convergence).                                                    */
int cpfft_pcg(cpfft_ctx* ctx, const double* b, double* x, double tol, int* iters,
              double* relres);

/* drive_eps_sig (drive_eps_sig.f:16-45): from FN, FN1 and STATE_N compute
 * PN1, K4 and STATE_NP1.  iiter = 0 selects the explicit elastic predictor
 * (rstgp1.f:870-876).  Returns the number of voxels whose local solve
 * failed after all sub-step cuts (0 normally).                            */
int cpfft_drive_eps_sig(cpfft_ctx* ctx, int step, int iiter);

/* tangent_homo (tangent_homo.f:11-73): homogenised tangent C_homo(81) on
 * the host, C_homo[9(j-1)+i] = d mean(P_j) / d Fbar_i.                    */
int cpfft_tangent_homo(cpfft_ctx* ctx, double tol, double C_homo[81]);

/* update (update.f:16-193) together with the roll-over of FFT_nr3.f:173-175:
 * state n <- n+1, Fn <- Fn1, Pn <- Pn1, and CPFFT_EPS += the new state n's
 * strain increment (eps_n <- eps_n1 = eps_n + uddt).                      */
int cpfft_update(cpfft_ctx* ctx);

/* ---- BLAS-1 on device vectors of 9*V doubles (FFT_nr3.f:86-114) ------ */
int cpfft_dnrm2(cpfft_ctx* ctx, const double* x, double* out);       /* dnrm2 */
int cpfft_daxpy(cpfft_ctx* ctx, double a, const double* x, double* y); /* daxpy */
int cpfft_dscal(cpfft_ctx* ctx, double a, double* x);                  /* dscal */
int cpfft_dcopy(cpfft_ctx* ctx, const double* x, double* y);           /* dcopy */
/* dFm(1:N3,ii) = DbarF(ii) (FFT_nr3.f:85-87) */
int cpfft_fill9(cpfft_ctx* ctx, double* x, const double v[9]);
/* P_bar(ii) = sum_jj Pn1(jj,ii) / N3 (FFT_nr3.f:127-134) */
int cpfft_mean9(cpfft_ctx* ctx, const double* x, double out[9]);

/* ---- displacement output (f2disp.f:16-175) -----------------------------
 * Nodal displacements of the (N+1)^3-node brick mesh from the deformation
 * gradient F (device, [9][V], component 3i+j = F_ij), box edge lengths
 * lengths[3] (the deck's sizes l_x, l_y, l_z): the reference's system
 * A x = b (form_BTB 2x2x2 Gauss, one-point right-hand side, node 1 held at
 * zero), solved in closed form by discrete cosine transforms instead of
 * PARDISO; u = x - X.  u (host) receives 3 (N+1)^3 doubles, node-major
 * (u[3 n + i], n = ix + (N+1) iy + (N+1)^2 iz), as the reference's u
 * array.  Uses the context's spectrum buffers as scratch.  -1 on a bad
 * argument.                                                                */
int cpfft_f2disp(cpfft_ctx* ctx, const double* F, const double lengths[3], double* u);

/* ---- G_K_dF for callers that cannot allocate --------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Computes G_K_dF exactly as cpfft_gkdf
 * does (the same kernels; bit-identical results) without a context:
 * Gamma:(K4:F), or Gamma:F when K4 is NULL, for the N^3 grid, on
 * caller-owned device arrays F and GKF ([9][N^3], may not alias) and K4
 * ([81][N^3]), enqueued on `stream` (a hipStream_t; NULL: the null stream)
 * and returning without waiting.  The spectrum scratch is one static
 * device array of 16.6 MB per GPU (N <= CPFFT_NOALLOC_MAX_N), held until
 * the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU (different streams or threads) race, so callers
 * serialise them.  Its address is resolved once per device (by
 * cpfft_create, or by the first call there); the call itself makes no
 * allocation.  Returns 0; -1 for a bad argument, N above the maximum
 * included; -3 for an N the operator cannot serve (prime factor above 16);
 * -100 or below for a HIP error (cpfft_last_error() describes each).     */
#define CPFFT_NOALLOC_MAX_N 48
int cpfft_gkdf_noalloc(int N, int gamma_mode, const double* F, const double* K4, double* GKF,
                       void* stream);

/* ---- execution options (no counterpart in the reference) -------------- */
/* on = 1 or 2: cpfft_pcg and cpfft_tangent_homo replay their chunks of
 * eight CG iterations from HIP graphs (one per solution pointer for
 * cpfft_pcg), one graph launch per host round trip instead of about 70
 * kernel launches.  Results are unchanged.  With 1 the graphs the load-step
 * driver uses are captured now (and again after a change of launch shapes);
 * with 2 each is captured on its first use.  Captures are serialised over
 * the process and checked (kernel nodes only, the same count as the first
 * capture of that chunk; -101 otherwise).  While kernel timing is on
 * (cpfft_set_timing) the kernels are launched one by one.  0 turns graphs
 * off; other values: -1.                                                  */
int cpfft_set_graphs(cpfft_ctx* ctx, int on);

/* Iteration limit of cpfft_pcg and of tangent_homo's solves, 1000 by
 * default as fftPcg's maxIter (FFT_nr3.f:224).  A solve that needs the
 * limit's number of iterations fails (-2) even when the last one passed
 * the convergence test, as fftPcg aborts on itercount >= maxIter
 * (FFT_nr3.f:335-339).  -1 for a limit below 1.                           */
int cpfft_set_cg_maxit(cpfft_ctx* ctx, int maxit);

/* Autotuner of the compile-time-planned FFT path (N in 16, 32, 48, 64, 96,
 * 128, 256): times the candidate launch shapes of each of pass A with K4,
 * pass A without K4, the y passes and pass C on this context's buffers
 * (reps launches each) and keeps the fastest.  The choice is cached per
 * device and N for every context of the process (force != 0 measures
 * again).  Only shapes that compute bit-identical results are candidates
 * (which block transforms which lines; never the radix order or a
 * reduction), so tuning changes speed, not results.  Returns 1 (nothing to
 * tune) for other N.                                                      */
int cpfft_autotune(cpfft_ctx* ctx, int reps, int force);
/* shapes in use: out[0] pencils per block of pass A with K4, out[1]
 * without K4, out[2] lines per block of the y passes, out[3] component
 * lines per block of pass C: 24, 12 or 6 (one tensor row of 8, 4 or 2 y
 * lines, transforms in registers) for N = 64, 128, 256, else 18 or 9 (all
 * nine components of 2 or 1 y lines); 0 where N has no planned path      */
int cpfft_get_tune(cpfft_ctx* ctx, int out[4]);
/* set them (each must be one of the candidates; -1 otherwise)             */
int cpfft_set_tune(cpfft_ctx* ctx, const int in[4]);

/* ---- diagnostics ------------------------------------------------------ */
/* counters since creation: [0] G_K_dF applications, [1] CG iterations,
 * [2] drive_eps_sig calls, [3] local predictor iterations, [4] local
 * Newton iterations, [5] sub-step cuts, [6] failed voxels, [7] residual
 * evaluations in line searches                                            */
int cpfft_counters(cpfft_ctx* ctx, long long out[8]);
/* device-time measurement with HIP events recorded on the context's stream
 * (no host synchronisation while recording).  out[0..3]: average ms per
 * launch group since cpfft_set_timing of [0] G_K_dF with K4 (five kernels),
 * [1] drive_eps_sig, [2] CG vector update, [3] G_K_dF without K4;
 * out[4..7]: number of launch groups measured.
 * cpfft_set_timing(ctx, k): 0 off; k > 0 records every k-th launch group
 * of each family.  Each recorded pair costs the stream about 10 us of idle
 * time around it, so a long run is timed with a sampling stride (bench.py
 * uses 7, coprime to the CG chunk of 8).  Negative k: -1.                  */
int cpfft_kernel_times(cpfft_ctx* ctx, double out[8]);
int cpfft_set_timing(cpfft_ctx* ctx, int on);

/* ---- learned warm start ------------------------------------------------
 * No counterpart in the reference.  An nin -> 128 -> 128 -> nout tanh
 * network, trained and evaluated with BF16 matrix-core GEMMs (FP32
 * accumulation), predicts each voxel's slip increments from its trial
 * state.  The prediction only sets the starting point of the local Newton
 * solve (mm10_a.f:2945-3023): the FP64 iteration and its tolerances are
 * unchanged, and a warm-started solve that fails is redone from the plain
 * start.  For fcc (ns = 12 slip systems) and bcc24 (ns = 24).
 * Inputs (nin = 18 + 2 ns: 42 for fcc, 66 for bcc24): trial stress s_n + C:D
 * in the lattice frame (3x3), the slip resistance once per system (ns), the
 * lattice-frame strain increment (3x3), the previous slip increments (ns).
 * Outputs (nout = ns): slip increments.  Sizes as functions of ns:        */
#define CPFFT_MLP_HID 128
#define CPFFT_MLP_NIN(ns) (18 + 2 * (ns))
#define CPFFT_MLP_NOUT(ns) (ns)
/* W1 (128 x nin) b1 W2 (128 x 128) b2 W3 (nout x 128) b3, rows = outputs  */
#define CPFFT_MLP_NPARAMS(ns) \
  (CPFFT_MLP_HID * (CPFFT_MLP_NIN(ns) + 1) + CPFFT_MLP_HID * (CPFFT_MLP_HID + 1) + \
   (ns) * (CPFFT_MLP_HID + 1))
/* mean_x[nin] sd_x[nin] mean_t[nout] sd_t[nout]                            */
#define CPFFT_MLP_NNORMS(ns) (2 * CPFFT_MLP_NIN(ns) + 2 * (ns))
/* the fcc sizes                                                            */
#define CPFFT_MLP_IN 42
#define CPFFT_MLP_OUT 12
#define CPFFT_MLP_NPARAM 23564
#define CPFFT_MLP_NNORM 108
/* allocate the network (Glorot-uniform weights from seed), a log of
 * `capacity` training pairs and work buffers for batches of `batch` rows  */
int cpfft_mlp_init(cpfft_ctx* ctx, uint64_t seed, long capacity, int batch);
/* append (inputs, converged slip increments) of every stride-th voxel of
 * the current update (FN, FN1, STATE_N -> STATE_NP1) to the log; the log is
 * a ring of `capacity` pairs, so once full the newest pairs replace the
 * oldest (online retraining on recent load steps)                         */
int cpfft_mlp_log(cpfft_ctx* ctx, int step, long stride);
/* normalise the log, then `iters` Adam steps (learning rate lr) on batches
 * drawn with replacement; *loss = mean squared error (normalised units) of
 * the last batch                                                          */
int cpfft_mlp_train(cpfft_ctx* ctx, int iters, double lr, uint64_t seed, double* loss);
/* Warm start of the local solves of drive_eps_sig calls with iiter > 0
 * (no counterpart in the reference; off by default, when results are those
 * of the reference's start).  mode is a bit set of the candidate starts:
 * CPFFT_WARM_NETWORK the network's predicted slip increments (needs
 * cpfft_mlp_init; 12 or 24 slip systems), CPFFT_WARM_PREVIOUS the step's previous
 * global iterate held in STATE_NP1.  With both, the network serves only
 * iiter == 1 (where the previous iterate is the explicit update of iiter
 * 0) and the previous iterate the later calls.  Per voxel the candidate of smaller
 * full residual replaces the reference's predictor stage when its residual
 * is below the predictor's starting residual; the Newton tolerances are
 * unchanged and a failed warm solve is redone from the reference's start.
 * Returns -1 on unknown bits or a network start without a network.       */
#define CPFFT_WARM_NETWORK 1
#define CPFFT_WARM_PREVIOUS 2
int cpfft_set_warm_start(cpfft_ctx* ctx, int mode);
/* on != 0 adds CPFFT_WARM_NETWORK to the warm-start mode, on == 0 removes it */
int cpfft_mlp_warm_start(cpfft_ctx* ctx, int on);
/* parameters then normalisation (CPFFT_MLP_NPARAMS(ns) + CPFFT_MLP_NNORMS(ns)
 * floats): get != 0 copies them out, get == 0 sets them                   */
int cpfft_mlp_params(cpfft_ctx* ctx, float* buf, int get);
/* network output Y[n][nout] for normalised inputs Xn[n][nin] (n <= batch) */
int cpfft_mlp_forward(cpfft_ctx* ctx, const float* Xn, int n, float* Y);
/* gradient (CPFFT_MLP_NPARAMS(ns) floats) and value of mean((Y - Tn)^2)
 * over a batch of normalised inputs and targets (n <= batch)              */
int cpfft_mlp_grad(cpfft_ctx* ctx, const float* Xn, const float* Tn, int n, float* grad,
                   double* loss);
/* the warm start of every voxel for the current state (FN, FN1, STATE_N)
 * of load step `step`: X (if not NULL) receives the raw network inputs
 * [V][nin], ws (if not NULL) the predicted slip increments, as the
 * normalised resolved shears the local solve starts from, [nout][V] --
 * what cpfft_drive_eps_sig uses when the warm start is on                 */
int cpfft_mlp_predict(cpfft_ctx* ctx, int step, float* X, float* ws);
/* Divergence of the local solves since creation (no counterpart in the
 * reference): [0] local iterations of the voxels (predictor + full Newton),
 * [1] local iterations issued, i.e. per wavefront of 64 voxels 64 times its
 * slowest voxel's count, [2] line-search residual evaluations, [3] sub-step
 * cuts.  [0] / [1] is the fraction of the issued iterations doing work.   */
int cpfft_local_stats(cpfft_ctx* ctx, long long out[4]);
/* [0] logged pairs, [1] training steps, [2] warm-started voxel updates,
 * [3] warm-started solves redone from the plain start                     */
int cpfft_mlp_stats(cpfft_ctx* ctx, long long out[4]);

#ifdef __cplusplus
}
#endif
#endif
