/* cpfft_driver.h -- C-ABI of libcpfft_driver.so: the load-step driver
 * (FFT_nr3) written in Fortran over the operators of include/cpfft.h, and
 * a runner that drives many independent RVEs at once.
 */
#ifndef CPFFT_DRIVER_H
#define CPFFT_DRIVER_H

#include "cpfft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* compute + FFT_nr3 (FFT_finite_3d.f:138-147, FFT_nr3.f:14-200) for load
 * steps first_step .. first_step+nstep-1 on one context.  first_step = 1
 * starts a run (elastic call and initial tangent); a later first_step
 * continues the run.  BC_all[nstep][9]: boundary values as inlod fills
 * them (inlod.f:57-63); isNBC[9]: 1 where the component of P is
 * prescribed.  Outputs: Pbar_out[nstep][9]; counters[nstep][4] = Newton
 * iterations, CG iterations, mixed-BC loops, drive_eps_sig calls;
 * times[3] = pcg, sig-eps, total wall seconds (thyme.f:15-47).  Steps up to
 * mlp_log_last log warm-start training pairs (every mlp_log_stride-th
 * voxel); 0 turns that off.  Returns 0, or the negative code of the
 * reference's die_abort paths: -2 CG not converged (FFT_nr3.f:335-339),
 * -3 NBC_update singular (:418-421), -4 Newton maxIter (:116-119), -5
 * mixed-BC maxIter (:153-156), other negatives from the operators.
 * Holds no state outside the context (cpfft_driver_state), so calls on
 * different contexts may run at the same time on different threads.    */
int cpfft_compute(cpfft_ctx* ctx, int first_step, int nstep, double* BC_all, int* isNBC,
                  double tolPCG, double tolNR, int maxIter, int verbose, double* Pbar_out,
                  int* counters, double* times, int mlp_log_last, int mlp_log_stride);

/* One RVE of a batch: the arguments of cpfft_compute (verbose off, no
 * warm-start logging) plus its results. */
typedef struct {
  cpfft_ctx* ctx;
  int first_step, nstep;
  double* BC_all;     /* [nstep][9] */
  int* isNBC;         /* [9] */
  double tolPCG, tolNR;
  int maxIter;
  double weight;      /* expected cost (e.g. N^3 * nstep); heavier RVEs start first */
  double* Pbar_out;   /* [nstep][9] */
  int* counters;      /* [nstep][4] */
  double times[3];    /* out: as cpfft_compute */
  int rc;             /* out: cpfft_compute's return code */
} cpfft_batch_item;

/* Batched uneven RVEs (no counterpart in the reference, which solves one
 * RVE per run): runs cpfft_compute on every item, nthreads host threads
 * taking items from a shared queue in order of decreasing weight (longest
 * first).  Each context has its own HIP stream, so the RVEs' kernels
 * overlap on the GPU.  Items must use distinct contexts.  Returns 0 when
 * every item returned 0, else the first nonzero rc in item order (every
 * item is still run).                                                    */
int cpfft_compute_batch(cpfft_batch_item* items, int n, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* CPFFT_DRIVER_H */
