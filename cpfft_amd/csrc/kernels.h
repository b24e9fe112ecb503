// kernels.h -- launch wrappers for the CDNA4 kernels (defined in
// cpfft_kernels.hip), called by the context code in cpfft_api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_device.h"
#include "cpfft_common.h"
#include "green.h"
#include "mm10_core.h"
#include "mlp.h"

namespace cpfft {

// Launch shapes of the compile-time-planned FFT path chosen by the
// autotuner, as candidate indices (0 = the default shape): a, a0 pass A
// with / without K4, b the y passes, c pass C.  All candidates compute
// bit-identical results.
struct PencilTune {
  int a = 0, a0 = 0, b = 0, c = 0;
};
constexpr int kPencilCands = 5;  // candidate indices per pass are below this

struct SpecGeom {
  int N, Nz;        // grid edge, number of z frequencies kept (N/2+1)
  long V;           // N^3
  long S1;          // complex entries per spectral component: N*N*Nz
  FftPlan plan;     // radix plan for length N
  int gamma_mode;   // 0 reference operator, 1 textbook projection
  int ty;           // kz columns per block in the y pass
  int tz;           // kz columns per block in the x pass
  const int* skip;  // if non-null and *skip != 0, every block returns at once
  PencilTune tune;  // launch shapes of the planned path
};

// Geometry of the operator for edge N and Gamma variant gamma_mode: the
// spectrum sizes, the radix plan and the generic path's tiles, sized for
// min(lds, 64 KiB) bytes of shared memory per block (the same tiles on
// every device, so results do not depend on the device).  Returns nullptr,
// or why N cannot be served.
const char* make_spec_geom(int N, int gamma_mode, int lds, SpecGeom& g);

// Displacement reconstruction (f2disp.hip): the cosine table (of
// f2disp_cos_doubles(N) doubles) once, then u (device, [3 (N+1)^3]
// node-major) from F; work holds f2disp_scratch_doubles(N) doubles and must
// not overlap u.
size_t f2disp_scratch_doubles(int N);
size_t f2disp_cos_doubles(int N);
void launch_f2disp_cos(int N, double* C, hipStream_t s);
void launch_f2disp(int N, const double* F, const double len[3], const double* C, double* work,
                   double* u, hipStream_t s);

// Gamma:(K4:F) or Gamma:F (K4 == nullptr).  spec: 2*9*S1 complex scratch.
// tw: the N twiddles exp(-2 pi i k/N).  Uses the compile-time-planned path
// of fft_pencil.hip where it exists for N, else the generic runtime plan.
// CG work fused into the operator's passes (compile-time-planned path only):
// dir_p != nullptr: pass A first sets dir_p = dir_r + cs->beta * dir_p (the
//   CG direction, F must be dir_p) and transforms the new direction;
// dot_p != nullptr: pass D writes per-block partial sums of dot_p . out
//   into dot_part (gkdf_dot_parts(N) values, fixed order);
// dc != nullptr: pass C stores the voxel sums of the nine components of its
//   input (F, or K4:F) in dc[0..9);
// pq_part != nullptr (with dir_p): pass A writes the partial sums of
//   p . (K4:p) over each pencil pair (y, y+1) of each x into
//   pq_part[x*N/2 + y/2] (N*N/2 values, the same under every tile).  For a
//   symmetric projection G (the textbook operator) and p in its range, as
//   every CG direction of a right-hand side G(.) is, p . G(K4:p) equals
//   p . (K4:p), so pass D then needs no dot product.
// alpha_np > 0: right after pass A, alpha = cs->rr / (sum of the alpha_np
//   partial sums in pq_part);
// upd_r != nullptr (alpha already in cs): pass D does not store the
//   transform q but updates upd_r -= alpha q, with upd_x != nullptr also
//   upd_x += alpha dir_p, and writes per-block partial sums of the new
//   upd_r . upd_r into rr_part (gkdf_dot_parts(N) values).
struct CGFuse {
  double* dir_p;
  const double* dir_r;
  CGScalars* cs;
  const double* dot_p;
  double* dot_part;
  double* dc;
  double* pq_part;
  int alpha_np;
  double* upd_r;
  double* upd_x;
  double* rr_part;
};

void launch_gkdf(const SpecGeom& g, const double* F, const double* K4, double* out, cplx* spec,
                 const cplx* tw, hipStream_t s);
bool gkdf_pencil_supported(int N);
// returns false (and launches nothing) when N has no compile-time plan
bool launch_gkdf_pencil(const SpecGeom& g, const double* F, const double* K4, double* out,
                        cplx* spec, const cplx* tw, hipStream_t s, const CGFuse* fuse = nullptr);
// number of partial sums pass D writes for a fused dot product
int gkdf_dot_parts(int N);
// autotuner support: the value (pencils or lines per block) of candidate
// `variant` of pass `pass` (0 A with K4, 1 A without, 2 B, 3 C) for edge N,
// 0 if there is none; and one pass launched by itself with that candidate
int pencil_candidate(int N, int pass, int variant);
bool launch_pencil_pass(const SpecGeom& g, int pass, int variant, const double* F,
                        const double* K4, cplx* spec, const cplx* tw, hipStream_t s);

// alpha = rr / sum(part[0..np)) (p.q partial sums); for nsys systems at
// once, system b reading part + b pstride and cs[b] (one block each)
void launch_cg_alpha(const double* part, int np, CGScalars* cs, hipStream_t s, int nsys = 1,
                     long pstride = 0);
// convergence test and beta from the r.r partial sums part[0..np); with
// m != nullptr also m[c] += alpha * dc[c] * scale for c < 9 (the running sum
// over iterations of alpha times the mean of K4:p); nsys systems as for
// launch_cg_alpha, with dc and m advanced by 9 per system
void launch_cg_check(const double* part, int np, CGScalars* cs, hipStream_t s,
                     const double* dc = nullptr, double* m = nullptr, double scale = 0.0,
                     int nsys = 1, long pstride = 0);
// x += alpha p (xupd; without it x and p are not touched); r -= alpha q
// with the r.r partial sums, then the convergence test and beta (no p.q, no
// direction update)
void launch_cg_update_check(long n, CGScalars* cs, const double* p, const double* q, double* x,
                            double* r, double* partials, hipStream_t s, bool xupd = true,
                            const double* dc = nullptr, double* m = nullptr,
                            double scale = 0.0);

// One CG iteration after q = A p: alpha = rr / (p.q); x += alpha p;
// r -= alpha q; convergence test on |r|; p = r + beta p.  All scalars stay
// in *cs; nothing happens once cs->done is set.
void launch_cg_tail(long n, CGScalars* cs, const double* p, const double* q, double* x,
                    double* r, double* partials, hipStream_t s);
// p = r + beta p with beta from *cs (skipped once cs->done is set)
void launch_cg_dir_dev(long n, const CGScalars* cs, const double* r, double* p, hipStream_t s);

// up to nine field pointers (multi-RHS solves)
struct Ptr9 {
  const double* p[9];
  double* q[9];
};
// q_i = K4 : p_i (ddot42n) for every i < nrhs with cs[i].done == 0; with
// r != nullptr the direction is first updated, p_i = r_i + cs[i].beta p_i
struct Ptr9r {
  const double* r[9];
};
// pqpart != nullptr: also p_i . (K4:p_i) of each running right-hand side,
// per block, at pqpart[i * k4_contract_multi_blocks(V) + block]
void launch_k4_contract_multi(long V, const double* K4, const Ptr9& pp, const CGScalars* cs,
                              int nrhs, hipStream_t s, const Ptr9r* rr = nullptr,
                              double* pqpart = nullptr);
int k4_contract_multi_blocks(long V);

// reductions: result -> d_out[0..nres) on device (two-stage, fixed order)
void launch_dot(const double* x, const double* y, long n, double* partials, double* d_out,
                hipStream_t s);
void launch_mean9(const double* x, long V, double* partials, double* d_out, hipStream_t s);
void launch_axpy(long n, double a, const double* x, double* y, hipStream_t s);
void launch_scal(long n, double a, double* x, hipStream_t s);
void launch_copy(long n, const double* x, double* y, hipStream_t s);
void launch_fill9(long V, const double* v9_host, double* x, hipStream_t s);
// x[0..n) = v (a store, unlike a scaling by zero, which keeps NaN or Inf)
void launch_fill(long n, double v, double* x, hipStream_t s);
void launch_fill_f32(long n, float v, float* x, hipStream_t s);
// initial material state [W][V] (mm10_init_cc_hist0, mm10_a.f:473-573, with
// mm10_init_voche, :2042-2055): zero, except Rp = I and tau_tilde = tt0
void launch_state_init(long V, int W, double tt0, double* state, hipStream_t s);
void launch_add_identity_comp(long V, int comp, double* x, hipStream_t s);
// CG vector update: x += a p; r -= a q; partial r.r into partials
void launch_cg_update(long n, double a, const double* p, const double* q, double* x, double* r,
                      double* partials, double* d_out, hipStream_t s);
// p = r + b p
void launch_cg_dir(long n, double b, const double* r, double* p, hipStream_t s);
// C[j] = K4[(9 j + i) rows] copy (tangent_homo column i)
void launch_gather_k4_col(long V, int i, const double* K4, double* C, hipStream_t s);
// out[j] = sum_e sum_m K4[9j+m][e] * C[m][e] for j=0..8 (tangent_homo.f:62-66)
void launch_k4_dot9(long V, const double* K4, const double* C, double* partials, double* d_out,
                    hipStream_t s);

// per-voxel update: stress update (state n+1, P, material tangent into
// tan6 [36][V]) followed by its pull-back to K4 = dP/dF
// ws: predicted slip increments [NS][V] for the warm start, or nullptr;
// warm_prev != 0: the state n+1 the call overwrites (this step's previous
// global iterate) is a warm-start candidate too
template <int NS>
void launch_drive_eps_sig(long V, const MatParams& P, const Orient<NS>* orients,
                          const int32_t* orient_id, int step, int iter, const double* Fn,
                          const double* Fn1, const double* sn, double* snp1, double* Pn1,
                          double* tan6, double* K4, const float* ws, int warm_prev,
                          long long* counters, hipStream_t s);

// warm-start network inputs of voxels e0 + r*stride (r < nrows) into
// X[r][18 + 2 NS]; with T also their converged slip increments (state n+1)
template <int NS>
void launch_mlp_features(long V, const MatParams& P, const Orient<NS>* orients, const int32_t* oid,
                         int step, const double* Fn, const double* Fn1, const double* sn,
                         const double* snp1, long e0, long stride, long nrows, float* X, float* T,
                         hipStream_t s);
// every voxel's inputs normalised with nrm ([mean | sd]) as BF16 rows of
// ninp columns (zero-padded) into Xv
template <int NS>
void launch_mlp_features_norm(long V, const MatParams& P, const Orient<NS>* orients,
                              const int32_t* oid, int step, const double* Fn, const double* Fn1,
                              const double* sn, const float* nrm, int ninp, uint16_t* Xv,
                              hipStream_t s);

int reduce_partials_count(long n);

// error text for cpfft_last_error (of the calling thread)
void set_last_error(const char* what);
// cpfft_gkdf_noalloc's scratch: the device address of the current GPU's
// copy of its static array (resolved on the first call per device; also
// from cpfft_create).  0, or a negative code with the error text set.
int noalloc_scratch_current_device(cplx** out);

}  // namespace cpfft
