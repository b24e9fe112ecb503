// cpfft_api.hip -- context, host-side setup and the C-ABI entry points of
// libcpfft_amd.so (declared in include/cpfft.h).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cpfft.h"
#include "kernels.h"
#include "setup_host.h"

using namespace cpfft;

static thread_local std::string g_err;

static int fail(int code, const char* what) {
  g_err = what;
  return code;
}

void cpfft::set_last_error(const char* what) { g_err = what; }

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
      return -100;                                                                \
    }                                                                             \
  } while (0)

struct cpfft_ctx {
  int N = 0, Nz = 0;
  long V = 0;
  int nslip = 0;
  int dev = 0;
  hipStream_t stream = nullptr;
  SpecGeom geom{};
  MatParams mp{};
  double* fields[CPFFT_NFIELDS] = {};
  long field_len[CPFFT_NFIELDS] = {};
  cplx* spec = nullptr;
  cplx* tw = nullptr;
  double* r = nullptr;   // CG residual
  double* p = nullptr;   // CG direction
  double* q = nullptr;   // CG operator image
  double* tmp9 = nullptr;  // tangent_homo right-hand side
  double* tan6 = nullptr;  // per-voxel material tangent [36][V]
  double* partials = nullptr;
  double* dpart = nullptr;  // p.q partial sums of the fused passes A or D
  double* rpart = nullptr;  // r.r partial sums of pass D in its update mode
  double* dres = nullptr;  // device results of reductions (16 doubles)
  double* hres = nullptr;  // pinned host mirror
  CGScalars* dcg = nullptr;  // device scalars of the running CG solve
  CGScalars* hcg = nullptr;  // pinned host mirror
  // tangent_homo's nine simultaneous solves: x, r, p, q per right-hand side
  double* th[4][9] = {};
  CGScalars* dcg9 = nullptr;
  // on the planned path tangent_homo keeps no x: per right-hand side i,
  // hdc[9i..9i+9) receives the voxel sums of K4:p_i from pass C and
  // hm[9i..9i+9) the running sum of alpha * mean(K4:p_i) = mean(K4:x_i)
  double* hdc = nullptr;
  double* hm = nullptr;
  double* hpq = nullptr;  // p_i . (K4:p_i) block sums of the K4 pass [9][blocks]
  double* f2d_cos = nullptr;  // cosine table of cpfft_f2disp
  CGScalars* hcg9 = nullptr;
  // learned warm start
  MlpBuffers mlp;
  bool mlp_ready = false;  // buffers allocated
  int warm = 0;  // warm-start sources: CPFFT_WARM_NETWORK | CPFFT_WARM_PREVIOUS
  long mlp_count = 0;      // pairs held in the log (at most its capacity)
  long mlp_total = 0;      // pairs ever logged; the log is a ring over the last ones
  long mlp_cap = 0;        // ring size (the capacity of the last cpfft_mlp_init)
  long mlp_steps = 0;      // Adam steps taken
  long long ws_updates = 0;
  void* orients = nullptr;
  int32_t* oid = nullptr;
  long long* dcnt = nullptr;
  // pinned host mirrors of device results read back on the context's
  // stream: counters before [0..8) and after [8..16) a stress update, and
  // tangent_homo's accumulated means (81)
  long long* hcnt = nullptr;
  double* hhm = nullptr;
  long long cg_iters = 0, gkdf_calls = 0, des_calls = 0;
  double drv[99] = {};  // driver state: barF_t, P_bar, C_homo
  // CG iteration chunks replayed from HIP graphs (cpfft_set_graphs): one
  // per solution pointer for cpfft_pcg, one for tangent_homo
  int use_graphs = 0;
  int cg_maxit = 1000;  // fftPcg's maxIter (FFT_nr3.f:224)
  std::vector<std::pair<const double*, hipGraphExec_t>> pcg_graphs;
  hipGraphExec_t homo_graph = nullptr;
  // kernel timing: event pairs recorded on the stream without host
  // synchronisation, resolved by cpfft_kernel_times; timing = k > 0 records
  // every k-th launch group of each family (ev_seq counts them), ev_it the
  // CG iteration a recorded group belongs to (ev_cur, -1 outside cpfft_pcg)
  static constexpr int kFam = 4, kRing = 8192;
  int timing = 0;
  long ev_seq[kFam] = {0, 0, 0, 0};
  int ev_cur = -1;
  std::vector<hipEvent_t> ev[kFam][2];
  std::vector<int> ev_it[kFam];
  int ev_n[kFam] = {0, 0, 0, 0};
  double t_ms[kFam] = {0, 0, 0, 0};
  long t_n[kFam] = {0, 0, 0, 0};
  int open_fam = -1;
};

template <int NS>
static int upload_orients(cpfft_ctx* c, const std::vector<std::array<double, 3>>& uniq,
                          const double stiff[6][6], const double b[kMaxSlip][3],
                          const double n[kMaxSlip][3]) {
  std::vector<Orient<NS>> h(uniq.size());
  for (size_t k = 0; k < uniq.size(); ++k) build_orient<NS>(stiff, b, n, uniq[k].data(), h[k]);
  HIP_OK(hipMalloc(&c->orients, h.size() * sizeof(Orient<NS>)));
  HIP_OK(hipMemcpy(c->orients, h.data(), h.size() * sizeof(Orient<NS>), hipMemcpyHostToDevice));
  return 0;
}

static void drop_graphs(cpfft_ctx* c);
static void enqueue_homo_chunk(cpfft_ctx* c);
static int prepare_graphs(cpfft_ctx* c);

extern "C" const char* cpfft_last_error(void) { return g_err.c_str(); }

extern "C" int cpfft_destroy(cpfft_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int f = 0; f < CPFFT_NFIELDS; ++f)
    if (c->fields[f]) (void)hipFree(c->fields[f]);
  void* bufs[] = {c->spec, c->tw, c->r, c->p, c->q, c->tmp9, c->tan6, c->partials, c->dpart, c->rpart, c->dres, c->orients, c->oid, c->dcnt, c->dcg};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->hres) (void)hipHostFree(c->hres);
  if (c->hcg) (void)hipHostFree(c->hcg);
  if (c->hcg9) (void)hipHostFree(c->hcg9);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  if (c->hhm) (void)hipHostFree(c->hhm);
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 9; ++i)
      if (c->th[a][i]) (void)hipFree(c->th[a][i]);
  if (c->dcg9) (void)hipFree(c->dcg9);
  if (c->hdc) (void)hipFree(c->hdc);
  if (c->hm) (void)hipFree(c->hm);
  if (c->hpq) (void)hipFree(c->hpq);
  if (c->f2d_cos) (void)hipFree(c->f2d_cos);
  {
    MlpBuffers& b = c->mlp;
    void* mb[] = {b.params, b.pbf, b.grads, b.m, b.v, b.norm, b.logX, b.logT, b.Xb, b.Xf, b.Tb,
                  b.H1, b.H2, b.Y, b.dY, b.dZ1, b.dZ2, b.H1b, b.H2b, b.dYb, b.dZ1b, b.dZ2b,
                  b.part, b.Xv, b.ws};
    for (void* q : mb)
      if (q) (void)hipFree(q);
  }
  for (int f = 0; f < cpfft_ctx::kFam; ++f)
    for (int k = 0; k < 2; ++k)
      for (hipEvent_t e : c->ev[f][k]) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int cpfft_reset(cpfft_ctx* c);

extern "C" int cpfft_create(int N, const cpfft_material* m, const double* angles, double dt,
                            int gamma_mode, cpfft_ctx** out) {
  *out = nullptr;
  if (N < 2 || N > 256) return fail(-1, "cpfft_create: N must be in [2, 256]");
  if (!m || !angles) return fail(-1, "cpfft_create: null material or angles");
  if (slip_count(m->slip_set) == 0)
    return fail(-1, "cpfft_create: slip_set must be 1 (fcc), 8 (bcc48) or 80 (bcc24)");
  if (gamma_mode != 0 && gamma_mode != 1) return fail(-1, "cpfft_create: bad gamma_mode");
  if (!(dt > 0.0)) return fail(-1, "cpfft_create: dt must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(-2, "cpfft_create: no HIP device available (this library has no CPU path)");
  // set-up of cpfft_gkdf_noalloc's scratch on this device, so that a later
  // call there (perhaps under graph capture) only launches kernels
  if (int rc = noalloc_scratch_current_device(nullptr)) return rc;
  cpfft_ctx* c = new cpfft_ctx();
  HIP_OK(hipGetDevice(&c->dev));
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->N = N;
  c->Nz = N / 2 + 1;
  c->V = (long)N * N * N;
  const long V = c->V;
  // geometry
  int maxlds = 65536;
  (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->dev);
  if (const char* why = make_spec_geom(N, gamma_mode, maxlds, c->geom)) {
    cpfft_destroy(c);
    return fail(-3, (std::string("cpfft_create: ") + why).c_str());
  }
  // material
  double b[kMaxSlip][3], nn[kMaxSlip][3];
  slip_systems(m->slip_set, c->nslip, b, nn);
  double st[6][6];
  if (!cubic_stiffness(*m, st)) {
    cpfft_destroy(c);
    return fail(-1, "cpfft_create: singular compliance");
  }
  c->mp = make_params(*m, dt);
  // orientations: deduplicate the per-voxel angles
  std::map<std::tuple<double, double, double>, int32_t> ids;
  std::vector<std::array<double, 3>> uniq;
  std::vector<int32_t> hoid(V);
  for (long e = 0; e < V; ++e) {
    auto key = std::make_tuple(angles[e], angles[V + e], angles[2 * V + e]);
    auto it = ids.find(key);
    if (it == ids.end()) {
      int32_t id = (int32_t)uniq.size();
      ids.emplace(key, id);
      uniq.push_back({angles[e], angles[V + e], angles[2 * V + e]});
      hoid[e] = id;
    } else {
      hoid[e] = it->second;
    }
  }
  int rc = c->nslip == 12   ? upload_orients<12>(c, uniq, st, b, nn)
           : c->nslip == 24 ? upload_orients<24>(c, uniq, st, b, nn)
                            : upload_orients<48>(c, uniq, st, b, nn);
  if (rc) { cpfft_destroy(c); return rc; }
  HIP_OK(hipMalloc(&c->oid, V * sizeof(int32_t)));
  HIP_OK(hipMemcpy(c->oid, hoid.data(), V * sizeof(int32_t), hipMemcpyHostToDevice));
  // fields
  const int W = CPFFT_NWORDS(c->nslip);
  long lens[CPFFT_NFIELDS] = {9 * V, 9 * V, 9 * V, 9 * V, 9 * V, 9 * V, 81 * V, W * V, W * V, 9 * V,
                              6 * V};
  for (int f = 0; f < CPFFT_NFIELDS; ++f) {
    c->field_len[f] = lens[f];
    HIP_OK(hipMalloc(&c->fields[f], lens[f] * sizeof(double)));
  }
  HIP_OK(hipMalloc(&c->spec, 2 * 9 * c->geom.S1 * sizeof(cplx)));
  HIP_OK(hipMalloc(&c->r, 9 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->p, 9 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->q, 9 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->tmp9, 9 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->tan6, 36 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->partials, reduce_partials_count(9 * V) * sizeof(double)));
  // p.q partial sums: pass D's (gkdf_dot_parts) or pass A's (N*N/2, one per
  // pencil pair)
  HIP_OK(hipMalloc(&c->dpart, std::max({1, gkdf_dot_parts(N), N * N}) * sizeof(double)));
  // r.r partial sums of up to nine systems (tangent_homo's right-hand sides)
  HIP_OK(hipMalloc(&c->rpart, 9 * (size_t)std::max(1, gkdf_dot_parts(N)) * sizeof(double)));
  HIP_OK(hipMalloc(&c->dres, 16 * sizeof(double)));
  HIP_OK(hipHostMalloc(&c->hres, 16 * sizeof(double), hipHostMallocDefault));
  HIP_OK(hipMalloc(&c->dcg, sizeof(CGScalars)));
  HIP_OK(hipHostMalloc(&c->hcg, sizeof(CGScalars), hipHostMallocDefault));
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 9; ++i) HIP_OK(hipMalloc(&c->th[a][i], 9 * V * sizeof(double)));
  HIP_OK(hipMalloc(&c->dcg9, 9 * sizeof(CGScalars)));
  HIP_OK(hipMalloc(&c->hdc, 81 * sizeof(double)));
  HIP_OK(hipMalloc(&c->hm, 81 * sizeof(double)));
  HIP_OK(hipMalloc(&c->hpq, 9L * k4_contract_multi_blocks(V) * sizeof(double)));
  HIP_OK(hipHostMalloc(&c->hcg9, 9 * sizeof(CGScalars), hipHostMallocDefault));
  HIP_OK(hipMalloc(&c->dcnt, 8 * sizeof(long long)));
  HIP_OK(hipMemsetAsync(c->dcnt, 0, 8 * sizeof(long long), c->stream));
  HIP_OK(hipHostMalloc(&c->hcnt, 16 * sizeof(long long), hipHostMallocDefault));
  HIP_OK(hipHostMalloc(&c->hhm, 81 * sizeof(double), hipHostMallocDefault));
  // twiddles exp(-2 pi i k/N), computed in long double
  std::vector<cplx> htw(N);
  const long double pil = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < N; ++k) {
    long double a = 2.0L * pil * (long double)k / (long double)N;
    htw[k] = cplx{(double)cosl(a), (double)-sinl(a)};
  }
  HIP_OK(hipMalloc(&c->tw, N * sizeof(cplx)));
  HIP_OK(hipMemcpyAsync(c->tw, htw.data(), N * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
  rc = cpfft_reset(c);
  if (rc) { cpfft_destroy(c); return rc; }
  *out = c;
  return 0;
}

extern "C" int cpfft_nslip(const cpfft_ctx* c) { return c ? c->nslip : -1; }

extern "C" double* cpfft_field(cpfft_ctx* c, int f) {
  if (!c || f < 0 || f >= CPFFT_NFIELDS) return nullptr;
  return c->fields[f];
}

extern "C" int cpfft_upload(cpfft_ctx* c, int f, const double* h) {
  if (!c || f < 0 || f >= CPFFT_NFIELDS) return fail(-1, "cpfft_upload: bad field");
  HIP_OK(hipMemcpyAsync(c->fields[f], h, c->field_len[f] * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int cpfft_download(cpfft_ctx* c, int f, double* h) {
  if (!c || f < 0 || f >= CPFFT_NFIELDS) return fail(-1, "cpfft_download: bad field");
  HIP_OK(hipMemcpyAsync(h, c->fields[f], c->field_len[f] * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int cpfft_download_voxels(cpfft_ctx* c, int f, long e0, long n, double* h) {
  if (!c || f < 0 || f >= CPFFT_NFIELDS) return fail(-1, "cpfft_download_voxels: bad field");
  if (e0 < 0 || n < 1 || e0 + n > c->V || !h) return fail(-1, "cpfft_download_voxels: bad range");
  const long ncomp = c->field_len[f] / c->V;
  HIP_OK(hipMemcpy2DAsync(h, n * sizeof(double), c->fields[f] + e0, c->V * sizeof(double),
                          n * sizeof(double), ncomp, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int cpfft_sync(cpfft_ctx* c) {
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

static void driver_state_init(cpfft_ctx* c) {
  std::fill(c->drv, c->drv + 99, 0.0);
  c->drv[0] = c->drv[4] = c->drv[8] = 1.0;
}

extern "C" double* cpfft_driver_state(cpfft_ctx* c) { return c ? c->drv : nullptr; }

extern "C" int cpfft_reset(cpfft_ctx* c) {
  if (!c) return fail(-1, "cpfft_reset: null context");
  driver_state_init(c);
  const long V = c->V;
  const int W = CPFFT_NWORDS(c->nslip);
  // device fills on the context's stream: ordered after the work already
  // queued there, and no host copy of the fields
  const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  launch_fill9(V, id, c->fields[CPFFT_FN], c->stream);
  launch_fill9(V, id, c->fields[CPFFT_FN1], c->stream);
  for (int f : {CPFFT_PN, CPFFT_PN1, CPFFT_DFM, CPFFT_B, CPFFT_WORK, CPFFT_K4, CPFFT_EPS})
    launch_fill(c->field_len[f], 0.0, c->fields[f], c->stream);
  // tau_tilde starts at tau_y + 1e-5 (mm10_init_voche, mm10_a.f:2042-2055)
  for (int f : {CPFFT_STATE_N, CPFFT_STATE_NP1})
    launch_state_init(V, W, c->mp.tau_y + 1.0e-5, c->fields[f], c->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

// ------------------------------------------------------------ timing
// Families: 0 G_K_dF with K4, 1 drive_eps_sig, 2 CG vector update,
// 3 G_K_dF without K4.
static void flush_events(cpfft_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  for (int f = 0; f < cpfft_ctx::kFam; ++f) {
    for (int k = 0; k < c->ev_n[f]; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev[f][0][k], c->ev[f][1][k]) == hipSuccess) {
        c->t_ms[f] += ms;
        c->t_n[f] += 1;
      }
    }
    c->ev_n[f] = 0;
  }
}
static void tstart(cpfft_ctx* c, int fam) {
  if (!c->timing) return;
  if (c->ev_seq[fam]++ % c->timing != 0) return;  // not sampled: tstop does nothing
  if (c->ev_n[fam] == cpfft_ctx::kRing) flush_events(c);
  int k = c->ev_n[fam];
  if ((int)c->ev[fam][0].size() <= k) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    c->ev[fam][0].push_back(a);
    c->ev[fam][1].push_back(b);
    c->ev_it[fam].push_back(-1);
  }
  (void)hipEventRecord(c->ev[fam][0][k], c->stream);
  c->ev_it[fam][k] = c->ev_cur;
  c->open_fam = fam;
}
static void tstop(cpfft_ctx* c, int fam) {
  if (!c->timing || c->open_fam != fam) return;
  (void)hipEventRecord(c->ev[fam][1][c->ev_n[fam]], c->stream);
  c->ev_n[fam]++;
  c->open_fam = -1;
}

extern "C" int cpfft_set_timing(cpfft_ctx* c, int on) {
  if (!c) return fail(-1, "cpfft_set_timing: null context");
  if (on < 0) return fail(-1, "cpfft_set_timing: negative sampling stride");
  flush_events(c);
  c->timing = on;
  for (int i = 0; i < cpfft_ctx::kFam; ++i) {
    c->t_ms[i] = 0;
    c->t_n[i] = 0;
    c->ev_seq[i] = 0;
  }
  return 0;
}

// average device time (ms) per launch group of each family since
// cpfft_set_timing, followed by the counts in out[4..7]
extern "C" int cpfft_kernel_times(cpfft_ctx* c, double out[8]) {
  flush_events(c);
  for (int i = 0; i < cpfft_ctx::kFam; ++i) {
    out[i] = c->t_n[i] ? c->t_ms[i] / c->t_n[i] : 0.0;
    out[4 + i] = (double)c->t_n[i];
  }
  return 0;
}

// ------------------------------------------------------------ operators
static int check_ptr(const void* p) { return p ? 0 : fail(-1, "null device pointer"); }

extern "C" int cpfft_gkdf(cpfft_ctx* c, const double* F, double* GKF, int flgK) {
  if (!c || check_ptr(F) || check_ptr(GKF)) return -1;
  if (F == GKF) return fail(-1, "cpfft_gkdf: F and GKF must not alias");
  const int fam = flgK ? 0 : 3;
  tstart(c, fam);
  launch_gkdf(c->geom, F, flgK ? c->fields[CPFFT_K4] : nullptr, GKF, c->spec, c->tw, c->stream);
  tstop(c, fam);
  HIP_OK(hipGetLastError());
  c->gkdf_calls++;
  return 0;
}

static int read_res(cpfft_ctx* c, int n) {
  HIP_OK(hipMemcpyAsync(c->hres, c->dres, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

static int dot_host(cpfft_ctx* c, const double* x, const double* y, double* out) {
  launch_dot(x, y, 9 * c->V, c->partials, c->dres, c->stream);
  int rc = read_res(c, 1);
  if (rc) return rc;
  *out = c->hres[0];
  return 0;
}

extern "C" int cpfft_dnrm2(cpfft_ctx* c, const double* x, double* out) {
  double d;
  int rc = dot_host(c, x, x, &d);
  if (rc) return rc;
  *out = sqrt(d);
  return 0;
}

extern "C" int cpfft_daxpy(cpfft_ctx* c, double a, const double* x, double* y) {
  launch_axpy(9 * c->V, a, x, y, c->stream);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int cpfft_dscal(cpfft_ctx* c, double a, double* x) {
  launch_scal(9 * c->V, a, x, c->stream);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int cpfft_dcopy(cpfft_ctx* c, const double* x, double* y) {
  launch_copy(9 * c->V, x, y, c->stream);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int cpfft_fill9(cpfft_ctx* c, double* x, const double v[9]) {
  launch_fill9(c->V, v, x, c->stream);
  HIP_OK(hipGetLastError());
  return 0;
}
extern "C" int cpfft_mean9(cpfft_ctx* c, const double* x, double out[9]) {
  launch_mean9(x, c->V, c->partials, c->dres, c->stream);
  int rc = read_res(c, 9);
  if (rc) return rc;
  for (int i = 0; i < 9; ++i) out[i] = c->hres[i];
  return 0;
}

// f2disp (f2disp.f:16-175): the spectrum buffers (2 * 9 * S1 complex) hold
// the right-hand side, the transforms and u before the copy to the host
extern "C" int cpfft_f2disp(cpfft_ctx* c, const double* F, const double lengths[3], double* u) {
  if (!c || check_ptr(F)) return -1;
  if (!lengths || !u) return fail(-1, "cpfft_f2disp: null lengths or output");
  for (int a = 0; a < 3; ++a)
    if (!(lengths[a] > 0.0)) return fail(-1, "cpfft_f2disp: box lengths must be positive");
  const int N = c->N;
  const long P = N + 1, P3 = P * P * P;
  const size_t have = 2 * 9 * (size_t)c->geom.S1 * 16;
  if (N < 2 || f2disp_scratch_doubles(N) * sizeof(double) > have)
    return fail(-1, "cpfft_f2disp: grid too small");
  if (!c->f2d_cos) {
    HIP_OK(hipMalloc(&c->f2d_cos, f2disp_cos_doubles(N) * sizeof(double)));
    launch_f2disp_cos(N, c->f2d_cos, c->stream);
  }
  double* work = reinterpret_cast<double*>(c->spec);
  // the last transform leaves x in the first half of the scratch, so u
  // goes to the second
  double* ud = work + 3 * P3;
  launch_f2disp(N, F, lengths, c->f2d_cos, work, ud, c->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(u, ud, 3 * P3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

// CG iterations enqueued per host round trip
static constexpr int kChunk = 8;

// kChunk CG iterations of cpfft_pcg on the solution x: q = A p (pass A also
// forms p = r + beta p, pass D the p.q partial sums on the planned path),
// alpha, x and r updates, convergence test.  Every kernel reads the scalars
// in c->dcg and returns at once after convergence, so the chunk has no
// host-side control flow and can be replayed from a graph.
// base: the solve's iteration number of the chunk's first iteration (only
// labels timing events)
static void enqueue_pcg_chunk(cpfft_ctx* c, double* x, int base = 0) {
  const long n = 9 * c->V;
  SpecGeom g = c->geom;
  g.skip = &c->dcg->done;
  const bool fused = gkdf_pencil_supported(c->N);
  const int ndp = gkdf_dot_parts(c->N);
  // The textbook operator is a symmetric projection and p lies in its
  // range, so p.q = p.(K4:p): pass A forms it, alpha follows pass A, and
  // pass D applies x += alpha p, r -= alpha q without storing q.  The
  // reference's even-N operator is not a projection: there pass D forms p.q
  // and a separate kernel updates x and r.
  const bool apq = fused && c->geom.gamma_mode == 1;
  CGFuse fz{c->p, c->r, c->dcg, c->p, c->dpart, nullptr, nullptr, 0, nullptr, nullptr, nullptr};
  if (apq) {
    fz.dot_p = nullptr;
    fz.dot_part = nullptr;
    fz.pq_part = c->dpart;
    fz.alpha_np = c->N * c->N / 2;
    fz.upd_r = c->r;
    fz.upd_x = x;
    fz.rr_part = c->rpart;
  }
  for (int j = 0; j < kChunk; ++j) {
    c->ev_cur = base + j;
    tstart(c, 0);
    if (fused) {
      launch_gkdf_pencil(g, c->p, c->fields[CPFFT_K4], c->q, c->spec, c->tw, c->stream, &fz);
    } else {
      launch_gkdf(g, c->p, c->fields[CPFFT_K4], c->q, c->spec, c->tw, c->stream);
    }
    tstop(c, 0);
    tstart(c, 2);
    if (apq) {
      launch_cg_check(c->rpart, ndp, c->dcg, c->stream);
    } else if (fused) {
      launch_cg_alpha(c->dpart, ndp, c->dcg, c->stream);
      launch_cg_update_check(n, c->dcg, c->p, c->q, x, c->r, c->partials, c->stream);
    } else {
      launch_cg_tail(n, c->dcg, c->p, c->q, x, c->r, c->partials, c->stream);
    }
    tstop(c, 2);
    if (!fused) launch_cg_dir_dev(n, c->dcg, c->r, c->p, c->stream);
  }
  c->ev_cur = -1;
}

// Record what enqueue() launches on the context's stream into an
// executable graph.  Everything the runtime does for a graph happens here,
// under one process-wide lock, before the graph is first launched:
// - the stream must not be capturing already;
// - kernel timing is off during the capture (a recorded event would become
//   a node, and a full event ring synchronises the stream, which ends the
//   capture);
// - capture in thread-local mode; the graph must consist of kernel nodes
//   only, as many as the first capture of the same chunk (kind, edge,
//   operator and launch shapes) held, so that a capture that took in a
//   foreign launch or lost one of its own is an error, not a wrong result;
// - instantiation, then the upload of the executable graph (its kernel
//   arguments included) completed by a stream synchronisation.
static std::mutex g_capture_mu;
static std::map<std::array<int, 7>, size_t> g_chunk_nodes;

template <class F>
static int capture_chunk(cpfft_ctx* c, int kind, F&& enqueue, hipGraphExec_t* out) {
  std::lock_guard<std::mutex> lk(g_capture_mu);
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  HIP_OK(hipStreamIsCapturing(c->stream, &cst));
  if (cst != hipStreamCaptureStatusNone) return fail(-101, "graph capture: stream already capturing");
  const int timing = c->timing;
  c->timing = 0;
  hipGraph_t gr = nullptr;
  hipError_t be = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (be == hipSuccess) enqueue();
  const hipError_t le = hipGetLastError();
  const hipError_t ee = be == hipSuccess ? hipStreamEndCapture(c->stream, &gr) : be;
  c->timing = timing;
  HIP_OK(be);
  HIP_OK(le);
  HIP_OK(ee);
  size_t nn = 0;
  hipError_t ne = hipGraphGetNodes(gr, nullptr, &nn);
  std::vector<hipGraphNode_t> nodes(nn);
  if (ne == hipSuccess && nn) ne = hipGraphGetNodes(gr, nodes.data(), &nn);
  bool kernels_only = ne == hipSuccess;
  for (size_t i = 0; i < nn && kernels_only; ++i) {
    hipGraphNodeType t;
    kernels_only = hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t == hipGraphNodeTypeKernel;
  }
  const PencilTune& tu = c->geom.tune;
  const std::array<int, 7> key{kind, c->N, c->geom.gamma_mode, tu.a, tu.a0, tu.b, tu.c};
  auto it = g_chunk_nodes.find(key);
  const bool count_ok = it == g_chunk_nodes.end() || it->second == nn;
  if (!kernels_only || !count_ok || nn == 0) {
    (void)hipGraphDestroy(gr);
    return fail(-101, "graph capture: the captured graph is not the chunk's kernel sequence");
  }
  if (it == g_chunk_nodes.end()) g_chunk_nodes.emplace(key, nn);
  const hipError_t ie = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  HIP_OK(ie);
  HIP_OK(hipGraphUpload(*out, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

static void drop_graphs(cpfft_ctx* c) {
  for (auto& pg : c->pcg_graphs) (void)hipGraphExecDestroy(pg.second);
  c->pcg_graphs.clear();
  if (c->homo_graph) (void)hipGraphExecDestroy(c->homo_graph);
  c->homo_graph = nullptr;
}

// ---------------------------------------------------------- autotuner
// Shapes chosen per (device, N), shared by every context of the process.
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, PencilTune> g_tune;

static void apply_tune(cpfft_ctx* c, const PencilTune& t) {
  const bool changed = t.a != c->geom.tune.a || t.a0 != c->geom.tune.a0 || t.b != c->geom.tune.b ||
                       t.c != c->geom.tune.c;
  c->geom.tune = t;
  if (changed) {  // captured launches carry the old shapes
    drop_graphs(c);
    (void)prepare_graphs(c);
  }
}

extern "C" int cpfft_autotune(cpfft_ctx* c, int reps, int force) {
  if (!c) return fail(-1, "cpfft_autotune: null context");
  if (!gkdf_pencil_supported(c->N)) return 1;
  if (reps < 1) reps = 10;
  const std::pair<int, int> key(c->dev, c->N);
  if (!force) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      apply_tune(c, it->second);
      return 0;
    }
  }
  PencilTune t;
  int* knob[4] = {&t.a, &t.a0, &t.b, &t.c};
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  const double* F = c->fields[CPFFT_DFM];
  const double* K4 = c->fields[CPFFT_K4];
  // passes in pipeline order, so each times its variants on the spectrum
  // the previous pass wrote
  for (int pass = 0; pass < 4; ++pass) {
    float best = 0.f;
    int bi = 0;
    for (int v = 0; v < kPencilCands; ++v) {
      if (!pencil_candidate(c->N, pass, v)) continue;
      launch_pencil_pass(c->geom, pass, v, F, K4, c->spec, c->tw, c->stream);  // warm-up
      HIP_OK(hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r)
        launch_pencil_pass(c->geom, pass, v, F, K4, c->spec, c->tw, c->stream);
      HIP_OK(hipEventRecord(e1, c->stream));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      if (v == 0 || ms < best) {
        best = ms;
        bi = v;
      }
    }
    *knob[pass] = bi;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_OK(hipGetLastError());
  apply_tune(c, t);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = t;
  return 0;
}

extern "C" int cpfft_get_tune(cpfft_ctx* c, int out[4]) {
  if (!c) return fail(-1, "cpfft_get_tune: null context");
  const PencilTune& t = c->geom.tune;
  const int idx[4] = {t.a, t.a0, t.b, t.c};
  for (int p = 0; p < 4; ++p) out[p] = pencil_candidate(c->N, p, idx[p]);
  return 0;
}

extern "C" int cpfft_set_tune(cpfft_ctx* c, const int in[4]) {
  if (!c) return fail(-1, "cpfft_set_tune: null context");
  if (!gkdf_pencil_supported(c->N)) return fail(-1, "cpfft_set_tune: no tunable shapes for this N");
  PencilTune t;
  int* knob[4] = {&t.a, &t.a0, &t.b, &t.c};
  for (int p = 0; p < 4; ++p) {
    int found = -1;
    for (int v = 0; v < kPencilCands && found < 0; ++v)
      if (pencil_candidate(c->N, p, v) == in[p] && in[p] > 0) found = v;
    if (found < 0) return fail(-1, "cpfft_set_tune: shape is not among the candidates for this N");
    *knob[p] = found;
  }
  apply_tune(c, t);
  return 0;
}

// the chunk graphs the load-step driver uses (cpfft_pcg on the DFM field,
// tangent_homo), captured now rather than on first use, so that a batch of
// contexts run from several threads only launches graphs
static int prepare_graphs(cpfft_ctx* c) {
  if (c->use_graphs != 1) return 0;
  double* x = c->fields[CPFFT_DFM];
  bool have = false;
  for (auto& pg : c->pcg_graphs) have |= pg.first == x;
  int rc;
  if (!have) {
    hipGraphExec_t ge = nullptr;
    if ((rc = capture_chunk(c, 0, [&] { enqueue_pcg_chunk(c, x); }, &ge))) return rc;
    c->pcg_graphs.emplace_back(x, ge);
  }
  if (!c->homo_graph &&
      (rc = capture_chunk(c, 1, [&] { enqueue_homo_chunk(c); }, &c->homo_graph)))
    return rc;
  return 0;
}

extern "C" int cpfft_set_cg_maxit(cpfft_ctx* c, int maxit) {
  if (!c) return fail(-1, "cpfft_set_cg_maxit: null context");
  if (maxit < 1) return fail(-1, "cpfft_set_cg_maxit: the limit must be positive");
  c->cg_maxit = maxit;
  return 0;
}

extern "C" int cpfft_set_graphs(cpfft_ctx* c, int on) {
  if (!c) return fail(-1, "cpfft_set_graphs: null context");
  if (on < 0 || on > 2) return fail(-1, "cpfft_set_graphs: mode must be 0, 1 or 2");
  c->use_graphs = on;
  return prepare_graphs(c);
}

// fftPcg (FFT_nr3.f:214-360): textbook unpreconditioned CG, x0 = 0
extern "C" int cpfft_pcg(cpfft_ctx* c, const double* b, double* x, double tol, int* iters,
                         double* relres) {
  if (!c) return fail(-1, "cpfft_pcg: null context");
  const long n = 9 * c->V;
  const int maxIter = c->cg_maxit;
  const double eps = 2.220446049250313e-16;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  launch_fill(n, 0.0, x, c->stream);
  double bb;
  int rc = dot_host(c, b, b, &bb);
  if (rc) return rc;
  double n2b = sqrt(bb), tolb = tol * n2b;
  if (tol <= eps || tol >= 1.0) return fail(-1, "cpfft_pcg: improper tolerance");
  if (n2b <= eps) return 1;
  launch_copy(n, b, c->r, c->stream);
  launch_copy(n, b, c->p, c->stream);
  // Iterations run on the device with their scalars in c->dcg; the host
  // enqueues them in chunks and reads the state once per chunk.  Kernels of
  // iterations enqueued after convergence return at once.
  CGScalars h{};
  h.rr = bb;
  h.tolb = tolb;
  h.tol = tol;
  h.resnorm = n2b;
  h.k = 0;
  h.maxit = maxIter;
  h.done = 0;
  *c->hcg = h;
  HIP_OK(hipMemcpyAsync(c->dcg, c->hcg, sizeof(CGScalars), hipMemcpyHostToDevice, c->stream));
  const bool graph = c->use_graphs && !c->timing;
  hipGraphExec_t ge = nullptr;
  if (graph) {
    for (auto& pg : c->pcg_graphs)
      if (pg.first == x) ge = pg.second;
    if (!ge) {
      if ((rc = capture_chunk(c, 0, [&] { enqueue_pcg_chunk(c, x); }, &ge))) return rc;
      c->pcg_graphs.emplace_back(x, ge);
    }
  }
  int launched = 0;
  for (;;) {
    if (graph) {
      HIP_OK(hipGraphLaunch(ge, c->stream));
    } else {
      if (c->timing && (c->ev_n[0] + kChunk > cpfft_ctx::kRing || c->ev_n[2] + kChunk > cpfft_ctx::kRing))
        flush_events(c);
      enqueue_pcg_chunk(c, x, launched);
    }
    launched += kChunk;
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(c->hcg, c->dcg, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (c->hcg->done) break;
  }
  const int k = c->hcg->k;
  if (c->timing)  // drop the events of the no-op iterations at the end
    for (int f : {0, 2})
      while (c->ev_n[f] > 0 && c->ev_it[f][c->ev_n[f] - 1] >= k) c->ev_n[f]--;
  c->cg_iters += k;
  c->gkdf_calls += k;
  if (iters) *iters = k;
  if (relres) *relres = c->hcg->resnorm / n2b;
  if (c->hcg->done == 2) return fail(-2, "cpfft_pcg: no convergence within the iteration limit");
  return 0;
}

// predicted slip increments of every voxel into c->mlp.ws: the normalised
// inputs of all voxels, then one inference launch
template <int NS>
static void mlp_predict(cpfft_ctx* c, int step) {
  MlpBuffers& b = c->mlp;
  mlp_weights_to_bf16(b, c->stream);
  launch_mlp_features_norm<NS>(c->V, c->mp, (const Orient<NS>*)c->orients, c->oid, step,
                               c->fields[CPFFT_FN], c->fields[CPFFT_FN1],
                               c->fields[CPFFT_STATE_N], b.norm, mlp_nin_pad(b.d.nin), b.Xv,
                               c->stream);
  mlp_infer(b, c->V, c->stream);
}

extern "C" int cpfft_drive_eps_sig(cpfft_ctx* c, int step, int iiter) {
  if (!c) return fail(-1, "cpfft_drive_eps_sig: null context");
  long long* before = c->hcnt;
  long long* after = c->hcnt + 8;
  HIP_OK(hipMemcpyAsync(before, c->dcnt, 8 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  tstart(c, 1);
  const float* ws = nullptr;
  const int wp = (c->warm & CPFFT_WARM_PREVIOUS) && iiter > 0;
  // with the previous iterate on, the network serves the first Newton
  // iteration of a mixed-BC loop only, where the previous iterate is the
  // explicit update of iiter 0; later iterates are better starts than a
  // prediction
  const bool net = (c->warm & CPFFT_WARM_NETWORK) && iiter > 0 && (!wp || iiter == 1);
  if (net && c->nslip <= kMlpMaxOut) {
    if (c->nslip == 12)
      mlp_predict<12>(c, step);
    else
      mlp_predict<24>(c, step);
    ws = c->mlp.ws;
  }
  if (ws || wp) c->ws_updates += c->V;
  if (c->nslip == 12)
    launch_drive_eps_sig<12>(c->V, c->mp, (const Orient<12>*)c->orients, c->oid, step, iiter,
                             c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                             c->fields[CPFFT_STATE_NP1], c->fields[CPFFT_PN1], c->tan6,
                             c->fields[CPFFT_K4], ws, wp, c->dcnt, c->stream);
  else if (c->nslip == 24)
    launch_drive_eps_sig<24>(c->V, c->mp, (const Orient<24>*)c->orients, c->oid, step, iiter,
                             c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                             c->fields[CPFFT_STATE_NP1], c->fields[CPFFT_PN1], c->tan6,
                             c->fields[CPFFT_K4], ws, wp, c->dcnt, c->stream);
  else
    launch_drive_eps_sig<48>(c->V, c->mp, (const Orient<48>*)c->orients, c->oid, step, iiter,
                             c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                             c->fields[CPFFT_STATE_NP1], c->fields[CPFFT_PN1], c->tan6,
                             c->fields[CPFFT_K4], nullptr, wp, c->dcnt, c->stream);
  tstop(c, 1);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(after, c->dcnt, 8 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->des_calls++;
  return (int)(after[3] - before[3]);
}

// tangent_homo (tangent_homo.f:11-73).  Column i of C_homo is
// mean_e K4(e) : (E_i + dF_i(e)) with dF_i the CG solution of
// G(K4 : dF_i) = -G(K4 : E_i).  The nine solves run together: one pass over
// K4 forms K4 : p_i for every solve still running, then each solve applies
// the projection and its own CG update.  Each solve's arithmetic is that of
// fftPcg on its own, so iterates and iteration counts are unchanged.
// kChunk iterations of tangent_homo's nine CG solves (scalars in c->dcg9):
// one K4 pass forms K4 : p_i for every running solve (and on the planned
// path first p_i = r_i + beta_i p_i), then each solve applies the operator
// without K4 in place and its own alpha, update and test.  Kernels of
// finished solves return at once, so the chunk can be replayed from a graph.
static void enqueue_homo_chunk(cpfft_ctx* c) {
  const long V = c->V, n = 9 * V;
  double* const* X = c->th[0];
  double* const* R = c->th[1];
  double* const* Pd = c->th[2];
  double* const* Q = c->th[3];
  const double* K4 = c->fields[CPFFT_K4];
  Ptr9 pp;
  Ptr9r rr;
  for (int i = 0; i < 9; ++i) {
    pp.p[i] = Pd[i];
    pp.q[i] = Q[i];
    rr.r[i] = R[i];
  }
  const bool fused = gkdf_pencil_supported(c->N);
  const int ndp = gkdf_dot_parts(c->N);
  // textbook operator: p_i . q_i = p_i . (K4:p_i) from the K4 pass, alpha
  // before the operator, pass D updates r_i without storing q_i (see
  // enqueue_pcg_chunk); the reference's even-N operator: p.q from pass D
  const bool apq = fused && c->geom.gamma_mode == 1;
  const int nb = k4_contract_multi_blocks(V);
  for (int j = 0; j < kChunk; ++j) {
    launch_k4_contract_multi(V, K4, pp, c->dcg9, 9, c->stream, fused ? &rr : nullptr,
                             apq ? c->hpq : nullptr);
    // the nine alphas in one launch (one block per right-hand side)
    if (apq) launch_cg_alpha(c->hpq, nb, c->dcg9, c->stream, 9, nb);
    for (int i = 0; i < 9; ++i) {
      SpecGeom g = c->geom;
      g.skip = &c->dcg9[i].done;
      // in place: pass A reads all of Q[i] before pass D writes it
      if (fused) {
        // pass C: voxel sums of K4:p_i, from which the check accumulates
        // mean(K4:x_i); x_i itself is never formed
        double* const dc = c->hdc + 9 * i;
        double* const m = c->hm + 9 * i;
        const double sc = 1.0 / (double)V;
        if (apq) {
          const CGFuse fz{nullptr, nullptr, c->dcg9 + i, nullptr, nullptr, dc, nullptr, 0,
                          R[i], nullptr, c->rpart + (long)i * ndp};
          launch_gkdf_pencil(g, Q[i], nullptr, Q[i], c->spec, c->tw, c->stream, &fz);
        } else {
          const CGFuse fz{nullptr, nullptr, c->dcg9 + i, Pd[i], c->dpart, dc, nullptr, 0,
                          nullptr, nullptr, nullptr};
          launch_gkdf_pencil(g, Q[i], nullptr, Q[i], c->spec, c->tw, c->stream, &fz);
          launch_cg_alpha(c->dpart, ndp, c->dcg9 + i, c->stream);
          launch_cg_update_check(n, c->dcg9 + i, Pd[i], Q[i], X[i], R[i], c->partials, c->stream,
                                 false, dc, m, sc);
        }
      } else {
        launch_gkdf(g, Q[i], nullptr, Q[i], c->spec, c->tw, c->stream);
        launch_cg_tail(n, c->dcg9 + i, Pd[i], Q[i], X[i], R[i], c->partials, c->stream);
        launch_cg_dir_dev(n, c->dcg9 + i, R[i], Pd[i], c->stream);
      }
    }
    // the nine convergence tests in one launch, after all nine operators
    if (apq)
      launch_cg_check(c->rpart, ndp, c->dcg9, c->stream, c->hdc, c->hm, 1.0 / (double)V, 9, ndp);
  }
}

extern "C" int cpfft_tangent_homo(cpfft_ctx* c, double tol, double C_homo[81]) {
  if (!c) return fail(-1, "cpfft_tangent_homo: null context");
  const long V = c->V, n = 9 * V;
  const int maxIter = c->cg_maxit;
  const double eps = 2.220446049250313e-16;
  if (tol <= eps || tol >= 1.0) return fail(-1, "cpfft_tangent_homo: improper tolerance");
  const bool fused = gkdf_pencil_supported(c->N);
  double* const* X = c->th[0];
  double* const* R = c->th[1];
  double* const* Pd = c->th[2];
  double* const* Q = c->th[3];
  const double* K4 = c->fields[CPFFT_K4];
  // right-hand sides b_i = -G(K4 : E_i) in R[i]
  for (int i = 0; i < 9; ++i) {
    launch_gather_k4_col(V, i, K4, Q[i], c->stream);
    tstart(c, 3);
    launch_gkdf(c->geom, Q[i], nullptr, R[i], c->spec, c->tw, c->stream);
    tstop(c, 3);
    c->gkdf_calls++;
    launch_scal(n, -1.0, R[i], c->stream);
    launch_dot(R[i], R[i], n, c->partials, c->dres + i, c->stream);
  }
  int rc = read_res(c, 9);
  if (rc) return rc;
  double n2b[9];
  for (int i = 0; i < 9; ++i) {
    const double bb = c->hres[i];
    n2b[i] = sqrt(bb);
    CGScalars h{};
    h.rr = bb;
    h.tolb = tol * n2b[i];
    h.tol = tol;
    h.resnorm = n2b[i];
    h.maxit = maxIter;
    h.done = (n2b[i] <= eps) ? 1 : 0;  // zero right-hand side: dF_i = 0
    c->hcg9[i] = h;
    if (!fused) launch_fill(n, 0.0, X[i], c->stream);
    launch_copy(n, R[i], Pd[i], c->stream);
  }
  HIP_OK(hipMemcpyAsync(c->dcg9, c->hcg9, 9 * sizeof(CGScalars), hipMemcpyHostToDevice, c->stream));
  if (fused) launch_fill(81, 0.0, c->hm, c->stream);
  const bool graph = c->use_graphs && !c->timing;
  if (graph && !c->homo_graph &&
      (rc = capture_chunk(c, 1, [&] { enqueue_homo_chunk(c); }, &c->homo_graph)))
    return rc;
  for (;;) {
    if (graph) {
      HIP_OK(hipGraphLaunch(c->homo_graph, c->stream));
    } else {
      enqueue_homo_chunk(c);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(c->hcg9, c->dcg9, 9 * sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    bool all = true;
    for (int i = 0; i < 9; ++i) all &= c->hcg9[i].done != 0;
    if (all) break;
  }
  for (int i = 0; i < 9; ++i) {
    c->cg_iters += c->hcg9[i].k;
    c->gkdf_calls += c->hcg9[i].k;
    if (c->hcg9[i].done == 2) return fail(-2, "cpfft_tangent_homo: CG did not converge within the iteration limit");
  }
  if (fused) {
    // C_homo(9(j-1)+i) = mean(K4 : E_i)_j + mean(K4 : x_i)_j, the first
    // term the mean of K4 component 9j+i, the second accumulated above
    double* hm = c->hhm;
    HIP_OK(hipMemcpyAsync(hm, c->hm, 81 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    for (int j = 0; j < 9; ++j) {
      launch_mean9(K4 + (long)9 * j * V, V, c->partials, c->dres, c->stream);
      if ((rc = read_res(c, 9))) return rc;
      for (int i = 0; i < 9; ++i) C_homo[9 * j + i] = c->hres[i] + hm[9 * i + j];
    }
    return 0;
  }
  for (int i = 0; i < 9; ++i) {
    launch_add_identity_comp(V, i, X[i], c->stream);
    launch_k4_dot9(V, K4, X[i], c->partials, c->dres, c->stream);
    if ((rc = read_res(c, 9))) return rc;
    for (int j = 0; j < 9; ++j) C_homo[9 * j + i] = c->hres[j];
  }
  return 0;
}

extern "C" int cpfft_update(cpfft_ctx* c) {
  if (!c) return fail(-1, "cpfft_update: null context");
  std::swap(c->fields[CPFFT_STATE_N], c->fields[CPFFT_STATE_NP1]);
  launch_copy(9 * c->V, c->fields[CPFFT_FN1], c->fields[CPFFT_FN], c->stream);
  launch_copy(9 * c->V, c->fields[CPFFT_PN1], c->fields[CPFFT_PN], c->stream);
  // eps_n1 = eps_n + uddt of the step's last update (rstgp1.f:1388-1393),
  // then eps_n <- eps_n1 (update.f:85); the words CPFFT_W_D..+5 are
  // contiguous [6][V]
  launch_axpy(6 * c->V, 1.0, c->fields[CPFFT_STATE_N] + (long)CPFFT_W_D * c->V,
              c->fields[CPFFT_EPS], c->stream);
  HIP_OK(hipGetLastError());
  return 0;
}

extern "C" int cpfft_counters(cpfft_ctx* c, long long out[8]) {
  if (!c) return fail(-1, "cpfft_counters: null context");
  long long* d = c->hcnt;
  HIP_OK(hipMemcpyAsync(d, c->dcnt, 8 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  out[0] = c->gkdf_calls;
  out[1] = c->cg_iters;
  out[2] = c->des_calls;
  out[3] = d[0];
  out[4] = d[1];
  out[5] = d[2];
  out[6] = d[3];
  out[7] = d[4];
  return 0;
}

// ------------------------------------------------------ learned warm start
extern "C" int cpfft_mlp_init(cpfft_ctx* c, uint64_t seed, long capacity, int batch) {
  if (!c) return fail(-1, "cpfft_mlp_init: null context");
  if (c->nslip > kMlpMaxOut)
    return fail(-1, "cpfft_mlp_init: the network serves crystals of 12 or 24 slip systems");
  if (capacity < 1 || batch < 16) return fail(-1, "cpfft_mlp_init: capacity >= 1 and batch >= 16 required");
  MlpBuffers& b = c->mlp;
  if (!c->mlp_ready) {
    const int rows = std::max(batch, 4096);
    b.d = mlp_dims(c->nslip);
    const MlpDims& d = b.d;
    b.rows = rows;
    b.cap = capacity;
    HIP_OK(hipMalloc(&b.params, d.nparam * sizeof(float)));
    HIP_OK(hipMalloc(&b.pbf, d.nparam * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&b.grads, d.nparam * sizeof(float)));
    HIP_OK(hipMalloc(&b.m, d.nparam * sizeof(float)));
    HIP_OK(hipMalloc(&b.v, d.nparam * sizeof(float)));
    HIP_OK(hipMalloc(&b.norm, d.nnorm * sizeof(float)));
    HIP_OK(hipMalloc(&b.logX, capacity * d.nin * sizeof(float)));
    HIP_OK(hipMalloc(&b.logT, capacity * d.nout * sizeof(float)));
    HIP_OK(hipMalloc(&b.Xb, (long)rows * d.nin * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&b.Xf, (long)rows * d.nin * sizeof(float)));
    HIP_OK(hipMalloc(&b.Tb, (long)rows * d.nout * sizeof(float)));
    float** f32[] = {&b.H1, &b.H2, &b.dZ1, &b.dZ2};
    for (float** q : f32) HIP_OK(hipMalloc(q, (long)rows * kMlpHid * sizeof(float)));
    HIP_OK(hipMalloc(&b.Y, (long)rows * d.nout * sizeof(float)));
    HIP_OK(hipMalloc(&b.dY, (long)rows * d.nout * sizeof(float)));
    uint16_t** b16[] = {&b.H1b, &b.H2b, &b.dZ1b, &b.dZ2b};
    for (uint16_t** q : b16) HIP_OK(hipMalloc(q, (long)rows * kMlpHid * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&b.dYb, (long)rows * d.nout * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&b.part, (long)mlp_loss_parts(b, rows) * sizeof(double)));
    HIP_OK(hipMalloc(&b.Xv, c->V * mlp_nin_pad(d.nin) * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&b.ws, (long)d.nout * c->V * sizeof(float)));
    c->mlp_ready = true;
  } else if (batch > b.rows || capacity > b.cap) {
    return fail(-1, "cpfft_mlp_init: a larger batch or log than the first call allocated");
  }
  const MlpDims& d = b.d;
  // Glorot-uniform weights, zero biases, identity normalisation
  std::vector<float> h(d.nparam, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](long off, int fan_out, int fan_in) {
    const double a = sqrt(6.0 / (double)(fan_in + fan_out));
    std::uniform_real_distribution<double> U(-a, a);
    for (long i = 0; i < (long)fan_out * fan_in; ++i) h[off + i] = (float)U(rng);
  };
  fill(d.offW1, kMlpHid, d.nin);
  fill(d.offW2, kMlpHid, kMlpHid);
  fill(d.offW3, d.nout, kMlpHid);
  HIP_OK(hipMemcpyAsync(b.params, h.data(), d.nparam * sizeof(float), hipMemcpyHostToDevice, c->stream));
  launch_fill_f32(d.nparam, 0.0f, b.m, c->stream);
  launch_fill_f32(d.nparam, 0.0f, b.v, c->stream);
  std::vector<float> nrm(d.nnorm, 0.0f);
  for (int i = 0; i < d.nin; ++i) nrm[d.nin + i] = 1.0f;
  for (int i = 0; i < d.nout; ++i) nrm[2 * d.nin + d.nout + i] = 1.0f;
  HIP_OK(hipMemcpyAsync(b.norm, nrm.data(), d.nnorm * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->mlp_count = 0;
  c->mlp_total = 0;
  c->mlp_cap = capacity;
  c->mlp_steps = 0;
  c->warm &= ~CPFFT_WARM_NETWORK;
  return 0;
}

extern "C" int cpfft_mlp_log(cpfft_ctx* c, int step, long stride) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_log: cpfft_mlp_init first");
  if (stride < 1) return fail(-1, "cpfft_mlp_log: stride must be >= 1");
  MlpBuffers& b = c->mlp;
  // every stride-th voxel (at most one log's worth), written at the ring
  // position after the last pair logged, wrapping over the oldest pairs
  const long cap = c->mlp_cap;
  const long rows = std::min((c->V + stride - 1) / stride, cap);
  long pos = c->mlp_total % cap, done = 0;
  while (done < rows) {
    const long n = std::min(rows - done, cap - pos);
    if (c->nslip == 12)
      launch_mlp_features<12>(c->V, c->mp, (const Orient<12>*)c->orients, c->oid, step,
                              c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                              c->fields[CPFFT_STATE_NP1], done * stride, stride, n,
                              b.logX + pos * b.d.nin, b.logT + pos * b.d.nout, c->stream);
    else
      launch_mlp_features<24>(c->V, c->mp, (const Orient<24>*)c->orients, c->oid, step,
                              c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                              c->fields[CPFFT_STATE_NP1], done * stride, stride, n,
                              b.logX + pos * b.d.nin, b.logT + pos * b.d.nout, c->stream);
    done += n;
    pos = (pos + n) % cap;
  }
  HIP_OK(hipGetLastError());
  c->mlp_total += rows;
  c->mlp_count = std::min(c->mlp_total, cap);
  return 0;
}

extern "C" int cpfft_mlp_train(cpfft_ctx* c, int iters, double lr, uint64_t seed, double* loss) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_train: cpfft_mlp_init first");
  if (c->mlp_count < 1) return fail(-1, "cpfft_mlp_train: no logged pairs");
  MlpBuffers& b = c->mlp;
  const int rows = std::min<long>(4096, b.rows);
  mlp_stats(b, c->mlp_count, c->stream);
  for (int it = 0; it < iters; ++it) {
    mlp_weights_to_bf16(b, c->stream);
    mlp_gather_batch(b, rows, c->mlp_count, (uint32_t)seed, (uint32_t)(c->mlp_steps + it), c->stream);
    mlp_forward(b, rows, b.Xb, b.Y, c->stream);
    mlp_backward(b, rows, b.Xb, b.Tb, b.grads, b.part, c->stream);
    mlp_adam(b, (float)lr, c->mlp_steps + it + 1, c->stream);
  }
  HIP_OK(hipGetLastError());
  c->mlp_steps += iters;
  const int np = mlp_loss_parts(b, rows);
  std::vector<double> hp(np);
  HIP_OK(hipMemcpyAsync(hp.data(), b.part, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  double acc = 0.0;
  for (double v : hp) acc += v;
  if (loss) *loss = acc / ((double)rows * b.d.nout);
  return 0;
}

extern "C" int cpfft_set_warm_start(cpfft_ctx* c, int mode) {
  if (!c) return fail(-1, "cpfft_set_warm_start: null context");
  if (mode & ~(CPFFT_WARM_NETWORK | CPFFT_WARM_PREVIOUS))
    return fail(-1, "cpfft_set_warm_start: unknown mode bits");
  if ((mode & CPFFT_WARM_NETWORK) && !c->mlp_ready)
    return fail(-1, "cpfft_set_warm_start: the network start needs cpfft_mlp_init first");
  c->warm = mode;
  return 0;
}

extern "C" int cpfft_mlp_warm_start(cpfft_ctx* c, int on) {
  if (!c) return fail(-1, "cpfft_mlp_warm_start: null context");
  return cpfft_set_warm_start(c, on ? (c->warm | CPFFT_WARM_NETWORK) : (c->warm & ~CPFFT_WARM_NETWORK));
}

extern "C" int cpfft_mlp_params(cpfft_ctx* c, float* buf, int get) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_params: cpfft_mlp_init first");
  MlpBuffers& b = c->mlp;
  const MlpDims& d = b.d;
  if (get) {
    HIP_OK(hipMemcpyAsync(buf, b.params, d.nparam * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(buf + d.nparam, b.norm, d.nnorm * sizeof(float), hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    HIP_OK(hipMemcpyAsync(b.params, buf, d.nparam * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(b.norm, buf + d.nparam, d.nnorm * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

static int mlp_upload_batch(cpfft_ctx* c, const float* Xn, const float* Tn, int n) {
  MlpBuffers& b = c->mlp;
  if (n < 1 || n > b.rows) return fail(-1, "cpfft_mlp: batch size out of range");
  HIP_OK(hipMemcpyAsync(b.Xf, Xn, (long)n * b.d.nin * sizeof(float), hipMemcpyHostToDevice, c->stream));
  mlp_to_bf16((long)n * b.d.nin, b.Xf, b.Xb, c->stream);
  if (Tn)
    HIP_OK(hipMemcpyAsync(b.Tb, Tn, (long)n * b.d.nout * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
  mlp_weights_to_bf16(b, c->stream);
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int cpfft_mlp_forward(cpfft_ctx* c, const float* Xn, int n, float* Y) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_forward: cpfft_mlp_init first");
  int rc = mlp_upload_batch(c, Xn, nullptr, n);
  if (rc) return rc;
  mlp_forward(c->mlp, n, c->mlp.Xb, c->mlp.Y, c->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipMemcpyAsync(Y, c->mlp.Y, (long)n * c->mlp.d.nout * sizeof(float), hipMemcpyDeviceToHost,
                        c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int cpfft_mlp_grad(cpfft_ctx* c, const float* Xn, const float* Tn, int n, float* grad,
                              double* loss) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_grad: cpfft_mlp_init first");
  int rc = mlp_upload_batch(c, Xn, Tn, n);
  if (rc) return rc;
  MlpBuffers& b = c->mlp;
  mlp_forward(b, n, b.Xb, b.Y, c->stream);
  mlp_backward(b, n, b.Xb, b.Tb, b.grads, b.part, c->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c->stream));
  const int np = mlp_loss_parts(b, n);
  std::vector<double> hp(np);
  HIP_OK(hipMemcpyAsync(grad, b.grads, b.d.nparam * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(hp.data(), b.part, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  double acc = 0.0;
  for (double v : hp) acc += v;
  if (loss) *loss = acc / ((double)n * b.d.nout);
  return 0;
}

extern "C" int cpfft_mlp_predict(cpfft_ctx* c, int step, float* X, float* ws) {
  if (!c || !c->mlp_ready) return fail(-1, "cpfft_mlp_predict: cpfft_mlp_init first");
  MlpBuffers& b = c->mlp;
  if (X) {
    float* d = nullptr;
    HIP_OK(hipMalloc(&d, c->V * b.d.nin * sizeof(float)));
    if (c->nslip == 12)
      launch_mlp_features<12>(c->V, c->mp, (const Orient<12>*)c->orients, c->oid, step,
                              c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                              nullptr, 0, 1, c->V, d, nullptr, c->stream);
    else
      launch_mlp_features<24>(c->V, c->mp, (const Orient<24>*)c->orients, c->oid, step,
                              c->fields[CPFFT_FN], c->fields[CPFFT_FN1], c->fields[CPFFT_STATE_N],
                              nullptr, 0, 1, c->V, d, nullptr, c->stream);
    const hipError_t e = hipMemcpyAsync(X, d, c->V * b.d.nin * sizeof(float),
                                        hipMemcpyDeviceToHost, c->stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    HIP_OK(e);
    HIP_OK(e2);
  }
  if (ws) {
    if (c->nslip == 12)
      mlp_predict<12>(c, step);
    else
      mlp_predict<24>(c, step);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(ws, b.ws, c->V * b.d.nout * sizeof(float), hipMemcpyDeviceToHost,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

extern "C" int cpfft_local_stats(cpfft_ctx* c, long long out[4]) {
  if (!c) return fail(-1, "cpfft_local_stats: null context");
  long long* d = c->hcnt;
  HIP_OK(hipMemcpyAsync(d, c->dcnt, 8 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  out[0] = d[0] + d[1];  // local iterations of the voxels (predictor + full Newton)
  out[1] = d[6];         // local iterations issued: per wavefront 64 x its slowest voxel's
  out[2] = d[4];         // residual evaluations of the line searches
  out[3] = d[2];         // sub-step cuts
  return 0;
}

extern "C" int cpfft_mlp_stats(cpfft_ctx* c, long long out[4]) {
  if (!c) return fail(-1, "cpfft_mlp_stats: null context");
  long long* d = c->hcnt;
  HIP_OK(hipMemcpyAsync(d, c->dcnt, 8 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  out[0] = c->mlp_count;
  out[1] = c->mlp_steps;
  out[2] = c->ws_updates;
  out[3] = d[5];
  return 0;
}
