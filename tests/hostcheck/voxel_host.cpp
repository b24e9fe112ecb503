// voxel_host.cpp -- test-only host build of the library's per-voxel update
// (cpfft_amd/csrc/mm10_core.h compiled for the CPU).  It lets the CPU test
// suite check the device code's algebra against the oracle without a GPU;
// the library itself has no CPU path and never loads this file.
#include <vector>

#include "../../cpfft_amd/csrc/setup_host.h"

using namespace cpfft;

template <int NS>
static void run(const cpfft_material& m, double dt, long V, const double* ang, int step,
                int iter, const double* Fn, const double* Fn1, const double* sn, double* snp1,
                double* Pn1, double* K4, long long* cnt, const float* ws = nullptr,
                int warm_prev = 0) {
  double b[kMaxSlip][3], nn[kMaxSlip][3], st[6][6];
  int ns = 0;
  slip_systems(m.slip_set, ns, b, nn);
  cubic_stiffness(m, st);
  const MatParams P = make_params(m, dt);
  long long tot[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tot[:6])
  for (long e = 0; e < V; ++e) {
    Orient<NS> o;
    const double a[3] = {ang[e], ang[V + e], ang[2 * V + e]};
    build_orient<NS>(st, b, nn, a, o);
    Counters c{0, 0, 0, 0, 0, 0};
    double tan6[36], xp[7];
    if (warm_prev) {  // snp1 holds the previous global iterate on entry
      for (int k = 0; k < 6; ++k) xp[k] = snp1[k * V + e];
      xp[6] = snp1[15 * V + e];
    }
    auto ldF = [&](int which, int k) { return which ? Fn1[k * V + e] : Fn[k * V + e]; };
    voxel_update<NS>(
        P, o, step, iter, ldF, [&](int w) { return sn[w * V + e]; },
        [&](int w, double v) { snp1[w * V + e] = v; },
        [&](int k, double v) { Pn1[k * V + e] = v; }, [&](int k, double v) { tan6[k] = v; }, c,
        ws ? ws + e : nullptr, V, warm_prev ? xp : nullptr);
    double t6[6];
    for (int k = 0; k < 6; ++k) t6[k] = snp1[k * V + e];
    voxel_dPdF(ldF, t6, [&](int k) { return tan6[k]; },
               [&](int k, double v) { K4[k * V + e] = v; });
    tot[0] += c.pred_it;
    tot[1] += c.upd_it;
    tot[2] += c.cuts;
    tot[3] += c.failed;
    tot[4] += c.ls_evals;
    tot[5] += c.ws_fallback;
  }
  for (int k = 0; k < 6; ++k) cnt[k] = tot[k];
}

extern "C" int cpfft_host_drive_eps_sig(const cpfft_material* m, double dt, long V,
                                        const double* angles, int step, int iter,
                                        const double* Fn, const double* Fn1, const double* sn,
                                        double* snp1, double* Pn1, double* K4, long long* cnt) {
  switch (slip_count(m->slip_set)) {
    case 12: run<12>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt); break;
    case 24: run<24>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt); break;
    case 48: run<48>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt); break;
    default: return -1;
  }
  return 0;
}

// the same update warm-started from slip increments ws[12][V] (fcc)
extern "C" int cpfft_host_drive_eps_sig_ws(const cpfft_material* m, double dt, long V,
                                           const double* angles, int step, int iter,
                                           const double* Fn, const double* Fn1, const double* sn,
                                           double* snp1, double* Pn1, double* K4, long long* cnt,
                                           const float* ws) {
  if (m->slip_set != SLIP_FCC) return -1;
  run<12>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt, ws);
  return 0;
}

// the same update warm-started from the previous global iterate that snp1
// holds on entry (any slip set)
extern "C" int cpfft_host_drive_eps_sig_prev(const cpfft_material* m, double dt, long V,
                                             const double* angles, int step, int iter,
                                             const double* Fn, const double* Fn1,
                                             const double* sn, double* snp1, double* Pn1,
                                             double* K4, long long* cnt) {
  switch (slip_count(m->slip_set)) {
    case 12: run<12>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt, nullptr, 1); break;
    case 24: run<24>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt, nullptr, 1); break;
    case 48: run<48>(*m, dt, V, angles, step, iter, Fn, Fn1, sn, snp1, Pn1, K4, cnt, nullptr, 1); break;
    default: return -1;
  }
  return 0;
}

// the library's crystal-frame slip table for slip set `set`: b[s][3] and
// n[s][3] for s < ns (48 at most); returns ns (0 for an unknown set)
extern "C" int cpfft_host_slip_systems(int set, double* b, double* n) {
  double bb[kMaxSlip][3], nn[kMaxSlip][3];
  int ns = 0;
  slip_systems(set, ns, bb, nn);
  for (int s = 0; s < ns; ++s)
    for (int k = 0; k < 3; ++k) {
      b[3 * s + k] = bb[s][k];
      n[3 * s + k] = nn[s][k];
    }
  return ns;
}

// the library's per-orientation invariants for Kocks angles ang (degrees):
// rotated stiffness C[36] (row-major Voigt, engineering shear), Schmid
// tensors ms[s][6] and spin vectors qs[s][3]; returns ns, or -1 when the
// compliance is singular
template <int NS>
static void orient_out(const double st[6][6], const double b[kMaxSlip][3],
                       const double n[kMaxSlip][3], const double* ang, double* C, double* ms,
                       double* qs) {
  Orient<NS> o;
  build_orient<NS>(st, b, n, ang, o);
  for (int k = 0; k < 36; ++k) C[k] = o.C[k];
  for (int s = 0; s < NS; ++s) {
    for (int k = 0; k < 6; ++k) ms[6 * s + k] = o.ms[s][k];
    for (int k = 0; k < 3; ++k) qs[3 * s + k] = o.qs[s][k];
  }
}

extern "C" int cpfft_host_orientation(const cpfft_material* m, const double* ang, double* C,
                                      double* ms, double* qs) {
  double b[kMaxSlip][3], nn[kMaxSlip][3], st[6][6];
  int ns = 0;
  slip_systems(m->slip_set, ns, b, nn);
  if (!cubic_stiffness(*m, st)) return -1;
  switch (ns) {
    case 12: orient_out<12>(st, b, nn, ang, C, ms, qs); break;
    case 24: orient_out<24>(st, b, nn, ang, C, ms, qs); break;
    case 48: orient_out<48>(st, b, nn, ang, C, ms, qs); break;
    default: return 0;
  }
  return ns;
}

// The network inputs of every voxel (voxel_features, FP32) beside the same
// quantities in FP64: step_kinematics's uddt, the trial stress s_n + C:D and
// D rotated to the lattice frame Rp^T X Rp, tt_n, and slip_to_u of the
// previous slip increments (words 23..).  f32 / f64 are [V][18 + 2 NS].
template <int NS>
static void features(const cpfft_material& m, double dt, long V, const double* ang, int step,
                     const double* Fn, const double* Fn1, const double* sn, float* f32,
                     double* f64) {
  double b[kMaxSlip][3], nn[kMaxSlip][3], st[6][6];
  int ns = 0;
  slip_systems(m.slip_set, ns, b, nn);
  cubic_stiffness(m, st);
  const MatParams P = make_params(m, dt);
  constexpr int NIN = 18 + 2 * NS;
#pragma omp parallel for schedule(dynamic, 64)
  for (long e = 0; e < V; ++e) {
    Orient<NS> o;
    const double a[3] = {ang[e], ang[V + e], ang[2 * V + e]};
    build_orient<NS>(st, b, nn, a, o);
    auto ldn = [&](int w) { return sn[w * V + e]; };
    voxel_features<NS>(
        P, o, step, [&](int which, int k) { return which ? Fn1[k * V + e] : Fn[k * V + e]; }, ldn,
        f32 + e * NIN);
    double F0[3][3], F1[3][3], rnh[3][3], rn1[3][3], fi[3][3], ud[6];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        F0[i][j] = Fn[(3 * i + j) * V + e];
        F1[i][j] = Fn1[(3 * i + j) * V + e];
      }
    step_kinematics(F0, F1, rnh, rn1, fi, ud);
    double s0[6] = {0, 0, 0, 0, 0, 0}, Rp[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double tt = P.tau_y + 1.0e-5;
    if (step > 1) {
      for (int k = 0; k < 6; ++k) s0[k] = ldn(k);
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) Rp[i][j] = ldn(6 + 3 * j + i);
      tt = ldn(15);
    }
    double strial[6];
    for (int i = 0; i < 6; ++i) {
      strial[i] = s0[i];
      for (int k = 0; k < 6; ++k) strial[i] += o.C[6 * i + k] * ud[k];
    }
    const double S[3][3] = {{strial[0], strial[3], strial[5]},
                            {strial[3], strial[1], strial[4]},
                            {strial[5], strial[4], strial[2]}};
    const double D[3][3] = {{ud[0], 0.5 * ud[3], 0.5 * ud[5]},
                            {0.5 * ud[3], ud[1], 0.5 * ud[4]},
                            {0.5 * ud[5], 0.5 * ud[4], ud[2]}};
    double* g = f64 + e * NIN;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double vs = 0.0, vd = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            vs += Rp[k][i] * S[k][l] * Rp[l][j];
            vd += Rp[k][i] * D[k][l] * Rp[l][j];
          }
        g[3 * i + j] = vs;
        g[9 + NS + 3 * i + j] = vd;
      }
    const double dg = P.alter_mode ? P.eps_dot_0_y * P.dt : 1.0;
    for (int s = 0; s < NS; ++s) {
      g[9 + s] = tt;
      g[18 + NS + s] = step == 1 ? 0.0 : slip_to_u(P, dg, ldn(23 + s));
    }
  }
}

extern "C" int cpfft_host_features(const cpfft_material* m, double dt, long V,
                                   const double* angles, int step, const double* Fn,
                                   const double* Fn1, const double* sn, float* f32, double* f64) {
  switch (slip_count(m->slip_set)) {
    case 12: features<12>(*m, dt, V, angles, step, Fn, Fn1, sn, f32, f64); break;
    case 24: features<24>(*m, dt, V, angles, step, Fn, Fn1, sn, f32, f64); break;
    default: return -1;
  }
  return 0;
}
