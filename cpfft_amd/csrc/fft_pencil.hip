// fft_pencil.hip -- G_K_dF (src/G_K_dF.f:11-87) for grid edges with a
// compile-time radix plan (16, 32, 48, 64, 96, 128, 256).
//
// Five passes over two half-spectrum buffers in HBM, each pass reading and
// writing contiguous runs (whole lines or >= 64-byte tiles):
//   A  z forward:  K4:F contraction (ddot42n order, G_K_dF.f:241-268) and a
//                  real-to-complex FFT along z.  Two real pencils y, y+1 share
//                  one complex FFT (a + i b) and are separated afterwards.
//                  F[c][x][y][z] -> T1[c][kz][x][y]
//   B  y forward:  complex FFT along y, transposing      T1 -> T2[c][kz][y][x]
//   C  x:          FFT along x of all nine components, Green projection at
//                  each frequency, inverse FFT along x, in place in T2
//                  (for N = 64, 128, 256 with the transforms in registers)
//   B' y inverse:  inverse FFT along y, transposing       T2 -> T1
//   D  z inverse:  complex-to-real FFT along z (two pencils per complex
//                  FFT), scale 1/N^3                      T1 -> out[c][x][y][z]
// In LDS every stage is an in-place Stockham radix pass: each thread loads
// its butterfly inputs into registers, the block synchronises, and the
// outputs are written back; radix 8/4/2 butterflies are explicit, 3/5/7 use
// the twiddle table.  Lines are padded by one element in eight so that the
// strided butterfly writes fall on distinct LDS banks.
#include "kernels.h"

// No implicit contraction into FMAs in this file: the compiler fused
// differently in different instantiations of the same transform (e.g. pass B
// with 2 or 8 lines per block), and every launch shape must give the same
// bits.  The transforms' FMAs are written out (cmul, butterfly).
#pragma clang fp contract(off)

namespace cpfft {
namespace fp {

constexpr int kT = 256;  // threads per block

struct Plan {
  int nst;
  int r[12];
};

constexpr Plan plan_for(int n) {
  Plan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n;
  while (m % 8 == 0) { p.r[p.nst++] = 8; m /= 8; }
  while (m % 4 == 0) { p.r[p.nst++] = 4; m /= 4; }
  while (m % 2 == 0) { p.r[p.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { p.r[p.nst++] = 3; m /= 3; }
  while (m % 5 == 0) { p.r[p.nst++] = 5; m /= 5; }
  while (m % 7 == 0) { p.r[p.nst++] = 7; m /= 7; }
  if (m != 1) p.nst = -1;
  return p;
}

// padded position of element n of a line, and the padded line stride
__device__ __forceinline__ constexpr int pidx(int n) { return n + (n >> 3); }
constexpr int lstride(int n) { return n + ((n - 1) >> 3); }

// pencils per block of the z passes: the largest even divisor TY of N with
// TY*N <= 1024 (at least 2)
constexpr int tile_z(int n) {
  int best = 2;
  for (int t = 2; t <= n; t += 2)
    if (n % t == 0 && t * n <= 1024) best = t;
  return best;
}
// Launch-shape candidates of the autotuner (PencilTune): the i-th largest
// (i = 0, 1, ...) of the admissible values, 0 when there is none.  Every
// candidate computes bit-identical results: the shapes only change which
// block transforms which lines, never the arithmetic of a line.
// pass A: pencils per block, even divisors t of N with t*N <= 1024
constexpr int tz_cand(int n, int i) {
  int k = 0;
  for (int t = n; t >= 2; --t)
    if (t % 2 == 0 && n % t == 0 && t * n <= 1024) {
      if (k == i) return t;
      ++k;
    }
  return 0;
}
// passes B, B': lines per block, divisors of N up to 8
constexpr int ty_cand(int n, int i) {
  int k = 0;
  for (int t = 8; t >= 1; --t)
    if (n % t == 0) {
      if (k == i) return t;
      ++k;
    }
  return 0;
}
// pass C in registers (kC_xr) needs N = 8 M with M a multiple of 8
constexpr bool regc_ok(int n) {
  const Plan p = plan_for(n);
  return p.nst >= 2 && p.r[0] == 8 && (n / 8) % 8 == 0;
}
constexpr bool lds_fits(int lines, int n) { return (lines * lstride(n) + n) * 16 <= 65536; }
// pass C: component lines per block, each candidate where its lines fit
// the 64 KiB of static LDS.  Where regc_ok: first kC_xr with the three
// components of one tensor row for 8, 4 or 2 y lines (24, 12, 6).  Then,
// for every N, kC_x with all nine components of 2 or 1 y lines (18, 9).
// Both compute the same bits; the autotuner picks (at 128^3 kC_xr takes
// 60-66 us against 85-88 us for kC_x).
constexpr int xl_cand(int n, int i) {
  int k = 0;
  if (regc_ok(n)) {
    for (int yb = 8; yb >= 2; yb /= 2)
      if ((3 * yb * lstride(n) + n) * 16 + 2 * yb * n * 8 <= 65536) {  // lines, twiddles, weights
        if (k == i) return 3 * yb;
        ++k;
      }
  }
  for (int yb = 2; yb >= 1; --yb)
    if (lds_fits(9 * yb, n)) {
      if (k == i) return 9 * yb;
      ++k;
    }
  return 0;
}
constexpr bool xl_reg(int v) { return v % 9 != 0; }  // kC_xr (else kC_x)
static_assert(kPencilCands >= 5, "candidate slots");

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
// explicit FMAs: fewer instructions than separate products and sums, and
// the same in every instantiation
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
// multiply by -i (forward) or +i (inverse)
template <bool FWD>
__device__ __forceinline__ cplx mul_mi(cplx a) {
  return FWD ? cplx{a.im, -a.re} : cplx{-a.im, a.re};
}
// table holds exp(-2 pi i e/N); the inverse transform uses its conjugate
template <bool FWD>
__device__ __forceinline__ cplx twid(const cplx* tw, int e) {
  cplx w = tw[e];
  if (!FWD) w.im = -w.im;
  return w;
}

template <bool FWD>
__device__ __forceinline__ void dft4(cplx& a0, cplx& a1, cplx& a2, cplx& a3) {
  cplx b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3), b3 = mul_mi<FWD>(csub(a1, a3));
  a0 = cadd(b0, b2);
  a1 = cadd(b1, b3);
  a2 = csub(b0, b2);
  a3 = csub(b1, b3);
}

// in-register DFT of size R on x[0..R)
template <int R, bool FWD, int N>
__device__ __forceinline__ void butterfly(cplx* x, const cplx* tw) {
  if constexpr (R == 2) {
    cplx a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4<FWD>(x[0], x[1], x[2], x[3]);
  } else if constexpr (R == 8) {
    // X_q = E_q + w8^q O_q, X_{q+4} = E_q - w8^q O_q with E, O the 4-point
    // DFTs of the even and odd inputs
    cplx e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    cplx o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<FWD>(e0, e1, e2, e3);
    dft4<FWD>(o0, o1, o2, o3);
    // X_1 = E_1 + h s1, X_3 = E_3 + h s3 with the sums s below, as FMAs
    const double h = 0.70710678118654752440;  // sqrt(2)/2
    const cplx s1 = FWD ? cplx{o1.re + o1.im, o1.im - o1.re} : cplx{o1.re - o1.im, o1.re + o1.im};
    const cplx s3 = FWD ? cplx{o3.im - o3.re, -(o3.re + o3.im)} : cplx{-(o3.re + o3.im), o3.re - o3.im};
    const cplx w2 = mul_mi<FWD>(o2);
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cplx{fma(h, s1.re, e1.re), fma(h, s1.im, e1.im)};
    x[5] = cplx{fma(-h, s1.re, e1.re), fma(-h, s1.im, e1.im)};
    x[2] = cadd(e2, w2);
    x[6] = csub(e2, w2);
    x[3] = cplx{fma(h, s3.re, e3.re), fma(h, s3.im, e3.im)};
    x[7] = cplx{fma(-h, s3.re, e3.re), fma(-h, s3.im, e3.im)};
  } else {
    // direct DFT with the R-th roots of unity from the N-th root table
    cplx y[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cplx acc = x[0];
#pragma unroll
      for (int t = 1; t < R; ++t) acc = cadd(acc, cmul(x[t], twid<FWD>(tw, ((q * t) % R) * (N / R))));
      y[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = y[q];
  }
}

// One in-place Stockham pass of radix R over L lines of length N in LDS
// (line l, element n at buf[l*LS + pidx(n)]); NS is the product of the
// radices of the earlier passes.  Starts and ends with all threads' LDS
// accesses complete.
template <int N, int R, int NS, int L, bool FWD, int T>
__device__ __forceinline__ void stage(cplx* buf, const cplx* tw) {
  constexpr int NB = N / R;             // butterflies per line
  constexpr int TOT = L * NB;
  constexpr int BPT = (TOT + T - 1) / T;
  constexpr int TS = N / (NS * R);      // twiddle exponent step
  constexpr int LS = lstride(N);
  cplx x[BPT][R];
#pragma unroll
  for (int u = 0; u < BPT; ++u) {
    const int b = threadIdx.x + u * T;
    if (TOT % T == 0 || b < TOT) {
      const int line = b / NB, j = b - line * NB;
      const cplx* s = buf + line * LS;
#pragma unroll
      for (int t = 0; t < R; ++t) x[u][t] = s[pidx(j + t * NB)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; ++u) {
    const int b = threadIdx.x + u * T;
    if (TOT % T == 0 || b < TOT) {
      const int line = b / NB, j = b - line * NB;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; ++t) x[u][t] = cmul(x[u][t], twid<FWD>(tw, t * k * TS));
      }
      butterfly<R, FWD, N>(x[u], tw);
      const int o = (j / NS) * NS * R + k;
      cplx* d = buf + line * LS;
#pragma unroll
      for (int q = 0; q < R; ++q) d[pidx(o + q * NS)] = x[u][q];
    }
  }
  __syncthreads();
}

// unscaled FFT (forward: exp(-2 pi i nk/N)) of L lines in place by a block
// of T threads
template <int N, int L, bool FWD, int T = kT, int ST = 0, int NS = 1>
__device__ __forceinline__ void fft(cplx* buf, const cplx* tw) {
  constexpr Plan P = plan_for(N);
  static_assert(P.nst > 0, "grid edge without a compile-time radix plan");
  if constexpr (ST < P.nst) {
    constexpr int R = P.r[ST];
    stage<N, R, NS, L, FWD, T>(buf, tw);
    fft<N, L, FWD, T, ST + 1, NS * R>(buf, tw);
  }
}

template <int N, int T = kT>
__device__ __forceinline__ void load_tw(cplx* tw, const cplx* __restrict__ twg) {
  for (int i = threadIdx.x; i < N; i += T) tw[i] = twg[i];
}

// threads per block of pass A for a tile of p voxels: two voxels per thread
// with K4 (its registers hold F and two K4 row blocks), four without
constexpr int threads_z(int p, bool hask) { return hask && p / 2 >= 512 ? 512 : 256; }

// ---- pass A: K4:F and real-to-complex FFT along z ------------------------
// grid (N/TY, N): blockIdx.x = y tile, blockIdx.y = x.  The nine output
// components are transformed one after the other; the K4 rows (or F
// component) of component c+1 are loaded into a second register set before
// component c is transformed, so the loads overlap the LDS work.
// DIR: F is the CG direction p, first replaced by p = R + cs->beta p (each
// element read and written by the same thread).
template <int N, bool HASK, bool DIR, int TY>
__global__ __launch_bounds__(threads_z(TY * N, HASK)) void kA_zfwd(const double* F,
                                                              const double* __restrict__ K4,
                                                              cplx* __restrict__ T1,
                                                              const cplx* __restrict__ twg,
                                                              const int* __restrict__ skip,
                                                              double* Pd,
                                                              const double* __restrict__ R,
                                                              const CGScalars* __restrict__ cs,
                                                              double* __restrict__ pqp) {
  if (skip && *skip) return;
  constexpr int TA = threads_z(TY * N, HASK);
  constexpr int P = TY * N, PT = (P + TA - 1) / TA, LP = TY / 2;
  constexpr int LS = lstride(N), NZ = N / 2 + 1;
  constexpr int NK = HASK ? 9 : 1;  // values per point loaded per component
  constexpr long V = (long)N * N * N;
  __shared__ cplx tw[N];
  __shared__ cplx buf[LP * LS];
  const int y0 = blockIdx.x * TY, x = blockIdx.y;
  load_tw<N, TA>(tw, twg);
  const long e0 = ((long)x * N + y0) * N;  // the tile is P consecutive voxels
  auto inb = [&](int p) { return P % TA == 0 || p < P; };
  double f[HASK ? PT : 1][9];
  if constexpr (HASK) {
    const double beta = DIR ? cs->beta : 0.0;
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int p = threadIdx.x + u * TA;
      if (inb(p)) {
#pragma unroll
        for (int d = 0; d < 9; ++d) {
          if constexpr (DIR) {
            const double v = __builtin_nontemporal_load(R + d * V + e0 + p) +
                             beta * __builtin_nontemporal_load(Pd + d * V + e0 + p);
            __builtin_nontemporal_store(v, Pd + d * V + e0 + p);
            f[u][d] = v;
          } else {
            f[u][d] = __builtin_nontemporal_load(F + d * V + e0 + p);
          }
        }
      }
    }
  }
  // K4 rows 9c..9c+8 (or component c of F) at this thread's points
  auto load_c = [&](int c, double (&k)[PT][NK]) {
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int p = threadIdx.x + u * TA;
      if (inb(p)) {
#pragma unroll
        for (int d = 0; d < NK; ++d)
          k[u][d] = HASK ? __builtin_nontemporal_load(K4 + (9 * c + d) * V + e0 + p)
                         : __builtin_nontemporal_load(F + c * V + e0 + p);
      }
    }
  };
  double kc[PT][NK], kn[PT][NK];
  load_c(0, kc);
  double* bd = reinterpret_cast<double*>(buf);
  double pq[PT];  // pqp: F . (K4:F) at this thread's points
#pragma unroll
  for (int u = 0; u < PT; ++u) pq[u] = 0.0;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int p = threadIdx.x + u * TA;
      if (inb(p)) {
        double v;
        if constexpr (HASK) {
          v = ddot42n_row(kc[u], f[u]);
          if (pqp) pq[u] += v * f[u][c];
        } else {
          v = kc[u][0];
        }
        const int yy = p / N, z = p - yy * N;
        bd[2 * ((yy >> 1) * LS + pidx(z)) + (yy & 1)] = v;
      }
    }
    if (c < 8) load_c(c + 1, kn);
    __syncthreads();
    fft<N, LP, true, TA>(buf, tw);
    // Z = FFT(a + i b): A_k = (Z_k + conj Z_{N-k})/2, B_k = -i (Z_k - conj Z_{N-k})/2
    for (int q = threadIdx.x; q < LP * NZ; q += TA) {
      const int k = q / LP, m = q - k * LP;
      const cplx* line = buf + m * LS;
      const cplx zk = line[pidx(k)], zn = line[pidx(k == 0 ? 0 : N - k)];
      cplx* o = T1 + (((long)c * NZ + k) * N + x) * N + y0 + 2 * m;
      o[0] = cplx{0.5 * (zk.re + zn.re), 0.5 * (zk.im - zn.im)};
      o[1] = cplx{0.5 * (zk.im + zn.im), -0.5 * (zk.re - zn.re)};
    }
    __syncthreads();
    if (c < 8) {
#pragma unroll
      for (int u = 0; u < PT; ++u)
#pragma unroll
        for (int d = 0; d < NK; ++d) kc[u][d] = kn[u][d];
    }
  }
  if constexpr (HASK) {
    // one partial sum per pencil pair (y lines y, y+1 of this x: 2N voxels,
    // local points [2Nm, 2N(m+1))), reduced by a fixed halving order of its
    // own, so the sums do not depend on the tile of the block
    if (pqp) {
      __shared__ double red[P];
#pragma unroll
      for (int u = 0; u < PT; ++u) {
        const int p = threadIdx.x + u * TA;
        if (inb(p)) red[p] = pq[u];
      }
      __syncthreads();
      for (int len = 2 * N; len > 1;) {
        const int half = (len + 1) / 2, n2 = len - half;
        for (int i = threadIdx.x; i < LP * n2; i += TA) {
          const int m = i / n2, j = i - m * n2;
          red[2 * N * m + j] += red[2 * N * m + j + half];
        }
        __syncthreads();
        len = half;
      }
      for (int m = threadIdx.x; m < LP; m += TA) pqp[(long)x * (N / 2) + y0 / 2 + m] = red[2 * N * m];
    }
  }
}

// Non-temporal 16-byte load of one complex value: every word of the half
// spectrum is read once per launch.  Its stores stay ordinary, so that the
// next pass finds what it reads in the memory-side cache.
typedef double cplx_nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cplx ld_nt(const cplx* a) {
  const cplx_nt v = __builtin_nontemporal_load(reinterpret_cast<const cplx_nt*>(a));
  return cplx{v.x, v.y};
}

// ---- passes B, B': FFT along y with a transpose ---------------------------
// FWD: T1[c][kz][x][y] -> T2[c][kz][y][x]; inverse: T2 -> T1.
// grid (N/TX, NZ, 9)
template <int N, bool FWD, int TX>
__global__ __launch_bounds__(kT) void kB_y(const cplx* __restrict__ src, cplx* __restrict__ dst,
                                           const cplx* __restrict__ twg,
                                           const int* __restrict__ skip) {
  if (skip && *skip) return;
  constexpr int LS = lstride(N), NZ = N / 2 + 1;
  __shared__ cplx tw[N];
  __shared__ cplx buf[TX * LS];
  const int x0 = blockIdx.x * TX, kz = blockIdx.y, c = blockIdx.z;
  const long plane = ((long)c * NZ + kz) * N * N;
  load_tw<N>(tw, twg);
  if constexpr (FWD) {
    for (int i = threadIdx.x; i < TX * N; i += kT) {
      const int t = i / N, y = i - t * N;
      buf[t * LS + pidx(y)] = ld_nt(src + plane + (long)(x0 + t) * N + y);
    }
  } else {
    for (int i = threadIdx.x; i < TX * N; i += kT) {
      const int y = i / TX, t = i - y * TX;
      buf[t * LS + pidx(y)] = ld_nt(src + plane + (long)y * N + x0 + t);
    }
  }
  __syncthreads();
  fft<N, TX, FWD>(buf, tw);
  if constexpr (FWD) {
    for (int i = threadIdx.x; i < TX * N; i += kT) {
      const int y = i / TX, t = i - y * TX;
      dst[plane + (long)y * N + x0 + t] = buf[t * LS + pidx(y)];
    }
  } else {
    for (int i = threadIdx.x; i < TX * N; i += kT) {
      const int t = i / N, y = i - t * N;
      dst[plane + (long)(x0 + t) * N + y] = buf[t * LS + pidx(y)];
    }
  }
}

// ---- pass C: FFT along x, Green projection, inverse FFT along x ------------
// YB consecutive y lines of all nine components per block.
// grid (N/YB, NZ): blockIdx.x = y / YB, blockIdx.y = kz
// dc != nullptr: the block of frequency (0, 0, 0) also stores the real parts
// of that frequency before the projection, i.e. the voxel sums of the nine
// input components (the transforms are unscaled).
template <int N, int YB>
__global__ __launch_bounds__(kT) void kC_x(cplx* __restrict__ T2, const cplx* __restrict__ twg,
                                           int gamma_mode, const int* __restrict__ skip,
                                           double* __restrict__ dc) {
  if (skip && *skip) return;
  constexpr int LS = lstride(N), NZ = N / 2 + 1;
  __shared__ cplx tw[N];
  __shared__ cplx buf[9 * YB * LS];
  const int y0 = blockIdx.x * YB, kz = blockIdx.y;
  load_tw<N>(tw, twg);
  // line (c*YB + t) holds component c of y line y0 + t
  for (int i = threadIdx.x; i < 9 * YB * N; i += kT) {
    const int l = i / N, x = i - l * N;
    const int c = l / YB, t = l - c * YB;
    buf[l * LS + pidx(x)] = T2[(((long)c * NZ + kz) * N + y0 + t) * N + x];
  }
  __syncthreads();
  fft<N, 9 * YB, true>(buf, tw);
  // thread 0 reads (t, x) = (0, 0) before it projects that same point
  if (dc && blockIdx.x == 0 && kz == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 9; ++c) dc[c] = buf[c * YB * LS + pidx(0)].re;
  }
  for (int i = threadIdx.x; i < YB * N; i += kT) {
    const int t = i / N, x = i - t * N;
    cplx A[9], out[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) A[c] = buf[(c * YB + t) * LS + pidx(x)];
    project9(N, gamma_mode, x, y0 + t, kz, A, out);
#pragma unroll
    for (int c = 0; c < 9; ++c) buf[(c * YB + t) * LS + pidx(x)] = out[c];
  }
  __syncthreads();
  fft<N, 9 * YB, false>(buf, tw);
  for (int i = threadIdx.x; i < 9 * YB * N; i += kT) {
    const int l = i / N, x = i - l * N;
    const int c = l / YB, t = l - c * YB;
    T2[(((long)c * NZ + kz) * N + y0 + t) * N + x] = buf[l * LS + pidx(x)];
  }
}

// ---- pass C with the transforms in registers -----------------------------
// For N = 8 M with M a multiple of 8 (64, 128, 256).  The Stockham passes
// of fft() are regrouped without changing their arithmetic: the radix-8
// first pass combines elements j + M t of column j (t < 8) and writes
// element 8 j + q; every later pass maps the elements of one residue class
// q mod 8 onto the same class.  So a thread first transforms columns
// j = g, g + 8, ... with inputs loaded straight from HBM, the block
// exchanges once through LDS, and thread q then runs all later passes on
// the M elements q + 8 u of its class in registers; the results land in
// natural order.  Per line: eight threads, two LDS exchanges per transform
// instead of one per pass.
// the Stockham passes ST.. of fft() on residue class q (v[u] = element
// q + 8 u), NS the product of the earlier radices
template <int N, bool FWD, int ST, int NS>
__device__ __forceinline__ void reg_passes(cplx* v, int q, const cplx* tw) {
  constexpr Plan P = plan_for(N);
  if constexpr (ST < P.nst) {
    constexpr int R1 = P.r[0], M = N / R1, R = P.r[ST], NBT = M / R, m = NS / R1;
    constexpr int TS = N / (NS * R);
    cplx w[M];
#pragma unroll
    for (int b = 0; b < NBT; ++b) {
      cplx x[R];
#pragma unroll
      for (int t = 0; t < R; ++t) x[t] = v[b + t * NBT];
      const int k = q + R1 * (b % m);  // the butterfly's j % NS in fft()
#pragma unroll
      for (int t = 1; t < R; ++t) x[t] = cmul(x[t], twid<FWD>(tw, t * k * TS));
      butterfly<R, FWD, N>(x, tw);
      const int o = (b / m) * m * R + b % m;
#pragma unroll
      for (int t = 0; t < R; ++t) w[o + t * m] = x[t];
    }
#pragma unroll
    for (int u = 0; u < M; ++u) v[u] = w[u];
    reg_passes<N, FWD, ST + 1, NS * R>(v, q, tw);
  }
}

// YB consecutive y lines of the three components of one tensor row per
// block (the projection of row i reads only row i); eight threads per line.
// grid (N/YB, NZ, 3): blockIdx.x = y / YB, blockIdx.y = kz, blockIdx.z = row.
// Same operations on the same operands as kC_x.
template <int N, int YB>
__global__ __launch_bounds__(24 * YB) void kC_xr(cplx* __restrict__ T2,
                                                 const cplx* __restrict__ twg, int gamma_mode,
                                                 const int* __restrict__ skip,
                                                 double* __restrict__ dc) {
  if (skip && *skip) return;
  constexpr int R1 = 8, M = N / R1, G = M / R1, LS = lstride(N), NZ = N / 2 + 1, T = 24 * YB;
  static_assert(regc_ok(N), "register pass C needs N = 8 M, M a multiple of 8");
  __shared__ cplx tw[N];
  __shared__ cplx buf[3 * YB * LS];
  __shared__ double sw[2][YB * N];  // projection weights at (t, x), shared by the row's lines
  const int y0 = blockIdx.x * YB, kz = blockIdx.y, row = blockIdx.z;
  const int L = threadIdx.x / R1, g = threadIdx.x % R1;  // line L = jc * YB + t
  const int jc = L / YB, t = L - jc * YB, c = 3 * row + jc;
  cplx* gl = T2 + (((long)c * NZ + kz) * N + y0 + t) * N;
  cplx* bl = buf + L * LS;
  load_tw<N, T>(tw, twg);
  cplx x[G][R1];
#pragma unroll
  for (int gc = 0; gc < G; ++gc)
#pragma unroll
    for (int u = 0; u < R1; ++u) x[gc][u] = ld_nt(gl + g + R1 * gc + M * u);
  // the weights, while the loads are in flight (read after the barriers)
  for (int i = threadIdx.x; i < YB * N; i += T) {
    const int tt = i / N, xx = i - tt * N;
    proj_weights(N, gamma_mode, xx, y0 + tt, kz, sw[0][i], sw[1][i]);
  }
  // forward: first pass on columns g + 8 gc, exchange, later passes
#pragma unroll
  for (int gc = 0; gc < G; ++gc) {
    butterfly<R1, true, N>(x[gc], tw);
#pragma unroll
    for (int q = 0; q < R1; ++q) bl[pidx((g + R1 * gc) * R1 + q)] = x[gc][q];
  }
  __syncthreads();
  cplx v[M];
#pragma unroll
  for (int u = 0; u < M; ++u) v[u] = bl[pidx(g + R1 * u)];
  reg_passes<N, true, 1, R1>(v, g, tw);
  if (dc && blockIdx.x == 0 && kz == 0 && t == 0 && g == 0) dc[c] = v[0].re;
  // each thread rewrites the elements it read: no barrier before this
#pragma unroll
  for (int u = 0; u < M; ++u) bl[pidx(g + R1 * u)] = v[u];
  __syncthreads();
  // projection at the inputs of the inverse first pass
  const cplx* r0 = buf + t * LS;
#pragma unroll
  for (int gc = 0; gc < G; ++gc)
#pragma unroll
    for (int u = 0; u < R1; ++u) {
      const int xx = g + R1 * gc + M * u;
      const cplx A3[3] = {r0[pidx(xx)], r0[YB * LS + pidx(xx)], r0[2 * YB * LS + pidx(xx)]};
      x[gc][u] = project1w(N, gamma_mode, xx, y0 + t, kz, A3, jc, sw[0][t * N + xx],
                           sw[1][t * N + xx]);
    }
#pragma unroll
  for (int gc = 0; gc < G; ++gc) butterfly<R1, false, N>(x[gc], tw);
  __syncthreads();
#pragma unroll
  for (int gc = 0; gc < G; ++gc)
#pragma unroll
    for (int q = 0; q < R1; ++q) bl[pidx((g + R1 * gc) * R1 + q)] = x[gc][q];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < M; ++u) v[u] = bl[pidx(g + R1 * u)];
  reg_passes<N, false, 1, R1>(v, g, tw);
#pragma unroll
  for (int u = 0; u < M; ++u) gl[g + R1 * u] = v[u];
}

// ---- pass D: complex-to-real FFT along z, two pencils per FFT --------------
// grid (N/TY, N, 9): blockIdx.x = y tile, blockIdx.y = x, blockIdx.z = c
// MODE 0: out = the transform.  MODE 1 (dot): also the block's partial sum
// of Pd . out in part[block].  MODE 2 (update, a CG iteration whose alpha is
// already in cs): the transform q is not stored; instead R -= alpha q, with
// X != nullptr also X += alpha Pd, and the block's partial sum of the new
// R . R goes to part[block].
template <int N, int MODE>
__global__ __launch_bounds__(kT) void kD_zinv(const cplx* __restrict__ T1, double* __restrict__ out,
                                              const cplx* __restrict__ twg,
                                              const int* __restrict__ skip,
                                              const double* __restrict__ Pd,
                                              double* __restrict__ part, double* __restrict__ R,
                                              double* __restrict__ X,
                                              const CGScalars* __restrict__ cs) {
  if (skip && *skip) return;
  constexpr int TY = tile_z(N), P = TY * N, LP = TY / 2, LS = lstride(N), NZ = N / 2 + 1;
  constexpr long V = (long)N * N * N;
  constexpr bool DOT = MODE == 1, UPD = MODE == 2;
  __shared__ cplx tw[N];
  __shared__ cplx buf[LP * LS];
  const int y0 = blockIdx.x * TY, x = blockIdx.y, c = blockIdx.z;
  load_tw<N>(tw, twg);
  // full line of pencil pair m: Z_k = A_k + i B_k, with A_{N-k} = conj A_k
  for (int q = threadIdx.x; q < LP * NZ; q += kT) {
    const int k = q / LP, m = q - k * LP;
    const cplx* s = T1 + (((long)c * NZ + k) * N + x) * N + y0 + 2 * m;
    const cplx A = s[0], B = s[1];
    cplx* line = buf + m * LS;
    line[pidx(k)] = cplx{A.re - B.im, A.im + B.re};
    if (k >= 1 && N - k >= NZ) line[pidx(N - k)] = cplx{A.re + B.im, B.re - A.im};
  }
  const long e0 = ((long)x * N + y0) * N;
  // this thread's vector values, loaded before the FFT so that their
  // latency overlaps it
  constexpr int PP = (P + kT - 1) / kT;
  double pv[DOT || UPD ? PP : 1], rv[UPD ? PP : 1], xv[UPD ? PP : 1];
  const bool xu = UPD && X != nullptr;
#pragma unroll
  for (int j = 0; j < PP; ++j) {
    const int p = threadIdx.x + j * kT;
    const long i = c * V + e0 + p;
    if constexpr (DOT) pv[j] = p < P ? __builtin_nontemporal_load(Pd + i) : 0.0;
    if constexpr (UPD) {
      rv[j] = p < P ? __builtin_nontemporal_load(R + i) : 0.0;
      pv[j] = xu && p < P ? __builtin_nontemporal_load(Pd + i) : 0.0;
      xv[j] = xu && p < P ? __builtin_nontemporal_load(X + i) : 0.0;
    }
  }
  __syncthreads();
  fft<N, LP, false>(buf, tw);
  const double sc = 1.0 / (double)V;
  const double a = UPD ? cs->alpha : 0.0;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < PP; ++j) {
    const int p = threadIdx.x + j * kT;
    if (p < P) {
      const int yy = p / N, z = p - yy * N;
      const cplx v = buf[(yy >> 1) * LS + pidx(z)];
      const double q = ((yy & 1) ? v.im : v.re) * sc;
      const long i = c * V + e0 + p;
      if constexpr (UPD) {
        if (xu) __builtin_nontemporal_store(xv[j] + a * pv[j], X + i);
        const double rn = rv[j] - a * q;
        R[i] = rn;  // an ordinary store: the next pass A reads r soon after
        acc += rn * rn;
      } else {
        __builtin_nontemporal_store(q, out + i);
        if constexpr (DOT) acc += q * pv[j];
      }
    }
  }
  if constexpr (DOT || UPD) {
    __shared__ double red[kT];
    const double sum = tree_sum0<kT>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = sum;
  }
}

// pass A with the I-th pencil-count candidate (nothing if there is none)
template <int N, int I>
void launch_a(const double* F, const double* K4, cplx* T1, const cplx* tw, const int* sk,
              hipStream_t s, const CGFuse* fz) {
  constexpr int TY = tz_cand(N, I);
  if constexpr (TY > 0) {
    const dim3 ga(N / TY, N);
    if (K4 && fz && fz->dir_p)
      kA_zfwd<N, true, true, TY><<<ga, threads_z(TY * N, true), 0, s>>>(
          F, K4, T1, tw, sk, fz->dir_p, fz->dir_r, fz->cs, fz->pq_part);
    else if (K4)
      kA_zfwd<N, true, false, TY><<<ga, threads_z(TY * N, true), 0, s>>>(
          F, K4, T1, tw, sk, nullptr, nullptr, nullptr, nullptr);
    else
      kA_zfwd<N, false, false, TY><<<ga, threads_z(TY * N, false), 0, s>>>(
          F, nullptr, T1, tw, sk, nullptr, nullptr, nullptr, nullptr);
  }
}
template <int N, int I>
void launch_b(bool fwd, const cplx* src, cplx* dst, const cplx* tw, const int* sk, hipStream_t s) {
  constexpr int TX = ty_cand(N, I), NZ = N / 2 + 1;
  if constexpr (TX > 0) {
    if (fwd)
      kB_y<N, true, TX><<<dim3(N / TX, NZ, 9), kT, 0, s>>>(src, dst, tw, sk);
    else
      kB_y<N, false, TX><<<dim3(N / TX, NZ, 9), kT, 0, s>>>(src, dst, tw, sk);
  }
}
template <int N, int I>
void launch_c(cplx* T2, const cplx* tw, int gamma_mode, const int* sk, hipStream_t s, double* dc) {
  constexpr int V = xl_cand(N, I), NZ = N / 2 + 1;
  if constexpr (V > 0 && xl_reg(V)) {
    if constexpr (regc_ok(N))
      kC_xr<N, V / 3><<<dim3(N / (V / 3), NZ, 3), 8 * V, 0, s>>>(T2, tw, gamma_mode, sk, dc);
  } else if constexpr (V > 0) {
    kC_x<N, V / 9><<<dim3(N / (V / 9), NZ), kT, 0, s>>>(T2, tw, gamma_mode, sk, dc);
  }
}
template <int N>
void pass_a(int i, const double* F, const double* K4, cplx* T1, const cplx* tw, const int* sk,
            hipStream_t s, const CGFuse* fz) {
  switch (i) {
    case 1: launch_a<N, 1>(F, K4, T1, tw, sk, s, fz); break;
    case 2: launch_a<N, 2>(F, K4, T1, tw, sk, s, fz); break;
    default: launch_a<N, 0>(F, K4, T1, tw, sk, s, fz); break;
  }
}
template <int N>
void pass_b(int i, bool fwd, const cplx* src, cplx* dst, const cplx* tw, const int* sk,
            hipStream_t s) {
  switch (i) {
    case 1: launch_b<N, 1>(fwd, src, dst, tw, sk, s); break;
    case 2: launch_b<N, 2>(fwd, src, dst, tw, sk, s); break;
    default: launch_b<N, 0>(fwd, src, dst, tw, sk, s); break;
  }
}
template <int N>
void pass_c(int i, cplx* T2, const cplx* tw, int gamma_mode, const int* sk, hipStream_t s,
            double* dc = nullptr) {
  switch (i) {
    case 1: launch_c<N, 1>(T2, tw, gamma_mode, sk, s, dc); break;
    case 2: launch_c<N, 2>(T2, tw, gamma_mode, sk, s, dc); break;
    case 3: launch_c<N, 3>(T2, tw, gamma_mode, sk, s, dc); break;
    case 4: launch_c<N, 4>(T2, tw, gamma_mode, sk, s, dc); break;
    default: launch_c<N, 0>(T2, tw, gamma_mode, sk, s, dc); break;
  }
}

// one pass by itself (autotuner timing): 0 A with K4, 1 A without, 2 B, 3 C
template <int N>
void launch_pass(const SpecGeom& g, int pass, int variant, const double* F, const double* K4,
                 cplx* T1, cplx* T2, const cplx* tw, hipStream_t s) {
  switch (pass) {
    case 0: pass_a<N>(variant, F, K4, T1, tw, nullptr, s, nullptr); break;
    case 1: pass_a<N>(variant, F, nullptr, T1, tw, nullptr, s, nullptr); break;
    case 2: pass_b<N>(variant, true, T1, T2, tw, nullptr, s); break;
    default: pass_c<N>(variant, T2, tw, g.gamma_mode, nullptr, s); break;
  }
}

template <int N>
void launch(const SpecGeom& g, const double* F, const double* K4, double* out, cplx* T1, cplx* T2,
            const cplx* tw, hipStream_t s, const CGFuse* fz) {
  // pass D keeps the default tile: its p.q partial sums, and so their order,
  // follow that tile
  constexpr int TY = tile_z(N);
  static_assert(N % TY == 0 && TY % 2 == 0 && tz_cand(N, 0) == TY, "tile shapes");
  const int* sk = g.skip;
  const PencilTune& t = g.tune;
  pass_a<N>(K4 ? t.a : t.a0, F, K4, T1, tw, sk, s, fz);
  if (fz && fz->alpha_np > 0) launch_cg_alpha(fz->pq_part, fz->alpha_np, fz->cs, s);
  pass_b<N>(t.b, true, T1, T2, tw, sk, s);
  pass_c<N>(t.c, T2, tw, g.gamma_mode, sk, s, fz ? fz->dc : nullptr);
  pass_b<N>(t.b, false, T2, T1, tw, sk, s);
  const dim3 gd(N / TY, N, 9);
  if (fz && fz->upd_r)
    kD_zinv<N, 2><<<gd, kT, 0, s>>>(T1, nullptr, tw, sk, fz->upd_x ? fz->dir_p : nullptr,
                                    fz->rr_part, fz->upd_r, fz->upd_x, fz->cs);
  else if (fz && fz->dot_p)
    kD_zinv<N, 1><<<gd, kT, 0, s>>>(T1, out, tw, sk, fz->dot_p, fz->dot_part, nullptr, nullptr,
                                    nullptr);
  else
    kD_zinv<N, 0><<<gd, kT, 0, s>>>(T1, out, tw, sk, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// candidate value v of pass p (0 A with K4, 1 A without, 2 B, 3 C), 0 if none
template <int N>
int cand(int p, int v) {
  switch (p) {
    case 0: case 1: return v == 0 ? tz_cand(N, 0) : v == 1 ? tz_cand(N, 1) : v == 2 ? tz_cand(N, 2) : 0;
    case 2: return v == 0 ? ty_cand(N, 0) : v == 1 ? ty_cand(N, 1) : v == 2 ? ty_cand(N, 2) : 0;
    default:
      return v == 0 ? xl_cand(N, 0) : v == 1 ? xl_cand(N, 1) : v == 2 ? xl_cand(N, 2)
           : v == 3 ? xl_cand(N, 3) : v == 4 ? xl_cand(N, 4) : 0;
  }
}

template <int N>
constexpr int dot_parts() { return (N / tile_z(N)) * N * 9; }

}  // namespace fp

bool gkdf_pencil_supported(int N) {
  switch (N) {
    case 16: case 32: case 48: case 64: case 96: case 128: case 256: return true;
    default: return false;
  }
}

bool launch_gkdf_pencil(const SpecGeom& g, const double* F, const double* K4, double* out,
                        cplx* spec, const cplx* tw, hipStream_t s, const CGFuse* fuse) {
  cplx* T1 = spec;
  cplx* T2 = spec + 9 * g.S1;
  switch (g.N) {
    case 16: fp::launch<16>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 32: fp::launch<32>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 48: fp::launch<48>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 64: fp::launch<64>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 96: fp::launch<96>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 128: fp::launch<128>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    case 256: fp::launch<256>(g, F, K4, out, T1, T2, tw, s, fuse); return true;
    default: return false;
  }
}

int pencil_candidate(int N, int pass, int variant) {
  if (pass < 0 || pass > 3 || variant < 0 || variant >= kPencilCands) return 0;
  switch (N) {
    case 16: return fp::cand<16>(pass, variant);
    case 32: return fp::cand<32>(pass, variant);
    case 48: return fp::cand<48>(pass, variant);
    case 64: return fp::cand<64>(pass, variant);
    case 96: return fp::cand<96>(pass, variant);
    case 128: return fp::cand<128>(pass, variant);
    case 256: return fp::cand<256>(pass, variant);
    default: return 0;
  }
}

bool launch_pencil_pass(const SpecGeom& g, int pass, int variant, const double* F,
                        const double* K4, cplx* spec, const cplx* tw, hipStream_t s) {
  if (pencil_candidate(g.N, pass, variant) == 0) return false;
  cplx* T1 = spec;
  cplx* T2 = spec + 9 * g.S1;
  switch (g.N) {
    case 16: fp::launch_pass<16>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 32: fp::launch_pass<32>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 48: fp::launch_pass<48>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 64: fp::launch_pass<64>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 96: fp::launch_pass<96>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 128: fp::launch_pass<128>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    case 256: fp::launch_pass<256>(g, pass, variant, F, K4, T1, T2, tw, s); return true;
    default: return false;
  }
}

int gkdf_dot_parts(int N) {
  switch (N) {
    case 16: return fp::dot_parts<16>();
    case 32: return fp::dot_parts<32>();
    case 48: return fp::dot_parts<48>();
    case 64: return fp::dot_parts<64>();
    case 96: return fp::dot_parts<96>();
    case 128: return fp::dot_parts<128>();
    case 256: return fp::dot_parts<256>();
    default: return 0;
  }
}

}  // namespace cpfft
