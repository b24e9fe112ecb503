// cpfft_kernels.hip -- CDNA4 (gfx950) kernels of the CPFFT solver step.
//
//  * G_K_dF (src/G_K_dF.f:11-87) as five passes over an R2C half spectrum
//    held in HBM, [9][N][N][N/2+1] complex (x slowest):
//      1. z: fused K4:F contraction (ddot42n, G_K_dF.f:241-268) + real-to-
//         complex FFT along z, one workgroup per (x,y) pencil
//      2. y: complex FFT along y, one workgroup per (c, x, kz tile)
//      3. x: forward FFT along x, Green projection for all nine
//         components, inverse FFT along x, one workgroup per (ky, kz tile)
//      4. y: inverse complex FFT
//      5. z: complex-to-real FFT with the 1/N^3 scale
//    FFTs run in LDS as mixed-radix (4,2,3,5,generic) Stockham stages with
//    ping-pong buffers; twiddles come from a table of N-th roots of unity.
//  * BLAS-1 and reductions for CG, Newton norms and volume averages
//    (FFT_nr3.f:85-147, tangent_homo.f:62-66): deterministic two-stage
//    block reductions (fixed order, no floating-point atomics).
//  * drive_eps_sig (src/drive_eps_sig.f:16-339): one thread per voxel,
//    body in mm10_core.h.
#include "../../include/cpfft.h"
#include "green.h"
#include "kernels.h"

namespace cpfft {

static constexpr int kBlock = 256;

// ------------------------------------------------------------------ FFT
// Twiddle for exponent e (0 <= e < N: every caller's exponent t*k*N/(Ns*r)
// with t < r, k < Ns, or ((q t) mod r) * N/r, stays below N) of the N-th
// root table, conjugated for the inverse transform.
__device__ inline cplx tw_get(const cplx* __restrict__ tw, int, int e, double sgn) {
  cplx w = tw[e];
  w.im *= sgn;  // table holds exp(-2 pi i k/N); sgn=+1 forward, -1 inverse
  return w;
}

__device__ inline cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Batched in-LDS complex FFT.  Line L, element n lives at
// buf[L*ls + n*es]; tmp has the same layout.  fwd: +1 forward, -1 inverse
// (unscaled).  Result in buf.  All threads of the block must call it.
__device__ void lds_fft(cplx* buf, cplx* tmp, int nlines, int ls, int es, const FftPlan& pl,
                        const cplx* __restrict__ tw, double fwd) {
  const int N = pl.n;
  const int T = blockDim.x, tid = threadIdx.x;
  int Ns = 1;
  cplx* src = buf;
  cplx* dst = tmp;
  for (int st = 0; st < pl.nst; ++st) {
    const int r = pl.rad[st];
    const int nb = N / r;
    const int total = nlines * nb;
    const int tstride = N / (Ns * r);  // twiddle exponent step: exp(-2pi i t k/(Ns r))
    for (int b = tid; b < total; b += T) {
      const int line = b / nb;
      const int j = b - line * nb;
      const int k = j % Ns;
      const cplx* s = src + line * ls;
      cplx* d = dst + line * ls;
      const int o = (j / Ns) * Ns * r + k;
      if (r == 2) {
        cplx a0 = s[j * es];
        cplx a1 = cmul(s[(j + nb) * es], tw_get(tw, N, k * tstride, fwd));
        d[o * es] = cplx{a0.re + a1.re, a0.im + a1.im};
        d[(o + Ns) * es] = cplx{a0.re - a1.re, a0.im - a1.im};
      } else if (r == 4) {
        cplx a0 = s[j * es];
        cplx a1 = cmul(s[(j + nb) * es], tw_get(tw, N, k * tstride, fwd));
        cplx a2 = cmul(s[(j + 2 * nb) * es], tw_get(tw, N, 2 * k * tstride, fwd));
        cplx a3 = cmul(s[(j + 3 * nb) * es], tw_get(tw, N, 3 * k * tstride, fwd));
        cplx b0{a0.re + a2.re, a0.im + a2.im}, b1{a0.re - a2.re, a0.im - a2.im};
        cplx b2{a1.re + a3.re, a1.im + a3.im}, b3{a1.re - a3.re, a1.im - a3.im};
        // multiply b3 by -i (forward) or +i (inverse)
        cplx b3r{fwd * b3.im, -fwd * b3.re};
        d[o * es] = cplx{b0.re + b2.re, b0.im + b2.im};
        d[(o + Ns) * es] = cplx{b1.re + b3r.re, b1.im + b3r.im};
        d[(o + 2 * Ns) * es] = cplx{b0.re - b2.re, b0.im - b2.im};
        d[(o + 3 * Ns) * es] = cplx{b1.re - b3r.re, b1.im - b3r.im};
      } else {
        // generic radix r (3, 5, or a prime factor): direct DFT of size r
        cplx x[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          if (t < r) {
            cplx v = s[(j + t * nb) * es];
            x[t] = t ? cmul(v, tw_get(tw, N, t * k * tstride, fwd)) : v;
          }
        }
        const int rstep = N / r;  // r-th roots from the N-th root table
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          if (q < r) {
            cplx acc = x[0];
#pragma unroll
            for (int t = 1; t < 16; ++t) {
              if (t < r) {
                cplx w = tw_get(tw, N, ((q * t) % r) * rstep, fwd);
                cplx m = cmul(x[t], w);
                acc.re += m.re;
                acc.im += m.im;
              }
            }
            d[(o + q * Ns) * es] = acc;
          }
        }
      }
    }
    __syncthreads();
    cplx* t2 = src;
    src = dst;
    dst = t2;
    Ns *= r;
  }
  if (src != buf) {
    for (int b = tid; b < nlines * N; b += T) {
      int line = b / N, n = b - line * N;
      buf[line * ls + n * es] = src[line * ls + n * es];
    }
    __syncthreads();
  }
}

// pass 1: (x,y) pencils. out spec[c][x][y][kz]
__global__ __launch_bounds__(kBlock) void k_fft_z_fwd(SpecGeom g, const double* __restrict__ F,
                                                      const double* __restrict__ K4,
                                                      cplx* __restrict__ spec,
                                                      const cplx* __restrict__ tw) {
  if (g.skip && *g.skip) return;
  extern __shared__ cplx lds[];
  const int N = g.N, Nz = g.Nz;
  const long V = g.V;
  const long line0 = blockIdx.x;  // x*N + y
  cplx* buf = lds;
  cplx* tmp = lds + 9 * N;
  const long e0 = line0 * N;
  for (int z = threadIdx.x; z < N; z += blockDim.x) {
    const long e = e0 + z;
    double f[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) f[c] = F[c * V + e];
    if (K4) {
      // ddot42n: C_r = t1 + (((t2+t6)+(t4+t8)) + ((t3+t7)+(t5+t9)))
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        double k[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) k[c] = K4[(9 * r + c) * V + e];
        buf[r * N + z] = cplx{ddot42n_row(k, f), 0.0};
      }
    } else {
#pragma unroll
      for (int r = 0; r < 9; ++r) buf[r * N + z] = cplx{f[r], 0.0};
    }
  }
  __syncthreads();
  lds_fft(buf, tmp, 9, N, 1, g.plan, tw, 1.0);
  for (int i = threadIdx.x; i < 9 * Nz; i += blockDim.x) {
    int c = i / Nz, kz = i - c * Nz;
    spec[c * g.S1 + line0 * Nz + kz] = buf[c * N + kz];
  }
}

// pass 2 / 4: y transforms. grid: (ceil(Nz/ty), N, 9)
__global__ __launch_bounds__(kBlock) void k_fft_y(SpecGeom g, cplx* __restrict__ spec,
                                                  const cplx* __restrict__ tw, double fwd) {
  if (g.skip && *g.skip) return;
  extern __shared__ cplx lds[];
  const int N = g.N, Nz = g.Nz, ty = g.ty;
  const int kz0 = blockIdx.x * ty;
  const int x = blockIdx.y, c = blockIdx.z;
  const int nt = min(ty, Nz - kz0);
  cplx* base = spec + c * g.S1 + (long)x * N * Nz;
  cplx* buf = lds;
  cplx* tmp = lds + ty * N;
  // line t = kz column, element n = y; layout buf[t*N + n]
  for (int i = threadIdx.x; i < N * nt; i += blockDim.x) {
    int y = i / nt, t = i - y * nt;
    buf[t * N + y] = base[(long)y * Nz + kz0 + t];
  }
  __syncthreads();
  lds_fft(buf, tmp, nt, N, 1, g.plan, tw, fwd);
  for (int i = threadIdx.x; i < N * nt; i += blockDim.x) {
    int y = i / nt, t = i - y * nt;
    base[(long)y * Nz + kz0 + t] = buf[t * N + y];
  }
}

// pass 3: x transforms + projection. grid: (ceil(Nz/tz), N)
__global__ __launch_bounds__(kBlock) void k_proj_x(SpecGeom g, cplx* __restrict__ spec,
                                                   const cplx* __restrict__ tw) {
  if (g.skip && *g.skip) return;
  extern __shared__ cplx lds[];
  const int N = g.N, Nz = g.Nz, tz = g.tz;
  const int kz0 = blockIdx.x * tz;
  const int y = blockIdx.y;
  const int nt = min(tz, Nz - kz0);
  const int nl = 9 * nt;
  cplx* buf = lds;
  cplx* tmp = lds + 9 * tz * N;
  // line (c*nt + t), element x; layout buf[(c*nt+t)*N + x]
  for (int i = threadIdx.x; i < 9 * N * nt; i += blockDim.x) {
    int t = i % nt;
    int rest = i / nt;
    int x = rest % N, c = rest / N;
    buf[(c * nt + t) * N + x] = spec[c * g.S1 + ((long)x * N + y) * Nz + kz0 + t];
  }
  __syncthreads();
  lds_fft(buf, tmp, nl, N, 1, g.plan, tw, 1.0);
  // projection at every (x, kz) point of the tile
  for (int i = threadIdx.x; i < N * nt; i += blockDim.x) {
    int x = i / nt, t = i - x * nt;
    cplx A[9], out[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) A[c] = buf[(c * nt + t) * N + x];
    project9(N, g.gamma_mode, x, y, kz0 + t, A, out);
#pragma unroll
    for (int c = 0; c < 9; ++c) buf[(c * nt + t) * N + x] = out[c];
  }
  __syncthreads();
  lds_fft(buf, tmp, nl, N, 1, g.plan, tw, -1.0);
  for (int i = threadIdx.x; i < 9 * N * nt; i += blockDim.x) {
    int t = i % nt;
    int rest = i / nt;
    int x = rest % N, c = rest / N;
    spec[c * g.S1 + ((long)x * N + y) * Nz + kz0 + t] = buf[(c * nt + t) * N + x];
  }
}

// pass 5: complex-to-real along z, scale 1/N^3
__global__ __launch_bounds__(kBlock) void k_fft_z_inv(SpecGeom g, const cplx* __restrict__ spec,
                                                      double* __restrict__ out,
                                                      const cplx* __restrict__ tw) {
  if (g.skip && *g.skip) return;
  extern __shared__ cplx lds[];
  const int N = g.N, Nz = g.Nz;
  const long V = g.V;
  const long line0 = blockIdx.x;
  cplx* buf = lds;
  cplx* tmp = lds + 9 * N;
  for (int i = threadIdx.x; i < 9 * N; i += blockDim.x) {
    int c = i / N, z = i - c * N;
    cplx v;
    if (z < Nz) {
      v = spec[c * g.S1 + line0 * Nz + z];
    } else {
      cplx w = spec[c * g.S1 + line0 * Nz + (N - z)];
      v = cplx{w.re, -w.im};
    }
    buf[c * N + z] = v;
  }
  __syncthreads();
  lds_fft(buf, tmp, 9, N, 1, g.plan, tw, -1.0);
  const double sc = 1.0 / (double)V;
  for (int i = threadIdx.x; i < 9 * N; i += blockDim.x) {
    int c = i / N, z = i - c * N;
    out[c * V + line0 * N + z] = buf[c * N + z].re * sc;
  }
}

const char* make_spec_geom(int N, int gamma_mode, int lds, SpecGeom& g) {
  const int maxlds = lds > 65536 ? 65536 : lds;
  g = SpecGeom{};
  g.N = N;
  g.Nz = N / 2 + 1;
  g.V = (long)N * N * N;
  g.S1 = (long)N * N * g.Nz;
  g.plan = make_plan(N);
  g.gamma_mode = gamma_mode;
  g.ty = 8;
  g.tz = 1;
  while (g.tz < 4 && 2 * 9 * (long)(2 * g.tz) * N * (long)sizeof(cplx) <= maxlds) g.tz *= 2;
  if (!gkdf_pencil_supported(N)) {  // limits of the generic runtime-plan path
    if (2 * 9 * (long)N * (long)sizeof(cplx) > maxlds)
      return "grid edge too large for the LDS of this device";
    for (int s = 0; s < g.plan.nst; ++s)
      if (g.plan.rad[s] > 16) return "N has a prime factor above 16";
  }
  return nullptr;
}

void launch_gkdf(const SpecGeom& g, const double* F, const double* K4, double* out, cplx* spec,
                 const cplx* tw, hipStream_t s) {
  if (launch_gkdf_pencil(g, F, K4, out, spec, tw, s)) return;
  const int N = g.N;
  size_t lz = 2 * 9 * (size_t)N * sizeof(cplx);
  k_fft_z_fwd<<<dim3((unsigned)((long)N * N)), kBlock, lz, s>>>(g, F, K4, spec, tw);
  size_t ly = 2 * (size_t)g.ty * N * sizeof(cplx);
  dim3 gy((g.Nz + g.ty - 1) / g.ty, N, 9);
  k_fft_y<<<gy, kBlock, ly, s>>>(g, spec, tw, 1.0);
  size_t lx = 2 * 9 * (size_t)g.tz * N * sizeof(cplx);
  dim3 gx((g.Nz + g.tz - 1) / g.tz, N);
  k_proj_x<<<gx, kBlock, lx, s>>>(g, spec, tw);
  k_fft_y<<<gy, kBlock, ly, s>>>(g, spec, tw, -1.0);
  k_fft_z_inv<<<dim3((unsigned)((long)N * N)), kBlock, lz, s>>>(g, spec, out, tw);
}

// ------------------------------------------------------------ reductions
static constexpr int kRedBlocks = 1024;

int reduce_partials_count(long) { return kRedBlocks * 9; }

// halving-tree sum over a block of kBlock threads (tree_sum0), valid in
// thread 0; the closing barrier lets a caller reuse sh in a loop
__device__ inline double block_sum(double v, double* sh) {
  const double r = tree_sum0<kBlock>(v, sh);
  __syncthreads();
  return r;
}

// fixed partition: block b handles the contiguous range [b*chunk, (b+1)*chunk)
__global__ __launch_bounds__(kBlock) void k_dot_part(const double* __restrict__ x,
                                                     const double* __restrict__ y, long n,
                                                     double* __restrict__ part) {
  __shared__ double sh[kBlock];
  long chunk = (n + gridDim.x - 1) / gridDim.x;
  long a = blockIdx.x * chunk, b = min(n, a + chunk);
  double acc = 0.0;
  for (long i = a + threadIdx.x; i < b; i += blockDim.x) acc += x[i] * y[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_parts(const double* __restrict__ part, int np,
                                                      int nres, double* __restrict__ out) {
  __shared__ double sh[kBlock];
  for (int r = 0; r < nres; ++r) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[r * np + i];
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[r] = s;
  }
}

void launch_dot(const double* x, const double* y, long n, double* partials, double* d_out,
                hipStream_t s) {
  k_dot_part<<<kRedBlocks, kBlock, 0, s>>>(x, y, n, partials);
  k_sum_parts<<<1, kBlock, 0, s>>>(partials, kRedBlocks, 1, d_out);
}

// mean of each of the 9 components: partial sums per component
__global__ __launch_bounds__(kBlock) void k_sum9_part(const double* __restrict__ x, long V,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock];
  long chunk = (V + gridDim.x - 1) / gridDim.x;
  long a = blockIdx.x * chunk, b = min(V, a + chunk);
  for (int c = 0; c < 9; ++c) {
    double acc = 0.0;
    for (long i = a + threadIdx.x; i < b; i += blockDim.x) acc += x[c * V + i];
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[c * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_scale9(double* out, double sc) {
  if (threadIdx.x < 9) out[threadIdx.x] *= sc;
}

void launch_mean9(const double* x, long V, double* partials, double* d_out, hipStream_t s) {
  k_sum9_part<<<kRedBlocks, kBlock, 0, s>>>(x, V, partials);
  k_sum_parts<<<1, kBlock, 0, s>>>(partials, kRedBlocks, 9, d_out);
  k_scale9<<<1, 64, 0, s>>>(d_out, 1.0 / (double)V);
}

// tangent_homo: out[j] = sum over (m,e) of K4[9j+m][e] C[m][e]
__global__ __launch_bounds__(kBlock) void k_k4dot9_part(const double* __restrict__ K4,
                                                        const double* __restrict__ C, long V,
                                                        double* __restrict__ part) {
  __shared__ double sh[kBlock];
  long n = 9 * V;
  long chunk = (n + gridDim.x - 1) / gridDim.x;
  long a = blockIdx.x * chunk, b = min(n, a + chunk);
  for (int j = 0; j < 9; ++j) {
    const double* Kj = K4 + (long)9 * j * V;
    double acc = 0.0;
    for (long i = a + threadIdx.x; i < b; i += blockDim.x) acc += Kj[i] * C[i];
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[j * gridDim.x + blockIdx.x] = s;
  }
}

void launch_k4_dot9(long V, const double* K4, const double* C, double* partials, double* d_out,
                    hipStream_t s) {
  k_k4dot9_part<<<kRedBlocks, kBlock, 0, s>>>(K4, C, V, partials);
  k_sum_parts<<<1, kBlock, 0, s>>>(partials, kRedBlocks, 9, d_out);
  k_scale9<<<1, 64, 0, s>>>(d_out, 1.0 / (double)V);
}

// --------------------------------------------------------------- BLAS-1
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}
__global__ void k_scal(long n, double a, double* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = a * x[i];
}
__global__ void k_copy(long n, const double* __restrict__ x, double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i];
}
struct V9 {
  double v[9];
};
__global__ void k_fill9(long V, V9 v, double* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < 9 * V; i += (long)gridDim.x * blockDim.x)
    x[i] = v.v[i / V];
}
__global__ void k_fill(long n, double v, double* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_state_init(long V, int W, double tt0, double* __restrict__ st) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (long)W * V;
       i += (long)gridDim.x * blockDim.x) {
    const long w = i / V;
    double v = 0.0;
    if (w == CPFFT_W_RP + 0 || w == CPFFT_W_RP + 4 || w == CPFFT_W_RP + 8) v = 1.0;
    if (w == CPFFT_W_TT) v = tt0;
    st[i] = v;
  }
}
__global__ void k_add1(long V, double* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < V; i += (long)gridDim.x * blockDim.x)
    x[i] = x[i] + 1.0;
}
__global__ void k_gather_col(long V, int col, const double* __restrict__ K4, double* __restrict__ C) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < 9 * V; i += (long)gridDim.x * blockDim.x) {
    long j = i / V, e = i - j * V;
    C[i] = K4[(9 * j + col) * V + e];
  }
}
__global__ __launch_bounds__(kBlock) void k_cg_update(long n, double a, const double* __restrict__ p,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock];
  long chunk = (n + gridDim.x - 1) / gridDim.x;
  long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double acc = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    x[i] = x[i] + a * p[i];
    double ri = r[i] - a * q[i];
    r[i] = ri;
    acc += ri * ri;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_cg_dir(long n, double b, const double* __restrict__ r, double* __restrict__ p) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = r[i] + b * p[i];
}

static inline unsigned grid_for(long n) {
  long g = (n + kBlock - 1) / kBlock;
  return (unsigned)(g > 8192 ? 8192 : (g < 1 ? 1 : g));
}

void launch_axpy(long n, double a, const double* x, double* y, hipStream_t s) {
  k_axpy<<<grid_for(n), kBlock, 0, s>>>(n, a, x, y);
}
void launch_scal(long n, double a, double* x, hipStream_t s) {
  k_scal<<<grid_for(n), kBlock, 0, s>>>(n, a, x);
}
void launch_copy(long n, const double* x, double* y, hipStream_t s) {
  k_copy<<<grid_for(n), kBlock, 0, s>>>(n, x, y);
}
void launch_fill9(long V, const double* v9, double* x, hipStream_t s) {
  V9 v;
  for (int i = 0; i < 9; ++i) v.v[i] = v9[i];
  k_fill9<<<grid_for(9 * V), kBlock, 0, s>>>(V, v, x);
}
void launch_fill(long n, double v, double* x, hipStream_t s) {
  k_fill<<<grid_for(n), kBlock, 0, s>>>(n, v, x);
}
__global__ void k_fill_f32(long n, float v, float* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}
void launch_fill_f32(long n, float v, float* x, hipStream_t s) {
  k_fill_f32<<<grid_for(n), kBlock, 0, s>>>(n, v, x);
}
void launch_state_init(long V, int W, double tt0, double* state, hipStream_t s) {
  k_state_init<<<grid_for((long)W * V), kBlock, 0, s>>>(V, W, tt0, state);
}
void launch_add_identity_comp(long V, int comp, double* x, hipStream_t s) {
  k_add1<<<grid_for(V), kBlock, 0, s>>>(V, x + (long)comp * V);
}
void launch_gather_k4_col(long V, int i, const double* K4, double* C, hipStream_t s) {
  k_gather_col<<<grid_for(9 * V), kBlock, 0, s>>>(V, i, K4, C);
}
void launch_cg_update(long n, double a, const double* p, const double* q, double* x, double* r,
                      double* partials, double* d_out, hipStream_t s) {
  k_cg_update<<<kRedBlocks, kBlock, 0, s>>>(n, a, p, q, x, r, partials);
  k_sum_parts<<<1, kBlock, 0, s>>>(partials, kRedBlocks, 1, d_out);
}
void launch_cg_dir(long n, double b, const double* r, double* p, hipStream_t s) {
  k_cg_dir<<<grid_for(n), kBlock, 0, s>>>(n, b, r, p);
}

// ------------------------------------------------- device-scalar CG
__global__ __launch_bounds__(kBlock) void k_dot_part_s(const double* __restrict__ x,
                                                       const double* __restrict__ y, long n,
                                                       double* __restrict__ part,
                                                       const CGScalars* __restrict__ cs) {
  if (cs->done) return;
  __shared__ double sh[kBlock];
  long chunk = (n + gridDim.x - 1) / gridDim.x;
  long a = blockIdx.x * chunk, b = min(n, a + chunk);
  double acc = 0.0;
  for (long i = a + threadIdx.x; i < b; i += blockDim.x) acc += x[i] * y[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// alpha = rr / (p.q) from the partial sums (same order as k_sum_parts)
// block b serves system b: partial sums part + b pstride, scalars cs + b
__global__ __launch_bounds__(kBlock) void k_cg_alpha(const double* __restrict__ part, int np,
                                                     CGScalars* __restrict__ cs, long pstride) {
  cs += blockIdx.x;
  part += blockIdx.x * pstride;
  if (cs->done) return;
  __shared__ double sh[kBlock];
  const double pq = ordered_sum<kBlock>(part, np, sh);
  if (threadIdx.x == 0) cs->alpha = cs->rr / pq;
}

template <bool XUPD>
__global__ __launch_bounds__(kBlock) void k_cg_update_s(long n, const CGScalars* __restrict__ cs,
                                                        const double* __restrict__ p,
                                                        const double* __restrict__ q,
                                                        double* __restrict__ x,
                                                        double* __restrict__ r,
                                                        double* __restrict__ part) {
  if (cs->done) return;
  const double a = cs->alpha;
  __shared__ double sh[kBlock];
  long chunk = (n + gridDim.x - 1) / gridDim.x;
  long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double acc = 0.0;
  // four elements per pass with their loads issued together (more bytes in
  // flight per wave); the sum keeps the element order
  long i = lo + threadIdx.x;
  for (; i + 3 * kBlock < hi; i += 4 * kBlock) {
    double pv[4], qv[4], xv[4], rv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (XUPD) {
        pv[u] = p[i + u * kBlock];
        xv[u] = x[i + u * kBlock];
      }
      qv[u] = q[i + u * kBlock];
      rv[u] = r[i + u * kBlock];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (XUPD) x[i + u * kBlock] = xv[u] + a * pv[u];
      const double ri = rv[u] - a * qv[u];
      r[i + u * kBlock] = ri;
      acc += ri * ri;
    }
  }
  for (; i < hi; i += kBlock) {
    if constexpr (XUPD) x[i] = x[i] + a * p[i];
    double ri = r[i] - a * q[i];
    r[i] = ri;
    acc += ri * ri;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// |r| test from the r.r partial sums (FFT_nr3.f:304-306); with m, first
// m[c] += alpha * dc[c] * scale (this iteration's alpha)
// block b serves system b (part + b pstride, cs + b, dc and m + 9 b)
__global__ __launch_bounds__(kBlock) void k_cg_check(const double* __restrict__ part, int np,
                                                     CGScalars* __restrict__ cs,
                                                     const double* __restrict__ dc,
                                                     double* __restrict__ m, double scale,
                                                     long pstride) {
  cs += blockIdx.x;
  part += blockIdx.x * pstride;
  if (dc) dc += 9 * blockIdx.x;
  if (m) m += 9 * blockIdx.x;
  if (cs->done) return;
  __shared__ double sh[kBlock];
  const double rr1 = ordered_sum<kBlock>(part, np, sh);
  if (threadIdx.x == 0) {
    if (m)
      for (int c = 0; c < 9; ++c) m[c] += cs->alpha * (dc[c] * scale);
    cg_check(cs, rr1);
  }
}
__global__ void k_cg_dir_s(long n, const CGScalars* __restrict__ cs, const double* __restrict__ r,
                           double* __restrict__ p) {
  if (cs->done) return;
  const double b = cs->beta;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = r[i] + b * p[i];
}

void launch_cg_update_check(long n, CGScalars* cs, const double* p, const double* q, double* x,
                            double* r, double* partials, hipStream_t s, bool xupd,
                            const double* dc, double* m, double scale) {
  if (xupd)
    k_cg_update_s<true><<<kRedBlocks, kBlock, 0, s>>>(n, cs, p, q, x, r, partials);
  else
    k_cg_update_s<false><<<kRedBlocks, kBlock, 0, s>>>(n, cs, p, q, x, r, partials);
  k_cg_check<<<1, kBlock, 0, s>>>(partials, kRedBlocks, cs, dc, m, scale, 0);
}
void launch_cg_alpha(const double* part, int np, CGScalars* cs, hipStream_t s, int nsys,
                     long pstride) {
  k_cg_alpha<<<nsys, kBlock, 0, s>>>(part, np, cs, pstride);
}
void launch_cg_check(const double* part, int np, CGScalars* cs, hipStream_t s, const double* dc,
                     double* m, double scale, int nsys, long pstride) {
  k_cg_check<<<nsys, kBlock, 0, s>>>(part, np, cs, dc, m, scale, pstride);
}

void launch_cg_tail(long n, CGScalars* cs, const double* p, const double* q, double* x,
                    double* r, double* partials, hipStream_t s) {
  k_dot_part_s<<<kRedBlocks, kBlock, 0, s>>>(p, q, n, partials, cs);
  k_cg_alpha<<<1, kBlock, 0, s>>>(partials, kRedBlocks, cs, 0);
  k_cg_update_s<true><<<kRedBlocks, kBlock, 0, s>>>(n, cs, p, q, x, r, partials);
  k_cg_check<<<1, kBlock, 0, s>>>(partials, kRedBlocks, cs, nullptr, nullptr, 0.0, 0);
}
void launch_cg_dir_dev(long n, const CGScalars* cs, const double* r, double* p, hipStream_t s) {
  k_cg_dir_s<<<grid_for(n), kBlock, 0, s>>>(n, cs, r, p);
}

// ------------------------------------------ multi-RHS K4 contraction
// q_i = K4 : p_i for the right-hand sides i < nrhs whose solve is still
// running (tangent_homo's nine solves share K4, which is read once).
template <bool DIR>
__global__ __launch_bounds__(kBlock) void k_k4_contract_multi(long V, const double* __restrict__ K4,
                                                              Ptr9 pp, const CGScalars* __restrict__ cs,
                                                              int nrhs, Ptr9r rr,
                                                              double* __restrict__ pqpart) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool valid = e < V;  // no early return: pqpart sums use a block barrier
  bool any = false;
  for (int i = 0; i < nrhs; ++i) any |= cs[i].done == 0;
  if (!any) return;  // the same in every thread of the launch
  double k[81];
#pragma unroll
  for (int j = 0; j < 81; ++j) k[j] = valid ? __builtin_nontemporal_load(K4 + j * V + e) : 0.0;
  // r and p of right-hand side i in buffer i % 2: those of the next running
  // one are loaded before the current one is computed and stored, so their
  // latency overlaps the work (the done flags are the same in every thread)
  auto running = [&](int i) { return i < nrhs && cs[i].done == 0; };
  double rb[2][9], pb[2][9];
  auto load = [&](int i, double (&r9)[9], double (&p9)[9]) {
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      if (DIR) r9[c] = valid ? __builtin_nontemporal_load(rr.r[i] + c * V + e) : 0.0;
      p9[c] = valid ? __builtin_nontemporal_load(pp.p[i] + c * V + e) : 0.0;
    }
  };
  // pqpart: per right-hand side, the wave sums of p_i . (K4:p_i)
  __shared__ double wred[9][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (running(0)) load(0, rb[0], pb[0]);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    if (i + 1 < 9 && running(i + 1)) load(i + 1, rb[(i + 1) & 1], pb[(i + 1) & 1]);
    if (!running(i)) continue;
    double f[9];
    if (DIR) {
      const double beta = cs[i].beta;
      double* pw = const_cast<double*>(pp.p[i]);
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        f[c] = rb[i & 1][c] + beta * pb[i & 1][c];
        if (valid) __builtin_nontemporal_store(f[c], pw + c * V + e);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 9; ++c) f[c] = pb[i & 1][c];
    }
    double pq = 0.0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const double q = ddot42n_row(k + 9 * r, f);
      if (valid) __builtin_nontemporal_store(q, pp.q[i] + r * V + e);
      pq += q * f[r];
    }
    if (pqpart) {
      // butterfly over the wave: every lane ends with the same sum
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) pq += __shfl_xor(pq, off);
      if (lane == 0) wred[i][wave] = pq;
    }
  }
  if (pqpart) {
    __syncthreads();
    const int i = threadIdx.x;
    if (running(i))
      pqpart[(long)i * gridDim.x + blockIdx.x] = (wred[i][0] + wred[i][1]) + (wred[i][2] + wred[i][3]);
  }
}

int k4_contract_multi_blocks(long V) { return (int)((V + kBlock - 1) / kBlock); }

void launch_k4_contract_multi(long V, const double* K4, const Ptr9& pp, const CGScalars* cs,
                              int nrhs, hipStream_t s, const Ptr9r* rr, double* pqpart) {
  const unsigned g = (unsigned)k4_contract_multi_blocks(V);
  if (rr)
    k_k4_contract_multi<true><<<g, kBlock, 0, s>>>(V, K4, pp, cs, nrhs, *rr, pqpart);
  else
    k_k4_contract_multi<false><<<g, kBlock, 0, s>>>(V, K4, pp, cs, nrhs, Ptr9r{}, pqpart);
}

// ------------------------------------------------------- stress update
template <int NS>
__global__ __launch_bounds__(64) void k_drive_eps_sig(long V, MatParams P,
                                                      const Orient<NS>* __restrict__ orients,
                                                      const int32_t* __restrict__ oid, int step,
                                                      int iter, const double* __restrict__ Fn,
                                                      const double* __restrict__ Fn1,
                                                      const double* __restrict__ sn,
                                                      double* __restrict__ snp1,
                                                      double* __restrict__ Pn1,
                                                      double* __restrict__ tan6,
                                                      const float* __restrict__ ws,
                                                      int warm_prev,
                                                      long long* __restrict__ counters) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  Counters cnt{0, 0, 0, 0, 0, 0};
  if (e < V) {
    const Orient<NS>& o = orients[oid[e]];
    // this step's previous global iterate (stress, slip resistance), read
    // before the update overwrites it
    double xp[7];
    if (warm_prev) {
#pragma unroll
      for (int k = 0; k < 6; ++k) xp[k] = snp1[k * V + e];
      xp[6] = snp1[15 * V + e];
    }
    voxel_update<NS>(
        P, o, step, iter,
        [&](int which, int c) { return which ? Fn1[c * V + e] : Fn[c * V + e]; },
        [&](int w) { return sn[w * V + e]; },
        [&](int w, double v) { snp1[w * V + e] = v; },
        [&](int c, double v) { Pn1[c * V + e] = v; },
        [&](int c, double v) { tan6[c * V + e] = v; }, cnt, ws ? ws + e : nullptr, V,
        warm_prev ? xp : nullptr);
  }
  // wavefront-level sums of the counters, one integer atomic per wavefront
  long long c[6] = {cnt.pred_it, cnt.upd_it, cnt.cuts, cnt.failed, cnt.ls_evals, cnt.ws_fallback};
  for (int k = 0; k < 6; ++k) {
    long long v = c[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long*)&counters[k], (unsigned long long)v);
  }
  // local iterations the wavefront issues: 64 lanes times its slowest lane's
  // count (counters[6]; against counters[0] + counters[1], the lanes' own
  // iterations, it measures the divergence of the local solves)
  long long m = cnt.pred_it + cnt.upd_it;
  for (int off = 32; off > 0; off >>= 1) m = max(m, (long long)__shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicAdd((unsigned long long*)&counters[6], (unsigned long long)(64 * m));
}

// K4 = dP/dF from the tangent of k_drive_eps_sig (cep2A)
__global__ __launch_bounds__(64) void k_cep2A(long V, const double* __restrict__ Fn,
                                              const double* __restrict__ Fn1,
                                              const double* __restrict__ snp1,
                                              const double* __restrict__ tan6,
                                              double* __restrict__ K4) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= V) return;
  double t6[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) t6[k] = snp1[k * V + e];
  voxel_dPdF([&](int which, int c) { return which ? Fn1[c * V + e] : Fn[c * V + e]; }, t6,
             [&](int c) { return tan6[c * V + e]; },
             [&](int c, double v) { K4[c * V + e] = v; });
}

template <int NS>
void launch_drive_eps_sig(long V, const MatParams& P, const Orient<NS>* orients,
                          const int32_t* orient_id, int step, int iter, const double* Fn,
                          const double* Fn1, const double* sn, double* snp1, double* Pn1,
                          double* tan6, double* K4, const float* ws, int warm_prev,
                          long long* counters, hipStream_t s) {
  const int bs = 64;
  unsigned g = (unsigned)((V + bs - 1) / bs);
  k_drive_eps_sig<NS><<<g, bs, 0, s>>>(V, P, orients, orient_id, step, iter, Fn, Fn1, sn, snp1,
                                       Pn1, tan6, ws, warm_prev, counters);
  k_cep2A<<<g, bs, 0, s>>>(V, Fn, Fn1, snp1, tan6, K4);
}

// ----------------------------------------------- warm-start features
// rows r < nrows of voxels e = e0 + r * stride: raw network inputs into
// X[r][18 + 2 NS]; with T != nullptr also the converged slip increments of
// state n+1, as u_s (slip_to_u), into T[r][NS] (training pairs)
template <int NS>
__global__ __launch_bounds__(64) void k_mlp_features(long V, MatParams P,
                                                     const Orient<NS>* __restrict__ orients,
                                                     const int32_t* __restrict__ oid, int step,
                                                     const double* __restrict__ Fn,
                                                     const double* __restrict__ Fn1,
                                                     const double* __restrict__ sn,
                                                     const double* __restrict__ snp1, long e0,
                                                     long stride, long nrows,
                                                     float* __restrict__ X, float* __restrict__ T) {
  constexpr int NIN = 18 + 2 * NS;
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const long e = e0 + r * stride;
  if (e >= V) return;
  float f[NIN];
  voxel_features<NS>(P, orients[oid[e]], step,
                     [&](int which, int c) { return which ? Fn1[c * V + e] : Fn[c * V + e]; },
                     [&](int w) { return sn[w * V + e]; }, f);
  for (int k = 0; k < NIN; ++k) X[r * NIN + k] = f[k];
  if (T) {
    const double dg = P.alter_mode ? P.eps_dot_0_y * P.dt : 1.0;
    for (int k = 0; k < NS; ++k) T[r * NS + k] = (float)slip_to_u(P, dg, snp1[(23 + k) * V + e]);
  }
}

// every voxel's network inputs, normalised ((x - mean) / sd, nrm =
// [mean | sd]) and rounded to BF16, into Xv[e][NINP] with the pad columns
// zero: the A operand of the inference kernel
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (uint16_t)(u >> 16);
}
template <int NS>
__global__ __launch_bounds__(64) void k_mlp_features_norm(long V, MatParams P,
                                                          const Orient<NS>* __restrict__ orients,
                                                          const int32_t* __restrict__ oid,
                                                          int step, const double* __restrict__ Fn,
                                                          const double* __restrict__ Fn1,
                                                          const double* __restrict__ sn,
                                                          const float* __restrict__ nrm, int ninp,
                                                          uint16_t* __restrict__ Xv) {
  constexpr int NIN = 18 + 2 * NS;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= V) return;
  float f[NIN];
  voxel_features<NS>(P, orients[oid[e]], step,
                     [&](int which, int c) { return which ? Fn1[c * V + e] : Fn[c * V + e]; },
                     [&](int w) { return sn[w * V + e]; }, f);
  uint16_t* x = Xv + e * ninp;
  // NIN is even; compile-time indexes keep f in registers
#pragma unroll
  for (int k = 0; k < NIN; k += 2) {
    const uint16_t lo = f32_to_bf16((f[k] - nrm[k]) / nrm[NIN + k]);
    const uint16_t hi = f32_to_bf16((f[k + 1] - nrm[k + 1]) / nrm[NIN + k + 1]);
    *reinterpret_cast<uint32_t*>(x + k) = (uint32_t)lo | ((uint32_t)hi << 16);
  }
  for (int k = NIN; k < ninp; k += 2) *reinterpret_cast<uint32_t*>(x + k) = 0u;
}

template <int NS>
void launch_mlp_features(long V, const MatParams& P, const Orient<NS>* orients, const int32_t* oid,
                         int step, const double* Fn, const double* Fn1, const double* sn,
                         const double* snp1, long e0, long stride, long nrows, float* X, float* T,
                         hipStream_t s) {
  k_mlp_features<NS><<<(unsigned)((nrows + 63) / 64), 64, 0, s>>>(
      V, P, orients, oid, step, Fn, Fn1, sn, snp1, e0, stride, nrows, X, T);
}
template <int NS>
void launch_mlp_features_norm(long V, const MatParams& P, const Orient<NS>* orients,
                              const int32_t* oid, int step, const double* Fn, const double* Fn1,
                              const double* sn, const float* nrm, int ninp, uint16_t* Xv,
                              hipStream_t s) {
  k_mlp_features_norm<NS><<<(unsigned)((V + 63) / 64), 64, 0, s>>>(V, P, orients, oid, step, Fn,
                                                                   Fn1, sn, nrm, ninp, Xv);
}
#define CPFFT_INST_FEATURES(NS)                                                                  \
  template void launch_mlp_features<NS>(long, const MatParams&, const Orient<NS>*, const int32_t*, \
                                        int, const double*, const double*, const double*,         \
                                        const double*, long, long, long, float*, float*,          \
                                        hipStream_t);                                             \
  template void launch_mlp_features_norm<NS>(long, const MatParams&, const Orient<NS>*,           \
                                             const int32_t*, int, const double*, const double*,   \
                                             const double*, const float*, int, uint16_t*,         \
                                             hipStream_t);
CPFFT_INST_FEATURES(12)
CPFFT_INST_FEATURES(24)
#undef CPFFT_INST_FEATURES

template void launch_drive_eps_sig<12>(long, const MatParams&, const Orient<12>*, const int32_t*,
                                       int, int, const double*, const double*, const double*,
                                       double*, double*, double*, double*, const float*, int,
                                       long long*, hipStream_t);
template void launch_drive_eps_sig<24>(long, const MatParams&, const Orient<24>*, const int32_t*,
                                       int, int, const double*, const double*, const double*,
                                       double*, double*, double*, double*, const float*, int,
                                       long long*, hipStream_t);
template void launch_drive_eps_sig<48>(long, const MatParams&, const Orient<48>*, const int32_t*,
                                       int, int, const double*, const double*, const double*,
                                       double*, double*, double*, double*, const float*, int,
                                       long long*, hipStream_t);

}  // namespace cpfft
