// f2disp.hip -- nodal displacements from the deformation gradient
// (src/f2disp.f:16-175), without a sparse factorisation.
//
// The reference meshes the N^3 voxels with eight-node bricks on an
// (N+1)^3-node grid (oumodel.f blkgen), assembles A = sum_e int B^T B with
// 2x2x2 Gauss points (form_BTB, f2disp.f:202-430) and b = sum_e w detJ
// B(0)^T F^T with the one-point rule (f2disp.f:97-147), holds node 1 at zero
// and solves A x = b for the three position components with PARDISO
// (f2disp.f:163-165); u = x - X (f2disp.f:169-174).
//
// On the regular brick grid A is a sum of Kronecker products of the 1-D
// stiffness K = tridiag(-1, 2, -1) (end diagonal 1) and mass
// M = tridiag(1, 4, 1)/6 (end diagonal 2): A = a_x K(x)M(x)M + a_y M(x)K(x)M +
// a_z M(x)M(x)K with a_x = h_y h_z / h_x etc.  With W = diag(1/2, 1, .., 1,
// 1/2) and V the DCT-I matrix V[j][k] = cos(pi j k / N), K V = W V Lk and
// M V = W V Lm, Lk = 4 sin^2(pi k / 2N), Lm = (2 + cos(pi k / N))/3, and
// V^T W V = diag(c_k) with c_0 = c_N = N, else N/2.  So
//   x = V3 [ (V3^T b) / (Lambda c) ],  Lambda = a_x Lk Lm Lm + ...,
// the (0,0,0) term dropped (A's null space is the constants and b sums to
// zero), then x -= x(node 1): the same x the reduced system gives.  Each
// V along an axis is a dense (N+1)-point transform of every line, done as a
// small matrix product per block with 4 x 4 register blocking (FP64 vector
// units).
//
// Scratch layout [3][P^3] with z fastest, m = iz + P iy + P^2 ix (P = N + 1),
// like F's voxels, so the right-hand side gathers F with contiguous loads;
// the output kernel transposes to the reference's node numbering
// n = ix + P iy + P^2 iz.
#include "kernels.h"

namespace cpfft {
namespace {

constexpr int kFT = 256;  // threads per block (elementwise kernels)
constexpr int kB = 4;     // register block of the line transforms (k and t)

// padded row length of the cosine table: a multiple of kB
inline int cos_stride(int P) { return (P + kB - 1) / kB * kB; }

// C[j][k] = cos(pi j k / N) for k < P, 0 in the padding; the argument is
// reduced exactly mod 2N
__global__ __launch_bounds__(kFT) void k_f2d_cos(int N, int Pc, double* __restrict__ C) {
  const int P = N + 1;
  const long i = (long)blockIdx.x * kFT + threadIdx.x;
  if (i >= (long)P * Pc) return;
  const int j = (int)(i / Pc), k = (int)(i - (long)j * Pc);
  C[i] = k < P ? cospi((double)((j * k) % (2 * N)) / (double)N) : 0.0;
}

// b[i][m] = sum over the bricks around node (ix, iy, iz) of
// sum_k s_k coef_k F_ik, s_k = +1 where the node is the brick's upper corner
// along k, else -1; coef_k = (h_x h_y h_z) / (4 h_k) (the one-point rule:
// w det J dN/dX_k).  Consecutive threads take consecutive iz.
__global__ __launch_bounds__(kFT) void k_f2d_rhs(int N, const double* __restrict__ F, double c0,
                                                 double c1, double c2, double* __restrict__ b) {
  const int P = N + 1;
  const long P3 = (long)P * P * P, V = (long)N * N * N;
  const long m = (long)blockIdx.x * kFT + threadIdx.x;
  if (m >= P3) return;
  const int iz = (int)(m % P), iy = (int)((m / P) % P), ix = (int)(m / ((long)P * P));
  double acc[3] = {0.0, 0.0, 0.0};
  for (int ex = ix - 1; ex <= ix; ++ex) {
    if (ex < 0 || ex >= N) continue;
    const double sx = (ex < ix ? 1.0 : -1.0) * c0;
    for (int ey = iy - 1; ey <= iy; ++ey) {
      if (ey < 0 || ey >= N) continue;
      const double sy = (ey < iy ? 1.0 : -1.0) * c1;
      for (int ez = iz - 1; ez <= iz; ++ez) {
        if (ez < 0 || ez >= N) continue;
        const double sz = (ez < iz ? 1.0 : -1.0) * c2;
        const long v = ((long)ex * N + ey) * N + ez;
#pragma unroll
        for (int i = 0; i < 3; ++i)
          acc[i] += sx * F[(3 * i + 0) * V + v] + sy * F[(3 * i + 1) * V + v] +
                    sz * F[(3 * i + 2) * V + v];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i * P3 + m] = acc[i];
}

// out = V applied along scratch axis AX (0: stride 1 (z), 1: stride P (y),
// 2: stride P^2 (x)) to every line of the three components:
// out(k) = sum_j C[j][k] in(j).  A block takes TL lines (blockIdx.z =
// component): for AX 1, 2 the TL consecutive lines at one value of the
// remaining index (blockIdx.y), for AX 0 TL consecutive lines.  The lines
// are staged in LDS as xs[j][t]; each thread forms a kB x kB block of
// outputs (kB lines t, kB outputs k).  Threads run over t blocks (AX 1, 2)
// or over k blocks (AX 0), so the HBM loads and stores are contiguous.
template <int AX>
__global__ __launch_bounds__(1024) void k_f2d_dct(int P, int Pc, int TL,
                                                  const double* __restrict__ in,
                                                  double* __restrict__ out,
                                                  const double* __restrict__ C) {
  extern __shared__ double xs[];
  const long P2 = (long)P * P, P3 = P2 * P;
  const double* src = in + blockIdx.z * P3;
  double* dst = out + blockIdx.z * P3;
  const long l0 = (long)blockIdx.x * TL;  // first line of the block
  const long nl = AX == 0 ? P2 : P;       // lines along blockIdx.x
  auto addr = [&](long t, int j) -> long {  // element j of line t
    if constexpr (AX == 0) return t * P + j;
    else if constexpr (AX == 1) return t + (long)P * j + P2 * blockIdx.y;
    else return t + (long)P * blockIdx.y + P2 * j;
  };
  if constexpr (AX == 0) {
    for (int i = threadIdx.x; i < TL * P; i += blockDim.x) {
      const int t = i / P, j = i - t * P;
      xs[j * TL + t] = l0 + t < nl ? src[addr(l0 + t, j)] : 0.0;
    }
  } else {
    for (int i = threadIdx.x; i < TL * P; i += blockDim.x) {
      const int j = i / TL, t = i - j * TL;
      xs[j * TL + t] = l0 + t < nl ? src[addr(l0 + t, j)] : 0.0;
    }
  }
  __syncthreads();
  const int nt = TL / kB, nk = Pc / kB;
  const int it = threadIdx.x;
  if (it >= nt * nk) return;
  int tb, kb;
  if constexpr (AX == 0) {
    kb = it % nk;
    tb = it / nk;
  } else {
    tb = it % nt;
    kb = it / nt;
  }
  const int t0 = tb * kB, k0 = kb * kB;
  double acc[kB][kB];
#pragma unroll
  for (int a = 0; a < kB; ++a)
#pragma unroll
    for (int c = 0; c < kB; ++c) acc[a][c] = 0.0;
  for (int j = 0; j < P; ++j) {
    double xv[kB], cv[kB];
#pragma unroll
    for (int a = 0; a < kB; ++a) xv[a] = xs[j * TL + t0 + a];
#pragma unroll
    for (int c = 0; c < kB; ++c) cv[c] = C[(long)j * Pc + k0 + c];
#pragma unroll
    for (int a = 0; a < kB; ++a)
#pragma unroll
      for (int c = 0; c < kB; ++c) acc[a][c] += cv[c] * xv[a];
  }
#pragma unroll
  for (int a = 0; a < kB; ++a) {
    if (l0 + t0 + a >= nl) continue;
#pragma unroll
    for (int c = 0; c < kB; ++c)
      if (k0 + c < P) dst[addr(l0 + t0 + a, k0 + c)] = acc[a][c];
  }
}

// X[i][m] /= Lambda(k) c(k), k = (kx, ky, kz) of scratch index m; the
// (0,0,0) term is zero
__global__ __launch_bounds__(kFT) void k_f2d_div(int N, double ax, double ay, double az,
                                                 double* __restrict__ X) {
  const int P = N + 1;
  const long P3 = (long)P * P * P;
  const long m = (long)blockIdx.x * kFT + threadIdx.x;
  if (m >= P3) return;
  const int k[3] = {(int)(m / ((long)P * P)), (int)((m / P) % P), (int)(m % P)};  // x, y, z
  double lk[3], lm[3], cc = 1.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double s = sinpi((double)k[a] / (2.0 * N));
    lk[a] = 4.0 * s * s;
    lm[a] = (2.0 + cospi((double)k[a] / (double)N)) / 3.0;
    cc *= (k[a] == 0 || k[a] == N) ? (double)N : 0.5 * N;
  }
  const double lam = ax * lk[0] * lm[1] * lm[2] + ay * lm[0] * lk[1] * lm[2] + az * lm[0] * lm[1] * lk[2];
  const double inv = m == 0 ? 0.0 : 1.0 / (lam * cc);
#pragma unroll
  for (int i = 0; i < 3; ++i) X[i * P3 + m] *= inv;
}

// u[3 n + i] = (x_i(n) - x_i(node 1)) - X_i(n), n = ix + P iy + P^2 iz:
// 32 x 32 (iz, ix) tiles at one iy transposed through LDS
constexpr int kTT = 32;
__global__ __launch_bounds__(kFT) void k_f2d_out(int N, double hx, double hy, double hz,
                                                 const double* __restrict__ x, double* __restrict__ u) {
  __shared__ double tile[3][kTT][kTT + 1];
  const int P = N + 1;
  const long P2 = (long)P * P, P3 = P2 * P;
  const int iz0 = blockIdx.x * kTT, ix0 = blockIdx.y * kTT, iy = blockIdx.z;
  for (int i = threadIdx.x; i < kTT * kTT; i += kFT) {
    const int zl = i % kTT, xl = i / kTT;
    if (iz0 + zl < P && ix0 + xl < P) {
      const long m = (iz0 + zl) + (long)P * iy + P2 * (ix0 + xl);
#pragma unroll
      for (int c = 0; c < 3; ++c) tile[c][xl][zl] = x[c * P3 + m] - x[c * P3];
    }
  }
  __syncthreads();
  const double h[3] = {hx, hy, hz};
  for (int i = threadIdx.x; i < kTT * kTT; i += kFT) {
    const int xl = i % kTT, zl = i / kTT;
    const int ix = ix0 + xl, iz = iz0 + zl;
    if (ix < P && iz < P) {
      const long n = ix + (long)P * iy + P2 * iz;
      const int id[3] = {ix, iy, iz};
#pragma unroll
      for (int c = 0; c < 3; ++c) u[3 * n + c] = tile[c][xl][zl] - id[c] * h[c];
    }
  }
}

unsigned nblk(long n) { return (unsigned)((n + kFT - 1) / kFT); }

template <int AX>
void dct_axis(int P, const double* in, double* out, const double* C, hipStream_t s) {
  // lines per block: a multiple of kB that 64 KiB of LDS holds, at most 32
  int TL = 32;
  while (TL > kB && (long)TL * P * 8 > 65536) TL /= 2;
  const int Pc = cos_stride(P);
  const int items = (TL / kB) * (Pc / kB);
  const int threads = (items + 63) / 64 * 64;
  const long nl = AX == 0 ? (long)P * P : P;
  const dim3 g((unsigned)((nl + TL - 1) / TL), AX == 0 ? 1 : P, 3);
  k_f2d_dct<AX><<<g, threads, (size_t)TL * P * 8, s>>>(P, Pc, TL, in, out, C);
}

}  // namespace

size_t f2disp_scratch_doubles(int N) { return 6 * (size_t)(N + 1) * (N + 1) * (N + 1); }
size_t f2disp_cos_doubles(int N) { return (size_t)(N + 1) * cos_stride(N + 1); }

void launch_f2disp_cos(int N, double* C, hipStream_t s) {
  k_f2d_cos<<<nblk((long)f2disp_cos_doubles(N)), kFT, 0, s>>>(N, cos_stride(N + 1), C);
}

void launch_f2disp(int N, const double* F, const double len[3], const double* C, double* work,
                   double* u, hipStream_t s) {
  const int P = N + 1;
  const long P3 = (long)P * P * P;
  const double h[3] = {len[0] / N, len[1] / N, len[2] / N};
  const double vol = h[0] * h[1] * h[2];
  double* A = work;
  double* B = work + 3 * P3;
  k_f2d_rhs<<<nblk(P3), kFT, 0, s>>>(N, F, vol / (4.0 * h[0]), vol / (4.0 * h[1]),
                                     vol / (4.0 * h[2]), A);
  dct_axis<0>(P, A, B, C, s);
  dct_axis<1>(P, B, A, C, s);
  dct_axis<2>(P, A, B, C, s);
  k_f2d_div<<<nblk(P3), kFT, 0, s>>>(N, h[1] * h[2] / h[0], h[0] * h[2] / h[1],
                                     h[0] * h[1] / h[2], B);
  dct_axis<2>(P, B, A, C, s);
  dct_axis<1>(P, A, B, C, s);
  dct_axis<0>(P, B, A, C, s);
  const unsigned nt = (unsigned)((P + kTT - 1) / kTT);
  k_f2d_out<<<dim3(nt, nt, P), kFT, 0, s>>>(N, h[0], h[1], h[2], A, u);
}

}  // namespace cpfft
