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
// small matrix product per block (FP64 vector units; the three components
// are the grid's z dimension).
//
// Node numbering as the reference's: n = ix + P iy + P^2 iz, P = N + 1.
#include "kernels.h"

namespace cpfft {
namespace {

constexpr int kFT = 256;  // threads per block

// C[j][k] = cos(pi j k / N), the argument reduced exactly mod 2N
__global__ __launch_bounds__(kFT) void k_f2d_cos(int N, double* __restrict__ C) {
  const int P = N + 1;
  const long i = (long)blockIdx.x * kFT + threadIdx.x;
  if (i >= (long)P * P) return;
  const int j = (int)(i / P), k = (int)(i - (long)j * P);
  C[i] = cospi((double)((j * k) % (2 * N)) / (double)N);
}

// b[i][n] = sum over the bricks around node n of sum_k s_k coef_k F_ik,
// s_k = +1 where the node is the brick's upper corner along k, else -1;
// coef_k = (h_x h_y h_z) / (4 h_k) (the one-point rule: w detJ dN/dX_k)
__global__ __launch_bounds__(kFT) void k_f2d_rhs(int N, const double* __restrict__ F, double c0,
                                                 double c1, double c2, double* __restrict__ b) {
  const int P = N + 1;
  const long P3 = (long)P * P * P, V = (long)N * N * N;
  const long n = (long)blockIdx.x * kFT + threadIdx.x;
  if (n >= P3) return;
  const int ix = (int)(n % P), iy = (int)((n / P) % P), iz = (int)(n / ((long)P * P));
  const double coef[3] = {c0, c1, c2};
  double acc[3] = {0.0, 0.0, 0.0};
  for (int ex = ix - 1; ex <= ix; ++ex) {
    if (ex < 0 || ex >= N) continue;
    const double sx = ex < ix ? 1.0 : -1.0;
    for (int ey = iy - 1; ey <= iy; ++ey) {
      if (ey < 0 || ey >= N) continue;
      const double sy = ey < iy ? 1.0 : -1.0;
      for (int ez = iz - 1; ez <= iz; ++ez) {
        if (ez < 0 || ez >= N) continue;
        const double sz = ez < iz ? 1.0 : -1.0;
        const long v = ((long)ex * N + ey) * N + ez;
#pragma unroll
        for (int i = 0; i < 3; ++i)
          acc[i] += sx * coef[0] * F[(3 * i + 0) * V + v] + sy * coef[1] * F[(3 * i + 1) * V + v] +
                    sz * coef[2] * F[(3 * i + 2) * V + v];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i * P3 + n] = acc[i];
}

// out = V applied along axis AX (0 x, 1 y, 2 z) to every line of the three
// components [3][P^3]: out(k) = sum_j C[j][k] in(j).  A block takes TL
// lines (blockIdx.z = component): for AX 1, 2 consecutive ix at one
// (iz or iy) = blockIdx.y, for AX 0 consecutive line numbers iy + P iz.
// The lines are staged in LDS as xs[j][t].  Lanes run over the line (AX 1,
// 2) or over k (AX 0), so that the loads and stores of HBM are contiguous.
template <int AX>
__global__ __launch_bounds__(kFT) void k_f2d_dct(int P, int TL, const double* __restrict__ in,
                                                 double* __restrict__ out,
                                                 const double* __restrict__ C) {
  extern __shared__ double xs[];
  const long P2 = (long)P * P, P3 = P2 * P;
  const double* src = in + blockIdx.z * P3;
  double* dst = out + blockIdx.z * P3;
  const long l0 = (long)blockIdx.x * TL;  // first line of the block
  const long nl = AX == 0 ? P2 : P;       // lines along blockIdx.x
  // element j of line t
  auto addr = [&](long t, int j) -> long {
    if constexpr (AX == 0) return t * P + j;
    else if constexpr (AX == 1) return t + (long)P * j + P2 * blockIdx.y;
    else return t + (long)P * blockIdx.y + P2 * j;
  };
  if constexpr (AX == 0) {
    for (int i = threadIdx.x; i < TL * P; i += kFT) {
      const int t = i / P, j = i - t * P;
      xs[j * TL + t] = l0 + t < nl ? src[addr(l0 + t, j)] : 0.0;
    }
  } else {
    for (int i = threadIdx.x; i < TL * P; i += kFT) {
      const int j = i / TL, t = i - j * TL;
      xs[j * TL + t] = l0 + t < nl ? src[addr(l0 + t, j)] : 0.0;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < TL * P; i += kFT) {
    int t, k;
    if constexpr (AX == 0) {
      t = i / P;
      k = i - t * P;
    } else {
      k = i / TL;
      t = i - k * TL;
    }
    if (l0 + t >= nl) continue;
    double acc = 0.0;
    for (int j = 0; j < P; ++j) acc += C[(long)j * P + k] * xs[j * TL + t];
    dst[addr(l0 + t, k)] = acc;
  }
}

// X[i][k] /= Lambda(k) c(k) (k = (kx, ky, kz)); the (0,0,0) term is zero
__global__ __launch_bounds__(kFT) void k_f2d_div(int N, double ax, double ay, double az,
                                                 double* __restrict__ X) {
  const int P = N + 1;
  const long P3 = (long)P * P * P;
  const long n = (long)blockIdx.x * kFT + threadIdx.x;
  if (n >= P3) return;
  const int k[3] = {(int)(n % P), (int)((n / P) % P), (int)(n / ((long)P * P))};
  double lk[3], lm[3], cc = 1.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double s = sinpi((double)k[a] / (2.0 * N));
    lk[a] = 4.0 * s * s;
    lm[a] = (2.0 + cospi((double)k[a] / (double)N)) / 3.0;
    cc *= (k[a] == 0 || k[a] == N) ? (double)N : 0.5 * N;
  }
  const double lam = ax * lk[0] * lm[1] * lm[2] + ay * lm[0] * lk[1] * lm[2] + az * lm[0] * lm[1] * lk[2];
  const double inv = n == 0 ? 0.0 : 1.0 / (lam * cc);
#pragma unroll
  for (int i = 0; i < 3; ++i) X[i * P3 + n] *= inv;
}

// u[3 n + i] = (x_i(n) - x_i(node 1)) - X_i(n)
__global__ __launch_bounds__(kFT) void k_f2d_out(int N, double hx, double hy, double hz,
                                                 const double* __restrict__ x, double* __restrict__ u) {
  const int P = N + 1;
  const long P3 = (long)P * P * P;
  const long n = (long)blockIdx.x * kFT + threadIdx.x;
  if (n >= P3) return;
  const double X[3] = {(n % P) * hx, ((n / P) % P) * hy, (n / ((long)P * P)) * hz};
#pragma unroll
  for (int i = 0; i < 3; ++i) u[3 * n + i] = (x[i * P3 + n] - x[i * P3]) - X[i];
}

unsigned nblk(long n) { return (unsigned)((n + kFT - 1) / kFT); }

template <int AX>
void dct_axis(int P, const double* in, double* out, const double* C, hipStream_t s) {
  // lines per block: as many as 64 KiB of LDS holds, at most 32
  int TL = 32;
  while (TL > 1 && (long)TL * P * 8 > 65536) TL /= 2;
  const long nl = AX == 0 ? (long)P * P : P;
  const dim3 g((unsigned)((nl + TL - 1) / TL), AX == 0 ? 1 : P, 3);
  k_f2d_dct<AX><<<g, kFT, (size_t)TL * P * 8, s>>>(P, TL, in, out, C);
}

}  // namespace

size_t f2disp_scratch_doubles(int N) { return 6 * (size_t)(N + 1) * (N + 1) * (N + 1); }

void launch_f2disp_cos(int N, double* C, hipStream_t s) {
  const long P = N + 1;
  k_f2d_cos<<<nblk(P * P), kFT, 0, s>>>(N, C);
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
  k_f2d_out<<<nblk(P3), kFT, 0, s>>>(N, h[0], h[1], h[2], A, u);
}

}  // namespace cpfft
