// cpfft_noalloc.hip -- cpfft_gkdf_noalloc: the G_K_dF operator for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It runs the kernels of
// cpfft_gkdf on caller-owned inputs and outputs, on the caller's stream,
// with its spectrum scratch and twiddle table in one statically sized
// device array instead of a context's buffers.  It calls no allocating
// routine: host-side set-up, kernel launches and hipGetLastError only.
#include <math.h>

#include <mutex>
#include <string>

#include "../../include/cpfft.h"
#include "kernels.h"

using namespace cpfft;

static constexpr int kMaxN = CPFFT_NOALLOC_MAX_N;
// two half spectra of nine components (2 * 9 * N * N * (N/2+1) complex)
// plus the N twiddles, for N = kMaxN: 16.6 MB
static constexpr long kSpecMax = 2L * 9 * kMaxN * kMaxN * (kMaxN / 2 + 1);
static constexpr long kScratchCplx = kSpecMax + 64;

// The scratch array, one copy per GPU, held until the process ends.  In
// host code this name is only a stand-in for the symbol: each GPU's copy is
// used through the address hipGetSymbolAddress returns for it (g_scratch).
static __device__ __align__(256) cplx scratch[kScratchCplx];

static constexpr int kMaxDev = 64;
static std::mutex g_mu;
static cplx* g_scratch[kMaxDev] = {};  // device address of `scratch` on each GPU

static int nfail(int code, const char* what) {
  set_last_error((std::string("cpfft_gkdf_noalloc: ") + what).c_str());
  return code;
}

// Resolve (once per device, with that device current, which it is when
// this runs) the address of the current GPU's copy of the scratch array.
int cpfft::noalloc_scratch_current_device(cplx** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev)
    return nfail(-2, "no current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nfail(-100, (std::string("hipGetSymbolAddress: ") + hipGetErrorString(e)).c_str());
    g_scratch[dev] = static_cast<cplx*>(p);
  }
  if (out) *out = g_scratch[dev];
  return 0;
}

// the twiddle table travels by value in the kernel's arguments
struct NoallocTwiddles {
  cplx w[kMaxN];
};

static __global__ void k_put_twiddles(NoallocTwiddles t, int n, cplx* __restrict__ dst) {
  if ((int)threadIdx.x < n) dst[threadIdx.x] = t.w[threadIdx.x];
}

extern "C" int cpfft_gkdf_noalloc(int N, int gamma_mode, const double* F, const double* K4,
                                  double* GKF, void* stream) {
  if (N < 2 || N > kMaxN) return nfail(-1, "N must be in [2, CPFFT_NOALLOC_MAX_N = 48]");
  if (gamma_mode != 0 && gamma_mode != 1) return nfail(-1, "bad gamma_mode");
  if (!F || !GKF) return nfail(-1, "null field pointer");
  if (F == GKF) return nfail(-1, "F and GKF must not alias");
  SpecGeom g;
  // the geometry of a context (cpfft_create clamps the LDS per block to
  // 64 KiB), so the results are bit-identical to cpfft_gkdf's
  if (const char* why = make_spec_geom(N, gamma_mode, 65536, g)) return nfail(-3, why);
  if (2 * 9 * g.S1 > kSpecMax) return nfail(-1, "spectrum exceeds the scratch array");
  cplx* sp = nullptr;
  if (int rc = noalloc_scratch_current_device(&sp)) return rc;
  cplx* tw = sp + kSpecMax;
  // twiddles exp(-2 pi i k/N) in long double, as cpfft_create computes them
  NoallocTwiddles t{};
  const long double pil = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < N; ++k) {
    const long double a = 2.0L * pil * (long double)k / (long double)N;
    t.w[k] = cplx{(double)cosl(a), (double)-sinl(a)};
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_put_twiddles<<<1, 64, 0, s>>>(t, N, tw);
  launch_gkdf(g, F, K4, GKF, sp, tw, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nfail(-100, hipGetErrorString(e));
  return 0;
}
