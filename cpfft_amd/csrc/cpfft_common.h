// cpfft_common.h -- shared definitions for the MI355X kernels and their
// host-side unit-test build.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CPFFT_HD __host__ __device__
#else
#include <cmath>
#define CPFFT_HD
using std::atan2;
using std::copysign;
using std::cos;
using std::fabs;
using std::fmax;
using std::pow;
using std::sin;
using std::sqrt;
#endif

#include <stdint.h>

namespace cpfft {

struct cplx {
  double re, im;
};

// radix plan of a length-N transform (N <= 256)
struct FftPlan {
  int n;
  int nst;
  int rad[12];
};

CPFFT_HD inline FftPlan make_plan(int n) {
  FftPlan p;
  p.n = n;
  p.nst = 0;
  int m = n;
  while (m % 4 == 0 && m > 4 && p.nst < 12) { p.rad[p.nst++] = 4; m /= 4; }
  const int primes[6] = {4, 2, 3, 5, 7, 11};
  for (int k = 0; k < 6 && m > 1; ++k)
    while (m % primes[k] == 0 && m > 1) { p.rad[p.nst++] = primes[k]; m /= primes[k]; }
  if (m > 1) p.rad[p.nst++] = m;  // remaining prime factor: generic radix
  return p;
}

}  // namespace cpfft
