// batch_runner.cpp -- cpfft_compute_batch (include/cpfft_driver.h): many
// independent RVEs driven at once.  A pool of host threads takes RVEs from
// a shared queue ordered by decreasing weight (longest-processing-time
// first, which keeps the pool busy to the end when sizes are uneven); each
// thread runs the Fortran load-step driver on its RVE's context, whose
// kernels go to that context's own HIP stream, so RVEs overlap on the GPU
// while each one's host-side control flow (Newton, mixed-BC loop) stays its
// own.  No counterpart in the reference, which solves one RVE per run.
#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/cpfft_driver.h"

extern "C" int cpfft_compute_batch(cpfft_batch_item* items, int n, int nthreads) {
  if (n < 0 || (n > 0 && !items)) return -1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (items[i].ctx == items[j].ctx) return -1;  // one context per RVE
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return items[a].weight > items[b].weight; });
  std::atomic<int> next{0};
  auto worker = [&] {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= n) return;
      cpfft_batch_item& it = items[order[k]];
      it.rc = cpfft_compute(it.ctx, it.first_step, it.nstep, it.BC_all, it.isNBC, it.tolPCG,
                            it.tolNR, it.maxIter, 0, it.Pbar_out, it.counters, it.times, 0, 1);
    }
  };
  const int nt = std::max(1, std::min(nthreads, n));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  for (int i = 0; i < n; ++i)
    if (items[i].rc != 0) return items[i].rc;
  return 0;
}
