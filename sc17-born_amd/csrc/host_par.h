// host_par.h -- a contiguous split of a host loop over a few std::threads.
//
// Each index is handled by exactly one thread and the body of an index does
// the same floating-point operations in the same order as the serial loop,
// so results do not depend on the number of threads.
#ifndef BORN_HOST_PAR_H
#define BORN_HOST_PAR_H

#include <algorithm>
#include <thread>
#include <vector>

namespace born_par {

// at most 16 threads: the host share of one GPU
inline unsigned threads() {
  static const unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  return t;
}

// f(lo, hi) over contiguous blocks covering [0, n); blocks shorter than
// `grain` indices are not split further
template <class F>
void blocks(long n, long grain, F &&f) {
  const long nt = std::min<long>(threads(), (n + grain - 1) / std::max(grain, 1L));
  if (nt <= 1) {
    if (n > 0) f(0L, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  const long step = (n + nt - 1) / nt;
  for (long t = 1; t < nt; t++) {
    const long lo = t * step, hi = std::min(n, lo + step);
    if (lo < hi) pool.emplace_back([&f, lo, hi] { f(lo, hi); });
  }
  f(0L, std::min(n, step));
  for (auto &th : pool) th.join();
}

}  // namespace born_par

#endif
