// Host thread pool for the work-list generators: a dynamic loop over [0, n) that keeps
// the first exception and rethrows it on the calling thread.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace chm {

// CHM_THREADS, else OMP_NUM_THREADS, else the hardware's count; at most 64.
inline int host_threads() {
  for (const char* name : {"CHM_THREADS", "OMP_NUM_THREADS"}) {
    const char* v = std::getenv(name);
    if (v && std::atoi(v) > 0) return std::min(64, std::atoi(v));
  }
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(64u, h));
}

// f(i, thread) for every i in [0, n), items handed out one at a time.
template <class F>
void parallel_for(int64_t n, int nthreads, F&& f) {
  nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n));
  if (nthreads <= 1) {
    for (int64_t i = 0; i < n; i++) f(i, 0);
    return;
  }
  std::atomic<int64_t> next{0};
  std::exception_ptr err;
  std::mutex m;
  auto work = [&](int tid) {
    try {
      for (;;) {
        const int64_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i, tid);
      }
    } catch (...) {
      std::lock_guard<std::mutex> g(m);
      if (!err) err = std::current_exception();
      next.store(n);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace chm
