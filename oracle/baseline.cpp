// Oracle (test infrastructure only): CPU-baseline timing of the reference traversal under a
// FIFO thread pool with the semantics of myrt::thread_pool (src/thread_pool.hpp:34-162,
// src/thread_pool.cpp:1-50): a fixed set of workers, one mutex-guarded deque of jobs, run_async
// returning a future that carries the job's result or exception.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

class ThreadPool {
 public:
  explicit ThreadPool(unsigned n) {
    for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~ThreadPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  template <typename F>
  auto run_async(F&& f) -> std::future<decltype(f())> {
    using R = decltype(f());
    auto prom = std::make_shared<std::promise<R>>();
    std::future<R> fut = prom->get_future();
    {
      std::lock_guard<std::mutex> lk(m_);
      jobs_.push_back([prom, fn = std::forward<F>(f)]() mutable {
        try {
          if constexpr (std::is_void_v<R>) {
            fn();
            prom->set_value();
          } else {
            prom->set_value(fn());
          }
        } catch (...) {
          prom->set_exception(std::current_exception());
        }
      });
    }
    cv_.notify_one();
    return fut;
  }
  unsigned concurrency() const { return unsigned(workers_.size()); }

 private:
  void loop() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (stop_) return;
      auto job = std::move(jobs_.front());
      jobs_.pop_front();
      lk.unlock();
      job();
      lk.lock();
    }
  }
  std::vector<std::thread> workers_;
  std::deque<std::function<void()>> jobs_;
  std::mutex m_;
  std::condition_variable cv_;
  bool stop_ = false;
};

}  // namespace

extern "C" double orc_reference_time(orc_accel a, const orc_ray* rays, size_t n, int any_hit, int threads,
                                     size_t chunk, int reps, int* threads_used) {
  unsigned nt = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
  if (chunk == 0) chunk = 65536;
  if (reps < 1) reps = 1;
  std::vector<orc_hit> hits(n);
  ThreadPool pool(nt);
  if (threads_used) *threads_used = int(pool.concurrency());
  const auto pass = [&]() {
    std::vector<std::future<void>> futs;
    for (size_t b = 0; b < n; b += chunk) {
      const size_t e = std::min(n, b + chunk);
      futs.push_back(pool.run_async([=, &hits] { orc_reference_trace(a, rays + b, e - b, any_hit, hits.data() + b); }));
    }
    for (auto& f : futs) f.get();
  };
  pass();  // warm-up
  std::vector<double> secs;
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    pass();
    const auto t1 = std::chrono::steady_clock::now();
    secs.push_back(std::chrono::duration<double>(t1 - t0).count());
  }
  std::sort(secs.begin(), secs.end());
  const double med = secs[secs.size() / 2];
  return med > 0 ? double(n) / med : 0.0;
}
