// A small persistent thread pool for the per-instance host work of a batched
// prove (transcripts, proof assembly): instances are independent, so their
// host work runs concurrently on the CPU cores next to the GPU.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace lg {

class ThreadPool {
 public:
  static int hardware_threads() {
    int hw = (int)std::thread::hardware_concurrency();
    if (const char* e = getenv("LOOMGPU_HOST_THREADS")) hw = atoi(e);
    return std::max(1, std::min(hw, 16));
  }
  static ThreadPool& get() {
    static ThreadPool p(hardware_threads() - 1);
    return p;
  }
  // The pool of prover lane `lane` when `lanes` lanes share the host: the
  // hardware threads are divided between the lanes so that their host work
  // runs concurrently.
  static ThreadPool& for_lane(int lane, int lanes) {
    if (lanes <= 1) return get();
    static std::mutex mu;
    static std::unique_ptr<ThreadPool> pools[8][9];
    std::lock_guard<std::mutex> lk(mu);
    auto& p = pools[lane & 7][std::min(lanes, 8)];
    if (!p) p.reset(new ThreadPool(std::max(1, hardware_threads() / lanes) - 1));
    return *p;
  }

  // Runs fn(i) for i in [0, n); rethrows the first exception.
  void parallel_for(int n, const std::function<void(int)>& fn) {
    if (n <= 1 || workers_.empty()) {
      for (int i = 0; i < n; ++i) fn(i);
      return;
    }
    std::unique_lock<std::mutex> job_lock(job_mu_);  // one parallel_for at a time
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      done_ = 0;
      err_ = nullptr;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ == (int)workers_.size(); });
    fn_ = nullptr;
    if (err_) std::rethrow_exception(err_);
  }

  ~ThreadPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  explicit ThreadPool(int n_workers) {
    for (int i = 0; i < n_workers; ++i) workers_.emplace_back([this] { loop(); });
  }

  void work() {
    for (;;) {
      int i = next_.fetch_add(1);
      if (i >= n_) break;
      try {
        (*fn_)(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!err_) err_ = std::current_exception();
      }
    }
  }

  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      work();
      {
        std::lock_guard<std::mutex> lk(mu_);
        ++done_;
      }
      done_cv_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex mu_, job_mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  int n_ = 0;
  std::atomic<int> next_{0};
  int done_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr err_;
};

// The pool parallel_for uses on this thread (set by a prover lane), or null
// for the shared one.
inline ThreadPool*& current_pool() {
  static thread_local ThreadPool* p = nullptr;
  return p;
}

inline void parallel_for(int n, const std::function<void(int)>& fn) {
  ThreadPool* p = current_pool();
  (p ? *p : ThreadPool::get()).parallel_for(n, fn);
}

}  // namespace lg
