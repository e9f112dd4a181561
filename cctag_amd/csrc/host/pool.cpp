// TaskPool (see pool.hpp).
#include "pool.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>

namespace cctagx {
namespace {

inline void cpu_relax() {
#if defined(__x86_64__) || defined(__i386__)
  __builtin_ia32_pause();
#endif
}

// Spins (pause only: an idle sibling hardware thread should cost the busy
// one as little as possible) until done() or `micros` have passed.
// Returns done().
template <class F>
bool spin_until(F&& done, int micros) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  for (int i = 0;; ++i) {
    if (done()) return true;
    if ((i & 255) == 255 &&
        std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - t0).count() >= micros)
      return done();
    cpu_relax();
  }
}

}  // namespace

TaskPool::TaskPool(int threads) {
  for (int i = 1; i < threads; ++i) workers_.emplace_back([this] { worker(); });
}

TaskPool::~TaskPool() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_.store(true);
  }
  wake_.notify_all();
  for (std::thread& t : workers_) t.join();
}

int TaskPool::default_threads() {
  if (const char* s = std::getenv("CCTAGX_HOST_THREADS")) {
    const int v = std::atoi(s);
    if (v >= 1) return std::min(v, 256);
  }
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(hw, 16));
}

void TaskPool::drain(const std::function<void(int)>& fn, int n) {
  for (;;) {
    const int i = next_.fetch_add(1, std::memory_order_relaxed);
    if (i >= n) return;
    try {
      fn(i);
    } catch (...) {
      std::lock_guard<std::mutex> lk(m_);
      if (!error_) error_ = std::current_exception();
    }
  }
}

void TaskPool::worker() {
  uint64_t seen = 0;
  for (;;) {
    if (!spin_until([&] { return stop_.load() || gen_.load() != seen; }, kIdleSpinMicros)) {
      std::unique_lock<std::mutex> lk(m_);
      wake_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
    }
    if (stop_.load()) return;
    seen = gen_.load();
    busy_.fetch_add(1);
    const std::function<void(int)>* fn = fn_.load();
    if (fn) drain(*fn, n_.load());
    busy_.fetch_sub(1);
    if (caller_waits_.load()) {
      { std::lock_guard<std::mutex> lk(m_); }
      done_.notify_all();
    }
  }
}

void TaskPool::run(int n, const std::function<void(int)>& fn) {
  if (n <= 0) return;
  if (workers_.empty() || n == 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  // a worker still leaving the previous job must not see this one's fields
  // half written
  spin_until([&] { return busy_.load() == 0; }, kWaitSpinMicros);
  while (busy_.load() != 0) std::this_thread::yield();
  n_.store(n);
  next_.store(0);
  error_ = nullptr;
  fn_.store(&fn);
  {
    std::lock_guard<std::mutex> lk(m_);
    gen_.fetch_add(1);
  }
  wake_.notify_all();
  drain(fn, n);
  // every index is handed out; wait for the workers still inside the job
  if (!spin_until([&] { return busy_.load() == 0; }, kWaitSpinMicros)) {
    std::unique_lock<std::mutex> lk(m_);
    caller_waits_.store(true);
    done_.wait(lk, [&] { return busy_.load() == 0; });
    caller_waits_.store(false);
  }
  fn_.store(nullptr);
  std::exception_ptr err;
  {
    std::lock_guard<std::mutex> lk(m_);
    err = error_;
    error_ = nullptr;
  }
  if (err) std::rethrow_exception(err);
}

void parallel_for(TaskPool* pool, int n, const std::function<void(int)>& fn) {
  if (pool) {
    pool->run(n, fn);
  } else {
    for (int i = 0; i < n; ++i) fn(i);
  }
}

}  // namespace cctagx
