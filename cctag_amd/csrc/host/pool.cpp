// TaskPool (see pool.hpp).
#include "pool.hpp"

#include <algorithm>
#include <cstdlib>

namespace cctagx {

TaskPool::TaskPool(int threads) {
  for (int i = 1; i < threads; ++i) workers_.emplace_back([this] { worker(); });
}

TaskPool::~TaskPool() {
  {
    std::lock_guard<std::mutex> lk(m_);
    stop_ = true;
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
    int i;
    {
      std::lock_guard<std::mutex> lk(m_);
      if (next_ >= n) return;
      i = next_++;
    }
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
    const std::function<void(int)>* fn;
    int n;
    {
      std::unique_lock<std::mutex> lk(m_);
      wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      // a job the caller already finished alone: nothing to join
      if (!fn_ || next_ >= n_) continue;
      fn = fn_;
      n = n_;
      ++busy_;
    }
    drain(*fn, n);
    {
      std::lock_guard<std::mutex> lk(m_);
      --busy_;
    }
    done_.notify_all();
  }
}

void TaskPool::run(int n, const std::function<void(int)>& fn) {
  if (n <= 0) return;
  if (workers_.empty() || n == 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(m_);
    fn_ = &fn;
    n_ = n;
    next_ = 0;
    error_ = nullptr;
    ++gen_;
  }
  wake_.notify_all();
  drain(fn, n);
  std::exception_ptr err;
  {
    std::unique_lock<std::mutex> lk(m_);
    // every index handed out and every worker that joined this job gone
    done_.wait(lk, [&] { return next_ >= n_ && busy_ == 0; });
    fn_ = nullptr;
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
