// A small persistent thread pool for the host candidate stage: run(n, fn)
// calls fn(0..n-1) on the workers and the calling thread and returns when all
// calls have ended.  The reference parallelises the same loops with
// tbb::parallel_for (Detection.cpp:616,644,680).
//
// A detection issues dozens of short parallel loops separated by short
// serial steps, so idle workers first spin on the job generation for a
// bounded time (kIdleSpinMicros) and only then sleep on the condition variable;
// indices are handed out by an atomic counter.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace cctagx {

class TaskPool {
 public:
  // `threads` counts the caller: threads - 1 workers are started.
  explicit TaskPool(int threads);
  ~TaskPool();
  TaskPool(const TaskPool&) = delete;
  TaskPool& operator=(const TaskPool&) = delete;

  int threads() const { return (int)workers_.size() + 1; }
  // Rethrows the first exception a call raised, after every call has ended.
  void run(int n, const std::function<void(int)>& fn);

  // min(16, hardware threads), or CCTAGX_HOST_THREADS when set.
  static int default_threads();

  // How long an idle worker spins for the next job, and a caller for the
  // workers still inside its job, before sleeping.
  static constexpr int kIdleSpinMicros = 1000;
  static constexpr int kWaitSpinMicros = 2000;

 private:
  void worker();
  void drain(const std::function<void(int)>& fn, int n);

  // A worker joins a job by raising busy_ and then reading fn_; run()
  // publishes a job (n_, next_, then fn_) only once busy_ is back to zero
  // and clears fn_ before it returns, so a worker sees either no job or a
  // whole one, which stays valid while it is counted in busy_.
  std::vector<std::thread> workers_;
  std::mutex m_;  // sleeping and error capture only
  std::condition_variable wake_, done_;
  std::atomic<const std::function<void(int)>*> fn_{nullptr};
  std::atomic<int> n_{0};
  std::atomic<int> next_{0};          // next index to hand out
  std::atomic<int> busy_{0};          // workers that joined and have not left
  std::atomic<uint64_t> gen_{0};      // job generation (raised under m_)
  std::atomic<bool> stop_{false};     // set under m_
  std::atomic<bool> caller_waits_{false};
  std::exception_ptr error_;          // under m_
};

// Runs fn(0..n-1) on `pool`, or serially in index order when pool is null.
void parallel_for(TaskPool* pool, int n, const std::function<void(int)>& fn);

}  // namespace cctagx
