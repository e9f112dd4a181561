// A small persistent thread pool for the host candidate stage: run(n, fn)
// calls fn(0..n-1) on the workers and the calling thread and returns when all
// calls have ended.  The reference parallelises the same loops with
// tbb::parallel_for (Detection.cpp:616,644,680).
#pragma once

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

 private:
  void worker();
  void drain(const std::function<void(int)>& fn, int n);

  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable wake_, done_;
  const std::function<void(int)>* fn_ = nullptr;
  int n_ = 0;
  int next_ = 0;      // next index to hand out (under m_)
  int busy_ = 0;      // workers inside the current job
  uint64_t gen_ = 0;  // job generation
  bool stop_ = false;
  std::exception_ptr error_;
};

// Runs fn(0..n-1) on `pool`, or serially in index order when pool is null.
void parallel_for(TaskPool* pool, int n, const std::function<void(int)>& fn);

}  // namespace cctagx
