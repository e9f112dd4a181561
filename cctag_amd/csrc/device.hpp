// Small HIP runtime helpers shared by the host code of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace cctagx {

// Last error text of the calling thread, for cctagx_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define CCTAGX_HIP(call)                                                          \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::cctagx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));     \
      return -2;                                                                  \
    }                                                                             \
  } while (0)

// Waits for the work queued on `s` by polling it.  The runtime's blocking
// wait may put the thread to sleep, and a sleeping thread wakes well after a
// short launch sequence has ended; the detection path waits on the GPU
// several times per frame, each time for less than a millisecond.
inline hipError_t stream_wait(hipStream_t s) {
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e != hipErrorNotReady) return e;
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
}

// One device allocation carved into 256-byte aligned sub-arrays.
class DeviceArena {
 public:
  DeviceArena() = default;
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  ~DeviceArena() { release(); }

  // Plan phase: reserve bytes, returns offset.
  size_t reserve(size_t bytes) {
    const size_t off = size_;
    size_ += (bytes + 255) & ~size_t(255);
    return off;
  }
  int allocate() {
    release();
    if (size_ == 0) return 0;
    CCTAGX_HIP(hipMalloc(&base_, size_));
    return 0;
  }
  void release() {
    if (base_) hipFree(base_);
    base_ = nullptr;
  }
  template <class T>
  T* at(size_t off) const { return reinterpret_cast<T*>(static_cast<char*>(base_) + off); }
  size_t size() const { return size_; }
  void clear_plan() { size_ = 0; }

 private:
  void* base_ = nullptr;
  size_t size_ = 0;
};

// Device buffer that only grows.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  int ensure(size_t bytes) {
    if (bytes <= cap_) return 0;
    release();
    CCTAGX_HIP(hipMalloc(&p_, bytes));
    cap_ = bytes;
    return 0;
  }
  void release() {
    if (p_) hipFree(p_);
    p_ = nullptr;
    cap_ = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t cap_ = 0;
};

// Pinned host buffer, also mapped for the device: kernels may read it (jobs
// the host wrote before the launch) or write it (results the host reads
// after synchronizing; allocate such buffers with hipHostMallocCoherent).
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  PinnedBuffer(PinnedBuffer&& o) noexcept : p_(o.p_), d_(o.d_), cap_(o.cap_) {
    o.p_ = o.d_ = nullptr;
    o.cap_ = 0;
  }
  PinnedBuffer& operator=(PinnedBuffer&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      d_ = o.d_;
      cap_ = o.cap_;
      o.p_ = o.d_ = nullptr;
      o.cap_ = 0;
    }
    return *this;
  }
  ~PinnedBuffer() { release(); }
  int ensure(size_t bytes, unsigned flags = hipHostMallocDefault) {
    if (bytes <= cap_) return 0;
    release();
    CCTAGX_HIP(hipHostMalloc(&p_, bytes, flags));
    cap_ = bytes;
    CCTAGX_HIP(hipHostGetDevicePointer(&d_, p_, 0));
    return 0;
  }
  void release() {
    if (p_) hipHostFree(p_);
    p_ = d_ = nullptr;
    cap_ = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p_); }
  // The same memory as the device addresses it.
  template <class T>
  T* dev() const { return static_cast<T*>(d_); }
  size_t capacity() const { return cap_; }

 private:
  void* p_ = nullptr;
  void* d_ = nullptr;
  size_t cap_ = 0;
};

}  // namespace cctagx
