// Device state of the front end for one plane set.
#pragma once

#include "device.hpp"
#include "frontend_launch.hpp"
#include "host/pool.hpp"

namespace cctagx {

// Copies `rows` rows of `row_bytes` bytes (source rows `stride` apart) to
// the packed destination, the rows shared among the pool's threads.
void copy_rows(TaskPool* pool, uint8_t* dst, const uint8_t* src, size_t row_bytes, int rows,
               size_t stride);

class FrontEndDevice {
 public:
  FrontEndDevice() = default;
  FrontEndDevice(const FrontEndDevice&) = delete;
  FrontEndDevice& operator=(const FrontEndDevice&) = delete;
  ~FrontEndDevice();
  int init(const PlaneSet& ps, const FrontEndParams& prm);
  // Copies a row-major 8-bit frame into the level-0 plane of `frame`.  With
  // a pool, the pool's threads copy the rows into the frame's region of a
  // pinned buffer of the library and one transfer takes it over (the
  // runtime's own copy from pageable memory stages it through one thread);
  // a region is reused once its last transfer has ended.
  int upload(int frame, const uint8_t* gray, size_t stride, hipStream_t s,
             TaskPool* pool = nullptr);
  // The same from an 8-bit BGR (channels 3) or BGRA (4) frame, converted to
  // gray on the device; the first call grows a staging buffer to the frame
  // (with a pool, the rows go through a pinned buffer as for gray frames).
  int upload_color(int frame, const uint8_t* bgr, size_t stride, int channels, hipStream_t s,
                   TaskPool* pool = nullptr);
  // Enqueues the whole front end (no host synchronization).
  int run(hipStream_t s, StageTimer* timer);

  const PlaneSet& planes() const { return ps_; }
  const FrontEndBuffers& buffers() const { return launch_.buf; }
  const FrontEndLaunch& launch() const { return launch_; }
  // Work partition of the launch sequence (filter tile variant, blocks per
  // plane of the edge-indexed kernels); every choice gives the same output.
  void set_partition(int filter_variant, int edge_blocks) {
    launch_.filter_variant = filter_variant;
    launch_.edge_blocks = edge_blocks;
  }

 private:
  PlaneSet ps_;
  DeviceArena arena_;
  FrontEndLaunch launch_;
  DeviceBuffer color_;  // staged color frame
  PinnedBuffer color_pinned_;  // color frame on its way to color_
  hipEvent_t color_sent_ = nullptr;
  bool color_in_flight_ = false;
  PinnedBuffer staged_;  // gray frames on their way to the device, one region per frame
  std::vector<size_t> staged_off_;       // per frame: its region's offset
  std::vector<hipEvent_t> staged_sent_;  // per frame: the last transfer out of its region
  std::vector<uint8_t> staged_in_flight_;
};

}  // namespace cctagx
