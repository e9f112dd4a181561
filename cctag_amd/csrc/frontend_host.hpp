// Device state of the front end for one plane set.
#pragma once

#include "device.hpp"
#include "frontend_launch.hpp"

namespace cctagx {

class FrontEndDevice {
 public:
  int init(const PlaneSet& ps, const FrontEndParams& prm);
  // Copies a row-major 8-bit frame into the level-0 plane of `frame`.
  int upload(int frame, const uint8_t* gray, size_t stride, hipStream_t s);
  // The same from an 8-bit BGR (channels 3) or BGRA (4) frame, converted to
  // gray on the device; the first call grows a staging buffer to the frame.
  int upload_color(int frame, const uint8_t* bgr, size_t stride, int channels, hipStream_t s);
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
};

}  // namespace cctagx
