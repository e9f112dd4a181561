// Device-to-host staging of what the host candidate stage reads of every
// plane: the counterpart of TagPipe::convertToHost (cuda/frame_export.cu:
// 23-165).  One counts read-back, then one gather kernel packs every
// plane's arrays (exact sizes) straight into one mapped pinned host buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"
#include "frontend_launch.hpp"
#include "host/candidate.hpp"

namespace cctagx {

// One contiguous piece to gather: `bytes` from `src` to packed offset `dst`
// (a multiple of 16).
struct CopySeg {
  const void* src;
  uint64_t dst;
  uint64_t bytes;
};

int launch_pack(const CopySeg* segs, int n_segs, uint8_t* dst, hipStream_t s);

class LevelStager {
 public:
  // Stages planes [first, first + count) of the plane set (all when count
  // < 0) and fills one view per plane.  Synchronizes the stream twice.
  int stage(const PlaneSet& ps, const FrontEndBuffers& b, hipStream_t s,
            std::vector<LevelEdges>& views, int first = 0, int count = -1);
  // Counts of the planes last staged (indexed like the views).
  const std::vector<PlaneCounts>& counts() const { return counts_; }
  size_t last_bytes() const { return last_bytes_; }

 private:
  PinnedBuffer h_counts_, h_segs_, h_packed_;
  DeviceBuffer d_segs_;
  std::vector<PlaneCounts> counts_;
  size_t last_bytes_ = 0;
};

}  // namespace cctagx
