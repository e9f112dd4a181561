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

// The device-planned form: the gather kernel itself lays the planes' arrays
// out from the counts the front end left on the device, so the gather needs
// no counts read-back and can be captured with the front end.  The buffer
// starts with a header: the bytes in all, whether they fit, and per plane
// its counts and each array's offset from the data start (kStageDataStart).
constexpr int kStageArrays = 11;  // bits, x, y, gx, gy, before, after, vindex, vlist, flow, seeds
constexpr int kMaxPlannedPlanes = 16;
struct StagePlane {
  PlaneCounts counts;
  uint64_t off[kStageArrays];
};
struct StageHeader {
  uint64_t total;     // data bytes
  uint32_t overflow;  // the data did not fit: nothing but this header was written
  uint32_t n_planes;
  StagePlane planes[kMaxPlannedPlanes];
};
constexpr uint64_t kStageDataStart = (sizeof(StageHeader) + 255) & ~uint64_t(255);

class LevelStager {
 public:
  // Stages planes [first, first + count) of the plane set (all when count
  // < 0) and fills one view per plane.  Synchronizes the stream twice.
  int stage(const PlaneSet& ps, const FrontEndBuffers& b, hipStream_t s,
            std::vector<LevelEdges>& views, int first = 0, int count = -1);
  // Counts of the planes last staged (indexed like the views).
  const std::vector<PlaneCounts>& counts() const { return counts_; }
  size_t last_bytes() const { return last_bytes_; }

  // The device-planned form over all planes of ps (at most
  // kMaxPlannedPlanes): reserve the buffer (outside any stream capture),
  // enqueue the gather (no host synchronization), then wait once and fill
  // the views from the header.  finish_planned returns 1 when the data did
  // not fit: it has grown the buffer, the caller stages this frame with
  // stage(), and a gather that was captured must be captured again (the
  // buffer moved).
  int reserve_planned(const PlaneSet& ps);
  int enqueue_planned(const PlaneSet& ps, const PlaneDesc* dev_planes, const FrontEndBuffers& b,
                      hipStream_t s);
  int finish_planned(const PlaneSet& ps, hipStream_t s, std::vector<LevelEdges>& views);
  int grow_planned(size_t data_bytes);

 private:
  PinnedBuffer h_counts_, h_segs_, h_packed_, h_planned_;
  DeviceBuffer d_segs_;
  std::vector<PlaneCounts> counts_;
  size_t last_bytes_ = 0;
};

}  // namespace cctagx
