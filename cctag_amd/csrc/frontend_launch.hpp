// Host-side description of one front-end launch sequence over a plane set.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "frontend.hpp"

namespace cctagx {

// Records HIP events between stages on the launch stream.
struct StageTimer {
  std::vector<hipEvent_t> events;
  std::vector<const char*> names;
  size_t used = 0;
  // every kernel marked, or only the sequence's first and last marks (each
  // mark is an event record on the stream)
  bool per_kernel = false;
  void reset() {
    used = 0;
    names.clear();
  }
  void mark(const char* name, hipStream_t s) {
    if (used == events.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      events.push_back(e);
    }
    hipEventRecord(events[used++], s);
    names.push_back(name);
  }
  ~StageTimer() {
    for (hipEvent_t e : events) hipEventDestroy(e);
  }
};

struct FrontEndLaunch {
  FrontEndBuffers buf;
  const PlaneDesc* dev_planes = nullptr;
  int n_planes = 0;
  int n_levels = 0;
  int n_tiles_p = 0;   // pyramid blocks (level-0 64x64 tiles)
  int n_tiles_f = 0;   // filter blocks of the default variant
  int filter_variant = 1;                     // index into kFilterTileHs
  const int32_t* dev_filter_starts[kFilterVariants] = {};  // per variant: first block of each plane
  int n_tiles_fv[kFilterVariants] = {};       // per variant: filter blocks
  int n_tiles_s = 0;   // stencil blocks
  int n_tiles_c = 0;   // compaction chunks
  int edge_blocks = 0; // blocks per plane of the edge-indexed kernels
  FrontEndParams prm{};
};

// Enqueues the full front end on stream s; returns 0 or a negative error.
int launch_front_end(const FrontEndLaunch& L, hipStream_t s, StageTimer* timer);

// Plane set for frames of given sizes, n_levels each, frame-major.
struct PlaneSet {
  std::vector<PlaneDesc> planes;
  int n_levels = 0;
  int total_pix = 0;   // pixel / edge slots, plane starts aligned to 256
  int n_tiles_p = 0, n_tiles_f = 0, n_tiles_s = 0, n_tiles_c = 0;
  std::vector<int32_t> filter_starts[kFilterVariants];  // per variant, one entry per plane
  int n_tiles_fv[kFilterVariants] = {};
};

// Returns false when a level is not an exact half of its parent (the fast
// 2x2 area path is the only resampling built).
bool plan_planes(const std::vector<std::pair<int, int>>& frame_sizes, int n_levels,
                 PlaneSet& out);

}  // namespace cctagx
