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
  // Level-by-level resampling, used when some level is not an exact half of
  // its parent: per level l >= 1, its planes (indices into dev_planes, at
  // dev_level_planes + level_plane_start[l]) and their blocks.
  bool resample = false;
  const int32_t* dev_level_planes = nullptr;
  std::vector<int> level_plane_start, level_plane_count, n_tiles_r;
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
  // true when every level of every frame is an exact half of its parent
  // (the pyramid then runs as one fused 2x2 kernel)
  bool exact_halves = true;
  std::vector<std::vector<int32_t>> level_planes;  // per level: its planes in plane order
  std::vector<int> n_tiles_r;                       // per level: resampling blocks
};

// Levels are floor(w/2) x floor(h/2) of their parent (ImagePyramid.cpp:23-33).
// Returns false when a level would be empty.
bool plan_planes(const std::vector<std::pair<int, int>>& frame_sizes, int n_levels,
                 PlaneSet& out);

}  // namespace cctagx
