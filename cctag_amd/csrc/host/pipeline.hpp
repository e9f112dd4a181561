// Host driver of one detection after the front end: candidate stage per
// level (coarse to fine), level-0 re-fit, identification, de-duplication.
// Detection.cpp:773-999, Multiresolution.cpp:268-413.
#pragma once

#include <vector>

#include "candidate.hpp"
#include "ident.hpp"

namespace cctagx {

// Host memory reused from one detection to the next (edge sets per level).
struct FrameWorkspace {
  std::vector<EdgeSet> levels;
};

// Fills `out` with every marker found, identified or not (status 1 means
// identified), sorted by id like the reference's markers.sort().  The
// candidate stage runs its tasks on `pool` (serially when null); the result
// is the same either way.
void detect_markers(const std::vector<LevelEdges>& levels, const HostParams& p,
                    const std::vector<std::vector<float>>& bank, IdentBackend* ident,
                    std::vector<Marker>& out, TaskPool* pool = nullptr,
                    HostTimes* times = nullptr, FrameWorkspace* ws = nullptr);

// The stages of detect_markers, for callers that batch identification over
// frames: candidates of one frame (all levels, level-0 re-fit), then the
// double de-duplication and the sort by id (Detection.cpp:966-999).
void frame_candidates(const std::vector<LevelEdges>& levels, const HostParams& p,
                      TaskPool* pool, std::vector<Marker>& markers, HostTimes* times = nullptr,
                      FrameWorkspace* ws = nullptr);
// (`markers` is consumed.)
void finish_markers(std::vector<Marker>& markers, std::vector<Marker>& out);

}  // namespace cctagx
