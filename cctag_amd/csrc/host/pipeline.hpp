// Host driver of one detection after the front end: candidate stage per
// level (coarse to fine), level-0 re-fit, identification, de-duplication.
// Detection.cpp:773-999, Multiresolution.cpp:268-413.
#pragma once

#include <vector>

#include "candidate.hpp"
#include "ident.hpp"

namespace cctagx {

// Fills `out` with every marker found, identified or not (status 1 means
// identified), sorted by id like the reference's markers.sort().  The
// candidate stage runs its tasks on `pool` (serially when null); the result
// is the same either way.
void detect_markers(const std::vector<LevelEdges>& levels, const HostParams& p,
                    const std::vector<std::vector<float>>& bank, IdentBackend* ident,
                    std::vector<Marker>& out, TaskPool* pool = nullptr);

}  // namespace cctagx
