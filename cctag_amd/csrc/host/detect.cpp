// Host driver of one detection (Detection.cpp:773-999): the candidate
// stage of every level, identification, then de-duplication.  Kept apart
// from pipeline.cpp, whose stages the oracle also runs around its own
// identification.
#include <chrono>

#include "pipeline.hpp"

namespace cctagx {

namespace {
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}
}  // namespace

void detect_markers(const std::vector<LevelEdges>& levels, const HostParams& p,
                    const std::vector<std::vector<float>>& bank, IdentBackend* ident,
                    std::vector<Marker>& out, TaskPool* pool, HostTimes* times,
                    FrameWorkspace* ws) {
  std::vector<Marker> markers;
  frame_candidates(levels, p, pool, markers, times, ws);
  auto t = Clock::now();
  if (p.do_identification && ident) identify_markers(markers, p, bank, *ident, pool, times);
  if (times) times->ident += ms_since(t);
  t = Clock::now();
  finish_markers(markers, out);
  if (times) times->finish += ms_since(t);
}

}  // namespace cctagx
