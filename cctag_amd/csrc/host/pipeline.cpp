// Host detection driver (see pipeline.hpp).
#include "pipeline.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace cctagx {

namespace {
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}
}  // namespace

void frame_candidates(const std::vector<LevelEdges>& levels, const HostParams& p,
                      TaskPool* pool, std::vector<Marker>& markers, HostTimes* times,
                      FrameWorkspace* ws) {
  markers.clear();
  auto t = Clock::now();
  FrameWorkspace local;
  std::vector<EdgeSet>& es = (ws ? ws : &local)->levels;
  const int nl = (int)levels.size();
  es.resize(nl);
  std::vector<int> level_h(nl);
  parallel_for(pool, nl, [&](int l) {
    es[l].prepare(levels[l]);
    level_h[l] = levels[l].h;
  });
  constexpr int kFillChunk = 16384;
  std::vector<std::pair<int, int>> chunks;
  for (int l = 0; l < nl; ++l)
    for (int b = 0; b < levels[l].n; b += kFillChunk) chunks.emplace_back(l, b);
  parallel_for(pool, (int)chunks.size(), [&](int c) {
    const int l = chunks[c].first, b = chunks[c].second;
    es[l].fill(levels[l], b, std::min(b + kFillChunk, levels[l].n));
  });
  if (times) times->load += ms_since(t);
  detect_candidates(es, level_h, p, pool, markers, times);
  t = Clock::now();
  rescale_markers(markers, es[0], p, pool);
  if (times) times->rescale += ms_since(t);
}

void finish_markers(std::vector<Marker>& markers, std::vector<Marker>& out) {
  out.clear();
  // both de-duplication passes on indices: a marker's comparison circle is
  // built once, and only the kept markers are moved out (`markers` is
  // consumed)
  std::vector<Ellipse> half(markers.size());
  std::vector<int> all(markers.size());
  for (size_t i = 0; i < markers.size(); ++i) {
    half[i] = markers[i].half_circle();
    all[i] = (int)i;
  }
  const std::vector<int> prelim = update_markers(markers, half, all);
  const std::vector<int> kept = update_markers(markers, half, prelim);
  // an index kept twice (one arrival replacing two entries) is copied,
  // its last occurrence moved
  std::vector<int> left(markers.size(), 0);
  for (const int i : kept) ++left[i];
  out.reserve(kept.size());
  for (const int i : kept) {
    if (--left[i] > 0) out.push_back(markers[i]);
    else out.push_back(std::move(markers[i]));
  }
  std::stable_sort(out.begin(), out.end(), [](const Marker& a, const Marker& b) { return a.id < b.id; });
}

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
