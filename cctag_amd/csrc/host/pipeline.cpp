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
  // the word ranks (edge counts of word ranges, their running sum, the
  // ranks) and the points, in pieces small enough to spread over the pool
  constexpr size_t kRankWords = 8192;
  constexpr int kFillChunk = 4096;
  struct WordRange {
    int l;
    size_t w0, w1;
    int32_t before;
  };
  std::vector<WordRange> ranges;
  for (int l = 0; l < nl; ++l)
    for (size_t w0 = 0; w0 < es[l].n_words(); w0 += kRankWords)
      ranges.push_back({l, w0, std::min(w0 + kRankWords, es[l].n_words()), 0});
  parallel_for(pool, (int)ranges.size(), [&](int r) {
    ranges[r].before = es[ranges[r].l].word_edges(ranges[r].w0, ranges[r].w1);
  });
  int32_t run = 0;
  for (size_t r = 0; r < ranges.size(); ++r) {
    if (r == 0 || ranges[r].l != ranges[r - 1].l) run = 0;
    const int32_t n = ranges[r].before;
    ranges[r].before = run;
    run += n;
  }
  std::vector<std::pair<int, int>> chunks;
  for (int l = 0; l < nl; ++l)
    for (int b = 0; b < levels[l].n; b += kFillChunk) chunks.emplace_back(l, b);
  const int nr = (int)ranges.size();
  parallel_for(pool, nr + (int)chunks.size(), [&](int c) {
    if (c < nr) {
      const WordRange& r = ranges[c];
      es[r.l].rank_words(r.w0, r.w1, r.before);
      return;
    }
    const int l = chunks[c - nr].first, b = chunks[c - nr].second;
    es[l].fill(levels[l], b, std::min(b + kFillChunk, levels[l].n),
               p.min_votes_to_select_candidate);
  });
  if (times) times->load += ms_since(t);
  detect_candidates(es, level_h, p, pool, markers, times);
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

}  // namespace cctagx
