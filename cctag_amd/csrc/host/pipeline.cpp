// Host detection driver (see pipeline.hpp).
#include "pipeline.hpp"

#include <algorithm>
#include <cmath>

namespace cctagx {

void frame_candidates(const std::vector<LevelEdges>& levels, const HostParams& p,
                      TaskPool* pool, std::vector<Marker>& markers) {
  markers.clear();
  const int nl = (int)levels.size();
  std::vector<EdgeSet> es(nl);
  std::vector<int> level_h(nl);
  parallel_for(pool, nl, [&](int l) {
    es[l].load(levels[l]);
    level_h[l] = levels[l].h;
  });
  detect_candidates(es, level_h, p, pool, markers);
  rescale_markers(markers, es[0], p, pool);
}

void finish_markers(const std::vector<Marker>& markers, std::vector<Marker>& out) {
  out.clear();
  std::vector<Marker> prelim;
  for (const Marker& m : markers) update_markers(prelim, m);
  for (const Marker& m : prelim) update_markers(out, m);
  std::stable_sort(out.begin(), out.end(), [](const Marker& a, const Marker& b) { return a.id < b.id; });
}

void detect_markers(const std::vector<LevelEdges>& levels, const HostParams& p,
                    const std::vector<std::vector<float>>& bank, IdentBackend* ident,
                    std::vector<Marker>& out, TaskPool* pool) {
  std::vector<Marker> markers;
  frame_candidates(levels, p, pool, markers);
  if (p.do_identification && ident) identify_markers(markers, p, bank, *ident);
  finish_markers(markers, out);
}

}  // namespace cctagx
