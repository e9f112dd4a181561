// Host detection driver (see pipeline.hpp).
#include "pipeline.hpp"

#include <algorithm>
#include <cmath>

namespace cctagx {

void detect_markers(const std::vector<LevelEdges>& levels, const HostParams& p,
                    const std::vector<std::vector<float>>& bank, IdentBackend* ident,
                    std::vector<Marker>& out) {
  out.clear();
  // One rand_5_k stream per detection (the reference keeps a thread_local
  // stream seeded 271828 for the life of the thread, Statistic.cpp:19).
  Pcg32 rng(271828u);
  const int nl = (int)levels.size();
  std::vector<EdgeSet> es(nl);
  for (int l = 0; l < nl; ++l) es[l].load(levels[l]);
  std::vector<Marker> markers;
  for (int l = nl - 1; l >= 0; --l) {
    std::vector<Marker> level_markers;
    detection_from_edges(level_markers, es[l], levels[l].h, l, (float)std::pow(2.0, l), p, rng);
    markers.insert(markers.end(), level_markers.begin(), level_markers.end());
  }
  rescale_markers(markers, es[0], p, rng);
  if (p.do_identification && ident) identify_markers(markers, p, bank, *ident);
  std::vector<Marker> prelim;
  for (const Marker& m : markers) update_markers(prelim, m);
  for (const Marker& m : prelim) update_markers(out, m);
  std::stable_sort(out.begin(), out.end(), [](const Marker& a, const Marker& b) { return a.id < b.id; });
}

}  // namespace cctagx
