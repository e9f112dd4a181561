// ORACLE -- test infrastructure only (see oracle.hpp).
// Serial CPU restatement of the whole identification of a frame's markers
// (Detection.cpp:873-965): identify_step_1 then identify_step_2 for every
// marker in list order, with the oracle's own geometry (no product code).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace cctagx {
class TaskPool;
}

namespace oracle {

// DirectedPoint2d (x, y, gradient).
struct DPt {
  float x = 0, y = 0, dx = 0, dy = 0;
};

// The identification parameters (Params.hpp:108-199, defaults Params.hpp:27-66).
struct IdentParams {
  int n_crowns = 3;
  int n_samples_outer_ellipse = 150;
  int num_cuts_in_ident_step = 22;
  int num_samples_outer_edge_points_refinement = 20;
  int sample_cut_length = 100;
  int imaged_center_n_grid_sample = 5;
  float imaged_center_neighbour_size = 0.2f;
  float min_ident_proba = 1e-6f;
};

// The CCTag fields identification reads (in) and writes (out).
struct IdentMarker {
  // in: rescaledOuterEllipse as (centre, a, b, angle), its points, the
  // initial centerImg and the marker's pyramid scale
  float cx = 0, cy = 0, a = 0, b = 0, angle = 0;
  std::vector<DPt> points;
  float center_x = 0, center_y = 0;
  float scale = 1.f;
  // out
  int status = 0;
  int id = -1;
  float quality = 0.f;     // 1 / normalised residual (setQuality)
  float residual = 0.f;    // normalised residual sqrt(res) / (max - min)
  float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // row-major, image <- marker plane
  std::vector<DPt> stops;  // the cuts identify_step_1 selected (their stops)
};

// Identifies `markers` in list order on the 8-bit frame img (w x h,
// row-major, stride w) with the bank rows `bank` (radius ratios, innermost
// first).  conditionerFromEllipse caches (a + b) / 2 of the first ellipse
// it conditions in the process (conditioner.hpp:27); here the cache lives
// for one call: the first marker that reaches the grid search sets it.
// With a pool, both steps run the markers on its threads: the first step's
// markers are independent, and the second step's share only that cache,
// which is set first from the marker the serial loop would reach the grid
// search with first -- the same results as one thread.
void identify(const uint8_t* img, int w, int h, const IdentParams& p,
              const std::vector<std::vector<float>>& bank, std::vector<IdentMarker>& markers,
              cctagx::TaskPool* pool = nullptr);

// Ellipse.cpp:259-298: the points sorted by their angle about the ellipse
// centre, then sub-sampled to at most `requested`.
void sorted_outer_points(float cx, float cy, const std::vector<DPt>& points,
                         std::vector<DPt>& out, size_t requested);

// computeHomographyFromEllipseAndImagedCenter (Identification.cpp:739-798)
// for the ellipse (cx, cy, a, b, angle) and the imaged centre (x, y).
void homography_from_center(float cx, float cy, float a, float b, float angle, float x, float y,
                            float H[9]);

// orazioDistanceRobust's distance of one cut's signal (sampled from `begin`
// to `end` of the unit radius) to every bank row (Identification.cpp:50-160).
void cut_distances(const std::vector<float>& sig, float begin, float end,
                   const std::vector<std::vector<float>>& bank, std::vector<float>& out);

// Where identify_step_1's cuts start, as a fraction of the unit radius
// (Identification.cpp:1221-1235).
float start_signal(int n_crowns);

// One cut's signal: with `homography`, extractSignalUsingHomography
// (Identification.cpp:216-280) through the 3x3 row-major H (marker plane ->
// image) towards `stop`; else cutInterpolated (:396-459) from the start
// point (H[0], H[1]) to `stop`.  n samples into sig; true when a sample
// fell outside the image.
bool sample_cut(const uint8_t* img, int w, int h, const float* H, const DPt& stop, float begin,
                float end, int n, bool homography, float* sig);

// outerEdgeRefinement (Identification.cpp:604-657) of one cut stop; false
// when its refinement cut leaves the image.
bool refine_stop(const uint8_t* img, int w, int h, DPt& stop, float scale, int n_ref);

// The conditioner of conditionerFromEllipse (conditioner.hpp:22-46) when
// the ellipse with semi-axes a, b set its cached scale, and the grid
// search's iteration count for the ellipse (Identification.cpp:869-890).
void conditioner(float cx, float cy, float a, float b, float T[9]);
int grid_iterations(float a, float b, float neighbour_size, int grid_n);

}  // namespace oracle
