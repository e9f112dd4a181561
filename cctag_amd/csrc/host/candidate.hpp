// Candidate stage: from one level's edge points, links, votes and seeds to
// marker candidates (outer ellipse + ring points), then the level-0 re-fit of
// markers found at coarser levels.  Host code, task-parallel over seeds'
// flow components, candidates and markers of all levels at once; the front
// end (GPU, or the oracle's CPU restatement) provides its input.
//
// Reference: Detection.cpp:69-693, Vote.cpp:249-742, EllipseGrowing.cpp,
// Fitting.cpp:228-364, Multiresolution.cpp:43-193,268-413, CCTag.{hpp,cpp}.
#pragma once

#include <cstdint>
#include <vector>

#include "geometry.hpp"
#include "pool.hpp"

namespace cctagx {

// Borrowed view of one level's front-end output (raster-order points).
struct LevelEdges {
  int w = 0, h = 0;              // level size
  int shape_w = 0, shape_h = 0;  // EdgePointCollection shape (level-0 size)
  // ceil(w*h/32) words, bit i of word j set when pixel 32j+i (raster order)
  // is an edge point; a point's index is its rank among the set bits.
  const uint32_t* bits = nullptr;
  int n = 0;
  const int16_t *x = nullptr, *y = nullptr, *gx = nullptr, *gy = nullptr;
  const int32_t *before = nullptr, *after = nullptr;
  const int32_t *vindex = nullptr, *vlist = nullptr;  // CSR voter lists
  const float* flow = nullptr;
  int n_seeds = 0;
  const int32_t* seeds = nullptr;  // in the order they reached min votes
};

// EdgePoint (EdgePoint.hpp:24-86) with the candidate stage's mutable state.
struct EdgePoint {
  int x, y;
  float gx, gy, norm;  // gradient and its norm
  float flow_length;
  int32_t is_max;
  uint64_t processed;  // per-candidate bit masks (EllipseGrowing.cpp:332)
  int32_t n_segment_out;
};

// EdgePointCollection (Types.hpp:19-158).
class EdgeSet {
 public:
  // Loading a level, in steps: sizes and flags; the point index of the first edge of
  // every 32-pixel word of the bit mask, for words [w0, w1) given the edges
  // before w0 (word_edges counts the edges of a word range); the points
  // [begin, end).  Independent ranges may be done concurrently.
  void prepare(const LevelEdges& L);
  size_t n_words() const { return rank_.size(); }
  int32_t word_edges(size_t w0, size_t w1) const;
  void rank_words(size_t w0, size_t w1, int32_t before);
  // Points [begin, end): is_max is the point's number of voters when it has
  // at least min_votes, else -1 (Vote.cpp:236-240), read off the voter lists.
  void fill(const LevelEdges& L, int begin, int end, int min_votes);
  // Point index at (x, y) or -1; the caller guarantees 0 <= x < shape_w and
  // 0 <= y < shape_h (the collection's shape is the level-0 size).
  int at(int x, int y) const {
    if (x >= w_ || y >= h_) return -1;
    const size_t i = (size_t)y * w_ + x;
    const uint32_t word = bits_[i >> 5], bit = 1u << (i & 31);
    if (!(word & bit)) return -1;
    return rank_[i >> 5] + __builtin_popcount(word & (bit - 1));
  }
  int shape_w() const { return shape_w_; }
  int shape_h() const { return shape_h_; }
  int size() const { return (int)pts.size(); }
  int before(int i) const { return before_[i]; }
  int after(int i) const { return after_[i]; }
  int voters_size(int i) const { return vindex_[i + 1] - vindex_[i]; }
  const int32_t* voters_begin(int i) const { return vlist_ + vindex_[i]; }
  const int32_t* voters_end(int i) const { return vlist_ + vindex_[i + 1]; }
  // Types.hpp:142-155: an invalid index throws std::out_of_range.
  void set_processed_in(int i, bool f) { pin_.at(check(i)) = f; }
  bool test_processed_in(int i) const { return pin_.at(check(i)); }
  std::vector<EdgePoint> pts;
  std::vector<int32_t> seeds;

 private:
  static int check(int i) {
    if (i < 0) throw std::out_of_range("EdgePointCollection: null point");
    return i;
  }
  int w_ = 0, h_ = 0, shape_w_ = 0, shape_h_ = 0;
  const uint32_t* bits_ = nullptr;
  std::vector<int32_t> rank_;  // set bits before each word
  const int32_t *before_ = nullptr, *after_ = nullptr, *vindex_ = nullptr, *vlist_ = nullptr;
  std::vector<uint8_t> pin_;
};

// The detection parameters the host stages read (Params.hpp:108-199).
struct HostParams {
  float average_vote_min = 0.f;
  float thr_median_distance_ellipse = 3.f;
  int maximum_nb_seeds = 500;
  int maximum_nb_candidates_loop_two = 40;
  int n_crowns = 3;
  int min_points_segment_candidate = 10;
  int min_votes_to_select_candidate = 3;
  float thresh_robust_estimation_of_outer_ellipse = 30.f;
  float ellipse_growing_elliptic_hull_width = 2.3f;
  int window_size_on_inner_elliptic_segment = 20;
  int n_samples_outer_ellipse = 150;
  int num_cuts_in_ident_step = 22;
  int num_samples_outer_edge_points_refinement = 20;
  int sample_cut_length = 100;
  int imaged_center_n_grid_sample = 5;
  float imaged_center_neighbour_size = 0.2f;
  float min_ident_proba = 1e-6f;
  bool search_for_another_segment = true;
  bool do_identification = true;
};

// Marker record (CCTag.hpp:49-409, the fields detection and identification use).
struct Marker {
  int id = -1;
  int status = 0;
  float quality = 0.f;
  int pyramid_level = 0;
  float scale = 1.f;
  Pt2 center_img;                       // _centerImg
  Ellipse outer_ellipse;                // _outerEllipse (level coordinates, +0.5 shift)
  Ellipse rescaled_outer_ellipse;       // _rescaledOuterEllipse (level-0 coordinates)
  std::vector<std::vector<Pt2>> points; // _points, one list per circle, outer last
  std::vector<Pt2> rescaled_outer_ellipse_points;
  Mat3 homography = Mat3::zero();
  std::vector<float> radius_ratios;
  std::vector<Pt2> cut_stops;           // stops of the cuts identify_step_1 selected

  // CCTag.hpp:68-93 constructor.
  static Marker make(const Ellipse& outer, std::vector<std::vector<Pt2>> pts, int level,
                     float scale, float quality);
  // CCTag.cpp:42-59: the outer ellipses shrunk to circles of half their
  // minor axis overlap.
  bool is_equal(const Marker& other) const;
  Ellipse half_circle() const;
};

// cctagDetectionFromEdges (Detection.cpp:568-693) for every level at once,
// appending markers coarse level first.  Loop one (seed -> flow component)
// runs per level in seed order; loop two (completeFlowComponent), the final
// candidate loop and the level-0 re-fit run as independent tasks over all
// levels on `pool` (serially when pool is null), each task with its own
// rand_5_k stream seeded 271828 and its own auxiliary point flags; a level's
// segment labels are assigned in candidate order once its loop two is done.
// The result does not depend on the number of threads.
// Wall time of the host stages of one detection, milliseconds.  Loop two,
// loop three and the re-fit overlap: loop_two runs until the last loop-two
// task ends, loop_three from there until the last loop-three task ends,
// rescale the rest; labels is the label assignment's own time inside them.
struct HostTimes {
  double load = 0, loop_one = 0, loop_two = 0, labels = 0, loop_three = 0, rescale = 0,
         ident = 0, finish = 0;
  // inside `ident`: host preparation and result reading, and the backend
  // calls (transfers and kernels included; selection and refinement in one
  // call count as refinement)
  double ident_prep = 0, ident_select = 0, ident_refine = 0, ident_post = 0;
};

// The candidate loops of every level (Detection.cpp:603-688) and the level-0
// re-fit of the markers found above level 0 (Multiresolution.cpp:319-401):
// the markers in level order (coarsest first), then candidate order.
void detect_candidates(std::vector<EdgeSet>& levels, const std::vector<int>& level_h,
                       const HostParams& p, TaskPool* pool, std::vector<Marker>& markers,
                       HostTimes* times = nullptr);

// The edge linking's decisions on gradient angles (Vote.cpp:290-340), as
// functions of the gradients: the first neighbour direction to try,
// round(((angle + pi/4) / 2pi) * 8) - 1 modulo 8, and the convexity test
// between the window's oldest (f) and newest (b) gradient direction (with a
// full window dir * sin(b - f) < 0, otherwise s < -0.707 with c > 0 or s < 0
// with c < 0 for s = dir * sin(b - f), c = cos(b - f)), angles being
// gradient_angle_of in float.  They give the float expressions' results.
// fmod(atan2(gy, gx) + 2 pi, 2 pi) (Vote.cpp:296).
float gradient_angle_of(float gx, float gy);
int linking_shift(float gx, float gy);
bool convexity_lost(float fx, float fy, float bx, float by, int dir, bool full);

// Multiresolution.cpp:172-193 over a whole list: `order` names markers in
// arrival order; the result names the kept ones (an entry replaced by a
// better-quality overlapping arrival takes that arrival's index).  `half`
// holds each marker's comparison circle (Marker::half_circle).
std::vector<int> update_markers(const std::vector<Marker>& markers,
                                const std::vector<Ellipse>& half, const std::vector<int>& order);

}  // namespace cctagx
