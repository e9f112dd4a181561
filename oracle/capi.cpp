// ORACLE -- test infrastructure only (see oracle.hpp).
// ctypes-facing wrappers around the CPU restatement.  The level view struct
// is the product's (include/cctag_amd.h) so tests compare like with like.
#include "../include/cctag_amd.h"
#include "oracle.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <exception>
#include <stdexcept>
#include <memory>
#include <vector>

namespace {
struct OracleRun {
  oracle::Params p;
  std::vector<oracle::LevelResult> levels;
};

template <class T>
const T* ptr(const std::vector<T>& v) { return v.empty() ? nullptr : v.data(); }
}  // namespace

extern "C" {

int oracle_front_end(const uint8_t* img, int w, int h, int n_levels, void** out) {
  try {
    auto* r = new OracleRun();
    r->p.n_levels = n_levels;
    oracle::front_end(img, w, h, r->p, r->levels);
    *out = r;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int oracle_level_view(void* handle, int level, cctagx_level_view* v) {
  auto* r = static_cast<OracleRun*>(handle);
  if (!r || level < 0 || level >= (int)r->levels.size()) return -1;
  const oracle::LevelResult& L = r->levels[level];
  v->width = L.w; v->height = L.h; v->shape_w = L.W0; v->shape_h = L.H0;
  v->img = ptr(L.img); v->dx = ptr(L.dx); v->dy = ptr(L.dy); v->edges = ptr(L.edges);
  v->edge_map = ptr(L.edge_map);
  v->n_edges = (int32_t)L.ex.size();
  v->edge_x = ptr(L.ex); v->edge_y = ptr(L.ey); v->edge_dx = ptr(L.edx); v->edge_dy = ptr(L.edy);
  v->before = ptr(L.before); v->after = ptr(L.after);
  v->chosen = ptr(L.chosen); v->total_dist = ptr(L.total_dist);
  v->voters_index = ptr(L.voters_index); v->voters_list = ptr(L.voters_list);
  v->flow_length = ptr(L.flow_length); v->is_max = ptr(L.is_max);
  v->n_seeds = (int32_t)L.seeds.size(); v->seeds = ptr(L.seeds);
  return 0;
}

void oracle_free(void* handle) { delete static_cast<OracleRun*>(handle); }

// Edge points, descent links, votes and seeds (Canny.cpp:15-34, Vote.cpp:
// 58-247) of one level given its thinned edge map and derivatives.
int oracle_edges_and_vote(const uint8_t* edges, const int16_t* dx, const int16_t* dy, int w, int h,
                          int dist_search, int thr_grad, float ratio_voting, int n_crowns,
                          int min_votes, void** out) {
  try {
    auto* r = new OracleRun();
    r->p.dist_search = dist_search;
    r->p.thr_grad_vote = thr_grad;
    r->p.ratio_voting = ratio_voting;
    r->p.n_crowns = n_crowns;
    r->p.min_votes = min_votes;
    r->levels.resize(1);
    oracle::LevelResult& L = r->levels[0];
    L.w = L.W0 = w;
    L.h = L.H0 = h;
    const size_t N = (size_t)w * h;
    L.img.assign(N, 0);
    L.edges.assign(edges, edges + N);
    L.dx.assign(dx, dx + N);
    L.dy.assign(dy, dy + N);
    oracle::edges_and_vote(L, r->p);
    *out = r;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

void oracle_derivatives(const uint8_t* img, int w, int h, int16_t* dx, int16_t* dy) {
  oracle::derivatives(img, w, h, dx, dy);
}

void oracle_derivatives_float(const uint8_t* img, int w, int h, float* dx, float* dy) {
  oracle::derivatives_float(img, w, h, dx, dy);
}

void oracle_canny(const int16_t* dx, const int16_t* dy, int w, int h, float low, float high,
                  uint8_t* edges) {
  oracle::canny(dx, dy, w, h, low, high, edges);
}

void oracle_thin(uint8_t* edges, int w, int h) { oracle::thin(edges, w, h); }

int oracle_thin_lut(int pass, int index) { return oracle::thin_lut(pass, index); }

void oracle_bgr_to_gray(const uint8_t* src, int w, int h, size_t stride, int channels,
                        uint8_t* dst) {
  oracle::bgr_to_gray(src, w, h, stride, channels, dst);
}

int oracle_downsample(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh) {
  try {
    oracle::downsample(src, sw, sh, dst, dw, dh);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Wall time of the whole front end on this host, for the CPU baseline.
double oracle_time_front_end(const uint8_t* img, int w, int h, int n_levels, int iters) {
  oracle::Params p;
  p.n_levels = n_levels;
  std::vector<oracle::LevelResult> levels;
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < iters; ++i) oracle::front_end(img, w, h, p, levels);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count() / iters;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Full CPU detection: the oracle front end, candidate stage, identification
// and de-duplication.  The library's host candidate stage (pipeline.hpp) is
// linked only as the thing under test of oracle_candidate_mismatch.
#include "../cctag_amd/csrc/host/cabi.hpp"
#include "../cctag_amd/csrc/host/pipeline.hpp"
#include "oracle_cand.hpp"
#include "oracle_ident.hpp"

namespace {
std::vector<cctagx::LevelEdges> level_edges(const std::vector<oracle::LevelResult>& levels) {
  std::vector<cctagx::LevelEdges> out;
  for (const oracle::LevelResult& L : levels) {
    cctagx::LevelEdges e;
    e.w = L.w; e.h = L.h; e.shape_w = L.W0; e.shape_h = L.H0;
    e.bits = L.edge_bits.data();
    e.n = (int)L.ex.size();
    e.x = L.ex.data(); e.y = L.ey.data(); e.gx = L.edx.data(); e.gy = L.edy.data();
    e.before = L.before.data(); e.after = L.after.data();
    e.vindex = L.voters_index.data(); e.vlist = L.voters_list.data();
    e.flow = L.flow_length.data();
    e.n_seeds = (int)L.seeds.size(); e.seeds = L.seeds.data();
    out.push_back(e);
  }
  return out;
}

oracle::CandidateParams candidate_params(const cctagx_params& c) {
  oracle::CandidateParams p;
  p.average_vote_min = c.average_vote_min;
  p.thr_median_distance_ellipse = c.thr_median_distance_ellipse;
  p.maximum_nb_seeds = c.maximum_nb_seeds;
  p.maximum_nb_candidates_loop_two = c.maximum_nb_candidates_loop_two;
  p.n_crowns = c.n_crowns;
  p.min_points_segment_candidate = c.min_points_segment_candidate;
  p.thresh_robust_estimation_of_outer_ellipse = c.thresh_robust_estimation_of_outer_ellipse;
  p.ellipse_growing_elliptic_hull_width = c.ellipse_growing_elliptic_hull_width;
  p.window_size_on_inner_elliptic_segment = c.window_size_on_inner_elliptic_segment;
  p.search_for_another_segment = c.search_for_another_segment != 0;
  return p;
}

oracle::Params front_params(const cctagx_params& c) {
  oracle::Params p;
  p.canny_low = c.canny_thr_low;
  p.canny_high = c.canny_thr_high;
  p.dist_search = c.dist_search;
  p.thr_grad_vote = c.thr_gradient_mag_in_vote;
  p.angle_voting = c.angle_voting;
  p.ratio_voting = c.ratio_voting;
  p.n_crowns = c.n_crowns;
  p.min_votes = c.min_votes_to_select_candidate;
  p.n_levels = c.number_of_processed_multires_layers;
  return p;
}

oracle::IdentParams ident_params(const cctagx_params& c) {
  oracle::IdentParams p;
  p.n_crowns = c.n_crowns;
  p.n_samples_outer_ellipse = c.n_samples_outer_ellipse;
  p.num_cuts_in_ident_step = c.num_cuts_in_ident_step;
  p.num_samples_outer_edge_points_refinement = c.num_samples_outer_edge_points_refinement;
  p.sample_cut_length = c.sample_cut_length;
  p.imaged_center_n_grid_sample = c.imaged_center_n_grid_sample;
  p.imaged_center_neighbour_size = c.imaged_center_neighbour_size;
  p.min_ident_proba = c.min_ident_proba;
  return p;
}

// The banks the oracle identifies with, set by the tests from the
// reference's tables (oracle_set_bank); indexed by ring count.
std::vector<std::vector<float>> g_bank[5];

const std::vector<std::vector<float>>& bank_of(int n_crowns) {
  if (n_crowns < 0 || n_crowns > 4 || g_bank[n_crowns].empty())
    throw std::runtime_error("oracle: no bank set for this ring count");
  return g_bank[n_crowns];
}

// The oracle identification of the candidate stage's markers, written back
// into them.
void identify_markers(std::vector<cctagx::Marker>& markers, const uint8_t* img, int w, int h,
                      const cctagx_params& params, const std::vector<std::vector<float>>& bank,
                      cctagx::TaskPool* pool = nullptr) {
  std::vector<oracle::IdentMarker> im(markers.size());
  for (size_t i = 0; i < markers.size(); ++i) {
    const cctagx::Marker& m = markers[i];
    oracle::IdentMarker& o = im[i];
    const cctagx::Ellipse& e = m.rescaled_outer_ellipse;
    o.cx = e.cx(); o.cy = e.cy(); o.a = e.a(); o.b = e.b(); o.angle = e.angle();
    for (const cctagx::Pt2& q : m.rescaled_outer_ellipse_points)
      o.points.push_back(oracle::DPt{q.x, q.y, q.dx, q.dy});
    o.center_x = m.center_img.x;
    o.center_y = m.center_img.y;
    o.scale = m.scale;
  }
  oracle::identify(img, w, h, ident_params(params), bank, im, pool);
  for (size_t i = 0; i < markers.size(); ++i) {
    cctagx::Marker& m = markers[i];
    const oracle::IdentMarker& o = im[i];
    m.status = o.status;
    m.id = o.id;
    m.center_img = cctagx::Pt2(o.center_x, o.center_y);
    if (o.status == 1 || o.status == -4 || o.status == -5 || o.status == -3) m.quality = o.quality;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m.homography.m[r][c] = o.H[r * 3 + c];
    m.cut_stops.clear();
    for (const oracle::DPt& q : o.stops) m.cut_stops.emplace_back(q.x, q.y, q.dx, q.dy);
  }
}
}  // namespace

extern "C" {

// The front end with the given parameters (a handle for oracle_level_view).
int oracle_front_end_params(const cctagx_params* params, const uint8_t* img, int w, int h,
                            void** out) {
  try {
    auto* r = new OracleRun();
    r->p = front_params(*params);
    oracle::front_end(img, w, h, r->p, r->levels);
    *out = r;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// The bank of n_crowns rings (rows x cols radius ratios, row-major) the
// oracle identifies with.
int oracle_set_bank(int n_crowns, const float* ratios, int rows, int cols) {
  if (n_crowns < 0 || n_crowns > 4 || !ratios || rows < 1 || cols < 1) return -1;
  g_bank[n_crowns].assign(rows, std::vector<float>(cols));
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) g_bank[n_crowns][i][j] = ratios[(size_t)i * cols + j];
  return 0;
}

namespace {
bool same_f(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }
bool same_ellipse(const cctagx::Ellipse& a, const cctagx::Ellipse& b) {
  return same_f(a.cx(), b.cx()) && same_f(a.cy(), b.cy()) && same_f(a.a(), b.a()) &&
         same_f(a.b(), b.b()) && same_f(a.angle(), b.angle());
}
bool same_points(const std::vector<cctagx::Pt2>& a, const std::vector<cctagx::Pt2>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!same_f(a[i].x, b[i].x) || !same_f(a[i].y, b[i].y) || !same_f(a[i].dx, b[i].dx) ||
        !same_f(a[i].dy, b[i].dy))
      return false;
  return true;
}
}  // namespace

// The library's host candidate stage (the thing under test, on `threads`
// threads) against the oracle's on the oracle front end's levels of one
// frame, before identification: the same markers in the same order with
// bit-identical ellipses, qualities and point lists.  0 when identical,
// 1 with a description in msg, -1 on an error.
int oracle_candidate_mismatch(const cctagx_params* params, const uint8_t* img, int w, int h,
                              int threads, int* n_markers, char* msg, int msg_len) {
  auto say = [&](const std::string& t) {
    if (msg && msg_len > 0) std::snprintf(msg, (size_t)msg_len, "%s", t.c_str());
  };
  try {
    std::vector<oracle::LevelResult> levels;
    oracle::front_end(img, w, h, front_params(*params), levels);
    std::unique_ptr<cctagx::TaskPool> pool;
    if (threads > 1) pool = std::make_unique<cctagx::TaskPool>(threads);
    std::vector<cctagx::Marker> lib, ref;
    cctagx::frame_candidates(level_edges(levels), cctagx::host_params(*params), pool.get(), lib);
    oracle::candidates(levels, candidate_params(*params), ref, nullptr);
    if (n_markers) *n_markers = (int)ref.size();
    if (lib.size() != ref.size()) {
      say("marker count " + std::to_string(lib.size()) + " vs oracle " + std::to_string(ref.size()));
      return 1;
    }
    for (size_t i = 0; i < ref.size(); ++i) {
      const cctagx::Marker &a = lib[i], &b = ref[i];
      std::string what;
      if (a.pyramid_level != b.pyramid_level || !same_f(a.scale, b.scale)) what = "level";
      else if (!same_ellipse(a.outer_ellipse, b.outer_ellipse)) what = "outer ellipse";
      else if (!same_ellipse(a.rescaled_outer_ellipse, b.rescaled_outer_ellipse)) what = "rescaled ellipse";
      else if (!same_f(a.center_img.x, b.center_img.x) || !same_f(a.center_img.y, b.center_img.y)) what = "centre";
      else if (!same_f(a.quality, b.quality)) what = "quality";
      else if (a.points.size() != b.points.size()) what = "circle count";
      else if (!same_points(a.rescaled_outer_ellipse_points, b.rescaled_outer_ellipse_points)) what = "rescaled points";
      else
        for (size_t k = 0; k < a.points.size() && what.empty(); ++k)
          if (!same_points(a.points[k], b.points[k])) what = "points of circle " + std::to_string(k);
      if (!what.empty()) {
        char buf[256];
        std::snprintf(buf, sizeof buf, " (level %d, centre %.3f %.3f vs %.3f %.3f)", b.pyramid_level,
                      a.outer_ellipse.cx(), a.outer_ellipse.cy(), b.outer_ellipse.cx(), b.outer_ellipse.cy());
        say("marker " + std::to_string(i) + ": " + what + buf);
        return 1;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    say(e.what());
    return -1;
  }
}

// The library's ordering of a marker's outer points for the cut selection
// (cctagx::sorted_outer_points, a radix sort of the angles) against the
// oracle's (a sort of (angle, index) pairs) on the same points about the
// centre (cx, cy): 0 when the two pick the same points in the same order.
int oracle_outer_points_mismatch(const float* x, const float* y, int n, float cx, float cy,
                                 int requested) {
  std::vector<cctagx::Pt2> lib_in;
  std::vector<oracle::DPt> ref_in;
  for (int i = 0; i < n; ++i) {
    lib_in.emplace_back(x[i], y[i]);
    oracle::DPt d;
    d.x = x[i];
    d.y = y[i];
    ref_in.push_back(d);
  }
  const cctagx::Ellipse e(cx, cy, 10.f, 5.f, 0.f);
  if (e.cx() != cx || e.cy() != cy) return -1;
  std::vector<cctagx::Pt2> lib_out;
  std::vector<oracle::DPt> ref_out;
  cctagx::sorted_outer_points(e, lib_in, lib_out, (size_t)requested);
  oracle::sorted_outer_points(cx, cy, ref_in, ref_out, (size_t)requested);
  if (lib_out.size() != ref_out.size()) return 1;
  for (size_t i = 0; i < lib_out.size(); ++i)
    if (lib_out[i].x != ref_out[i].x || lib_out[i].y != ref_out[i].y) return 1;
  return 0;
}

// cctagDetection on the CPU; returns the marker count or -1.  `threads` > 1
// runs the independent tasks of the candidate stage (loops two and three,
// the level-0 re-fit) and the identification of the markers on a pool of
// that many threads (the reference's tbb::parallel_for loops), 1 serially;
// the front end and loop one are serial either way.
int oracle_detect_threads(const cctagx_params* params, const uint8_t* img, int w, int h,
                          cctagx_marker* out, int max_out, int threads) {
  try {
    const auto& bank = bank_of(params->n_crowns);
    std::vector<oracle::LevelResult> levels;
    oracle::front_end(img, w, h, front_params(*params), levels);
    std::unique_ptr<cctagx::TaskPool> pool;
    if (threads > 1) pool = std::make_unique<cctagx::TaskPool>(threads);
    std::vector<cctagx::Marker> markers, kept;
    oracle::candidates(levels, candidate_params(*params), markers, pool.get());
    if (params->do_identification)
      identify_markers(markers, levels[0].img.data(), w, h, *params, bank, pool.get());
    oracle::deduplicate(markers, kept);
    for (size_t i = 0; i < kept.size() && (int)i < max_out; ++i) cctagx::to_c_marker(kept[i], out[i]);
    return (int)kept.size();
  } catch (const std::exception&) {
    return -1;
  }
}

// The oracle identification of given markers on an 8-bit frame (w x h,
// stride w), the counterpart of cctagx_identify.  Returns 0, or -1 on an
// exception.
int oracle_identify_given(const cctagx_params* params, const uint8_t* img, int w, int h,
                          const cctagx_ident_marker* in, int n, cctagx_ident_result* out) {
  try {
    const auto& bank = bank_of(params->n_crowns);
    std::vector<oracle::IdentMarker> im(n);
    for (int k = 0; k < n; ++k) {
      const cctagx_ident_marker& a = in[k];
      oracle::IdentMarker& o = im[k];
      o.cx = a.ellipse[0]; o.cy = a.ellipse[1]; o.a = a.ellipse[2]; o.b = a.ellipse[3];
      o.angle = a.ellipse[4];
      for (int j = 0; j < a.n_points; ++j)
        o.points.push_back(oracle::DPt{a.points[4 * j], a.points[4 * j + 1], a.points[4 * j + 2],
                                       a.points[4 * j + 3]});
      o.center_x = a.center[0];
      o.center_y = a.center[1];
      o.scale = a.scale;
    }
    oracle::identify(img, w, h, ident_params(*params), bank, im);
    for (int k = 0; k < n; ++k) {
      const oracle::IdentMarker& o = im[k];
      cctagx_ident_result& r = out[k];
      r.status = o.status;
      r.id = o.id;
      r.x = o.center_x;
      r.y = o.center_y;
      r.quality = o.quality;
      for (int i = 0; i < 9; ++i) r.homography[i] = o.H[i];
      r.n_cuts = (int)std::min<size_t>(o.stops.size(), 32);
      for (int c = 0; c < 32; ++c) {
        r.stops[c][0] = c < r.n_cuts ? o.stops[c].x : 0.f;
        r.stops[c][1] = c < r.n_cuts ? o.stops[c].y : 0.f;
      }
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// The oracle's closed forms, for the tests that pin them: the homography of
// an ellipse and an imaged centre, the conditioner and the grid search's
// iteration count, the sorted and sub-sampled outer points (indices into
// the input written to idx; returns their count).
void oracle_homography_from_center(const float* ell, float x, float y, float* H) {
  oracle::homography_from_center(ell[0], ell[1], ell[2], ell[3], ell[4], x, y, H);
}
void oracle_conditioner(float cx, float cy, float a, float b, float* T) {
  oracle::conditioner(cx, cy, a, b, T);
}
int oracle_sample_cut(const uint8_t* img, int w, int h, const float* H, const float* stop,
                      float begin, float end, int n, int homography, float* sig) {
  const oracle::DPt p{stop[0], stop[1], 0.f, 0.f};
  return oracle::sample_cut(img, w, h, H, p, begin, end, n, homography != 0, sig) ? 1 : 0;
}
int oracle_cut_distances(int n_crowns, const float* sig, int n, float begin, float end,
                         float* out) {
  try {
    std::vector<float> d;
    oracle::cut_distances(std::vector<float>(sig, sig + n), begin, end, bank_of(n_crowns), d);
    std::copy(d.begin(), d.end(), out);
    return (int)d.size();
  } catch (const std::exception&) {
    return -1;
  }
}
float oracle_start_signal(int n_crowns) { return oracle::start_signal(n_crowns); }
int oracle_refine_stop(const uint8_t* img, int w, int h, float* stop, float scale, int n_ref) {
  oracle::DPt p{stop[0], stop[1], stop[2], stop[3]};
  if (!oracle::refine_stop(img, w, h, p, scale, n_ref)) return 0;
  stop[0] = p.x;
  stop[1] = p.y;
  return 1;
}
int oracle_grid_iterations(float a, float b, float neighbour, int grid_n) {
  return oracle::grid_iterations(a, b, neighbour, grid_n);
}
int oracle_sorted_outer_points(float cx, float cy, const float* pts, int n, int requested,
                               int* idx) {
  std::vector<oracle::DPt> in(n), outp;
  for (int i = 0; i < n; ++i) in[i] = oracle::DPt{pts[2 * i], pts[2 * i + 1], (float)i, 0.f};
  oracle::sorted_outer_points(cx, cy, in, outp, (size_t)requested);
  for (size_t k = 0; k < outp.size(); ++k) idx[k] = (int)outp[k].dx;
  return (int)outp.size();
}

int oracle_detect(const cctagx_params* params, const uint8_t* img, int w, int h,
                  cctagx_marker* out, int max_out) {
  return oracle_detect_threads(params, img, w, h, out, max_out, 1);
}

// The edge linking's angle decisions of the shared candidate stage
// (cctagx::linking_shift, cctagx::convexity_lost) against the reference's
// float expressions on gradient angles (Vote.cpp:296-340), for n cases
// g[5i .. 5i+4] = (gx_front, gy_front, gx_back, gy_back, dir): the number of
// cases in which any of the three decisions differs.  The first neighbour
// direction is compared modulo 8, the only way the linking uses it.
int oracle_linking_mismatches(const float* g, int n) {
  const float pi = 3.14159265358979323846f;
  auto angle = [&](float gx, float gy) { return std::fmod(std::atan2(gy, gx) + 2.0f * pi, 2.0f * pi); };
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    const float fx = g[5 * i], fy = g[5 * i + 1], bx = g[5 * i + 2], by = g[5 * i + 3];
    const int dir = g[5 * i + 4] < 0.f ? -1 : 1;
    const float af = angle(fx, fy), ab = angle(bx, by);
    const int shifting = (int)std::round(((af + pi / 4.0f) / (2.0f * pi)) * 8.0f) - 1;
    const bool full = dir * std::sin(ab - af) < 0.0f;
    const float s = dir * std::sin(ab - af), c = std::cos(ab - af);
    const bool part = ((s < -0.707f) && (c > 0.f)) || ((s < 0.f) && (c < 0.f));
    if (cctagx::linking_shift(fx, fy) != ((shifting % 8) + 8) % 8 ||
        cctagx::convexity_lost(fx, fy, bx, by, dir, true) != full ||
        cctagx::convexity_lost(fx, fy, bx, by, dir, false) != part)
      ++bad;
  }
  return bad;
}

// The RANSAC trials' classification of the shared candidate stage
// (cctagx::conic_trial_class_fast) against the float derivation of
// Ellipse(Q) and its ratio test (Vote.cpp:486-503) for n conics
// t[5i .. 5i+4] with t0 t2 - t1^2 > 0: the number of conics the fast test
// decides differently; *undecided receives the number it leaves to the
// float derivation.
int oracle_conic_class_mismatches(const float* t, int n, float lo, float hi, int* undecided) {
  int bad = 0, open = 0;
  for (int i = 0; i < n; ++i) {
    const float* c = t + 5 * i;
    cctagx::Mat3 Q = cctagx::Mat3::zero();
    Q.m[0][0] = c[0];
    Q.m[0][1] = Q.m[1][0] = c[1];
    Q.m[0][2] = Q.m[2][0] = c[3];
    Q.m[1][1] = c[2];
    Q.m[1][2] = Q.m[2][1] = c[4];
    Q.m[2][2] = 1.f;
    const int fast = cctagx::conic_trial_class_fast(c, lo, hi);
    if (fast < 0) ++open;
    else if (fast != cctagx::conic_trial_class_exact(Q, lo, hi)) ++bad;
  }
  if (undecided) *undecided = open;
  return bad;
}

// Wall time of the full CPU detection (front end + candidates + identification)
// on `threads` threads.
double oracle_time_detect(const cctagx_params* params, const uint8_t* img, int w, int h, int iters,
                          int threads) {
  std::vector<cctagx_marker> out(256);
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < iters; ++i)
    oracle_detect_threads(params, img, w, h, out.data(), (int)out.size(), threads);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count() / iters;
}

}  // extern "C"
