// ORACLE -- test infrastructure only (see oracle.hpp).
// ctypes-facing wrappers around the CPU restatement.  The level view struct
// is the product's (include/cctag_amd.h) so tests compare like with like.
#include "../include/cctag_amd.h"
#include "oracle.hpp"

#include <chrono>
#include <cmath>
#include <exception>
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
// Full CPU detection: the oracle front end and identification around the
// shared host candidate stage.
#include "../cctag_amd/csrc/host/cabi.hpp"
#include "../cctag_amd/csrc/host/pipeline.hpp"
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
    e.flow = L.flow_length.data(); e.ismax = L.is_max.data();
    e.n_seeds = (int)L.seeds.size(); e.seeds = L.seeds.data();
    out.push_back(e);
  }
  return out;
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
}  // namespace

extern "C" {

// cctagDetection on the CPU; returns the marker count or -1.  `threads` > 1
// runs the shared host candidate stage and the per-candidate identification
// on a pool of that many threads (the reference's tbb::parallel_for loops),
// 1 serially; the front end is serial either way.
int oracle_detect_threads(const cctagx_params* params, const uint8_t* img, int w, int h,
                          cctagx_marker* out, int max_out, int threads) {
  try {
    std::vector<oracle::LevelResult> levels;
    oracle::front_end(img, w, h, front_params(*params), levels);
    const auto bank = cctagx::marker_bank(params->n_crowns);
    oracle::CpuIdentBackend backend(levels[0].img.data(), w, h, bank);
    std::vector<cctagx::Marker> markers;
    std::unique_ptr<cctagx::TaskPool> pool;
    if (threads > 1) pool = std::make_unique<cctagx::TaskPool>(threads);
    backend.set_pool(pool.get());
    cctagx::detect_markers(level_edges(levels), cctagx::host_params(*params), bank, &backend,
                           markers, pool.get());
    for (size_t i = 0; i < markers.size() && (int)i < max_out; ++i) cctagx::to_c_marker(markers[i], out[i]);
    return (int)markers.size();
  } catch (const std::exception&) {
    return -1;
  }
}

// The identification of given markers on an 8-bit frame (identify_markers
// with this backend, serially): marker k has the outer ellipse ell[5k ..
// 5k+4] (cx, cy, a, b, angle), npts[k] outer points pts[4j .. 4j+3]
// (x, y, gx, gy; the markers' points one after the other) and the starting
// centre c0[2k], c0[2k+1].  Writes each marker's status, ID and imaged
// centre.  Returns 0, or -1 on an exception.
int oracle_identify_given(const cctagx_params* params, const uint8_t* img, int w, int h, int n,
                          const float* ell, const int* npts, const float* pts, const float* c0,
                          int* status, int* id, float* center) {
  try {
    const auto bank = cctagx::marker_bank(params->n_crowns);
    oracle::CpuIdentBackend backend(img, w, h, bank);
    std::vector<cctagx::Marker> markers(n);
    size_t at = 0;
    for (int k = 0; k < n; ++k) {
      cctagx::Marker& m = markers[k];
      m.rescaled_outer_ellipse =
          cctagx::Ellipse(ell[5 * k], ell[5 * k + 1], ell[5 * k + 2], ell[5 * k + 3], ell[5 * k + 4]);
      m.outer_ellipse = m.rescaled_outer_ellipse;
      for (int j = 0; j < npts[k]; ++j, ++at)
        m.rescaled_outer_ellipse_points.emplace_back(pts[4 * at], pts[4 * at + 1], pts[4 * at + 2],
                                                     pts[4 * at + 3]);
      m.center_img = cctagx::Pt2(c0[2 * k], c0[2 * k + 1]);
    }
    cctagx::identify_markers(markers, cctagx::host_params(*params), bank, backend, nullptr);
    for (int k = 0; k < n; ++k) {
      status[k] = markers[k].status;
      id[k] = markers[k].id;
      center[2 * k] = markers[k].center_img.x;
      center[2 * k + 1] = markers[k].center_img.y;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
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
