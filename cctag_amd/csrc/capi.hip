// C ABI of the library (include/cctag_amd.h).
#include "../../include/cctag_amd.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "device.hpp"
#include "frontend_host.hpp"
#include "host/cabi.hpp"
#include "host/pipeline.hpp"
#include "ident_device.hpp"
#include "staging.hpp"
#include "capi_common.hpp"

using namespace cctagx;

namespace {

// Host copy of one level, backing a cctagx_level_view.
struct HostLevel {
  int w = 0, h = 0, W0 = 0, H0 = 0;
  std::vector<uint8_t> img, edges;
  std::vector<int16_t> dx, dy, ex, ey, egx, egy;
  std::vector<int32_t> emap, before, after, chosen, vindex, vlist, ismax, seeds;
  std::vector<float> tdist, flow;
  int n_edges = 0, n_seeds = 0;
};

template <class T>
int d2h(std::vector<T>& dst, const T* src, size_t n, hipStream_t s) {
  dst.resize(n);
  if (n == 0) return 0;
  CCTAGX_HIP(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s));
  return 0;
}

template <class T>
const T* ptr(const std::vector<T>& v) { return v.empty() ? nullptr : v.data(); }

}  // namespace

struct cctagx_handle {
  cctagx_params params{};
  HostParams host{};
  std::vector<std::vector<float>> bank;
  int width = 0, height = 0;
  hipStream_t stream = nullptr;
  FrontEndDevice fe;
  StageTimer timer;
  bool have_front_end = false;
  std::vector<HostLevel> levels;
  std::vector<bool> level_loaded;
  LevelStager stager;
  std::vector<LevelEdges> views;  // candidate-stage inputs of the last detection
  GpuIdentBackend ident;
  std::vector<Marker> markers;
  std::unique_ptr<TaskPool> pool;  // host candidate-stage workers
  double host_ms = 0;  // last detection: host wall time after the front end
  double stage_ms = 0; // last detection: wall time of the level staging copies
  bool staged_two_step = false;  // last detection: the levels did not fit the one-step gather
  HostTimes host_times;  // last detection: host stages
  FrameWorkspace workspace;  // host edge sets reused across detections
  bool have_frame = false;
  // the front end's launch sequence, captured at the first detection without
  // per-kernel timing and replayed (the frame size and buffers are fixed)
  hipGraph_t fe_graph = nullptr;
  hipGraphExec_t fe_exec = nullptr;
};

// Copies what the candidate stage reads of every level into pinned memory.
static int stage_levels(cctagx_handle* h) {
  const int rc = h->stager.stage(h->fe.planes(), h->fe.buffers(), h->stream, h->views);
  return rc == 0 ? CCTAGX_OK : (rc == -1 ? CCTAGX_ERR_ARG : CCTAGX_ERR_HIP);
}

extern "C" {

const char* cctagx_version(void) { return "cctag_amd 0.1 (gfx950)"; }

const char* cctagx_last_error(void) { return last_error(); }

int cctagx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void cctagx_params_default(cctagx_params* p, int n_crowns) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->canny_thr_low = 0.01f;
  p->canny_thr_high = 0.04f;
  p->dist_search = 30;
  p->thr_gradient_mag_in_vote = 2500;
  p->angle_voting = 0.0f;
  p->ratio_voting = 4.0f;
  p->average_vote_min = 0.0f;
  p->thr_median_distance_ellipse = 3.0f;
  p->maximum_nb_seeds = 500;
  p->maximum_nb_candidates_loop_two = 40;
  p->n_crowns = n_crowns;
  p->min_points_segment_candidate = 10;
  p->min_votes_to_select_candidate = 3;
  p->thresh_robust_estimation_of_outer_ellipse = 30.0f;
  p->ellipse_growing_elliptic_hull_width = 2.3f;
  p->window_size_on_inner_elliptic_segment = 20;
  p->number_of_multires_layers = 4;
  p->number_of_processed_multires_layers = 4;
  p->n_samples_outer_ellipse = 150;
  p->num_cuts_in_ident_step = 22;
  p->num_samples_outer_edge_points_refinement = 20;
  p->sample_cut_length = 100;
  p->imaged_center_n_grid_sample = 5;
  p->imaged_center_neighbour_size = 0.20f;
  p->min_ident_proba = 1e-6f;
  p->search_for_another_segment = 1;
  p->do_identification = 1;
}

int cctagx_create(const cctagx_params* params, int width, int height, cctagx_handle** out) {
  if (!out || width <= 0 || height <= 0) return CCTAGX_ERR_ARG;
  *out = nullptr;
  if (int rc = validate_params(params)) return rc;
  if (cctagx_device_count() <= 0) {
    set_error("no HIP device visible");
    return CCTAGX_ERR_NO_DEVICE;
  }
  std::unique_ptr<cctagx_handle> h(new (std::nothrow) cctagx_handle());
  if (!h) return CCTAGX_ERR_ARG;
  h->params = *params;
  h->width = width;
  h->height = height;
  PlaneSet ps;
  if (!plan_planes({{width, height}}, params->number_of_processed_multires_layers, ps)) {
    set_error("frame too small for the pyramid: a level would be empty");
    return CCTAGX_ERR_UNSUPPORTED;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return CCTAGX_ERR_HIP;
  if (int rc = h->fe.init(ps, front_end_params(*params))) return rc;
  h->levels.resize(ps.planes.size());
  h->level_loaded.assign(ps.planes.size(), false);
  h->host = host_params(*params);
  h->bank = marker_bank(params->n_crowns);
  try {
    if (int rc = h->ident.init(h->bank, h->host, h->stream)) return rc;
    const int threads = TaskPool::default_threads();
    if (threads > 1) h->pool = std::make_unique<TaskPool>(threads);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  *out = h.release();
  return CCTAGX_OK;
}

// A frame handed to the handle: non-null, the handle's size, a stride that
// covers a row.
static int check_frame(const cctagx_handle* h, const uint8_t* gray, int width, int height,
                       size_t stride) {
  if (!h || !gray) return CCTAGX_ERR_ARG;
  if (width != h->width || height != h->height) {
    set_error("frame is " + std::to_string(width) + "x" + std::to_string(height) +
              ", the handle was created for " + std::to_string(h->width) + "x" +
              std::to_string(h->height));
    return CCTAGX_ERR_ARG;
  }
  if (stride < (size_t)h->width) {
    set_error("stride shorter than a row");
    return CCTAGX_ERR_ARG;
  }
  return CCTAGX_OK;
}

int cctagx_load(cctagx_handle* h, const uint8_t* gray, int width, int height, size_t stride) {
  if (int rc = check_frame(h, gray, width, height, stride)) return rc;
  if (int rc = h->fe.upload(0, gray, stride, h->stream, h->pool.get())) return rc;
  h->have_frame = true;
  h->have_front_end = false;
  return CCTAGX_OK;
}

int cctagx_detect(cctagx_handle* h, const uint8_t* gray, int width, int height, size_t stride,
                  uint64_t frame, cctagx_marker* out, int max_out, int* n_out) {
  if ((max_out > 0 && !out) || !n_out) return CCTAGX_ERR_ARG;
  if (int rc = cctagx_load(h, gray, width, height, stride)) return rc;
  return cctagx_detect_loaded(h, frame, out, max_out, n_out);
}

int cctagx_load_bgr(cctagx_handle* h, const uint8_t* bgr, int width, int height, size_t stride,
                    int channels) {
  if (channels != 3 && channels != 4) {
    set_error("color frames have 3 (BGR) or 4 (BGRA) channels");
    return CCTAGX_ERR_ARG;
  }
  // stride / channels >= width exactly when a row of width pixels fits
  if (int rc = check_frame(h, bgr, width, height, stride / channels)) return rc;
  if (int rc = h->fe.upload_color(0, bgr, stride, channels, h->stream, h->pool.get())) return rc;
  h->have_frame = true;
  h->have_front_end = false;
  return CCTAGX_OK;
}

int cctagx_detect_bgr(cctagx_handle* h, const uint8_t* bgr, int width, int height, size_t stride,
                      int channels, uint64_t frame, cctagx_marker* out, int max_out, int* n_out) {
  if ((max_out > 0 && !out) || !n_out) return CCTAGX_ERR_ARG;
  if (int rc = cctagx_load_bgr(h, bgr, width, height, stride, channels)) return rc;
  return cctagx_detect_loaded(h, frame, out, max_out, n_out);
}

int cctagx_set_bank(cctagx_handle* h, const float* ratios, int rows, int cols) {
  if (!h || !ratios || rows < 1 || cols < 1) return CCTAGX_ERR_ARG;
  if (rows > kMaxBankRows || rows > kMaxBank || cols > kMaxRatios) {
    set_error("the identification kernels take at most 128 bank rows of at most 8 ratios");
    return CCTAGX_ERR_UNSUPPORTED;
  }
  std::vector<std::vector<float>> bank(rows, std::vector<float>(cols));
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      const float v = ratios[(size_t)i * cols + j];
      if (!(std::isfinite(v) && v > 0.f)) {
        set_error("bank ratios must be finite and positive");
        return CCTAGX_ERR_ARG;
      }
      bank[i][j] = v;
    }
  try {
    h->ident.set_bank(bank, h->host);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  h->bank = std::move(bank);
  return CCTAGX_OK;
}

int cctagx_identify(cctagx_handle* h, const cctagx_ident_marker* in, int n,
                    cctagx_ident_result* out) {
  if (!h || n < 0 || (n > 0 && (!in || !out))) return CCTAGX_ERR_ARG;
  if (!h->have_front_end) return CCTAGX_ERR_STATE;
  std::vector<Marker> markers(n);
  try {
    for (int k = 0; k < n; ++k) {
      const cctagx_ident_marker& a = in[k];
      if (a.n_points < 0 || (a.n_points > 0 && !a.points)) return CCTAGX_ERR_ARG;
      Marker& m = markers[k];
      m.rescaled_outer_ellipse =
          Ellipse(a.ellipse[0], a.ellipse[1], a.ellipse[2], a.ellipse[3], a.ellipse[4]);
      m.outer_ellipse = m.rescaled_outer_ellipse;
      m.center_img = Pt2(a.center[0], a.center[1]);
      m.scale = a.scale;
      m.rescaled_outer_ellipse_points.reserve(a.n_points);
      for (int j = 0; j < a.n_points; ++j) {
        const float* q = a.points + 4 * (size_t)j;
        m.rescaled_outer_ellipse_points.emplace_back(q[0], q[1], q[2], q[3]);
      }
    }
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_ARG;
  }
  try {
    const PlaneDesc& d0 = h->fe.planes().planes[0];
    h->ident.set_image(h->fe.buffers().quad + d0.pix, d0.w, d0.h);
    h->ident.clear_stats();
    identify_markers(markers, h->host, h->bank, h->ident, h->pool.get(), nullptr);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  for (int k = 0; k < n; ++k) {
    const Marker& m = markers[k];
    cctagx_ident_result& o = out[k];
    o.status = m.status;
    o.id = m.id;
    o.x = m.center_img.x;
    o.y = m.center_img.y;
    o.quality = m.quality;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) o.homography[i * 3 + j] = m.homography.m[i][j];
    o.n_cuts = (int32_t)std::min<size_t>(m.cut_stops.size(), 32);
    for (int c = 0; c < 32; ++c) {
      o.stops[c][0] = c < o.n_cuts ? m.cut_stops[c].x : 0.f;
      o.stops[c][1] = c < o.n_cuts ? m.cut_stops[c].y : 0.f;
    }
  }
  return CCTAGX_OK;
}

int cctagx_detect_loaded(cctagx_handle* h, uint64_t frame, cctagx_marker* out, int max_out,
                         int* n_out) {
  (void)frame;
  if (!h || (max_out > 0 && !out) || !n_out) return CCTAGX_ERR_ARG;
  if (!h->have_frame) return CCTAGX_ERR_STATE;
  *n_out = 0;
  h->timer.reset();
  // the staging gather lays the levels out from the device counts and
  // follows the front end without a host round trip (in its graph)
  const bool planned = (int)h->fe.planes().planes.size() <= kMaxPlannedPlanes;
  if (planned && h->stager.reserve_planned(h->fe.planes())) {
    set_error("pinned staging buffer allocation failed");
    return CCTAGX_ERR_HIP;
  }
  auto enqueue_gather = [&]() {
    return planned ? h->stager.enqueue_planned(h->fe.planes(), h->fe.launch().dev_planes,
                                               h->fe.buffers(), h->stream)
                   : 0;
  };
  if (h->timer.per_kernel) {
    if (int rc = h->fe.run(h->stream, &h->timer)) return rc;
    if (enqueue_gather()) return CCTAGX_ERR_HIP;
  } else {
    if (!h->fe_exec) {
      CCTAGX_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      int rc = h->fe.run(h->stream, nullptr);
      if (rc == 0 && enqueue_gather()) rc = CCTAGX_ERR_HIP;
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc != 0 || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        set_error(std::string("front-end graph capture failed: ") + hipGetErrorString(e));
        return CCTAGX_ERR_HIP;
      }
      h->fe_graph = g;
      CCTAGX_HIP(hipGraphInstantiate(&h->fe_exec, h->fe_graph, nullptr, nullptr, 0));
    }
    h->timer.mark("start", h->stream);
    CCTAGX_HIP(hipGraphLaunch(h->fe_exec, h->stream));
    h->timer.mark("front_end", h->stream);
  }
  h->have_front_end = true;
  h->level_loaded.assign(h->level_loaded.size(), false);
  const auto ts = std::chrono::steady_clock::now();
  int staged = planned ? h->stager.finish_planned(h->fe.planes(), h->stream, h->views) : 1;
  if (staged < 0) return CCTAGX_ERR_HIP;
  h->staged_two_step = staged == 1;
  if (staged == 1) {
    // levels larger than the planned buffer (or too many planes): the
    // two-step staging, and the graph captured again with the grown buffer
    if (planned && h->fe_exec) {
      hipGraphExecDestroy(h->fe_exec);
      hipGraphDestroy(h->fe_graph);
      h->fe_exec = nullptr;
      h->fe_graph = nullptr;
    }
    if (int rc = stage_levels(h)) return rc;
  }
  const auto t0 = std::chrono::steady_clock::now();
  h->stage_ms = std::chrono::duration<double, std::milli>(t0 - ts).count();
  h->ident.clear_stats();
  try {
    const PlaneDesc& d0 = h->fe.planes().planes[0];
    h->ident.set_image(h->fe.buffers().quad + d0.pix, d0.w, d0.h);
    h->host_times = HostTimes();
    detect_markers(h->views, h->host, h->bank, &h->ident, h->markers, h->pool.get(),
                   &h->host_times, &h->workspace);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  h->host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const int n = (int)h->markers.size();
  for (int i = 0; i < n && i < max_out; ++i) to_c_marker(h->markers[i], out[i]);
  *n_out = n;
  return CCTAGX_OK;
}

int cctagx_last_markers(cctagx_handle* h, cctagx_marker* out, int max_out, int* n_out) {
  if (!h || (max_out > 0 && !out) || !n_out) return CCTAGX_ERR_ARG;
  const int n = (int)h->markers.size();
  for (int i = 0; i < n && i < max_out; ++i) to_c_marker(h->markers[i], out[i]);
  *n_out = n;
  return CCTAGX_OK;
}

// Wall time of the host part and device time of the identification kernels
// of the last cctagx_detect, microseconds.
int cctagx_detect_times(cctagx_handle* h, float* host_us, float* select_us, float* refine_us) {
  if (!h || !host_us || !select_us || !refine_us) return CCTAGX_ERR_ARG;
  *host_us = (float)(h->host_ms * 1000.0);
  *select_us = h->ident.last_select_us();
  *refine_us = h->ident.last_refine_us();
  return CCTAGX_OK;
}

int cctagx_set_refinement(cctagx_handle* h, int gradient_steps) {
  if (!h || gradient_steps < 0 || gradient_steps > 10000) return CCTAGX_ERR_ARG;
  h->ident.set_gradient_steps(gradient_steps);
  return CCTAGX_OK;
}

int cctagx_set_host_threads(cctagx_handle* h, int threads) {
  if (!h || threads < 0) return CCTAGX_ERR_ARG;
  if (threads == 0) threads = TaskPool::default_threads();
  try {
    h->pool.reset();
    if (threads > 1) h->pool = std::make_unique<TaskPool>(threads);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_ARG;
  }
  return CCTAGX_OK;
}

int cctagx_get_host_threads(cctagx_handle* h) {
  if (!h) return CCTAGX_ERR_ARG;
  return h->pool ? h->pool->threads() : 1;
}

int cctagx_get_detect_stats(cctagx_handle* h, cctagx_detect_stats* st) {
  if (!h || !st) return CCTAGX_ERR_ARG;
  *st = cctagx_detect_stats{};
  if (h->timer.used >= 2) {
    float ms = 0.f;
    CCTAGX_HIP(hipEventElapsedTime(&ms, h->timer.events[0], h->timer.events[h->timer.used - 1]));
    st->front_end_us = ms * 1000.f;
  }
  st->stage_us = (float)(h->stage_ms * 1000.0);
  st->staged_bytes = (int64_t)h->stager.last_bytes();
  st->staged_two_step = h->staged_two_step ? 1 : 0;
  st->host_us = (float)(h->host_ms * 1000.0);
  st->select_us = h->ident.last_select_us();
  st->refine_us = h->ident.last_refine_us();
  st->n_select = h->ident.last_select_jobs();
  st->n_refine = h->ident.last_refine_jobs();
  st->select_samples = h->ident.last_select_samples();
  st->refine_samples = h->ident.last_refine_samples();
  st->refine_flops = h->ident.last_refine_flops();
  const HostTimes& t = h->host_times;
  st->load_us = (float)(t.load * 1e3);
  st->loop_one_us = (float)(t.loop_one * 1e3);
  st->loop_two_us = (float)(t.loop_two * 1e3);
  st->labels_us = (float)(t.labels * 1e3);
  st->loop_three_us = (float)(t.loop_three * 1e3);
  st->rescale_us = (float)(t.rescale * 1e3);
  st->ident_us = (float)(t.ident * 1e3);
  st->finish_us = (float)(t.finish * 1e3);
  st->ident_prep_us = (float)(t.ident_prep * 1e3);
  st->ident_select_us = (float)(t.ident_select * 1e3);
  st->ident_refine_us = (float)(t.ident_refine * 1e3);
  st->ident_post_us = (float)(t.ident_post * 1e3);
  return CCTAGX_OK;
}

int cctagx_destroy(cctagx_handle* h) {
  if (!h) return CCTAGX_ERR_ARG;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->fe_exec) hipGraphExecDestroy(h->fe_exec);
  if (h->fe_graph) hipGraphDestroy(h->fe_graph);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return CCTAGX_OK;
}

int cctagx_front_end(cctagx_handle* h, const uint8_t* gray, int width, int height,
                     size_t stride) {
  if (int rc = check_frame(h, gray, width, height, stride)) return rc;
  if (int rc = h->fe.upload(0, gray, stride, h->stream, h->pool.get())) return rc;
  h->have_frame = true;
  h->timer.reset();
  if (int rc = h->fe.run(h->stream, &h->timer)) return rc;
  CCTAGX_HIP(hipStreamSynchronize(h->stream));
  h->have_front_end = true;
  h->level_loaded.assign(h->level_loaded.size(), false);
  return CCTAGX_OK;
}

int cctagx_get_level(cctagx_handle* h, int level, cctagx_level_view* v) {
  if (!h || !v) return CCTAGX_ERR_ARG;
  if (!h->have_front_end) return CCTAGX_ERR_STATE;
  const PlaneSet& ps = h->fe.planes();
  if (level < 0 || level >= (int)ps.planes.size()) return CCTAGX_ERR_ARG;
  const PlaneDesc& d = ps.planes[level];
  HostLevel& L = h->levels[level];
  if (!h->level_loaded[level]) {
    const FrontEndBuffers& b = h->fe.buffers();
    hipStream_t s = h->stream;
    PlaneCounts c{};
    CCTAGX_HIP(hipMemcpy(&c, b.counts + level, sizeof(c), hipMemcpyDeviceToHost));
    const size_t N = (size_t)d.w * d.h, n = c.n_edges;
    L.w = d.w; L.h = d.h; L.W0 = d.shape_w; L.H0 = d.shape_h;
    L.n_edges = c.n_edges;
    L.n_seeds = c.n_seeds;
    int rc = 0;
    rc |= d2h(L.img, b.img + d.pix, N, s);
    rc |= d2h(L.dx, b.dx + d.pix, N, s);
    rc |= d2h(L.dy, b.dy + d.pix, N, s);
    rc |= d2h(L.edges, b.edges + d.pix, N, s);
    rc |= d2h(L.ex, b.ex + d.pix, n, s);
    rc |= d2h(L.ey, b.ey + d.pix, n, s);
    rc |= d2h(L.egx, b.egx + d.pix, n, s);
    rc |= d2h(L.egy, b.egy + d.pix, n, s);
    rc |= d2h(L.before, b.before + d.pix, n, s);
    rc |= d2h(L.after, b.after + d.pix, n, s);
    rc |= d2h(L.chosen, b.chosen + d.pix, n, s);
    rc |= d2h(L.tdist, b.tdist + d.pix, n, s);
    rc |= d2h(L.vindex, b.vindex + d.pix + level, n + 1, s);
    rc |= d2h(L.vlist, b.vlist + d.pix, (size_t)c.n_voters, s);
    rc |= d2h(L.flow, b.flow + d.pix, n, s);
    rc |= d2h(L.ismax, b.ismax + d.pix, n, s);
    rc |= d2h(L.seeds, b.seeds + d.pix, (size_t)c.n_seeds, s);
    if (rc) return CCTAGX_ERR_HIP;
    CCTAGX_HIP(hipStreamSynchronize(s));
    // the index map of the view: raster rank of every edge pixel, -1 elsewhere
    L.emap.resize(N);
    int32_t next = 0;
    for (size_t i = 0; i < N; ++i) L.emap[i] = L.edges[i] == 255 ? next++ : -1;
    h->level_loaded[level] = true;
  }
  v->width = L.w; v->height = L.h; v->shape_w = L.W0; v->shape_h = L.H0;
  v->img = ptr(L.img); v->dx = ptr(L.dx); v->dy = ptr(L.dy); v->edges = ptr(L.edges);
  v->edge_map = ptr(L.emap);
  v->n_edges = L.n_edges;
  v->edge_x = ptr(L.ex); v->edge_y = ptr(L.ey); v->edge_dx = ptr(L.egx); v->edge_dy = ptr(L.egy);
  v->before = ptr(L.before); v->after = ptr(L.after);
  v->chosen = ptr(L.chosen); v->total_dist = ptr(L.tdist);
  v->voters_index = ptr(L.vindex); v->voters_list = ptr(L.vlist);
  v->flow_length = ptr(L.flow); v->is_max = ptr(L.ismax);
  v->n_seeds = L.n_seeds; v->seeds = ptr(L.seeds);
  return CCTAGX_OK;
}

int cctagx_set_stage_timing(cctagx_handle* h, int per_kernel) {
  if (!h) return CCTAGX_ERR_ARG;
  h->timer.per_kernel = per_kernel != 0;
  return CCTAGX_OK;
}

int cctagx_stage_count(cctagx_handle* h) {
  if (!h) return CCTAGX_ERR_ARG;
  return h->timer.used > 0 ? (int)h->timer.used - 1 : 0;
}

int cctagx_stage_time(cctagx_handle* h, int i, const char** name, float* usec) {
  if (!h || !name || !usec || i < 0 || i + 1 >= (int)h->timer.used) return CCTAGX_ERR_ARG;
  float ms = 0.0f;
  CCTAGX_HIP(hipEventElapsedTime(&ms, h->timer.events[i], h->timer.events[i + 1]));
  *name = h->timer.names[i + 1];
  *usec = ms * 1000.0f;
  return CCTAGX_OK;
}

int cctagx_time_front_end(cctagx_handle* h, int iters, float* usec_per_frame) {
  if (!h || iters <= 0 || !usec_per_frame) return CCTAGX_ERR_ARG;
  if (!h->have_front_end) return CCTAGX_ERR_STATE;
  hipEvent_t a, b;
  CCTAGX_HIP(hipEventCreate(&a));
  CCTAGX_HIP(hipEventCreate(&b));
  hipEventRecord(a, h->stream);
  for (int i = 0; i < iters; ++i)
    if (int rc = h->fe.run(h->stream, nullptr)) return rc;
  hipEventRecord(b, h->stream);
  CCTAGX_HIP(hipEventSynchronize(b));
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  *usec_per_frame = ms * 1000.0f / iters;
  h->level_loaded.assign(h->level_loaded.size(), false);
  return CCTAGX_OK;
}

}  // extern "C"
