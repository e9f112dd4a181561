// Batched detection (include/cctag_amd.h, cctagx_batch_*): one plane set
// for every level of every frame, the front end captured once in a hipGraph,
// the host candidate stage parallel over frames, identification batched.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/cctag_amd.h"
#include "capi_common.hpp"
#include "device.hpp"
#include "frontend_host.hpp"
#include "host/cabi.hpp"
#include "host/pipeline.hpp"
#include "ident_device.hpp"
#include "staging.hpp"

using namespace cctagx;

namespace {
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}
constexpr int kEdgeBlockChoices[] = {64, 128, 256, 512, 1024, 2048};
}  // namespace

struct cctagx_batch {
  cctagx_params params{};
  HostParams host{};
  std::vector<std::vector<float>> bank;
  int n_frames = 0, n_levels = 0;
  std::vector<std::pair<int, int>> sizes;
  hipStream_t stream = nullptr;
  FrontEndDevice fe;
  GpuIdentBackend ident;
  std::unique_ptr<TaskPool> pool;
  LevelStager stager;
  std::vector<LevelEdges> views;
  std::vector<std::vector<Marker>> cand, found;
  std::vector<FrameWorkspace> workspace;  // per frame, reused across detections
  std::vector<bool> loaded;
  int filter_variant = -1, edge_blocks = 0;  // requested (-1 / 0: autotune)
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  cctagx_batch_stats stats{};

  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

// Times the launch sequence under each partition along one axis at a time
// (filter tiles, then edge-kernel blocks) on the loaded frames and keeps the
// fastest.  Every run recomputes the whole front end, so the buffers end in
// the same state whatever was tried.
int autotune(cctagx_batch* b) {
  StageTimer timer;
  timer.per_kernel = true;
  auto run = [&](int fv, int eb, const char* const* names, int n_names, float& us) -> int {
    b->fe.set_partition(fv, eb);
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      timer.reset();
      if (int rc = b->fe.run(b->stream, &timer)) return rc;
      CCTAGX_HIP(hipStreamSynchronize(b->stream));
      float t = 0.f;
      for (size_t i = 1; i < timer.used; ++i)
        for (int k = 0; k < n_names; ++k)
          if (std::strcmp(timer.names[i], names[k]) == 0) {
            float ms = 0.f;
            CCTAGX_HIP(hipEventElapsedTime(&ms, timer.events[i - 1], timer.events[i]));
            t += ms * 1000.f;
          }
      if (rep > 0) best = std::min(best, t);
    }
    us = best;
    return 0;
  };
  int fv = b->filter_variant >= 0 ? b->filter_variant : 1;
  int eb = b->edge_blocks > 0 ? b->edge_blocks : b->fe.launch().edge_blocks;
  if (b->filter_variant < 0) {
    static const char* const kF[] = {"filter"};
    float best = 1e30f;
    for (int v = 0; v < kFilterVariants; ++v) {
      float us;
      if (int rc = run(v, eb, kF, 1, us)) return rc;
      if (us < best) {
        best = us;
        fv = v;
      }
    }
  }
  if (b->edge_blocks <= 0) {
    static const char* const kE[] = {"descent", "vote", "voters_fill", "voters_finish"};
    float best = 1e30f;
    for (int e : kEdgeBlockChoices) {
      float us;
      if (int rc = run(fv, e, kE, 4, us)) return rc;
      if (us < best) {
        best = us;
        eb = e;
      }
    }
  }
  b->fe.set_partition(fv, eb);
  b->stats.filter_variant = fv;
  b->stats.edge_blocks = eb;
  b->tuned = true;
  return 0;
}

int capture(cctagx_batch* b) {
  b->drop_graph();
  CCTAGX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = b->fe.run(b->stream, nullptr);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != 0 || e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    set_error(std::string("front-end graph capture failed: ") + hipGetErrorString(e));
    return -2;
  }
  b->graph = g;
  CCTAGX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->stats.graph = 1;
  return 0;
}

}  // namespace

extern "C" {

int cctagx_batch_create(const cctagx_params* params, int n_frames, const int32_t* widths,
                        const int32_t* heights, cctagx_batch** out) {
  if (!out || n_frames <= 0 || !widths || !heights) return CCTAGX_ERR_ARG;
  *out = nullptr;
  if (int rc = validate_params(params)) return rc;
  if (cctagx_device_count() <= 0) {
    set_error("no HIP device visible");
    return CCTAGX_ERR_NO_DEVICE;
  }
  std::unique_ptr<cctagx_batch> b(new (std::nothrow) cctagx_batch());
  if (!b) return CCTAGX_ERR_ARG;
  b->params = *params;
  b->n_frames = n_frames;
  b->n_levels = params->number_of_processed_multires_layers;
  for (int f = 0; f < n_frames; ++f) {
    if (widths[f] <= 0 || heights[f] <= 0) return CCTAGX_ERR_ARG;
    b->sizes.emplace_back(widths[f], heights[f]);
  }
  PlaneSet ps;
  if (!plan_planes(b->sizes, b->n_levels, ps)) {
    set_error("a frame is too small for the pyramid: a level would be empty");
    return CCTAGX_ERR_UNSUPPORTED;
  }
  CCTAGX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  for (hipEvent_t& e : b->ev) CCTAGX_HIP(hipEventCreate(&e));
  if (int rc = b->fe.init(ps, front_end_params(*params))) return rc;
  b->host = host_params(*params);
  b->bank = marker_bank(params->n_crowns);
  try {
    if (int rc = b->ident.init(b->bank, b->host, b->stream)) return rc;
    std::vector<DeviceImage> images;
    for (const PlaneDesc& d : b->fe.planes().planes)
      if (d.level == 0) images.push_back(DeviceImage{b->fe.buffers().quad + d.pix, d.w, d.h});
    b->ident.set_images(images);
    const int threads = TaskPool::default_threads();
    if (threads > 1) b->pool = std::make_unique<TaskPool>(threads);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  b->loaded.assign(n_frames, false);
  b->cand.resize(n_frames);
  b->found.resize(n_frames);
  b->workspace.resize(n_frames);
  *out = b.release();
  return CCTAGX_OK;
}

int cctagx_batch_destroy(cctagx_batch* b) {
  if (!b) return CCTAGX_ERR_ARG;
  if (b->stream) hipStreamSynchronize(b->stream);
  b->drop_graph();
  for (hipEvent_t e : b->ev)
    if (e) hipEventDestroy(e);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return CCTAGX_OK;
}

int cctagx_batch_load(cctagx_batch* b, int frame, const uint8_t* gray, int width, int height,
                      size_t stride) {
  if (!b || !gray || frame < 0 || frame >= b->n_frames ||
      stride < (size_t)b->sizes[frame].first)
    return CCTAGX_ERR_ARG;
  if (width != b->sizes[frame].first || height != b->sizes[frame].second) {
    set_error("frame size differs from the one given at cctagx_batch_create");
    return CCTAGX_ERR_ARG;
  }
  if (int rc = b->fe.upload(frame, gray, stride, b->stream, b->pool.get())) return rc;
  b->loaded[frame] = true;
  return CCTAGX_OK;
}

int cctagx_batch_set_refinement(cctagx_batch* b, int gradient_steps) {
  if (!b || gradient_steps < 0 || gradient_steps > 10000) return CCTAGX_ERR_ARG;
  b->ident.set_gradient_steps(gradient_steps);
  return CCTAGX_OK;
}

int cctagx_batch_set_partition(cctagx_batch* b, int filter_variant, int edge_blocks) {
  if (!b || filter_variant < -1 || filter_variant >= kFilterVariants || edge_blocks < 0 ||
      edge_blocks > 65535)
    return CCTAGX_ERR_ARG;
  b->filter_variant = filter_variant;
  b->edge_blocks = edge_blocks;
  b->tuned = false;
  b->drop_graph();
  if (filter_variant >= 0 && edge_blocks > 0) {
    b->fe.set_partition(filter_variant, edge_blocks);
    b->stats.filter_variant = filter_variant;
    b->stats.edge_blocks = edge_blocks;
    b->tuned = true;
  }
  return CCTAGX_OK;
}

int cctagx_batch_detect(cctagx_batch* b, cctagx_marker* out, int max_out, int32_t* frame_first,
                        int* n_out) {
  if (!b || (max_out > 0 && !out) || !frame_first || !n_out) return CCTAGX_ERR_ARG;
  for (bool l : b->loaded)
    if (!l) return CCTAGX_ERR_STATE;
  *n_out = 0;
  b->stats.tune_us = 0.f;
  if (!b->tuned) {
    const auto t = Clock::now();
    if (int rc = autotune(b)) return rc;
    b->stats.tune_us = (float)(ms_since(t) * 1000.0);
  }
  if (!b->exec)
    if (int rc = capture(b)) return rc;
  CCTAGX_HIP(hipEventRecord(b->ev[0], b->stream));
  CCTAGX_HIP(hipGraphLaunch(b->exec, b->stream));
  CCTAGX_HIP(hipEventRecord(b->ev[1], b->stream));
  auto t = Clock::now();
  if (int rc = b->stager.stage(b->fe.planes(), b->fe.buffers(), b->stream, b->views))
    return rc == -1 ? CCTAGX_ERR_ARG : CCTAGX_ERR_HIP;
  float fe_ms = 0.f;
  CCTAGX_HIP(hipEventElapsedTime(&fe_ms, b->ev[0], b->ev[1]));
  b->stats.front_end_us = fe_ms * 1000.f;
  b->stats.stage_us = (float)(ms_since(t) * 1000.0);
  try {
    t = Clock::now();
    const int nl = b->n_levels;
    // frames in parallel, each frame's tasks serial on its worker; the
    // largest frames are handed out first (their candidate stages are the
    // longest, so starting them last would leave the other workers idle)
    std::vector<int> order(b->n_frames);
    for (int f = 0; f < b->n_frames; ++f) order[f] = f;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
      return (long long)b->sizes[a].first * b->sizes[a].second >
             (long long)b->sizes[c].first * b->sizes[c].second;
    });
    parallel_for(b->pool.get(), b->n_frames, [&](int k) {
      const int f = order[k];
      std::vector<LevelEdges> lv(b->views.begin() + (size_t)f * nl,
                                 b->views.begin() + (size_t)(f + 1) * nl);
      frame_candidates(lv, b->host, nullptr, b->cand[f], nullptr, &b->workspace[f]);
    });
    b->stats.candidates_us = (float)(ms_since(t) * 1000.0);
    t = Clock::now();
    b->ident.clear_stats();
    if (b->host.do_identification) {
      std::vector<IdentFrame> frames(b->n_frames);
      for (int f = 0; f < b->n_frames; ++f) {
        frames[f].markers = &b->cand[f];
        frames[f].image = f;
      }
      identify_frames(frames, b->host, b->bank, b->ident, b->pool.get());
    }
    parallel_for(b->pool.get(), b->n_frames,
                 [&](int f) { finish_markers(b->cand[f], b->found[f]); });
    b->stats.ident_us = (float)(ms_since(t) * 1000.0);
    b->stats.select_us = b->ident.last_select_us();
    b->stats.refine_us = b->ident.last_refine_us();
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  int n = 0;
  for (int f = 0; f < b->n_frames; ++f) {
    frame_first[f] = n;
    for (const Marker& m : b->found[f]) {
      if (n < max_out) to_c_marker(m, out[n]);
      ++n;
    }
  }
  frame_first[b->n_frames] = n;
  b->stats.n_markers = n;
  *n_out = n;
  return CCTAGX_OK;
}

int cctagx_batch_get_stats(cctagx_batch* b, cctagx_batch_stats* out) {
  if (!b || !out) return CCTAGX_ERR_ARG;
  *out = b->stats;
  return CCTAGX_OK;
}

}  // extern "C"
