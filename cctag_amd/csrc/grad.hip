// Differentiable centre refinement (include/cctag_amd.h, cctagx_refiner_*):
// host side.  Per marker the host derives the ellipse's canonic form (the
// closed form's constants) for the gradient kernels and the grid search's
// constants (make_refine_job) for the reference search on the same markers.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/cctag_amd.h"
#include "capi_common.hpp"
#include "device.hpp"
#include "grad_device.hpp"
#include "host/cabi.hpp"
#include "ident_device.hpp"
#include "staging.hpp"

using namespace cctagx;

struct cctagx_refiner {
  cctagx_params params{};
  HostParams host{};
  std::vector<std::vector<float>> bank;
  IdentConstants kc{};
  hipStream_t stream = nullptr;
  int max_markers = 0, n = 0;
  int w = 0, h = 0;
  DeviceBuffer img, quad, jobs, H, cost, grad, centers, coarse, start, coarse_of;
  PinnedBuffer h_jobs, h_io;
  GpuIdentBackend ident;
  std::vector<RefineJob> rjobs;
  std::vector<int> rjob_of;  // marker -> grid-search job or -1
  hipEvent_t ev[2] = {nullptr, nullptr};
  float descend_us = 0, grid_us = 0;
};

namespace {

template <class T>
int to_device(cctagx_refiner* r, DeviceBuffer& dst, const T* src, size_t count) {
  const size_t bytes = sizeof(T) * count;
  if (bytes == 0) return 0;
  if (dst.ensure(bytes) || r->h_io.ensure(bytes)) return CCTAGX_ERR_HIP;
  std::memcpy(r->h_io.as<void>(), src, bytes);
  CCTAGX_HIP(hipMemcpyAsync(dst.as<void>(), r->h_io.as<void>(), bytes, hipMemcpyHostToDevice,
                            r->stream));
  CCTAGX_HIP(hipStreamSynchronize(r->stream));
  return 0;
}

template <class T>
int to_host(cctagx_refiner* r, T* dst, const DeviceBuffer& src, size_t count) {
  const size_t bytes = sizeof(T) * count;
  if (bytes == 0) return 0;
  if (r->h_io.ensure(bytes)) return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipMemcpyAsync(r->h_io.as<void>(), src.as<void>(), bytes, hipMemcpyDeviceToHost,
                            r->stream));
  CCTAGX_HIP(hipStreamSynchronize(r->stream));
  std::memcpy(dst, r->h_io.as<void>(), bytes);
  return 0;
}

void copy9(float* dst, const Mat3& m) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dst[i * 3 + j] = m.m[i][j];
}

}  // namespace

extern "C" {

int cctagx_refiner_create(const cctagx_params* params, int max_markers, cctagx_refiner** out) {
  if (!out || max_markers <= 0) return CCTAGX_ERR_ARG;
  *out = nullptr;
  if (int rc = validate_params(params)) return rc;
  if (cctagx_device_count() <= 0) {
    set_error("no HIP device visible");
    return CCTAGX_ERR_NO_DEVICE;
  }
  std::unique_ptr<cctagx_refiner> r(new (std::nothrow) cctagx_refiner());
  if (!r) return CCTAGX_ERR_ARG;
  r->params = *params;
  r->host = host_params(*params);
  r->bank = marker_bank(params->n_crowns);
  r->kc = ident_constants(r->host, r->bank);
  r->max_markers = max_markers;
  CCTAGX_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  for (hipEvent_t& e : r->ev) CCTAGX_HIP(hipEventCreate(&e));
  try {
    if (int rc = r->ident.init(r->bank, r->host, r->stream)) return rc;
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  if (r->jobs.ensure(sizeof(GradJobDev) * max_markers) ||
      r->h_jobs.ensure(sizeof(GradJobDev) * max_markers) ||
      r->H.ensure(sizeof(float) * 9 * max_markers) || r->cost.ensure(sizeof(double) * max_markers) ||
      r->grad.ensure(sizeof(double) * 10 * max_markers) ||
      r->centers.ensure(sizeof(float) * 2 * max_markers))
    return CCTAGX_ERR_HIP;
  *out = r.release();
  return CCTAGX_OK;
}

int cctagx_refiner_destroy(cctagx_refiner* r) {
  if (!r) return CCTAGX_ERR_ARG;
  if (r->stream) hipStreamSynchronize(r->stream);
  for (hipEvent_t e : r->ev)
    if (e) hipEventDestroy(e);
  if (r->stream) hipStreamDestroy(r->stream);
  delete r;
  return CCTAGX_OK;
}

int cctagx_refiner_set_image(cctagx_refiner* r, const uint8_t* gray, int width, int height,
                             size_t stride) {
  if (!r || !gray || width < 2 || height < 2 || stride < (size_t)width) return CCTAGX_ERR_ARG;
  if ((size_t)width * height > (size_t)INT32_MAX / 4) return CCTAGX_ERR_ARG;
  if (r->img.ensure((size_t)width * height) || r->quad.ensure(4 * (size_t)width * height))
    return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipMemcpy2DAsync(r->img.as<uint8_t>(), width, gray, stride, width, height,
                              hipMemcpyHostToDevice, r->stream));
  if (launch_quads(r->img.as<uint8_t>(), width, height, r->quad.as<uint32_t>(), r->stream))
    return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipStreamSynchronize(r->stream));
  r->w = width;
  r->h = height;
  r->ident.set_image(r->quad.as<uint32_t>(), width, height);
  r->n = 0;
  return CCTAGX_OK;
}

int cctagx_refiner_set_markers(cctagx_refiner* r, const cctagx_refine_marker* markers, int n) {
  if (!r || (n > 0 && !markers) || n < 0 || n > r->max_markers) return CCTAGX_ERR_ARG;
  if (!r->w) return CCTAGX_ERR_STATE;
  GradJobDev* hj = r->h_jobs.as<GradJobDev>();
  r->rjobs.clear();
  r->rjob_of.assign(n, -1);
  bool have_mean_ab = false;
  float mean_ab = 0.f;
  for (int i = 0; i < n; ++i) {
    const cctagx_refine_marker& m = markers[i];
    if (m.n_cuts < 2 || m.n_cuts > kMaxCuts) return CCTAGX_ERR_ARG;
    try {
      const Ellipse e(m.ellipse[0], m.ellipse[1], m.ellipse[2], m.ellipse[3], m.ellipse[4]);
      Mat3 canonic, t_can, t_inv;
      e.canonic_form(canonic, t_can, t_inv);
      GradJobDev& d = hj[i];
      std::memset(&d, 0, sizeof(d));
      d.quad = r->quad.as<uint32_t>();
      d.iw = r->w;
      d.ih = r->h;
      d.Q11 = canonic.m[0][0];
      d.Q22 = canonic.m[1][1];
      d.Q33 = canonic.m[2][2];
      copy9(d.t_can, t_can);
      copy9(d.t_inv_can, t_inv);
      d.begin_sig = r->kc.start_sig;
      d.end_sig = 1.f;
      // a quarter of the grid search's first neighbourhood
      d.alpha0 = 0.25f * r->host.imaged_center_neighbour_size * std::max(e.a(), e.b());
      d.n_cuts = m.n_cuts;
      d.c0x = m.center0[0];
      d.c0y = m.center0[1];
      std::vector<Pt2> stops;
      for (int k = 0; k < m.n_cuts; ++k) {
        d.stops[k] = make_float2(m.stops[k][0], m.stops[k][1]);
        stops.emplace_back(m.stops[k][0], m.stops[k][1]);
      }
      if (!have_mean_ab) {
        mean_ab = (e.a() + e.b()) / 2.f;
        have_mean_ab = true;
      }
      RefineJob rj;
      if (make_refine_job(e, Pt2(m.center0[0], m.center0[1]), stops, r->host, r->kc, mean_ab, rj)) {
        r->rjob_of[i] = (int)r->rjobs.size();
        r->rjobs.push_back(std::move(rj));
      }
    } catch (const std::exception& ex) {
      set_error(std::string("marker ") + std::to_string(i) + ": " + ex.what());
      return CCTAGX_ERR_ARG;
    }
  }
  r->n = n;
  if (n > 0) {
    CCTAGX_HIP(hipMemcpyAsync(r->jobs.as<void>(), hj, sizeof(GradJobDev) * n, hipMemcpyHostToDevice,
                              r->stream));
    CCTAGX_HIP(hipStreamSynchronize(r->stream));
    if (int rc = to_device(r, r->coarse_of, r->rjob_of.data(), (size_t)n)) return rc;
  }
  return CCTAGX_OK;
}

int cctagx_refiner_cost_grad(cctagx_refiner* r, const float* H, int fp64, double* cost,
                             double* grad) {
  if (!r || !H || !cost) return CCTAGX_ERR_ARG;
  if (!r->n) return CCTAGX_ERR_STATE;
  if (int rc = to_device(r, r->H, H, (size_t)9 * r->n)) return rc;
  if (launch_grad_eval(r->jobs.as<GradJobDev>(), r->n, r->H.as<float>(), r->kc.n_samples, fp64 != 0,
                       r->cost.as<double>(), grad ? r->grad.as<double>() : nullptr, r->stream))
    return CCTAGX_ERR_HIP;
  if (int rc = to_host(r, cost, r->cost, r->n)) return rc;
  if (grad)
    if (int rc = to_host(r, grad, r->grad, (size_t)10 * r->n)) return rc;
  return CCTAGX_OK;
}

int cctagx_refiner_homography(cctagx_refiner* r, const float* centers, float* H) {
  if (!r || !centers || !H) return CCTAGX_ERR_ARG;
  if (!r->n) return CCTAGX_ERR_STATE;
  if (int rc = to_device(r, r->centers, centers, (size_t)2 * r->n)) return rc;
  if (launch_grad_homography(r->jobs.as<GradJobDev>(), r->n, r->centers.as<float>(), r->H.as<float>(),
                             r->stream))
    return CCTAGX_ERR_HIP;
  return to_host(r, H, r->H, (size_t)9 * r->n);
}

int cctagx_refiner_descend(cctagx_refiner* r, int steps, int fp64, float* centers, double* cost) {
  if (!r || steps < 1 || steps > 10000 || !centers || !cost) return CCTAGX_ERR_ARG;
  if (!r->n) return CCTAGX_ERR_STATE;
  CCTAGX_HIP(hipEventRecord(r->ev[0], r->stream));
  if (launch_grad_descend(r->jobs.as<GradJobDev>(), r->n, steps, r->kc.n_samples, fp64 != 0,
                          nullptr, r->centers.as<float>(), r->cost.as<double>(), r->stream))
    return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipEventRecord(r->ev[1], r->stream));
  CCTAGX_HIP(hipEventSynchronize(r->ev[1]));
  float ms = 0.f;
  CCTAGX_HIP(hipEventElapsedTime(&ms, r->ev[0], r->ev[1]));
  r->descend_us = ms * 1000.f;
  if (int rc = to_host(r, centers, r->centers, (size_t)2 * r->n)) return rc;
  return to_host(r, cost, r->cost, r->n);
}

int cctagx_refiner_refine(cctagx_refiner* r, int steps, int fp64, float* centers, double* cost) {
  if (!r || steps < 1 || steps > 10000 || !centers || !cost) return CCTAGX_ERR_ARG;
  if (!r->n) return CCTAGX_ERR_STATE;
  if (r->coarse.ensure(sizeof(float) * 2 * std::max<size_t>(r->rjobs.size(), 1)) ||
      r->start.ensure(sizeof(float) * 2 * r->n))
    return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipEventRecord(r->ev[0], r->stream));
  try {
    r->ident.coarse_centers(r->rjobs, r->host, r->coarse.as<float>());
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  if (launch_grad_start(r->jobs.as<GradJobDev>(), r->n, r->coarse_of.as<int>(), r->coarse.as<float>(),
                        r->start.as<float>(), r->stream) ||
      launch_grad_descend(r->jobs.as<GradJobDev>(), r->n, steps, r->kc.n_samples, fp64 != 0,
                          r->start.as<float>(), r->centers.as<float>(), r->cost.as<double>(),
                          r->stream))
    return CCTAGX_ERR_HIP;
  CCTAGX_HIP(hipEventRecord(r->ev[1], r->stream));
  CCTAGX_HIP(hipEventSynchronize(r->ev[1]));
  float ms = 0.f;
  CCTAGX_HIP(hipEventElapsedTime(&ms, r->ev[0], r->ev[1]));
  r->descend_us = ms * 1000.f;
  if (int rc = to_host(r, centers, r->centers, (size_t)2 * r->n)) return rc;
  return to_host(r, cost, r->cost, r->n);
}

int cctagx_refiner_grid(cctagx_refiner* r, float* centers, int32_t* ok) {
  if (!r || !centers || !ok) return CCTAGX_ERR_ARG;
  if (!r->n) return CCTAGX_ERR_STATE;
  std::vector<RefineResult> res;
  try {
    r->ident.refine(r->rjobs, r->host, res);
  } catch (const std::exception& e) {
    set_error(e.what());
    return CCTAGX_ERR_HIP;
  }
  r->grid_us = r->ident.last_refine_us();
  for (int i = 0; i < r->n; ++i) {
    const int j = r->rjob_of[i];
    ok[i] = j >= 0 && res[j].ok;
    centers[2 * i] = ok[i] ? res[j].center.x : 0.f;
    centers[2 * i + 1] = ok[i] ? res[j].center.y : 0.f;
  }
  return CCTAGX_OK;
}

int cctagx_refiner_times(cctagx_refiner* r, float* descend_us, float* grid_us) {
  if (!r || !descend_us || !grid_us) return CCTAGX_ERR_ARG;
  *descend_us = r->descend_us;
  *grid_us = r->grid_us;
  return CCTAGX_OK;
}

}  // extern "C"
