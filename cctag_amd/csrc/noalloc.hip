// The grid search of cctagx_refiner_grid as one call that allocates nothing
// (include/cctag_amd.h, cctagx_refine_grid_static).  Every device buffer the
// call needs -- the frame, its quad image, the jobs, the grid search's
// per-marker state, grid records and Gram matrices, the read-out's outputs
// -- lives in `scratch`, a fixed array of device memory that each GPU holds
// from the library's load until the process ends.  Host code reaches it
// through the address hipGetSymbolAddress returns on that device; the
// array's own name is a host-side stand-in and is never passed to a kernel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/cctag_amd.h"
#include "capi_common.hpp"
#include "device.hpp"
#include "grad_device.hpp"
#include "host/cabi.hpp"
#include "ident_device.hpp"

using namespace cctagx;

namespace {

constexpr size_t kMaxPixels = CCTAGX_STATIC_MAX_PIXELS;
constexpr int kMaxMarkers = CCTAGX_STATIC_MAX_MARKERS;
constexpr int kG2 = kMaxGrid * kMaxGrid;

constexpr size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

// Byte offsets of the buffers inside the array.
constexpr size_t kImg = 0;
constexpr size_t kQuad = up256(kImg + kMaxPixels);
constexpr size_t kJobs = up256(kQuad + 4 * kMaxPixels);
constexpr size_t kState = up256(kJobs + sizeof(RefineJobDev) * kMaxMarkers);
constexpr size_t kGrid = up256(kState + sizeof(RefineStateDev) * 2 * kMaxMarkers);
constexpr size_t kGram = up256(kGrid + sizeof(RefineGridDev) * 2 * kMaxMarkers * kG2);
constexpr size_t kOut = up256(kGram + sizeof(double) * 2 * (size_t)kMaxMarkers * kG2 * kMaxPairs);
constexpr size_t kDist = up256(kOut + sizeof(RefineOutDev) * kMaxMarkers);
constexpr size_t kScratchBytes = up256(kDist + sizeof(float));
static_assert(kScratchBytes <= 24309248, "the size the C ABI documents (24.3 MB)");

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The array's device address on the current device (looked up once per
// device).
int scratch_address(unsigned char** out) {
  int dev = -1;
  CCTAGX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index beyond the static entry point's table");
    return CCTAGX_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return CCTAGX_ERR_HIP;
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return CCTAGX_OK;
}

void copy9(float* dst, const Mat3& m) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dst[i * 3 + j] = m.m[i][j];
}

}  // namespace

extern "C" {

int cctagx_refine_grid_static(const cctagx_params* params, const uint8_t* gray, int width,
                              int height, size_t stride, const cctagx_refine_marker* markers,
                              int n, float* centers, int32_t* ok) {
  if (!gray || width < 2 || height < 2 || stride < (size_t)width || n < 0 ||
      (n > 0 && (!markers || !centers || !ok)))
    return CCTAGX_ERR_ARG;
  if (int rc = validate_params(params)) return rc;
  if ((size_t)width * height > kMaxPixels || n > kMaxMarkers) {
    set_error("the static entry point takes at most " + std::to_string(kMaxPixels) +
              " pixels and " + std::to_string(kMaxMarkers) + " markers");
    return CCTAGX_ERR_CAPACITY;
  }
  if (cctagx_device_count() <= 0) {
    set_error("no HIP device visible");
    return CCTAGX_ERR_NO_DEVICE;
  }
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  uint8_t* d_img = reinterpret_cast<uint8_t*>(base + kImg);
  uint32_t* d_quad = reinterpret_cast<uint32_t*>(base + kQuad);
  RefineJobDev* d_jobs = reinterpret_cast<RefineJobDev*>(base + kJobs);
  RefineWork work;
  work.state = reinterpret_cast<RefineStateDev*>(base + kState);
  work.grid = reinterpret_cast<RefineGridDev*>(base + kGrid);
  work.gram = reinterpret_cast<double*>(base + kGram);
  RefineOutDev* d_out = reinterpret_cast<RefineOutDev*>(base + kOut);
  float* d_dist = reinterpret_cast<float*>(base + kDist);

  // per-marker constants on the host, as cctagx_refiner_set_markers makes
  // them (the first marker's (a+b)/2 conditions every marker)
  const HostParams hp = host_params(*params);
  const IdentConstants kc = ident_constants(hp, std::vector<std::vector<float>>());
  RefineJobDev jobs[kMaxMarkers];
  int job_of[kMaxMarkers];
  int n_jobs = 0, max_iter = 0;
  float mean_ab = 0.f;
  const std::vector<Pt2> no_stops;
  for (int i = 0; i < n; ++i) {
    const cctagx_refine_marker& m = markers[i];
    if (m.n_cuts < 2 || m.n_cuts > kMaxCuts) return CCTAGX_ERR_ARG;
    job_of[i] = -1;
    try {
      const Ellipse e(m.ellipse[0], m.ellipse[1], m.ellipse[2], m.ellipse[3], m.ellipse[4]);
      if (i == 0) mean_ab = (e.a() + e.b()) / 2.f;
      RefineJob r;
      if (!make_refine_job(e, Pt2(m.center0[0], m.center0[1]), no_stops, hp, kc, mean_ab, r))
        continue;
      RefineJobDev& d = jobs[n_jobs];
      std::memset(&d, 0, sizeof(d));
      d.quad = d_quad;
      d.iw = width;
      d.ih = height;
      d.Q11 = r.canonic.m[0][0];
      d.Q22 = r.canonic.m[1][1];
      d.Q33 = r.canonic.m[2][2];
      copy9(d.t_can, r.t_can);
      copy9(d.t_inv_can, r.t_inv_can);
      copy9(d.cond, r.cond);
      copy9(d.cond_inv, r.cond_inv);
      d.grid_scale = r.grid_scale;
      d.neighbour0 = r.neighbour0;
      d.begin_sig = r.begin_sig;
      d.end_sig = r.end_sig;
      d.n_iter = r.n_iter;
      d.box_inside = box_inside(r, width, height) ? 1 : 0;
      d.n_cuts = m.n_cuts;
      d.c0x = m.center0[0];
      d.c0y = m.center0[1];
      d.dist_offset = 0;  // no bank rows: the read-out writes no distances
      d.sel_index = -1;
      for (int k = 0; k < m.n_cuts; ++k) d.stops[k] = make_float2(m.stops[k][0], m.stops[k][1]);
      max_iter = std::max(max_iter, r.n_iter);
      job_of[i] = n_jobs++;
    } catch (const std::exception& ex) {
      set_error(std::string("marker ") + std::to_string(i) + ": " + ex.what());
      return CCTAGX_ERR_ARG;
    }
  }

  hipStream_t s = nullptr;  // the current device's null stream
  CCTAGX_HIP(hipMemcpy2DAsync(d_img, width, gray, stride, width, height, hipMemcpyHostToDevice, s));
  if (launch_quads(d_img, width, height, d_quad, s)) return CCTAGX_ERR_HIP;
  RefineOutDev out[kMaxMarkers];
  if (n_jobs > 0) {
    CCTAGX_HIP(hipMemcpyAsync(d_jobs, jobs, sizeof(RefineJobDev) * n_jobs, hipMemcpyHostToDevice, s));
    if (launch_refine(d_jobs, n_jobs, max_iter, kc.n_samples, hp.imaged_center_n_grid_sample,
                      nullptr, 0, 0, work, d_out, d_dist, s))
      return CCTAGX_ERR_HIP;
    CCTAGX_HIP(hipMemcpyAsync(out, d_out, sizeof(RefineOutDev) * n_jobs, hipMemcpyDeviceToHost, s));
  }
  CCTAGX_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i) {
    const int j = job_of[i];
    ok[i] = j >= 0 && out[j].ok;
    centers[2 * i] = ok[i] ? out[j].cx : 0.f;
    centers[2 * i + 1] = ok[i] ? out[j].cy : 0.f;
  }
  return CCTAGX_OK;
}

}  // extern "C"
