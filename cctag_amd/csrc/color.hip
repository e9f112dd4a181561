// Input conversion: 8-bit BGR / BGRA frames to the gray level-0 plane.
#include <hip/hip_runtime.h>

#include "frontend_host.hpp"

namespace cctagx {
namespace {

// cv::cvtColor(src, gray, COLOR_BGR2GRAY) on 8U, as the reference's
// applications call it on every image and video frame before detection
// (applications/detection/main.cpp:357,427,515; regression/TestLog.cpp:93,
// 114): OpenCV's 14-bit fixed-point Rec. 601 weights,
//   gray = (1868 B + 9617 G + 4899 R + 2^13) >> 14,
// exact integer arithmetic (its table and vector forms give the same value).
// The weights sum to 2^14, so a gray pixel maps to itself; a fourth channel
// (BGRA) is ignored.
constexpr int kB2Y = 1868, kG2Y = 9617, kR2Y = 4899, kYShift = 14;
constexpr int kColorPixPerThread = 4;

template <int CN>
__global__ __launch_bounds__(256) void k_bgr_to_gray(const uint8_t* __restrict__ src, size_t pitch,
                                                      int w, uint8_t* __restrict__ dst) {
  const int y = blockIdx.y;
  const int x0 = (blockIdx.x * 256 + threadIdx.x) * kColorPixPerThread;
  if (x0 >= w) return;
  const uint8_t* s = src + (size_t)y * pitch + (size_t)x0 * CN;
  uint8_t* d = dst + (size_t)y * w + x0;
  const int n = min(kColorPixPerThread, w - x0);
#pragma unroll
  for (int k = 0; k < kColorPixPerThread; ++k)
    if (k < n)
      d[k] = (uint8_t)((kB2Y * s[k * CN] + kG2Y * s[k * CN + 1] + kR2Y * s[k * CN + 2] +
                        (1 << (kYShift - 1))) >> kYShift);
}

}  // namespace

int FrontEndDevice::upload_color(int frame, const uint8_t* bgr, size_t stride, int channels,
                                 hipStream_t s, TaskPool* pool) {
  if (channels != 3 && channels != 4) return -1;
  int p0 = -1;
  for (size_t i = 0; i < ps_.planes.size(); ++i)
    if (ps_.planes[i].frame == frame && ps_.planes[i].level == 0) {
      p0 = (int)i;
      break;
    }
  if (p0 < 0) return -1;
  const PlaneDesc& d = ps_.planes[p0];
  const size_t row = (size_t)d.w * channels;
  if (int rc = color_.ensure(row * d.h)) return rc;
  uint8_t* staged = color_.as<uint8_t>();
  if (pool) {
    if (!color_sent_) CCTAGX_HIP(hipEventCreateWithFlags(&color_sent_, hipEventDisableTiming));
    if (color_in_flight_) CCTAGX_HIP(hipEventSynchronize(color_sent_));
    color_in_flight_ = false;
    if (color_pinned_.ensure(row * d.h)) return -2;
    copy_rows(pool, color_pinned_.as<uint8_t>(), bgr, row, d.h, stride);
    CCTAGX_HIP(hipMemcpyAsync(staged, color_pinned_.as<uint8_t>(), row * d.h,
                              hipMemcpyHostToDevice, s));
    CCTAGX_HIP(hipEventRecord(color_sent_, s));
    color_in_flight_ = true;
  } else if (stride == row)
    CCTAGX_HIP(hipMemcpyAsync(staged, bgr, row * d.h, hipMemcpyHostToDevice, s));
  else
    CCTAGX_HIP(hipMemcpy2DAsync(staged, row, bgr, stride, row, d.h, hipMemcpyHostToDevice, s));
  const dim3 grid((d.w + 256 * kColorPixPerThread - 1) / (256 * kColorPixPerThread), d.h);
  uint8_t* gray = launch_.buf.img + d.pix;
  if (channels == 3) k_bgr_to_gray<3><<<grid, 256, 0, s>>>(staged, row, d.w, gray);
  else k_bgr_to_gray<4><<<grid, 256, 0, s>>>(staged, row, d.w, gray);
  CCTAGX_HIP(hipGetLastError());
  return 0;
}

}  // namespace cctagx
