// ORACLE -- test infrastructure only (see oracle.hpp).
//
// CPU restatement of the CCTag front end: pyramid, derivative filters, Canny
// (magnitude, NMS, hysteresis), thinning, edge-point collection, gradient
// descent and voting.  Reference: alicevision/CCTag v1.0.3, src/cctag/.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace oracle {

// ---------------------------------------------------------------------------
// a1: pyramid.  ImagePyramid.cpp:23-33 sizes level l at floor(w/2) x
// floor(h/2) of level l-1, and Level.cpp:67 fills it with cv::resize(src,
// dst, dst size), default INTER_LINEAR.  How OpenCV 4.x resolves that call
// is an assumption about its internals that no reference test pins:
//  * both factors exactly 2 (even parent): INTER_LINEAR is switched to the
//    fast 2x2 INTER_AREA path: SIMD lanes (8 outputs per step on the SSE
//    baseline) compute (a+b+c+d+2)>>2, the scalar tail computes
//    cvRound((a+b+c+d) * 0.25f) (round half to even);
//  * otherwise the generic fixed-point bilinear path (8U, 11 coefficient
//    bits): per output column the source column sx = floor(fx) with
//    fx = (float)((dx + 0.5) * scale_x - 0.5), scale_x = 1 / (dw / sw) in
//    double, weights round((1 - f) * 2048) and round(f * 2048) (half to
//    even); the horizontal pass is exact integer arithmetic, the vertical
//    pass computes ((b0 * (S0 >> 4)) >> 16) + ((b1 * (S1 >> 4)) >> 16),
//    plus 2, >> 2 for every column (the scalar loop of the 8U vertical
//    resampler repeats the SIMD formula); source rows clamp to the image.
static void downsample_area2(const uint8_t* src, int sw, uint8_t* dst, int dw, int dh) {
  const int vec_end = (dw / 8) * 8;
  for (int y = 0; y < dh; ++y) {
    const uint8_t* r0 = src + (size_t)(2 * y) * sw;
    const uint8_t* r1 = r0 + sw;
    for (int x = 0; x < dw; ++x) {
      int s = r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1];
      int v;
      if (x < vec_end) v = (s + 2) >> 2;
      else v = (int)std::nearbyint((float)s * 0.25f);
      dst[(size_t)y * dw + x] = (uint8_t)v;
    }
  }
}

// Source index and the two 11-bit weights of output coordinate d along an
// axis of source length sn resampled to dn.
void linear_tap(int d, int sn, int dn, int* s0, int* a0, int* a1) {
  const double inv_scale = (double)dn / (double)sn;
  const double scale = 1.0 / inv_scale;
  float f = (float)(((double)d + 0.5) * scale - 0.5);
  int s = (int)std::floor(f);
  f -= (float)s;
  if (s < 0) {
    f = 0.0f;
    s = 0;
  }
  if (s >= sn - 1) {  // the last source sample alone
    f = 0.0f;
    s = sn - 1;
  }
  *s0 = s;
  *a0 = (int)std::nearbyint((1.0f - f) * 2048.0f);
  *a1 = (int)std::nearbyint(f * 2048.0f);
}

static void downsample_linear(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh) {
  std::vector<int> xs(dw), xa0(dw), xa1(dw);
  for (int x = 0; x < dw; ++x) linear_tap(x, sw, dw, &xs[x], &xa0[x], &xa1[x]);
  std::vector<int> h0(dw), h1(dw);
  auto hpass = [&](int row, std::vector<int>& out) {
    const uint8_t* r = src + (size_t)row * sw;
    for (int x = 0; x < dw; ++x) {
      const int s = xs[x];
      out[x] = s + 1 < sw ? r[s] * xa0[x] + r[s + 1] * xa1[x] : r[s] * 2048;
    }
  };
  for (int y = 0; y < dh; ++y) {
    int sy, b0, b1;
    linear_tap(y, sh, dh, &sy, &b0, &b1);
    hpass(std::min(sy, sh - 1), h0);
    hpass(std::min(sy + 1, sh - 1), h1);
    for (int x = 0; x < dw; ++x)
      dst[(size_t)y * dw + x] =
          (uint8_t)((((b0 * (h0[x] >> 4)) >> 16) + ((b1 * (h1[x] >> 4)) >> 16) + 2) >> 2);
  }
}

// Input conversion of the reference's applications (detection/main.cpp:357,
// 427,515; regression/TestLog.cpp:93,114): cv::cvtColor(COLOR_BGR2GRAY) on 8U
// is OpenCV's fixed-point Rec. 601 luma with 14-bit weights B 1868, G 9617,
// R 4899 (0.114, 0.587, 0.299 times 2^14, rounded) and rounding 2^13.
void bgr_to_gray(const uint8_t* src, int w, int h, size_t stride, int channels, uint8_t* dst) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const uint8_t* p = src + (size_t)y * stride + (size_t)x * channels;
      dst[(size_t)y * w + x] = (uint8_t)((p[0] * 1868 + p[1] * 9617 + p[2] * 4899 + 8192) >> 14);
    }
}

void downsample(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh) {
  if (dw < 1 || dh < 1 || dw > sw || dh > sh)
    throw std::invalid_argument("oracle::downsample: the target must be a non-empty shrink");
  if (sw == 2 * dw && sh == 2 * dh) downsample_area2(src, sw, dst, dw, dh);
  else downsample_linear(src, sw, sh, dst, dw, dh);
}

// ---------------------------------------------------------------------------
// a2: derivative filters, cvRecode.cpp:74-92.  kerneldX is the 9x9 outer
// product of a sampled Gaussian (rows) and its derivative (columns); its
// centre column is zero.  kerneldY = kerneldX^T.  cv::filter2D (8U -> 16S,
// 81 taps < 130, so the direct path) accumulates the nonzero taps in
// row-major order with one fused multiply-add per tap starting from delta=0,
// then rounds half to even (FilterVec_8u16s; the v_muladd of the AVX2 build is
// an FMA).  BORDER_REPLICATE clamps source coordinates.
//
// The 20 distinct magnitudes, row r = distance-from-top (0..4), column c =
// distance-from-left (0..3); the literals are the reference's float literals.
static const float kTap[5][4] = {
    {0.000000143284235f, 0.000003558691641f, 0.000028902492951f, 0.000064765993382f},
    {0.000004744922188f, 0.000117847682078f, 0.000957119116802f, 0.002144755142391f},
    {0.000057804985902f, 0.001435678675203f, 0.011660097860113f, 0.026128466569370f},
    {0.000259063973527f, 0.006434265427174f, 0.052256933138740f, 0.117099663048638f},
    {0.000427124283626f, 0.010608310271112f, 0.086157117207395f, 0.193064705260108f},
};

static float kernel_dx(int i, int j) {  // row i, column j of kerneldX
  if (j == 4) return 0.0f;
  const int r = i <= 4 ? i : 8 - i;
  const int c = j < 4 ? j : 8 - j;
  return j < 4 ? -kTap[r][c] : kTap[r][c];
}

// Row-wise like OpenCV's FilterEngine: the 9 source rows of an output row are
// border-extended by 4 px and converted to float once, then every nonzero tap
// (row-major) does one FMA over the whole output row.  Each output pixel sees
// the same tap order as a per-pixel loop, so the result does not depend on
// how the x loop is vectorized.
template <class Out>
static void filter_impl(const uint8_t* img, int w, int h, Out* dx, Out* dy, bool round) {
  const int pw = w + 8;
  std::vector<float> rows((size_t)9 * pw), ax(w), ay(w);
  for (int y = 0; y < h; ++y) {
    for (int i = 0; i < 9; ++i) {
      const int yy = std::min(std::max(y + i - 4, 0), h - 1);
      const uint8_t* src = img + (size_t)yy * w;
      float* r = &rows[(size_t)i * pw];
      for (int x = 0; x < pw; ++x) r[x] = (float)src[std::min(std::max(x - 4, 0), w - 1)];
    }
    std::fill(ax.begin(), ax.end(), 0.0f);
    std::fill(ay.begin(), ay.end(), 0.0f);
    for (int i = 0; i < 9; ++i) {
      for (int j = 0; j < 9; ++j) {
        const float* r = &rows[(size_t)i * pw + j];
        const float kx = kernel_dx(i, j);
        if (kx != 0.0f)
          for (int x = 0; x < w; ++x) ax[x] = std::fma(r[x], kx, ax[x]);
        const float ky = kernel_dx(j, i);  // kerneldY(i, j)
        if (ky != 0.0f)
          for (int x = 0; x < w; ++x) ay[x] = std::fma(r[x], ky, ay[x]);
      }
    }
    Out* ox = dx + (size_t)y * w;
    Out* oy = dy + (size_t)y * w;
    for (int x = 0; x < w; ++x) {
      if (round) {
        ox[x] = (Out)std::nearbyint(ax[x]);
        oy[x] = (Out)std::nearbyint(ay[x]);
      } else {
        ox[x] = (Out)ax[x];
        oy[x] = (Out)ay[x];
      }
    }
  }
}

void derivatives(const uint8_t* img, int w, int h, int16_t* dx, int16_t* dy) {
  // |value| <= 255 * sum|taps| ~ 186, so saturate_cast never saturates.
  filter_impl<int16_t>(img, w, h, dx, dy, true);
}

void derivatives_float(const uint8_t* img, int w, int h, float* dx, float* dy) {
  filter_impl<float>(img, w, h, dx, dy, false);
}

// ---------------------------------------------------------------------------
// a3-a5: Canny, cvRecode.cpp:143-414, restated with the same buffers: a
// (w+2)x(h+2) map whose border is 1, a 3-row magnitude ring with zero rows
// above the first and below the last image row and zero columns at both
// ends, integer thresholds low = floor(0.01*256) = 2, high = floor(0.04*256)
// = 10 (Level.cpp:206-207, cvRecode.cpp:107-108), the prev_flag / upper
// neighbour tests of the NMS loop and the explicit stack flood fill.
void canny(const int16_t* imgDX, const int16_t* imgDY, int w, int h, float low_thresh,
           float high_thresh, uint8_t* edges) {
  if (low_thresh > high_thresh) std::swap(low_thresh, high_thresh);
  const int low = (int)std::floor(low_thresh);
  const int high = (int)std::floor(high_thresh);
  const ptrdiff_t mapstep = w + 2;
  std::vector<int> magbuf((size_t)(w + 2) * 3, 0);
  std::vector<uint8_t> mapbuf((size_t)(w + 2) * (h + 2));
  int* mag_buf[3] = {&magbuf[0], &magbuf[w + 2], &magbuf[2 * (w + 2)]};
  uint8_t* map = mapbuf.data();
  std::memset(map, 1, mapstep);
  std::memset(map + mapstep * (h + 1), 1, mapstep);
  std::vector<uint8_t*> stack;
  stack.reserve(std::max(1 << 10, w * h / 10));
  const int CANNY_SHIFT = 15;
  const int TG22 = (int)(0.4142135623730950488016887242097 * (1 << CANNY_SHIFT) + 0.5);

  for (int i = 0; i <= h; i++) {
    int* _mag = mag_buf[(i > 0) + 1] + 1;
    if (i < h) {
      const int16_t* rdx = imgDX + (size_t)i * w;
      const int16_t* rdy = imgDY + (size_t)i * w;
      _mag[-1] = _mag[w] = 0;
      for (int j = 0; j < w; j++) {
        const int x = rdx[j], y = rdy[j];
        _mag[j] = (int)rintf(std::sqrt((float)x * x + (float)y * y));
      }
    } else {
      std::memset(_mag - 1, 0, (w + 2) * sizeof(int));
    }
    if (i == 0) continue;

    uint8_t* _map = map + mapstep * i + 1;
    _map[-1] = _map[w] = 1;
    _mag = mag_buf[1] + 1;
    const int16_t* rdx = imgDX + (size_t)(i - 1) * w;
    const int16_t* rdy = imgDY + (size_t)(i - 1) * w;
    const ptrdiff_t magstep1 = mag_buf[2] - mag_buf[1];
    const ptrdiff_t magstep2 = mag_buf[0] - mag_buf[1];
    int prev_flag = 0;
    for (int j = 0; j < w; j++) {
      int x = rdx[j], y = rdy[j];
      int s = x ^ y;
      const int m = _mag[j];
      x = std::abs(x);
      y = std::abs(y);
      if (m > low) {
        const int tg22x = x * TG22;
        const int tg67x = tg22x + ((x + x) << CANNY_SHIFT);
        y <<= CANNY_SHIFT;
        bool pass;
        if (y < tg22x) {
          pass = m > _mag[j - 1] && m >= _mag[j + 1];
        } else if (y > tg67x) {
          pass = m > _mag[j + magstep2] && m >= _mag[j + magstep1];
        } else {
          s = s < 0 ? -1 : 1;
          pass = m > _mag[j + magstep2 - s] && m > _mag[j + magstep1 + s];
        }
        if (pass) {
          if (m > high && !prev_flag && _map[j - mapstep] != 2) {
            _map[j] = 2;
            stack.push_back(_map + j);
            prev_flag = 1;
          } else {
            _map[j] = 0;
          }
          continue;
        }
      }
      prev_flag = 0;
      _map[j] = 1;
    }
    int* t = mag_buf[0];
    mag_buf[0] = mag_buf[1];
    mag_buf[1] = mag_buf[2];
    mag_buf[2] = t;
  }

  while (!stack.empty()) {
    uint8_t* m = stack.back();
    stack.pop_back();
    const ptrdiff_t nb[8] = {-1, 1, -mapstep - 1, -mapstep, -mapstep + 1,
                             mapstep - 1, mapstep, mapstep + 1};
    for (ptrdiff_t d : nb) {
      if (!m[d]) {
        m[d] = 2;
        stack.push_back(m + d);
      }
    }
  }

  for (int i = 0; i < h; i++) {
    const uint8_t* _map = map + mapstep * (i + 1) + 1;
    uint8_t* out = edges + (size_t)i * w;
    for (int j = 0; j < w; j++) out[j] = (uint8_t)(-(_map[j] >> 1));
  }
}

// ---------------------------------------------------------------------------
// a6: thinning, thinning.cpp:12-51.  The tables are held as 512-bit masks
// (bit i set <=> table value 255).  Pass 1 writes the interior of a fresh
// temporary whose border stays 0 (Level.cpp:176, a zero-filled allocation);
// pass 2 writes the interior of the edge map, whose border keeps its Canny
// value.
static const uint32_t kLutThin[2][16] = {
    {0x337f0000u, 0x33330000u, 0x31ff0000u, 0x33ff0000u, 0xffff0000u, 0xffff0000u,
     0xffff0000u, 0xffff0000u, 0xffff0000u, 0x33230000u, 0xffff0000u, 0xffff0000u,
     0xffff0000u, 0xff220000u, 0xffff0000u, 0xffff0000u},
    {0xf7ff0000u, 0xffff0000u, 0xf5ff0000u, 0xffff0000u, 0xf4ff0000u, 0xecee0000u,
     0xf4ff0000u, 0xecee0000u, 0xffff0000u, 0xffef0000u, 0xffff0000u, 0xffff0000u,
     0xf4ff0000u, 0xecee0000u, 0xf4fe0000u, 0xecee0000u},
};

int thin_lut(int pass, int index) {
  return (kLutThin[pass][index >> 5] >> (index & 31)) & 1u ? 255 : 0;
}

static void image_iter(const uint8_t* in, uint8_t* out, int w, int h, int pass) {
  for (int y = 1; y < h - 1; ++y) {
    const uint8_t* pm = in + (size_t)(y - 1) * w;
    const uint8_t* p0 = in + (size_t)y * w;
    const uint8_t* pp = in + (size_t)(y + 1) * w;
    uint8_t* o = out + (size_t)y * w;
    for (int x = 1; x < w - 1; ++x) {
      if (p0[x] == 0) {
        o[x] = 0;
        continue;
      }
      const int ind = (pm[x - 1] == 255) + (pm[x] == 255) * 8 + (pm[x + 1] == 255) * 64 +
                      (p0[x - 1] == 255) * 2 + (p0[x] == 255) * 16 + (p0[x + 1] == 255) * 128 +
                      (pp[x - 1] == 255) * 4 + (pp[x] == 255) * 32 + (pp[x + 1] == 255) * 256;
      o[x] = (uint8_t)thin_lut(pass, ind);
    }
  }
}

void thin(uint8_t* edges, int w, int h) {
  std::vector<uint8_t> temp((size_t)w * h, 0);
  image_iter(edges, temp.data(), w, h, 0);
  image_iter(temp.data(), edges, w, h, 1);
}

// ---------------------------------------------------------------------------
// a8: Bresenham.cpp:18-209.  The walk is bounded by the collection shape
// (W0, H0) -- the level-0 size at every level -- and looks edges up in the
// level's map (points beyond the level size are absent).
static void update_xy(float dx, float dy, int& x, int& y, float& e, int& stpX, int& stpY) {
  const float a = std::fabs(dy / dx);
  stpX = (dx > 0) - (dx < 0);
  stpY = (dy > 0) - (dy < 0);
  e += a;
  x += stpX;
  if (e >= 0.5f) {
    y += stpY;
    e -= 1;
  }
}

static int descent(const LevelResult& L, int ip, int dir, int nmax, int thr) {
  const int px = L.ex[ip], py = L.ey[ip];
  const float gx = L.dx[(size_t)py * L.w + px], gy = L.dy[(size_t)py * L.w + px];
  float e = 0.0f;
  float dx = dir * gx, dy = dir * gy;
  const float dxRef = dx, dyRef = dy;
  const float adx = std::fabs(dx), ady = std::fabs(dy);
  int n = 0, stpX = 0, stpY = 0;
  int x = px, y = py;
  const int W = L.W0, H = L.H0;
  auto reorient = [&]() {
    if (dx * dx + dy * dy > thr) {
      const float dx2 = gx, dy2 = gy;
      const float d = dx2 * dxRef + dy2 * dyRef;
      const float sgn = (float)((d > 0) - (d < 0));
      dx = sgn * dx2;
      dy = sgn * dy2;
    }
  };
  if (ady > adx) {
    update_xy(dy, dx, y, x, e, stpY, stpX);
    n = n + 1;
    reorient();
    update_xy(dy, dx, y, x, e, stpY, stpX);
    n = n + 1;
    if (x >= 0 && x < W && y >= 0 && y < H) {
      const int r = L.edge_at(x, y);
      if (r >= 0) return r;
    } else {
      return -1;
    }
    while (n <= nmax) {
      update_xy(dy, dx, y, x, e, stpY, stpX);
      n = n + 1;
      if (x >= 0 && x < W && y >= 0 && y < H) {
        int r = L.edge_at(x, y);
        if (r >= 0) return r;
        if (x >= 0 && x < W && (y - stpY) >= 0 && (y - stpY) < H) {
          r = L.edge_at(x, y - stpY);
          if (r >= 0) return r;
        } else {
          return -1;
        }
      } else {
        return -1;
      }
    }
  } else {
    update_xy(dx, dy, x, y, e, stpX, stpY);
    n = n + 1;
    reorient();
    update_xy(dx, dy, x, y, e, stpX, stpY);
    n = n + 1;
    if (x >= 0 && x < W && y >= 0 && y < H) {
      const int r = L.edge_at(x, y);
      if (r >= 0) return r;
    } else {
      return -1;
    }
    while (n <= nmax) {
      update_xy(dx, dy, x, y, e, stpX, stpY);
      n = n + 1;
      if (x >= 0 && x < W && y >= 0 && y < H) {
        int r = L.edge_at(x, y);
        if (r >= 0) return r;
        if ((x - stpX) >= 0 && (x - stpX) < W && y >= 0 && y < H) {
          r = L.edge_at(x - stpX, y);
          if (r >= 0) return r;
        } else {
          return -1;
        }
      } else {
        return -1;
      }
    }
  }
  return -1;
}

static float dist2d(const LevelResult& L, int a, int b) {  // Distance.hpp:19-24
  const int ddx = L.ex[b] - L.ex[a], ddy = L.ey[b] - L.ey[a];
  return std::sqrt((float)(ddx * ddx) + (float)(ddy * ddy));
}

static float neg_dot(const LevelResult& L, int a, int b) {  // -grad(a).grad(b)
  return -((float)L.edx[a] * (float)L.edx[b] + (float)L.edy[a] * (float)L.edy[b]);
}

// a7 + a9 + a10: Canny.cpp:24-33 (raster order), Vote.cpp:58-247.
void edges_and_vote(LevelResult& L, const Params& p) {
  const int w = L.w, h = L.h;
  L.edge_map.assign((size_t)w * h, -1);
  L.edge_bits.assign(((size_t)w * h + 31) / 32, 0u);
  L.ex.clear(); L.ey.clear(); L.edx.clear(); L.edy.clear();
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const size_t o = (size_t)y * w + x;
      if (L.edges[o] == 255) {
        L.edge_map[o] = (int)L.ex.size();
        L.edge_bits[o >> 5] |= 1u << (o & 31);
        L.ex.push_back((int16_t)x);
        L.ey.push_back((int16_t)y);
        L.edx.push_back(L.dx[o]);
        L.edy.push_back(L.dy[o]);
      }
    }
  const int n = (int)L.ex.size();
  L.before.assign(n, -1);
  L.after.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    L.before[i] = descent(L, i, -1, p.dist_search, p.thr_grad_vote);
    L.after[i] = descent(L, i, 1, p.dist_search, p.thr_grad_vote);
  }
  if (p.angle_voting != 0.0f)
    throw std::domain_error("thrVotingAngle must be equal to 0");  // Vote.cpp:95-98

  L.chosen.assign(n, -1);
  L.total_dist.assign(n, 0.0f);
  L.flow_length.assign(n, 0.0f);
  L.is_max.assign(n, -1);
  L.seeds.clear();
  std::vector<std::vector<int>> voters(n);
  std::vector<float> vDist;
  for (int ip = 0; ip < n; ++ip) {
    int current = L.before[ip];
    int choosen = -1;
    vDist.clear();
    float totalDistance = 0.f;
    if (current >= 0) {
      float cosDiffTheta = neg_dot(L, ip, current);
      if (cosDiffTheta >= p.angle_voting) {
        float lastDist = dist2d(L, ip, current);
        vDist.push_back(lastDist);
        totalDistance += lastDist;
        int i = 1;
        while (i < p.n_crowns) {
          choosen = -1;
          int target = L.after[current];
          if (target < 0) break;
          cosDiffTheta = neg_dot(L, target, current);
          if (!(cosDiffTheta >= p.angle_voting)) break;
          float dist = dist2d(L, target, current);
          vDist.push_back(dist);
          totalDistance += dist;
          int flagDist = 1;
          if (vDist.size() > 1) {
            for (size_t a = 0; a < vDist.size(); ++a)
              for (size_t b = a + 1; b < vDist.size(); ++b)
                flagDist = (vDist[a] <= vDist[b] * p.ratio_voting) &&
                           (vDist[b] <= vDist[a] * p.ratio_voting) && flagDist;
          }
          if (!flagDist) break;
          lastDist = dist;
          current = target;
          target = L.before[current];
          if (target < 0) break;
          cosDiffTheta = neg_dot(L, target, current);
          if (!(cosDiffTheta >= p.angle_voting)) break;
          dist = dist2d(L, target, current);
          vDist.push_back(dist);
          totalDistance += dist;
          for (size_t a = 0; a < vDist.size(); ++a)
            for (size_t b = a + 1; b < vDist.size(); ++b)
              flagDist = (vDist[a] <= vDist[b] * p.ratio_voting) &&
                         (vDist[b] <= vDist[a] * p.ratio_voting) && flagDist;
          if (!flagDist) break;
          lastDist = dist;
          current = target;
          choosen = current;
          ++i;
        }
      }
    }
    L.total_dist[ip] = totalDistance;
    if (choosen >= 0) {
      L.chosen[ip] = choosen;
      std::vector<int>& vl = voters[choosen];
      vl.push_back(ip);
      const size_t nv = vl.size();
      L.flow_length[choosen] = (L.flow_length[choosen] * (nv - 1) + totalDistance) / nv;
      if (nv >= (size_t)p.min_votes) {
        if (L.is_max[choosen] == -1) L.seeds.push_back(choosen);
        L.is_max[choosen] = (int)nv;
      }
    }
  }
  L.voters_index.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) L.voters_index[i + 1] = L.voters_index[i] + (int)voters[i].size();
  L.voters_list.clear();
  for (int i = 0; i < n; ++i)
    L.voters_list.insert(L.voters_list.end(), voters[i].begin(), voters[i].end());
}

// ImagePyramid.cpp:23-52 (levels floor-halved), Level.cpp:193-217,
// Multiresolution.cpp:225-237.
void front_end(const uint8_t* img, int w, int h, const Params& p,
               std::vector<LevelResult>& levels) {
  levels.assign(p.n_levels, LevelResult());
  int lw = w, lh = h;
  for (int l = 0; l < p.n_levels; ++l) {
    LevelResult& L = levels[l];
    L.w = lw; L.h = lh; L.W0 = w; L.H0 = h;
    L.img.resize((size_t)lw * lh);
    if (l == 0) std::memcpy(L.img.data(), img, (size_t)lw * lh);
    else downsample(levels[l - 1].img.data(), levels[l - 1].w, levels[l - 1].h, L.img.data(), lw, lh);
    L.dx.resize((size_t)lw * lh);
    L.dy.resize((size_t)lw * lh);
    derivatives(L.img.data(), lw, lh, L.dx.data(), L.dy.data());
    L.edges.resize((size_t)lw * lh);
    canny(L.dx.data(), L.dy.data(), lw, lh, p.canny_low * 256, p.canny_high * 256, L.edges.data());
    thin(L.edges.data(), lw, lh);
    edges_and_vote(L, p);
    lw /= 2;
    lh /= 2;
  }
}

}  // namespace oracle
