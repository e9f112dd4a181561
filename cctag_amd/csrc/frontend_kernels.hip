// gfx950 kernels of the CCTag front end: pyramid, derivative filters, Canny
// (magnitude + NMS, union-find hysteresis), thinning, raster-order edge
// compaction, gradient descent, vote, voter lists, seeds.
//
// Every kernel covers all planes (levels x frames) of a plane set in one
// launch.  Float arithmetic reproduces the reference CPU path bit for bit:
// the build disables contraction, and every rounding step (FMA, division,
// square root, round-half-even) is spelled out with the IEEE intrinsics.
#include <hip/hip_runtime.h>

#include "frontend.hpp"
#include "frontend_launch.hpp"

namespace cctagx {
namespace {

// ---------------------------------------------------------------------------
// Plane lookup: the last plane whose first tile is <= blk (planes with no
// tiles in a grid share their start with the following plane).
template <int FIELD>
__device__ __forceinline__ int plane_of(const PlaneDesc* P, int np, int blk) {
  const int32_t* a = reinterpret_cast<const int32_t*>(P);
  constexpr int S = sizeof(PlaneDesc) / sizeof(int32_t);
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid * S + FIELD] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
constexpr int F_TILE_F = 8, F_TILE_S = 9, F_TILE_C = 10, F_TILE_P = 11;

// The same lookup over a plain array of per-plane first blocks.
__device__ __forceinline__ int plane_of_starts(const int32_t* a, int np, int blk) {
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
static_assert(sizeof(PlaneDesc) == 13 * sizeof(int32_t), "PlaneDesc layout");

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Correctly rounded fp32 square root.  On gfx950 the library's __fsqrt_rn
// lowers to a bare v_sqrt_f32 (not correctly rounded); sqrtf under the
// default correctly-rounded-sqrt mode adds the FMA correction step.
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }

// ---------------------------------------------------------------------------
// a1: pyramid (Level.cpp:67).  cv::resize to exactly half size on 8U runs
// OpenCV's fast 2x2 area path: vector lanes (the first floor(w/8)*8 outputs
// of a row) compute (a+b+c+d+2)>>2, the scalar tail rounds (a+b+c+d)*0.25f
// half to even.  One block turns a 64x64 level-0 tile into 32x32, 16x16 and
// 8x8 tiles of levels 1..3 through LDS.  It also writes the tile's quads:
// per level-0 pixel its 2x2 neighbourhood as one 32-bit word, so that the
// identification's bilinear samples take one aligned load instead of four
// byte loads.
__device__ __forceinline__ uint8_t area2(int s, int x, int vec_end) {
  return (uint8_t)(x < vec_end ? (s + 2) >> 2 : (int)rintf((float)s * 0.25f));
}

__global__ __launch_bounds__(256) void k_pyramid(const PlaneDesc* __restrict__ P, int np,
                                                  int n_levels, uint8_t* __restrict__ img,
                                                  uint32_t* __restrict__ quad) {
  __shared__ uint8_t s1[32][33];
  __shared__ uint8_t s2[16][17];
  const int p0 = plane_of<F_TILE_P>(P, np, blockIdx.x);
  const PlaneDesc d0 = P[p0];
  const int t = blockIdx.x - d0.tile_p;
  const int tx_n = (d0.w + kPyrTile - 1) / kPyrTile;
  const int X0 = (t % tx_n) * kPyrTile, Y0 = (t / tx_n) * kPyrTile;
  const int tid = threadIdx.x;
  for (int k = tid; k < kPyrTile * kPyrTile; k += 256) {
    const int x = X0 + (k % kPyrTile), y = Y0 + k / kPyrTile;
    if (x < d0.w && y < d0.h) {
      const uint8_t* r0 = img + d0.pix + y * d0.w + x;
      const int dx = x + 1 < d0.w ? 1 : 0, dy = y + 1 < d0.h ? d0.w : 0;
      quad[d0.pix + y * d0.w + x] = (uint32_t)r0[0] | (uint32_t)r0[dx] << 8 |
                                    (uint32_t)r0[dy] << 16 | (uint32_t)r0[dy + dx] << 24;
    }
  }
  if (n_levels < 2) return;
  // level 1: 32x32, 4 outputs per thread, straight from global level 0
  {
    const PlaneDesc d1 = P[p0 + 1];
    const int vec_end = (d1.w / 8) * 8;
    for (int k = tid; k < 32 * 32; k += 256) {
      const int ly = k >> 5, lx = k & 31;
      const int x = (X0 >> 1) + lx, y = (Y0 >> 1) + ly;
      uint8_t v = 0;
      if (x < d1.w && y < d1.h) {
        const uint8_t* r0 = img + d0.pix + (2 * y) * d0.w + 2 * x;
        const int s = r0[0] + r0[1] + r0[d0.w] + r0[d0.w + 1];
        v = area2(s, x, vec_end);
        img[d1.pix + y * d1.w + x] = v;
      }
      s1[ly][lx] = v;
    }
  }
  if (n_levels < 3) return;
  __syncthreads();
  {
    const PlaneDesc d2 = P[p0 + 2];
    const int vec_end = (d2.w / 8) * 8;
    const int ly = tid >> 4, lx = tid & 15;
    const int x = (X0 >> 2) + lx, y = (Y0 >> 2) + ly;
    uint8_t v = 0;
    if (x < d2.w && y < d2.h) {
      const int s = s1[2 * ly][2 * lx] + s1[2 * ly][2 * lx + 1] + s1[2 * ly + 1][2 * lx] +
                    s1[2 * ly + 1][2 * lx + 1];
      v = area2(s, x, vec_end);
      img[d2.pix + y * d2.w + x] = v;
    }
    s2[ly][lx] = v;
  }
  if (n_levels < 4) return;
  __syncthreads();
  if (tid < 64) {
    const PlaneDesc d3 = P[p0 + 3];
    const int vec_end = (d3.w / 8) * 8;
    const int ly = tid >> 3, lx = tid & 7;
    const int x = (X0 >> 3) + lx, y = (Y0 >> 3) + ly;
    if (x < d3.w && y < d3.h) {
      const int s = s2[2 * ly][2 * lx] + s2[2 * ly][2 * lx + 1] + s2[2 * ly + 1][2 * lx] +
                    s2[2 * ly + 1][2 * lx + 1];
      img[d3.pix + y * d3.w + x] = area2(s, x, vec_end);
    }
  }
}

// a1, general sizes: levels are floor-halved (ImagePyramid.cpp:23-33), so a
// level whose parent has an odd side is not an exact half of it, and
// cv::resize then takes OpenCV's generic fixed-point bilinear path (8U,
// 11-bit weights).  Along an axis of source length sn resampled to dn, output
// d reads source s0 = floor(f) and s0 + 1, f = (float)((d + 0.5) * scale -
// 0.5), scale = 1 / (dn / sn) in double, with weights round((1 - f) * 2048)
// and round(f * 2048), half to even; at the last source sample it reads that
// sample alone.  Rows are combined exactly, the two rows as
// ((b0 * (S0 >> 4)) >> 16) + ((b1 * (S1 >> 4)) >> 16), + 2, >> 2.
struct LinearTap {
  int s0, a0, a1;
};
__device__ __forceinline__ LinearTap linear_tap(int d, int sn, int dn) {
  const double scale = __ddiv_rn(1.0, __ddiv_rn((double)dn, (double)sn));
  float f = __double2float_rn(__dsub_rn(__dmul_rn(__dadd_rn((double)d, 0.5), scale), 0.5));
  int s = (int)floorf(f);
  f = __fsub_rn(f, (float)s);
  if (s < 0) {
    f = 0.0f;
    s = 0;
  }
  if (s >= sn - 1) {
    f = 0.0f;
    s = sn - 1;
  }
  return LinearTap{s, (int)rintf(__fmul_rn(__fsub_rn(1.0f, f), 2048.0f)),
                   (int)rintf(__fmul_rn(f, 2048.0f))};
}

// One level from its parent, for the planes of one level (an exact half
// takes the 2x2 area rule of k_pyramid).  64x16 outputs per block: 64 lanes
// across a row, each thread 4 rows.
__global__ __launch_bounds__(256) void k_resample(const PlaneDesc* __restrict__ P,
                                                   const int32_t* __restrict__ lplanes, int nlp,
                                                   uint8_t* __restrict__ img) {
  int lo = 0, hi = nlp - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P[lplanes[mid]].tile_r <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const PlaneDesc d = P[lplanes[lo]];
  const PlaneDesc s = P[d.parent];
  const int t = blockIdx.x - d.tile_r;
  const int tx_n = (d.w + kResTileW - 1) / kResTileW;
  const int x = (t % tx_n) * kResTileW + (threadIdx.x & 63);
  const int y0 = (t / tx_n) * kResTileH + (threadIdx.x >> 6) * (kResTileH / 4);
  if (x >= d.w) return;
  const uint8_t* src = img + s.pix;
  uint8_t* dst = img + d.pix;
  if (s.w == 2 * d.w && s.h == 2 * d.h) {
    const int vec_end = (d.w / 8) * 8;
    for (int k = 0; k < kResTileH / 4; ++k) {
      const int y = y0 + k;
      if (y >= d.h) break;
      const uint8_t* r0 = src + (2 * y) * s.w + 2 * x;
      dst[y * d.w + x] = area2(r0[0] + r0[1] + r0[s.w] + r0[s.w + 1], x, vec_end);
    }
    return;
  }
  const LinearTap cx = linear_tap(x, s.w, d.w);
  const int x1 = cx.s0 + 1 < s.w ? cx.s0 + 1 : cx.s0;
  const int a1 = cx.s0 + 1 < s.w ? cx.a1 : 0;
  const int a0 = cx.s0 + 1 < s.w ? cx.a0 : 2048;
  for (int k = 0; k < kResTileH / 4; ++k) {
    const int y = y0 + k;
    if (y >= d.h) break;
    const LinearTap cy = linear_tap(y, s.h, d.h);
    const int r0 = min(cy.s0, s.h - 1), r1 = min(cy.s0 + 1, s.h - 1);
    const int h0 = src[r0 * s.w + cx.s0] * a0 + src[r0 * s.w + x1] * a1;
    const int h1 = src[r1 * s.w + cx.s0] * a0 + src[r1 * s.w + x1] * a1;
    dst[y * d.w + x] =
        (uint8_t)((((cy.a0 * (h0 >> 4)) >> 16) + ((cy.a1 * (h1 >> 4)) >> 16) + 2) >> 2);
  }
}

// ---------------------------------------------------------------------------
// a2: derivative filters (cvRecode.cpp:74-92).  kerneldX(i, j) = -/+ tap for
// j < 4 / j > 4 (centre column zero); kerneldY = kerneldX^T.  The reference
// (cv::filter2D, 8U->16S direct path) accumulates the 72 nonzero taps of each
// kernel in row-major order with one FMA per tap from 0, then rounds half to
// even.  A thread owns one column of 8 output rows; it streams the 16 window
// rows from LDS once and applies each kernel row to the outputs that use it,
// in increasing kernel-row order, so every output keeps the reference order.
__host__ __device__ constexpr float tap(int r, int c) {
  return r == 0 ? (c == 0 ? 0.000000143284235f : c == 1 ? 0.000003558691641f
                 : c == 2 ? 0.000028902492951f : 0.000064765993382f)
       : r == 1 ? (c == 0 ? 0.000004744922188f : c == 1 ? 0.000117847682078f
                 : c == 2 ? 0.000957119116802f : 0.002144755142391f)
       : r == 2 ? (c == 0 ? 0.000057804985902f : c == 1 ? 0.001435678675203f
                 : c == 2 ? 0.011660097860113f : 0.026128466569370f)
       : r == 3 ? (c == 0 ? 0.000259063973527f : c == 1 ? 0.006434265427174f
                 : c == 2 ? 0.052256933138740f : 0.117099663048638f)
                : (c == 0 ? 0.000427124283626f : c == 1 ? 0.010608310271112f
                 : c == 2 ? 0.086157117207395f : 0.193064705260108f);
}
__host__ __device__ constexpr float kdx(int i, int j) {
  return j == 4 ? 0.0f : (j < 4 ? -tap(i <= 4 ? i : 8 - i, j) : tap(i <= 4 ? i : 8 - i, 8 - j));
}

// RPT rows per thread: a 64 x 4*RPT tile per 256-thread block.
template <int RPT>
__global__ __launch_bounds__(256) void k_filter(const PlaneDesc* __restrict__ P, int np,
                                                 const int32_t* __restrict__ starts,
                                                 const uint8_t* __restrict__ img,
                                                 int16_t* __restrict__ dxo,
                                                 int16_t* __restrict__ dyo) {
  constexpr int TW = kFilterTileW, TH = 4 * RPT, LW = TW + 8, LH = TH + 8;
  __shared__ float s[LH][LW + 1];
  const int p = plane_of_starts(starts, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int t = blockIdx.x - starts[p];
  const int tx_n = (d.w + TW - 1) / TW;
  const int x0 = (t % tx_n) * TW, y0 = (t / tx_n) * TH;
  const uint8_t* src = img + d.pix;
  for (int k = threadIdx.x; k < LH * LW; k += 256) {
    const int ly = k / LW, lx = k - ly * LW;
    const int gy = clampi(y0 + ly - 4, 0, d.h - 1), gx = clampi(x0 + lx - 4, 0, d.w - 1);
    s[ly][lx] = (float)src[gy * d.w + gx];
  }
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = (threadIdx.x >> 6) * RPT;
  float ax[RPT], ay[RPT];
#pragma unroll
  for (int o = 0; o < RPT; ++o) { ax[o] = 0.0f; ay[o] = 0.0f; }
#pragma unroll
  for (int r = 0; r < RPT + 8; ++r) {
    float v[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) v[j] = s[ty + r][tx + j];
#pragma unroll
    for (int o = 0; o < RPT; ++o) {
      const int i = r - o;
      if (i >= 0 && i < 9) {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
          if (kdx(i, j) != 0.0f) ax[o] = __builtin_fmaf(v[j], kdx(i, j), ax[o]);
          if (kdx(j, i) != 0.0f) ay[o] = __builtin_fmaf(v[j], kdx(j, i), ay[o]);
        }
      }
    }
  }
  const int x = x0 + tx;
  if (x < d.w) {
#pragma unroll
    for (int o = 0; o < RPT; ++o) {
      const int y = y0 + ty + o;
      if (y < d.h) {
        const int g = d.pix + y * d.w + x;
        dxo[g] = (int16_t)(int)rintf(ax[o]);
        dyo[g] = (int16_t)(int)rintf(ay[o]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// a3 + a4: magnitude and NMS (cvRecode.cpp:174-330).  Magnitude
// rint(sqrt(x*x + y*y)), zero outside the image; TG22 = 13573 in Q15; the
// horizontal / vertical tests are (>, >=), the diagonal test is (>, >).
// A candidate (bit 0) passed NMS with m > low; a strong pixel (bit 1) also
// has m > high.  The reference's prev_flag / upper-neighbour demotions only
// decide which strong pixels seed its flood fill, not which pixels it
// reaches, so the final map is: candidates 8-connected to a strong pixel.
__device__ __forceinline__ int magnitude(int x, int y) {
  return (int)rintf(sqrt_rn((float)x * (float)x + (float)y * (float)y));
}

// The class of pixel (tx, ty) of a tile from its gradient and the tile's
// magnitudes with a one-pixel ring.
constexpr int TG22 = 13573;
__device__ __forceinline__ uint8_t nms_class(int sx, int sy,
                                             const int (*smag)[kStencilTileW + 2], int tx,
                                             int ty, int low, int high) {
  const int m = smag[ty + 1][tx + 1];
  const int ax = sx < 0 ? -sx : sx;
  int ay = sy < 0 ? -sy : sy;
  if (m <= low) return 0;
  const int tg22x = ax * TG22;
  const int tg67x = tg22x + ((ax + ax) << 15);
  ay <<= 15;
  bool pass;
  if (ay < tg22x) {
    pass = m > smag[ty + 1][tx] && m >= smag[ty + 1][tx + 2];
  } else if (ay > tg67x) {
    pass = m > smag[ty][tx + 1] && m >= smag[ty + 2][tx + 1];
  } else {
    const int s = (sx ^ sy) < 0 ? -1 : 1;
    pass = m > smag[ty][tx + 1 - s] && m > smag[ty + 2][tx + 1 + s];
  }
  return pass ? (uint8_t)(m > high ? 3 : 1) : 0;
}

// ---------------------------------------------------------------------------
// a5: hysteresis as connected components over candidate pixels, with the
// lock-free union-find of Playne & Hawick (IEEE TPDS 2018): link the larger
// root under the smaller with atomicMin until both finds agree.
__device__ __forceinline__ int ld(const int32_t* L, int i) {
  return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int uf_find(const int32_t* L, int x) {
  int p = ld(L, x);
  while (p != x) {
    x = p;
    p = ld(L, x);
  }
  return x;
}
__device__ void uf_merge(int32_t* L, int a, int b) {
  bool done;
  do {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a < b) {
      const int old = atomicMin(L + b, a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(L + a, b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  } while (!done);
}

// Two levels: k_nms_local joins the candidates of one 64x16 tile in LDS and
// labels each with the global index of its tile-local root (the smallest
// index of its tile-local component, since local and global raster orders
// agree inside a tile); k_cc_border then joins across tile edges in global
// memory.  The roots that result are the components' smallest indices, as
// with a single global pass.
__device__ __forceinline__ int lds_ld(const int* L, int i) {
  return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lds_find(const int* L, int x) {
  int p = lds_ld(L, x);
  while (p != x) {
    x = p;
    p = lds_ld(L, x);
  }
  return x;
}
__device__ void lds_merge(int* L, int a, int b) {
  bool done;
  do {
    a = lds_find(L, a);
    b = lds_find(L, b);
    if (a < b) {
      const int old = atomicMin(L + b, a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(L + a, b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  } while (!done);
}

// Magnitude, NMS and the tile-local joins in one block per 64x16 tile: the
// classes (written, with the strong-root flags cleared) and then the joins
// by runs: a tile row is one wavefront's 64-bit candidate mask (a ballot); a
// candidate's parent is the first pixel of its run (its maximal horizontal
// segment of candidates), the run's root; then the first pixel of each run
// joins every run of the row above that touches it 8-connectedly (the row
// above masked to the run's columns widened by one on each side).  The
// components, and so their smallest indices, are those the per-pixel joins
// give.
static_assert(kStencilTileW == 64, "a tile row is one wavefront");
__global__ __launch_bounds__(256) void k_nms_local(const PlaneDesc* __restrict__ P, int np,
                                                    const int16_t* __restrict__ dx,
                                                    const int16_t* __restrict__ dy, int low,
                                                    int high, uint8_t* __restrict__ cls,
                                                    int32_t* __restrict__ lab,
                                                    uint8_t* __restrict__ root_strong) {
  constexpr int TW = kStencilTileW, TH = kStencilTileH, N = TW * TH;
  constexpr int kWaves = 256 / 64;
  __shared__ int smag[TH + 2][TW + 2];
  __shared__ int L[N];
  __shared__ uint64_t M[TH];
  const int p = plane_of<F_TILE_S>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int t = blockIdx.x - d.tile_s;
  const int tx_n = (d.w + TW - 1) / TW;
  const int x0 = (t % tx_n) * TW, y0 = (t / tx_n) * TH;
  for (int k = threadIdx.x; k < (TH + 2) * (TW + 2); k += 256) {
    const int ly = k / (TW + 2), lx = k - ly * (TW + 2);
    const int gx = x0 + lx - 1, gy = y0 + ly - 1;
    int m = 0;
    if (gx >= 0 && gx < d.w && gy >= 0 && gy < d.h) {
      const int g = d.pix + gy * d.w + gx;
      m = magnitude(dx[g], dy[g]);
    }
    smag[ly][lx] = m;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = x0 + lane;
  for (int ly = wave; ly < TH; ly += kWaves) {
    const int y = y0 + ly;
    bool c = false;
    if (x < d.w && y < d.h) {
      const int g = d.pix + y * d.w + x;
      const uint8_t k = nms_class(dx[g], dy[g], smag, lane, ly, low, high);
      cls[g] = k;
      root_strong[g] = 0;
      c = k & 1;
    }
    const uint64_t m = __ballot(c);
    if (lane == 0) M[ly] = m;
    // the run's first pixel: one past the highest non-candidate below
    const uint64_t gaps = ~m & ((1ull << lane) - 1ull);
    const int first = gaps ? 64 - __clzll(gaps) : 0;
    L[ly * TW + lane] = c ? ly * TW + first : -1;
  }
  __syncthreads();
  for (int ly = 1 + wave; ly < TH; ly += kWaves) {
    const uint64_t m = M[ly], up = M[ly - 1];
    const bool starts = ((m >> lane) & 1ull) && (lane == 0 || !((m >> (lane - 1)) & 1ull));
    if (!starts || !up) continue;
    const uint64_t rest = ~(m >> lane);  // bit j clear: lane + j is in the run
    const int len = rest ? __ffsll((long long)rest) - 1 : 64;
    const int lo = lane > 0 ? lane - 1 : 0, hi = min(TW - 1, lane + len);
    const int width = hi - lo + 1;
    const uint64_t window = width == 64 ? ~0ull : ((1ull << width) - 1ull) << lo;
    uint64_t touch = up & window;  // one segment per run above
    while (touch) {
      const int b = __ffsll((long long)touch) - 1;
      lds_merge(L, ly * TW + lane, (ly - 1) * TW + b);
      const uint64_t seg = ~(touch >> b);
      const int n = seg ? __ffsll((long long)seg) - 1 : 64 - b;
      touch = b + n >= 64 ? 0ull : touch & (~0ull << (b + n));
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < N; k += 256) {
    if (L[k] < 0) continue;
    const int r = lds_find(L, k);
    lab[d.pix + (y0 + k / TW) * d.w + x0 + k % TW] = d.pix + (y0 + r / TW) * d.w + x0 + r % TW;
  }
}

// The neighbour pairs that cross a tile edge, joined in global memory.  The
// tile's first row against the row above by runs, one wavefront: each run
// of the first row joins every run of the row above that touches it
// 8-connectedly (its columns widened by one, into the neighbouring tiles'
// columns at the ends), and the first pixel joins its left neighbour.  The
// other rows pixel by pixel, one lane each: left and up-left on the first
// column, up-right on the last.
__global__ __launch_bounds__(128) void k_cc_border(const PlaneDesc* __restrict__ P, int np,
                                                   const uint8_t* __restrict__ cls,
                                                   int32_t* __restrict__ lab) {
  constexpr int TW = kStencilTileW, TH = kStencilTileH;
  const int p = plane_of<F_TILE_S>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int t = blockIdx.x - d.tile_s;
  const int tx_n = (d.w + TW - 1) / TW;
  const int x0 = (t % tx_n) * TW, y0 = (t / tx_n) * TH;
  const int i = threadIdx.x;
  if (i < 64) {
    const int x = x0 + i;
    const int g = d.pix + y0 * d.w + x;  // row y0 < d.h: the tile exists
    const bool in = x < d.w;
    const bool c = in && (cls[g] & 1);
    if (i == 0 && c && x0 > 0 && (cls[g - 1] & 1)) uf_merge(lab, g, g - 1);
    if (y0 == 0) return;
    const uint64_t m = __ballot(c);
    const uint64_t up = __ballot(in && (cls[g - d.w] & 1));
    const int u0 = g - i - d.w;  // row above, column x0
    const bool up_left = x0 > 0 && (cls[u0 - 1] & 1);
    const bool up_right = x0 + TW < d.w && (cls[u0 + TW] & 1);
    const bool starts = c && (i == 0 || !((m >> (i - 1)) & 1ull));
    if (!starts) return;
    const uint64_t rest = ~(m >> i);  // bit j clear: i + j is in the run
    const int len = rest ? __ffsll((long long)rest) - 1 : 64;
    const int lo = i > 0 ? i - 1 : 0, hi = min(TW - 1, i + len);
    const int width = hi - lo + 1;
    const uint64_t window = width == 64 ? ~0ull : ((1ull << width) - 1ull) << lo;
    uint64_t touch = up & window;  // one segment per run above
    while (touch) {
      const int b = __ffsll((long long)touch) - 1;
      uf_merge(lab, g, u0 + b);
      const uint64_t seg = ~(touch >> b);
      const int n = seg ? __ffsll((long long)seg) - 1 : 64 - b;
      touch = b + n >= 64 ? 0ull : touch & (~0ull << (b + n));
    }
    if (i == 0 && up_left) uf_merge(lab, g, u0 - 1);
    if (i + len >= TW && up_right) uf_merge(lab, g, u0 + TW);
    return;
  }
  // rows 1 .. TH-1 of the first column (j < TH-1) and of the last (j >= TH-1)
  const int j = i - 64;
  if (j >= 2 * (TH - 1)) return;
  const bool first_col = j < TH - 1;
  const int lx = first_col ? 0 : TW - 1, ly = first_col ? j + 1 : j - (TH - 1) + 1;
  const int x = x0 + lx, y = y0 + ly;
  if (x >= d.w || y >= d.h) return;
  const int g = d.pix + y * d.w + x;
  if (!(cls[g] & 1)) return;
  if (first_col) {
    if (x > 0 && (cls[g - 1] & 1)) uf_merge(lab, g, g - 1);
    if (x > 0 && (cls[g - d.w - 1] & 1)) uf_merge(lab, g, g - d.w - 1);
  } else if (x + 1 < d.w && (cls[g - d.w + 1] & 1)) {
    uf_merge(lab, g, g - d.w + 1);
  }
}

__global__ __launch_bounds__(256) void k_cc_flatten(const PlaneDesc* __restrict__ P, int np,
                                                     const uint8_t* __restrict__ cls,
                                                     int32_t* __restrict__ lab,
                                                     uint8_t* __restrict__ root_strong) {
  constexpr int TW = kStencilTileW, TH = kStencilTileH;
  const int p = plane_of<F_TILE_S>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int t = blockIdx.x - d.tile_s;
  const int tx_n = (d.w + TW - 1) / TW;
  const int x = (t % tx_n) * TW + (threadIdx.x & 63);
  const int y0 = (t / tx_n) * TH;
  if (x >= d.w) return;
  for (int ty = threadIdx.x >> 6; ty < TH; ty += 4) {
    const int y = y0 + ty;
    if (y >= d.h) break;
    const int g = d.pix + y * d.w + x;
    const uint8_t c = cls[g];
    if (!(c & 1)) continue;
    const int r = uf_find(lab, g);
    lab[g] = r;
    if (c & 2) root_strong[r] = 1;
  }
}

// ---------------------------------------------------------------------------
// a5 final map + a6 thinning (cvRecode.cpp:401-414, thinning.cpp:12-51).
// The 512-entry tables as bit masks: bit i set <=> value 255.
__device__ __forceinline__ bool lut_bit(int pass, int idx) {
  constexpr uint32_t T0[16] = {0x337f0000u, 0x33330000u, 0x31ff0000u, 0x33ff0000u,
                               0xffff0000u, 0xffff0000u, 0xffff0000u, 0xffff0000u,
                               0xffff0000u, 0x33230000u, 0xffff0000u, 0xffff0000u,
                               0xffff0000u, 0xff220000u, 0xffff0000u, 0xffff0000u};
  constexpr uint32_t T1[16] = {0xf7ff0000u, 0xffff0000u, 0xf5ff0000u, 0xffff0000u,
                               0xf4ff0000u, 0xecee0000u, 0xf4ff0000u, 0xecee0000u,
                               0xffff0000u, 0xffef0000u, 0xffff0000u, 0xffff0000u,
                               0xf4ff0000u, 0xecee0000u, 0xf4fe0000u, 0xecee0000u};
  const uint32_t w = pass == 0 ? T0[idx >> 5] : T1[idx >> 5];
  return (w >> (idx & 31)) & 1u;
}

// 3x3 index of thinning.cpp:42-45 over a tile with a 1-pixel halo.
template <int LW>
__device__ __forceinline__ int thin_index(const uint8_t (*s)[LW], int ly, int lx) {
  return (s[ly - 1][lx - 1] == 255) + (s[ly - 1][lx] == 255) * 8 + (s[ly - 1][lx + 1] == 255) * 64 +
         (s[ly][lx - 1] == 255) * 2 + (s[ly][lx] == 255) * 16 + (s[ly][lx + 1] == 255) * 128 +
         (s[ly + 1][lx - 1] == 255) * 4 + (s[ly + 1][lx] == 255) * 32 +
         (s[ly + 1][lx + 1] == 255) * 256;
}

// Canny map e = 255 for candidates whose component holds a strong pixel,
// then both thinning passes in one block: pass 1 over the tile and a one-pixel
// ring around it (from the map with a two-pixel ring; interior only, the
// temporary's border is 0), pass 2 over the tile.  Border pixels keep the
// Canny map's value, interior pixels take pass 2's.
__global__ __launch_bounds__(256) void k_thin(const PlaneDesc* __restrict__ P, int np,
                                               const uint8_t* __restrict__ cls,
                                               const int32_t* __restrict__ lab,
                                               const uint8_t* __restrict__ root_strong,
                                               uint8_t* __restrict__ edges) {
  constexpr int TW = kStencilTileW, TH = kStencilTileH;
  constexpr int EW = TW + 4, EH = TH + 4;  // the map: tile + 2-pixel ring
  constexpr int TWW = TW + 2, THH = TH + 2;  // pass 1: tile + 1-pixel ring
  __shared__ uint8_t se[EH][EW];
  __shared__ uint8_t st[THH][TWW];
  const int p = plane_of<F_TILE_S>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int t = blockIdx.x - d.tile_s;
  const int tx_n = (d.w + TW - 1) / TW;
  const int x0 = (t % tx_n) * TW, y0 = (t / tx_n) * TH;
  for (int k = threadIdx.x; k < EH * EW; k += 256) {
    const int ly = k / EW, lx = k - ly * EW;
    const int gx = x0 + lx - 2, gy = y0 + ly - 2;
    uint8_t e = 0;
    if (gx >= 0 && gx < d.w && gy >= 0 && gy < d.h) {
      const int g = d.pix + gy * d.w + gx;
      if ((cls[g] & 1) && root_strong[lab[g]]) e = 255;
    }
    se[ly][lx] = e;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < THH * TWW; k += 256) {
    const int ly = k / TWW, lx = k - ly * TWW;
    const int gx = x0 + lx - 1, gy = y0 + ly - 1;
    uint8_t o = 0;
    if (gx > 0 && gy > 0 && gx < d.w - 1 && gy < d.h - 1 && se[ly + 1][lx + 1] != 0)
      o = lut_bit(0, thin_index<EW>(se, ly + 1, lx + 1)) ? 255 : 0;
    st[ly][lx] = o;
  }
  __syncthreads();
  const int tx = threadIdx.x & 63;
  const int x = x0 + tx;
  for (int ty = threadIdx.x >> 6; ty < TH; ty += 4) {
    const int y = y0 + ty;
    if (x >= d.w || y >= d.h) continue;
    uint8_t o = se[ty + 2][tx + 2];
    if (x > 0 && y > 0 && x < d.w - 1 && y < d.h - 1)
      o = st[ty + 1][tx + 1] != 0 && lut_bit(1, thin_index<TWW>(st, ty + 1, tx + 1)) ? 255 : 0;
    edges[d.pix + y * d.w + x] = o;
  }
}

// ---------------------------------------------------------------------------
// Block-wide exclusive scan of one int per thread (256 threads = 4 waves).
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int n = __shfl_up(incl, o, 64);
    if (lane >= o) incl += n;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wid) base += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return base + incl - v;
}

// a7: raster-order compaction of edges == 255 (Canny.cpp:24-33) in 4096-px
// chunks of the plane's linear pixel order.  Also writes the plane's edge
// bitmask (32 pixels per word), which is what the host reads to locate
// points by position.
__global__ __launch_bounds__(256) void k_compact_count(const PlaneDesc* __restrict__ P, int np,
                                                        const uint8_t* __restrict__ edges,
                                                        int32_t* __restrict__ tcount,
                                                        uint32_t* __restrict__ ebits) {
  const int p = plane_of<F_TILE_C>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int c = blockIdx.x - d.tile_c;
  const int n = d.w * d.h;
  const int b = c * kChunk + threadIdx.x * 16;
  uint32_t m16 = 0;  // bit q = pixel b + q
  if (b + 16 <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(edges + d.pix + b);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) m16 |= (uint32_t)(((w[k] >> (8 * q)) & 0xffu) == 255u) << (4 * k + q);
  } else {
    for (int k = b; k < n && k < b + 16; ++k) m16 |= (uint32_t)(edges[d.pix + k] == 255) << (k - b);
  }
  const int cnt = __builtin_popcount(m16);
  // even lanes own a 32-pixel word: their 16 bits and the next lane's
  const uint32_t hi = __shfl_down(m16, 1, 64);
  if ((threadIdx.x & 1) == 0 && b < n) ebits[(d.pix >> 5) + (b >> 5)] = m16 | (hi << 16);
  int total;
  block_exclusive_scan(cnt, &total);
  if (threadIdx.x == 0) tcount[blockIdx.x] = total;
}

// Per-plane exclusive scan of chunk counts (one 1024-thread block per plane).
__global__ __launch_bounds__(1024) void k_compact_scan(const PlaneDesc* __restrict__ P, int np,
                                                        int32_t* __restrict__ tcount,
                                                        PlaneCounts* __restrict__ counts) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int p = blockIdx.x;
  const PlaneDesc d = P[p];
  const int nchunks = (d.w * d.h + kChunk - 1) / kChunk;
  int32_t* a = tcount + d.tile_c;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < nchunks; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < nchunks ? a[i] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int n = __shfl_up(incl, o, 64);
      if (lane >= o) incl += n;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wb = 0, tot = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wid) wb += wsum[w];
      tot += wsum[w];
    }
    if (i < nchunks) a[i] = carry + wb + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[p].n_edges = carry;
}

// Writes the edge list and, per 32-pixel mask word, the index of its first
// edge pixel (a point's index is then that plus the edges before it in the
// word).
__global__ __launch_bounds__(256) void k_compact_write(
    const PlaneDesc* __restrict__ P, int np, const uint8_t* __restrict__ edges,
    const int16_t* __restrict__ dx, const int16_t* __restrict__ dy,
    const int32_t* __restrict__ tcount, int32_t* __restrict__ erank, int16_t* __restrict__ ex,
    int16_t* __restrict__ ey, int16_t* __restrict__ egx, int16_t* __restrict__ egy,
    int32_t* __restrict__ vcount, int32_t* __restrict__ seedmark) {
  const int p = plane_of<F_TILE_C>(P, np, blockIdx.x);
  const PlaneDesc d = P[p];
  const int c = blockIdx.x - d.tile_c;
  const int n = d.w * d.h;
  const int b = c * kChunk + threadIdx.x * 16;
  uint8_t e[16];
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    e[k] = (b + k < n) ? edges[d.pix + b + k] : 0;
    cnt += e[k] == 255;
  }
  int total;
  int idx = tcount[blockIdx.x] + block_exclusive_scan(cnt, &total);
  // even lanes start a 32-pixel word
  if ((threadIdx.x & 1) == 0 && b < n) erank[(d.pix >> 5) + (b >> 5)] = idx;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int li = b + k;
    if (li >= n) break;
    if (e[k] == 255) {
      const int g = d.pix + li;
      const int y = li / d.w, x = li - y * d.w;
      const int s = d.pix + idx;
      ex[s] = (int16_t)x;
      ey[s] = (int16_t)y;
      egx[s] = dx[g];
      egy[s] = dy[g];
      vcount[s] = 0;
      seedmark[s] = -1;
      ++idx;
    }
  }
}

// ---------------------------------------------------------------------------
// a8: gradientDirectionDescent (Bresenham.cpp:18-209).  One lane per
// (edge point, direction).  Bounds are the collection shape (the level-0
// size); lookups beyond the level size find nothing.  The walk tests the
// plane's 1-bit edge mask (1/32 of a word per pixel: the whole pyramid's,
// 344 KB at 1920x1080, stays in L2); at the edge it stops on, the point's
// index is the index of the mask word's first edge (a second table of one
// word per 32 pixels) plus the edges before it in the word.
struct EdgeLookup {
  const uint32_t* bits;  // the plane's mask words (plane starts are 256-aligned)
  const int32_t* rank;   // the plane's first index per mask word
  int w, h;
  __device__ __forceinline__ bool edge(int lin) const {
    return (bits[lin >> 5] >> (lin & 31)) & 1u;
  }
  // index of the edge point at linear position lin (an edge)
  __device__ __forceinline__ int index(int lin) const {
    return rank[lin >> 5] + __builtin_popcount(bits[lin >> 5] & ((1u << (lin & 31)) - 1u));
  }
  __device__ __forceinline__ int at(int x, int y) const {
    if (x >= w || y >= h) return -1;
    const int lin = y * w + x;
    return edge(lin) ? index(lin) : -1;
  }
};

__device__ __forceinline__ void update_xy(float dx, float dy, int& x, int& y, float& e, int& sx,
                                          int& sy) {
  const float a = fabsf(__fdiv_rn(dy, dx));
  sx = (dx > 0.0f) - (dx < 0.0f);
  sy = (dy > 0.0f) - (dy < 0.0f);
  e = __fadd_rn(e, a);
  x += sx;
  if (e >= 0.5f) {
    y += sy;
    e = __fsub_rn(e, 1.0f);
  }
}

constexpr int kDescentBatch = 8;

__device__ int descent(const EdgeLookup& L, int W, int H, int px, int py, float gx, float gy,
                       int dir, int nmax, int thr) {
  float e = 0.0f;
  float dx = (float)dir * gx, dy = (float)dir * gy;
  const float dxRef = dx, dyRef = dy;
  const bool ymajor = fabsf(dy) > fabsf(dx);
  int n = 0, sx = 0, sy = 0, x = px, y = py;
  // steps 1 and 2, re-orientation in between
  for (int s = 0; s < 2; ++s) {
    if (ymajor) update_xy(dy, dx, y, x, e, sy, sx);
    else update_xy(dx, dy, x, y, e, sx, sy);
    n = n + 1;
    if (s == 0 && __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)) > (float)thr) {
      const float dd = __fadd_rn(__fmul_rn(gx, dxRef), __fmul_rn(gy, dyRef));
      const float sg = (float)((dd > 0.0f) - (dd < 0.0f));
      dx = sg * gx;
      dy = sg * gy;
    }
  }
  if (x >= 0 && x < W && y >= 0 && y < H) {
    const int r = L.at(x, y);
    if (r >= 0) return r;
  } else {
    return -1;
  }
  // Later steps move along fixed (dx, dy): the positions follow from
  // (x, y, e) alone, so kDescentBatch steps are laid out first and their
  // mask reads issued together (a read outside the level reads its first
  // word and is ignored).  Slot 2j is step j's point, slot 2j+1 the point
  // beside it; the walk ends at the first slot that is an edge or that a
  // one-step-at-a-time walk would not reach (past nmax or off the image),
  // found with bit masks so that no read waits on another.
  const float a = ymajor ? fabsf(__fdiv_rn(dx, dy)) : fabsf(__fdiv_rn(dy, dx));
  sx = (dx > 0.0f) - (dx < 0.0f);
  sy = (dy > 0.0f) - (dy < 0.0f);
  static_assert(2 * kDescentBatch < 32, "slot masks are 32-bit");
  while (n <= nmax) {
    int l0[kDescentBatch], l1[kDescentBatch];
    uint32_t hit = 0, reach = ~0u << (2 * kDescentBatch);  // bits above the slots: reached
#pragma unroll
    for (int j = 0; j < kDescentBatch; ++j) {
      const bool live = n + j <= nmax;
      e = __fadd_rn(e, a);
      if (ymajor) {
        y += sy;
        if (e >= 0.5f) {
          x += sx;
          e = __fsub_rn(e, 1.0f);
        }
      } else {
        x += sx;
        if (e >= 0.5f) {
          y += sy;
          e = __fsub_rn(e, 1.0f);
        }
      }
      const int xs = ymajor ? x : x - sx, ys = ymajor ? y - sy : y;
      const bool in0 = x >= 0 && x < W && y >= 0 && y < H;
      const bool in1 = xs >= 0 && xs < W && ys >= 0 && ys < H;
      const bool m0 = in0 && x < L.w && y < L.h, m1 = in1 && xs < L.w && ys < L.h;
      l0[j] = m0 ? y * L.w + x : 0;
      l1[j] = m1 ? ys * L.w + xs : 0;
      const uint32_t w0 = L.bits[l0[j] >> 5], w1 = L.bits[l1[j] >> 5];
      hit |= ((uint32_t)m0 & (w0 >> (l0[j] & 31))) << (2 * j);
      hit |= ((uint32_t)m1 & (w1 >> (l1[j] & 31))) << (2 * j + 1);
      reach |= (uint32_t)(live && in0) << (2 * j);
      reach |= (uint32_t)(live && in0 && in1) << (2 * j + 1);
    }
    // slots reached in order: the run of ones at the bottom of reach
    const uint32_t upto = ~reach & (reach + 1u);  // lowest slot not reached (0 if all are)
    const uint32_t found = hit & reach & (upto - 1u);
    if (found) {
      const uint32_t first = found & (0u - found);
      int lin = 0;
#pragma unroll
      for (int j = 0; j < kDescentBatch; ++j) {
        lin = (first >> (2 * j)) & 1u ? l0[j] : lin;
        lin = (first >> (2 * j + 1)) & 1u ? l1[j] : lin;
      }
      return L.index(lin);
    }
    if (upto != 0u) return -1;
    n += kDescentBatch;
  }
  return -1;
}

__global__ __launch_bounds__(256) void k_descent(const PlaneDesc* __restrict__ P,
                                                  const PlaneCounts* __restrict__ counts,
                                                  const uint32_t* __restrict__ ebits,
                                                  const int32_t* __restrict__ erank,
                                                  const int16_t* __restrict__ ex,
                                                  const int16_t* __restrict__ ey,
                                                  const int16_t* __restrict__ egx,
                                                  const int16_t* __restrict__ egy, int nmax,
                                                  int thr, int32_t* __restrict__ before,
                                                  int32_t* __restrict__ after) {
  const PlaneDesc d = P[blockIdx.y];
  const int n2 = 2 * counts[blockIdx.y].n_edges;
  const EdgeLookup L{ebits + (d.pix >> 5), erank + (d.pix >> 5), d.w, d.h};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += gridDim.x * blockDim.x) {
    const int i = k >> 1, s = d.pix + i;
    const int dir = (k & 1) ? 1 : -1;
    const int r = descent(L, d.shape_w, d.shape_h, ex[s], ey[s], (float)egx[s], (float)egy[s],
                          dir, nmax, thr);
    if (dir < 0) before[s] = r;
    else after[s] = r;
  }
}

// ---------------------------------------------------------------------------
// a9: vote chain walk (Vote.cpp:101-243).  Gradients are small integers, so
// the dot products are exact; distances are sqrt of exact integer sums.
__device__ __forceinline__ float neg_dot(const int16_t* gx, const int16_t* gy, int a, int b) {
  return -__fadd_rn(__fmul_rn((float)gx[a], (float)gx[b]), __fmul_rn((float)gy[a], (float)gy[b]));
}
__device__ __forceinline__ float dist2d(const int16_t* ex, const int16_t* ey, int a, int b) {
  const int ddx = ex[b] - ex[a], ddy = ey[b] - ey[a];
  return sqrt_rn(__fadd_rn((float)(ddx * ddx), (float)(ddy * ddy)));
}

__global__ __launch_bounds__(256) void k_vote(const PlaneDesc* __restrict__ P,
                                               const PlaneCounts* __restrict__ counts,
                                               const int16_t* __restrict__ ex,
                                               const int16_t* __restrict__ ey,
                                               const int16_t* __restrict__ egx,
                                               const int16_t* __restrict__ egy,
                                               const int32_t* __restrict__ before,
                                               const int32_t* __restrict__ after, int n_crowns,
                                               float ratio, float angle,
                                               int32_t* __restrict__ chosen,
                                               float* __restrict__ tdist,
                                               int32_t* __restrict__ vcount) {
  const PlaneDesc d = P[blockIdx.y];
  const int n = counts[blockIdx.y].n_edges;
  const int16_t* X = ex + d.pix;
  const int16_t* Y = ey + d.pix;
  const int16_t* GX = egx + d.pix;
  const int16_t* GY = egy + d.pix;
  const int32_t* B = before + d.pix;
  const int32_t* A = after + d.pix;
  for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < n; ip += gridDim.x * blockDim.x) {
    float vd[16];
    int nv = 0;
    int current = B[ip];
    int choosen = -1;
    float total = 0.0f;
    if (current >= 0 && neg_dot(GX, GY, ip, current) >= angle) {
      float dist = dist2d(X, Y, ip, current);
      vd[nv++] = dist;
      total = __fadd_rn(total, dist);
      for (int i = 1; i < n_crowns; ++i) {
        choosen = -1;
        int target = A[current];
        if (target < 0) break;
        if (!(neg_dot(GX, GY, target, current) >= angle)) break;
        dist = dist2d(X, Y, target, current);
        vd[nv++] = dist;
        total = __fadd_rn(total, dist);
        bool ok = true;
        for (int a = 0; a < nv; ++a)
          for (int b = a + 1; b < nv; ++b)
            ok = ok && (vd[a] <= __fmul_rn(vd[b], ratio)) && (vd[b] <= __fmul_rn(vd[a], ratio));
        if (!ok) break;
        current = target;
        target = B[current];
        if (target < 0) break;
        if (!(neg_dot(GX, GY, target, current) >= angle)) break;
        dist = dist2d(X, Y, target, current);
        vd[nv++] = dist;
        total = __fadd_rn(total, dist);
        for (int a = 0; a < nv; ++a)
          for (int b = a + 1; b < nv; ++b)
            ok = ok && (vd[a] <= __fmul_rn(vd[b], ratio)) && (vd[b] <= __fmul_rn(vd[a], ratio));
        if (!ok) break;
        current = target;
        choosen = current;
      }
    }
    chosen[d.pix + ip] = choosen;
    tdist[d.pix + ip] = total;
    if (choosen >= 0) atomicAdd(vcount + d.pix + choosen, 1);
  }
}

// ---------------------------------------------------------------------------
// a10: voter lists (Types.cpp:66-84), flow length (Vote.cpp:232), seeds
// (Vote.cpp:236-240).  CSR offsets by a per-plane scan of vote counts; lists
// filled with atomics, then each list sorted ascending (the CPU pushes voters
// in index order), flow length accumulated in that order, and the point's
// seed key (the voter that brought it to min votes) recorded.
// Two-pass scans over a plane's edge points: kScanBlocks blocks per plane,
// block b owning one contiguous range of whole 1024-point steps.  Pass one
// stores each block's sum, pass two adds the sums of the blocks before it
// and scans its range.
constexpr int kScanStep = 1024;  // threads of a scan block = points per step

__device__ __forceinline__ void scan_range(int n, int& lo, int& hi) {
  const int per = (n + kScanBlocks - 1) / kScanBlocks;
  const int chunk = (per + kScanStep - 1) / kScanStep * kScanStep;
  lo = min(n, (int)blockIdx.x * chunk);
  hi = min(n, lo + chunk);
}

// Block-wide inclusive scan of v (1024 threads); *total gets the block sum.
__device__ __forceinline__ int block_incl_scan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int t = __shfl_up(incl, s, 64);
    if (lane >= s) incl += t;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int wb = 0, tot = 0;
  for (int w = 0; w < kScanStep / 64; ++w) {
    if (w < wid) wb += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return wb + incl;
}

// Sum of the blocks before this one (pass two's offset).
__device__ __forceinline__ int prior_parts(const int32_t* part) {
  __shared__ int off;
  if (threadIdx.x < 64) {
    int v = (int)threadIdx.x < (int)blockIdx.x ? part[threadIdx.x] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x == 0) off = v;
  }
  __syncthreads();
  return off;
}

// a10: voter lists (Types.cpp:66-84), flow length (Vote.cpp:232), seeds
// (Vote.cpp:236-240).  CSR offsets by a scan of the vote counts; lists
// filled with atomics, then each list sorted ascending (the CPU pushes voters
// in index order), flow length accumulated in that order, and the point's
// seed key (the voter that brought it to min votes) recorded.
__global__ __launch_bounds__(kScanStep) void k_voters_sum(const PlaneDesc* __restrict__ P,
                                                          const PlaneCounts* __restrict__ counts,
                                                          const int32_t* __restrict__ vcount,
                                                          int32_t* __restrict__ part) {
  __shared__ int wsum[kScanStep / 64];
  const int p = blockIdx.y;
  const int32_t* a = vcount + P[p].pix;
  int lo, hi;
  scan_range(counts[p].n_edges, lo, hi);
  int acc = 0;
  for (int i = lo + threadIdx.x; i < hi; i += kScanStep) acc += a[i];
  int tot;
  block_incl_scan(acc, wsum, &tot);
  if (threadIdx.x == 0) part[p * kScanBlocks + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanStep) void k_voters_scan(const PlaneDesc* __restrict__ P,
                                                           PlaneCounts* __restrict__ counts,
                                                           const int32_t* __restrict__ vcount,
                                                           const int32_t* __restrict__ part,
                                                           int32_t* __restrict__ vindex,
                                                           int32_t* __restrict__ cursor) {
  __shared__ int wsum[kScanStep / 64];
  const int p = blockIdx.y;
  const PlaneDesc d = P[p];
  const int n = counts[p].n_edges;
  const int32_t* a = vcount + d.pix;
  int32_t* o = vindex + d.pix + p;
  int lo, hi;
  scan_range(n, lo, hi);
  int carry = prior_parts(part + p * kScanBlocks);
  for (int base = lo; base < hi; base += kScanStep) {
    const int i = base + threadIdx.x;
    const int v = i < hi ? a[i] : 0;
    int tot;
    const int incl = block_incl_scan(v, wsum, &tot);
    if (i < hi) {
      o[i] = carry + incl - v;
      cursor[d.pix + i] = carry + incl - v;
    }
    carry += tot;
  }
  if (blockIdx.x == kScanBlocks - 1 && threadIdx.x == 0) {
    o[n] = carry;
    counts[p].n_voters = carry;
  }
}

__global__ __launch_bounds__(256) void k_voters_fill(const PlaneDesc* __restrict__ P,
                                                      const PlaneCounts* __restrict__ counts,
                                                      const int32_t* __restrict__ chosen,
                                                      int32_t* __restrict__ cursor,
                                                      int32_t* __restrict__ vlist) {
  const PlaneDesc d = P[blockIdx.y];
  const int n = counts[blockIdx.y].n_edges;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const int c = chosen[d.pix + v];
    if (c >= 0) {
      const int pos = atomicAdd(cursor + d.pix + c, 1);
      vlist[d.pix + pos] = v;
    }
  }
}

// The lists of a block's 256 consecutive points are one contiguous run of
// the voter array.  When the run fits in LDS: every voter finds its list's
// owner and its rank there (voter indices are distinct), lands at its place,
// the sorted run goes back and the voters' distances are gathered in one
// parallel pass; each thread then folds its own flow length from LDS.
// Longer runs sort in place.
constexpr int kFinishBlock = 256;
constexpr int kFinishStage = 2048;  // voters staged per block

__global__ __launch_bounds__(kFinishBlock) void k_voters_finish(const PlaneDesc* __restrict__ P,
                                                                const PlaneCounts* __restrict__ counts,
                                                                const int32_t* __restrict__ vindex,
                                                                const float* __restrict__ tdist,
                                                                int min_votes,
                                                                int32_t* __restrict__ vlist,
                                                                float* __restrict__ flow,
                                                                int32_t* __restrict__ ismax,
                                                                int32_t* __restrict__ seedmark) {
  __shared__ int32_t stage_in[kFinishStage];
  __shared__ int32_t stage_out[kFinishStage];
  __shared__ float dist_s[kFinishStage];
  __shared__ uint8_t owner[kFinishStage];
  __shared__ int seg_b[kFinishBlock], seg_e[kFinishBlock];
  const int p = blockIdx.y;
  const PlaneDesc d = P[p];
  const int n = counts[p].n_edges;
  const int32_t* vi = vindex + d.pix + p;
  int32_t* vl = vlist + d.pix;
  const float* td = tdist + d.pix;
  const int t = threadIdx.x;
  for (int c0 = blockIdx.x * kFinishBlock; c0 < n; c0 += gridDim.x * kFinishBlock) {
    const int c = c0 + t;
    const int c_end = min(n, c0 + kFinishBlock);
    const int run_b = vi[c0], run_e = vi[c_end], len = run_e - run_b;
    const int b = c < n ? vi[c] : run_e, e = c < n ? vi[c + 1] : run_e;
    const bool staged = len <= kFinishStage;
    if (staged) {
      for (int q = t; q < len; q += kFinishBlock) stage_in[q] = vl[run_b + q];
      seg_b[t] = b - run_b;
      seg_e[t] = e - run_b;
      for (int i = b; i < e; ++i) owner[i - run_b] = (uint8_t)t;
      __syncthreads();
      for (int q = t; q < len; q += kFinishBlock) {
        const int o = owner[q], key = stage_in[q];
        const int sb = seg_b[o], se = seg_e[o];
        int rank = 0;
        for (int j = sb; j < se; ++j) rank += stage_in[j] < key;
        stage_out[sb + rank] = key;
      }
      __syncthreads();
      for (int q = t; q < len; q += kFinishBlock) {
        const int v = stage_out[q];
        vl[run_b + q] = v;
        dist_s[q] = td[v];
      }
      __syncthreads();
    } else {
      // insertion sort in place
      for (int i = b + 1; i < e; ++i) {
        const int key = vl[i];
        int j = i - 1;
        while (j >= b && vl[j] > key) {
          vl[j + 1] = vl[j];
          --j;
        }
        vl[j + 1] = key;
      }
    }
    if (c < n) {
      const int nv = e - b;
      float fl = 0.0f;
      if (staged) {
        for (int k = 0; k < nv; ++k)
          fl = __fdiv_rn(__fadd_rn(__fmul_rn(fl, (float)k), dist_s[b - run_b + k]), (float)(k + 1));
      } else {
        for (int k = 0; k < nv; ++k)
          fl = __fdiv_rn(__fadd_rn(__fmul_rn(fl, (float)k), td[vl[b + k]]), (float)(k + 1));
      }
      flow[d.pix + c] = fl;
      ismax[d.pix + c] = nv >= min_votes ? nv : -1;
      if (nv >= min_votes)
        seedmark[d.pix + (staged ? stage_out[b - run_b + min_votes - 1] : vl[b + min_votes - 1])] = c;
    }
    __syncthreads();
  }
}

// Seeds in the order they reached min votes = ascending key voter index.
__global__ __launch_bounds__(kScanStep) void k_seeds_sum(const PlaneDesc* __restrict__ P,
                                                         const PlaneCounts* __restrict__ counts,
                                                         const int32_t* __restrict__ seedmark,
                                                         int32_t* __restrict__ part) {
  __shared__ int wsum[kScanStep / 64];
  const int p = blockIdx.y;
  const int32_t* a = seedmark + P[p].pix;
  int lo, hi;
  scan_range(counts[p].n_edges, lo, hi);
  int acc = 0;
  for (int i = lo + threadIdx.x; i < hi; i += kScanStep) acc += a[i] >= 0;
  int tot;
  block_incl_scan(acc, wsum, &tot);
  if (threadIdx.x == 0) part[p * kScanBlocks + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanStep) void k_seeds(const PlaneDesc* __restrict__ P,
                                                     PlaneCounts* __restrict__ counts,
                                                     const int32_t* __restrict__ seedmark,
                                                     const int32_t* __restrict__ part,
                                                     int32_t* __restrict__ seeds) {
  __shared__ int wsum[kScanStep / 64];
  const int p = blockIdx.y;
  const PlaneDesc d = P[p];
  int lo, hi;
  scan_range(counts[p].n_edges, lo, hi);
  int carry = prior_parts(part + p * kScanBlocks);
  for (int base = lo; base < hi; base += kScanStep) {
    const int i = base + threadIdx.x;
    const int s = i < hi ? seedmark[d.pix + i] : -1;
    const int v = s >= 0;
    int tot;
    const int incl = block_incl_scan(v, wsum, &tot);
    if (v) seeds[d.pix + carry + incl - 1] = s;
    carry += tot;
  }
  if (blockIdx.x == kScanBlocks - 1 && threadIdx.x == 0) counts[p].n_seeds = carry;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch sequence.  No host synchronization inside: data-dependent sizes are
// read on the device, so the whole sequence can be captured in a hipGraph.
int launch_front_end(const FrontEndLaunch& L, hipStream_t s, StageTimer* timer) {
  const FrontEndBuffers& b = L.buf;
  const PlaneDesc* P = L.dev_planes;
  const int np = L.n_planes;
  auto mark = [&](const char* name, bool edge = false) {
    if (timer && (edge || timer->per_kernel)) timer->mark(name, s);
  };

  mark("start", true);
  if (L.n_tiles_p > 0) {
    if (!L.resample) {
      k_pyramid<<<L.n_tiles_p, 256, 0, s>>>(P, np, L.n_levels, b.img, b.quad);
    } else {
      k_pyramid<<<L.n_tiles_p, 256, 0, s>>>(P, np, 1, b.img, b.quad);
      for (int l = 1; l < L.n_levels; ++l)
        k_resample<<<L.n_tiles_r[l], 256, 0, s>>>(P, L.dev_level_planes + L.level_plane_start[l],
                                                   L.level_plane_count[l], b.img);
    }
    mark("pyramid");
  }
  {
    const int v = L.filter_variant;
    const int32_t* fs = L.dev_filter_starts[v];
    const int nt = L.n_tiles_fv[v];
    if (kFilterTileHs[v] == 16) k_filter<4><<<nt, 256, 0, s>>>(P, np, fs, b.img, b.dx, b.dy);
    else if (kFilterTileHs[v] == 64) k_filter<16><<<nt, 256, 0, s>>>(P, np, fs, b.img, b.dx, b.dy);
    else k_filter<8><<<nt, 256, 0, s>>>(P, np, fs, b.img, b.dx, b.dy);
  }
  mark("filter");
  k_nms_local<<<L.n_tiles_s, 256, 0, s>>>(P, np, b.dx, b.dy, L.prm.low, L.prm.high, b.cls,
                                          b.lab, b.root_strong);
  mark("nms_local");
  k_cc_border<<<L.n_tiles_s, 128, 0, s>>>(P, np, b.cls, b.lab);
  mark("cc_border");
  k_cc_flatten<<<L.n_tiles_s, 256, 0, s>>>(P, np, b.cls, b.lab, b.root_strong);
  mark("cc_flatten");
  k_thin<<<L.n_tiles_s, 256, 0, s>>>(P, np, b.cls, b.lab, b.root_strong, b.edges);
  mark("thin");
  k_compact_count<<<L.n_tiles_c, 256, 0, s>>>(P, np, b.edges, b.tcount, b.ebits);
  mark("compact_count");
  k_compact_scan<<<np, 1024, 0, s>>>(P, np, b.tcount, b.counts);
  mark("compact_scan");
  k_compact_write<<<L.n_tiles_c, 256, 0, s>>>(P, np, b.edges, b.dx, b.dy, b.tcount, b.erank,
                                               b.ex, b.ey, b.egx, b.egy, b.vcount, b.seedmark);
  mark("compact_write");
  const dim3 g(L.edge_blocks, np);
  k_descent<<<g, 256, 0, s>>>(P, b.counts, b.ebits, b.erank, b.ex, b.ey, b.egx, b.egy,
                              L.prm.dist_search, L.prm.thr_grad, b.before, b.after);
  mark("descent");
  k_vote<<<g, 256, 0, s>>>(P, b.counts, b.ex, b.ey, b.egx, b.egy, b.before, b.after,
                           L.prm.n_crowns, L.prm.ratio_voting, L.prm.angle_voting, b.chosen,
                           b.tdist, b.vcount);
  mark("vote");
  const dim3 gs(kScanBlocks, np);
  k_voters_sum<<<gs, kScanStep, 0, s>>>(P, b.counts, b.vcount, b.scan_part);
  k_voters_scan<<<gs, kScanStep, 0, s>>>(P, b.counts, b.vcount, b.scan_part, b.vindex, b.ismax);
  mark("voters_scan");
  k_voters_fill<<<g, 256, 0, s>>>(P, b.counts, b.chosen, b.ismax, b.vlist);
  mark("voters_fill");
  k_voters_finish<<<g, kFinishBlock, 0, s>>>(P, b.counts, b.vindex, b.tdist, L.prm.min_votes,
                                              b.vlist, b.flow, b.ismax, b.seedmark);
  mark("voters_finish");
  k_seeds_sum<<<gs, kScanStep, 0, s>>>(P, b.counts, b.seedmark, b.scan_part);
  k_seeds<<<gs, kScanStep, 0, s>>>(P, b.counts, b.seedmark, b.scan_part, b.seeds);
  mark("seeds", true);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace cctagx
