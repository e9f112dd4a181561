// gfx950 kernels of CCTag identification.
//
// k_select_cuts  identify_step_1: collectCuts + selectCutCheapUniform
//                (Identification.cpp:396-709, 1172-1301); one block per
//                marker, one thread per outer point.  Bit-exact with the
//                reference's serial float arithmetic.
// k_refine_*     refineConicFamilyGlob (Identification.cpp:813-1158): one
//                launch per grid-search iteration, a block per (grid point,
//                marker), then the final signals and the per-cut distances
//                to every bank profile of orazioDistanceRobust (:50-160).
//                The pairwise sum of squared differences is taken from the
//                cuts' Gram matrix in double instead of the reference's
//                sequential float sum; everything else (grid, homographies,
//                sampling, the sticky out-of-bounds flag of a cut) follows
//                the reference's float operations.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ident_device.hpp"

namespace cctagx {
namespace {

__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }

// A pointer the compiler knows to be in device global memory (the loads
// become global rather than flat ones).
typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;

// The frame as quads: word (x, y) holds pixels (x, y), (x+1, y), (x, y+1),
// (x+1, y+1) in bytes 0..3, so a bilinear sample is one aligned load.
struct Img {
  const uint32_t* q;
  int w, h;
  __device__ __forceinline__ uint32_t quad(int x, int y) const { return ((gptr_u32)q)[x + y * w]; }
};

// Identification.hpp:197-222: half of the bilinear value at (x, y), from
// the quad of pixel (floor x, floor y).
__device__ __forceinline__ float bilinear_of(float x, float y, uint32_t quad) {
  const int px = (int)x, py = (int)y;
  const float fx = x - px, fy = y - py;
  const float fx1 = 1.0f - fx, fy1 = 1.0f - fy;
  const float w1 = fx1 * fy1, w2 = fx * fy1, w3 = fx1 * fy, w4 = fx * fy;
  const float a = (float)(quad & 255u), b = (float)((quad >> 8) & 255u),
              c = (float)((quad >> 16) & 255u), d = (float)(quad >> 24);
  return (a * w1 + b * w2 + c * w3 + d * w4) / 2;
}

__device__ __forceinline__ void apply_h(float& xr, float& yr, const float* H, float x, float y) {
  const float u = H[0] * x + H[1] * y + H[2];
  const float v = H[3] * x + H[4] * y + H[5];
  const float ww = H[6] * x + H[7] * y + H[8];
  xr = u / ww;
  yr = v / ww;
}

__device__ __forceinline__ float cof(const float* m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[i1 * 3 + j1] * m[i2 * 3 + j2] - m[i1 * 3 + j2] * m[i2 * 3 + j1];
}

// Adjugate inverse, same operation order as the host's inverse3.
__device__ void inverse3(const float* m, float* out) {
  const float c0 = cof(m, 0, 0), c1 = cof(m, 1, 0), c2 = cof(m, 2, 0);
  const float det = c0 * m[0] + c1 * m[3] + c2 * m[6];
  const float invdet = 1.0f / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i * 3 + j] = cof(m, j, i) * invdet;
}

__device__ __forceinline__ void mul3(const float* a, const float* b, float* r) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
}

// optimization/conditioner.hpp:66-72.
__device__ __forceinline__ void condition(float& x, float& y, const float* T) {
  const float c0 = T[0] * x + T[1] * y + T[2] * 1.f;
  const float c1 = T[3] * x + T[4] * y + T[5] * 1.f;
  const float c2 = T[6] * x + T[7] * y + T[8] * 1.f;
  x = c0 / c2;
  y = c1 / c2;
}

// boost::accumulators immediate mean and variance, one sample at a time.
// x / n, correctly rounded, for a float x and an integer 1 <= n <=
// kMaxSamples, from inv_n = 1/n rounded to double: the double product is
// within 2^-52 (relative) of x/n; x/n, a 24-bit significand over an
// integer below 2^8, is never a midpoint between two floats and lies more
// than 2^-33 (relative) away from every midpoint, so rounding the product to
// float rounds x/n.  The argument needs a result in the normal range: for a
// subnormal one (div_n_tiny) the caller takes the division.
__device__ __forceinline__ double div_n_product(float x, double inv_n) { return (double)x * inv_n; }
__device__ __forceinline__ bool div_n_tiny(double p) { return p != 0.0 && fabs(p) < 0x1p-126; }

// 1/n in double for n = 1 .. kMaxSamples (entry 0 unused), by the block.
__device__ __forceinline__ void fill_inv(double* inv) {
  for (int q = threadIdx.x; q <= kMaxSamples; q += blockDim.x) inv[q] = 1.0 / (double)q;
}

// Running mean and variance, one value at a time (at most kMaxSamples).
// The reciprocals a step needs are read from fill_inv's table one step
// ahead, so the table's latency stays off the chain of products.  With
// EXACT false a product in the subnormal range (where it need not round to
// the quotient) is only recorded in `tiny`, keeping branches off the chain:
// the caller then runs the same values through MeanVar<true>, which divides
// there, so the result is always the division's.
template <bool EXACT>
struct MeanVar {
  float mean = 0.f, var = 0.f;
  unsigned cnt = 0;
  bool tiny = false;
  const double* inv;
  double inv_cnt, inv_prev;  // 1/(cnt + 1) and 1/cnt for the next add
  __device__ explicit MeanVar(const double* table) : inv(table), inv_cnt(table[1]), inv_prev(0.0) {}
  __device__ __forceinline__ void add(float v) {
    ++cnt;
    const double inv_next = inv[min(cnt + 1, (unsigned)kMaxSamples)];
    const float before = (float)(cnt - 1);
    const float xm = mean * before + v;
    const double pm = div_n_product(xm, inv_cnt);
    mean = (float)pm;
    if (EXACT) {
      if (div_n_tiny(pm)) mean = xm / (float)cnt;
    } else {
      tiny |= div_n_tiny(pm);
    }
    if (cnt > 1) {
      const float t = v - mean;
      const float xa = var * before, xb = t * t;
      const double pa = div_n_product(xa, inv_cnt), pb = div_n_product(xb, inv_prev);
      float a = (float)pa, b = (float)pb;
      if (EXACT) {
        if (div_n_tiny(pa)) a = xa / (float)cnt;
        if (div_n_tiny(pb)) b = xb / before;
      } else {
        tiny |= div_n_tiny(pa) | div_n_tiny(pb);
      }
      var = a + b;
    }
    inv_prev = inv_cnt;
    inv_cnt = inv_next;
  }
};

// Identification.cpp:396-459 for a cut whose signal is only consumed
// sample by sample; returns false when a sample falls outside the image
// (the caller then discards what f saw).  Sample positions follow the
// reference's running sums; the image reads of kWalkBatch samples are issued
// together so their latencies overlap (a sample past the end re-reads the
// batch's first one, which is inside).
constexpr int kWalkBatch = 20;

template <class F>
__device__ bool cut_walk(const Img& im, float sx, float sy, float ex, float ey, float begin, float end,
                         int n, F&& f) {
  float xs, ys, xe, ye;
  const float diffX = ex - sx, diffY = ey - sy;
  if (begin != 0.f) {
    xs = sx + diffX * begin;
    ys = sy + diffY * begin;
  } else {
    xs = sx;
    ys = sy;
  }
  if (end != 1.f) {
    xe = sx + diffX * end;
    ye = sy + diffY * end;
  } else {
    xe = ex;
    ye = ey;
  }
  const float stepX = (xe - xs) / ((float)n - 1.f), stepY = (ye - ys) / ((float)n - 1.f);
  const float xmax = (float)(im.w - 1), ymax = (float)(im.h - 1);
  float x = xs, y = ys;
  for (int i0 = 0; i0 < n; i0 += kWalkBatch) {
    float px[kWalkBatch], py[kWalkBatch];
    bool inside = true;
#pragma unroll
    for (int t = 0; t < kWalkBatch; ++t) {
      px[t] = x;
      py[t] = y;
      if (i0 + t < n) inside &= (x >= 1.f && x < xmax && y >= 1.f && y < ymax);
      x += stepX;
      y += stepY;
    }
    if (!inside) return false;
    uint32_t qd[kWalkBatch];
#pragma unroll
    for (int t = 0; t < kWalkBatch; ++t) {
      const bool live = i0 + t < n;
      qd[t] = im.quad((int)(live ? px[t] : px[0]), (int)(live ? py[t] : py[0]));
    }
#pragma unroll
    for (int t = 0; t < kWalkBatch; ++t)
      if (i0 + t < n) f(i0 + t, bilinear_of(px[t], py[t], qd[t]));
  }
  return true;
}

__constant__ float kRefineKernels[3][9] = {
    {-0.0000f, -0.0003f, -0.1065f, -0.7863f, .0f, 0.7863f, 0.1065f, 0.0003f, 0.0000f},
    {-0.0044f, -0.0540f, -0.2376f, -0.3450f, .0f, 0.3450f, 0.2376f, 0.0540f, 0.0044f},
    {-0.0366f, -0.1113f, -0.1801f, -0.1594f, .0f, 0.1594f, 0.1801f, 0.1113f, 0.0366f}};

// One thread per outer point (n <= kMaxOuterPoints = blockDim.x).  The
// counts, the maximum variance and the candidates' ranks come from block
// reductions and a ballot scan; the uniform-stride pick of
// selectCutCheapUniform (Identification.cpp:546-601) is one lane per
// selected cut.
constexpr int kSelectBlock = 256;
static_assert(kSelectBlock >= kMaxOuterPoints, "one lane per outer point");

__global__ __launch_bounds__(kSelectBlock) void k_select_cuts(const SelectJobDev* __restrict__ jobs,
                                                              const float4* __restrict__ pts,
                                                              float start_sig, int n_samples, int n_ref,
                                                              int select_max, SelectOutDev* __restrict__ out) {
  constexpr int NW = kSelectBlock / 64;
  __shared__ float4 stop[kMaxOuterPoints];
  __shared__ int idx_of_rank[kMaxOuterPoints];
  __shared__ int w_cnt[NW], w_cand[NW];
  __shared__ float w_max[NW];
  __shared__ float ref_sig[kMaxRefineSamples][kSelectBlock];  // the thread's refinement cut
  __shared__ double inv[kMaxSamples + 1];
  fill_inv(inv);
  __syncthreads();
  const SelectJobDev J = jobs[blockIdx.x];
  const Img im{J.quad, J.iw, J.ih};
  const int n = J.n_outer;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  bool valid = false;
  float var = 0.f;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  if (t < n) {
    s = pts[J.pt_offset + t];
    MeanVar<false> mv(inv);
    valid = cut_walk(im, J.cx, J.cy, s.x, s.y, start_sig, 1.f, n_samples,
                     [&](int, float v) { mv.add(v); });
    var = mv.var;
    if (valid && mv.tiny) {  // the same walk again, dividing where a product was subnormal
      MeanVar<true> me(inv);
      cut_walk(im, J.cx, J.cy, s.x, s.y, start_sig, 1.f, n_samples,
               [&](int, float v) { me.add(v); });
      var = me.var;
    }
  }
  // number of collected cuts and their largest variance
  {
    int c = valid ? 1 : 0;
    float mx = valid ? var : -FLT_MAX;
    for (int o = 32; o > 0; o >>= 1) {
      c += __shfl_xor(c, o, 64);
      mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
      w_cnt[wid] = c;
      w_max[wid] = mx;
    }
  }
  __syncthreads();
  int nc = 0;
  float var_max = -FLT_MAX;
  for (int w = 0; w < NW; ++w) {
    nc += w_cnt[w];
    var_max = fmaxf(var_max, w_max[w]);
  }
  bool cand = false;
  if (valid) {
    // pointOnEllipse (EllipseFromPoints.cpp:29-49), double like the C cos/sin
    {
      const float x = s.x - J.cx, y = s.y - J.cy;
      float tx = (float)((double)x * J.ca + (double)y * J.sa);
      float ty = (float)(-(double)x * J.sa + (double)y * J.ca);
      const float cs = (float)sqrt((double)(tx * tx / (J.ea * J.ea) + ty * ty / (J.eb * J.eb)));
      tx /= cs;
      ty /= cs;
      s.x = (float)((double)tx * J.ca - (double)ty * J.sa + (double)J.cx);
      s.y = (float)((double)tx * J.sa + (double)ty * J.ca + (double)J.cy);
    }
    // outerEdgeRefinement (Identification.cpp:604-657)
    const float len = (float)(3.0 * 1.4142135623730951 * (double)J.scale);
    const float half = len / 2.f;
    const float gn = sqrt_rn(s.z * s.z + s.w * s.w);
    const float gdx = s.z / gn, gdy = s.w / gn;
    const float hx = half * gdx, hy = half * gdy;
    const float psx = s.x - hx, psy = s.y - hy;
    const float pex = s.x + half * gdx, pey = s.y + half * gdy;
    float (*sig)[kSelectBlock] = ref_sig;
    const bool ok =
        cut_walk(im, psx, psy, pex, pey, 0.f, 1.f, n_ref, [&](int i, float v) { sig[i][t] = v; });
    if (ok) {
      float key[3], loc[3];
      for (int k = 0; k < 3; ++k) {
        float best = 0.f;
        int bi = 0;
        for (int i = 0; i < n_ref; ++i) {
          float tmp = 0;
          for (int j = 0; j < 9; ++j) {
            const int r = i - 4 + j;
            const float v = r < 0 ? sig[0][t] : (r >= n_ref ? sig[n_ref - 1][t] : sig[r][t]);
            tmp += v * kRefineKernels[k][j];
          }
          if (i == 0 || tmp > best) {
            best = tmp;
            bi = i;
          }
        }
        key[k] = best;
        loc[k] = (float)bi;
      }
      // std::map<float,float> insert (first of equal keys kept), then the largest key
      int b = 0;
      for (int k = 1; k < 3; ++k)
        if (key[k] > key[b]) b = k;
      const float step = len / ((float)n_ref - 1.f);
      s.x = psx + step * loc[b] * gdx;
      s.y = psy + step * loc[b] * gdy;
      cand = var / var_max > 0.5f;
    }
  }
  // rank of each candidate among the collected cuts, in outer-point order
  const uint64_t ball = __ballot(cand);
  const int in_wave = __popcll(ball & ((1ull << lane) - 1ull));
  if (lane == 0) w_cand[wid] = __popcll(ball);
  if (t < n) stop[t] = s;
  __syncthreads();
  int before = 0, ni = 0;
  for (int w = 0; w < NW; ++w) {
    if (w < wid) before += w_cand[w];
    ni += w_cand[w];
  }
  if (cand) idx_of_rank[before + in_wave] = t;
  __syncthreads();
  // uniform stride over the candidates: pick k takes rank (int)(k * step)
  const int select = min(select_max, nc);
  const float step = fmaxf(1.f, (float)ni / (float)select);
  if (t < select) {
    const int at = (int)((float)t * step);
    if (at < ni) out[blockIdx.x].stops[t] = stop[idx_of_rank[at]];
  }
  // (out is host memory: written after the last barrier, which would
  // otherwise wait for the stores to reach the host)
  if (t == 0) {
    int nsel = 0;
    while (nsel < select && (int)((float)nsel * step) < ni) ++nsel;
    out[blockIdx.x].n_collected = nc;
    out[blockIdx.x].n_selected = nsel;
  }
}

// ---------------------------------------------------------------------------
// Centre refinement, one launch per grid-search iteration.  Block (g, m)
// of k_refine_iter evaluates grid point g of marker m: its homography, the
// signals of every selected cut (LDS), their out-of-bounds flags and the
// residual sum_{k<l} sum_i (s_ki - s_li)^2 over its valid cuts.  The
// selection that follows an iteration -- the sticky out-of-bounds flags in
// grid-point order, each grid point's residual over the cuts valid there,
// the move to the first minimum as the reference's loop does -- is redone
// by every block of the next launch (and by the final kernel), so an
// iteration costs one launch.
//
// A cut flagged at an earlier grid point of the same iteration is invalid
// at the later ones, which a block cannot see while it runs.  So block g
// first tests, for the grid points before it, whether each valid cut surely
// stays inside the image (cut_surely_inside).  When every one does -- the
// usual case -- the block knows its valid cuts and takes the residual as
// sum_i (m S2_i - S1_i^2), S1_i and S2_i the sum and the sum of squares of
// the valid cuts' samples (double, in cut order).  Otherwise it stores the
// Gram matrix C_kl = sum_i s_ki s_li in double, from which the selection
// recomputes the residual for any valid set as
// (m-1) sum_k C_kk - 2 sum_{k<l} C_kl.
constexpr int kRefineBlock = 256;
constexpr int kFinalBlock = 1024;  // one block per marker: every loop of the read-out in one pass
constexpr int kRefineChunk = 10;  // samples per sampling work item
constexpr int kLoadBatch = 10;    // samples whose image reads are issued together
static_assert(kRefineChunk % kLoadBatch == 0, "whole load batches per chunk");
constexpr int G2MAX = kMaxGrid * kMaxGrid;
constexpr int kFuseMaxBlocks = 2048;  // grid blocks up to which the selection is fused
constexpr int kXcds = 8;               // accelerator complex dies of the MI355X

// getNearbyPoints point g + computeHomographyFromEllipseAndImagedCenter
// (Identification.cpp:739-798, 1071-1109) and the inverse.
__device__ void grid_homography(const RefineJobDev& J, float cur_x, float cur_y, float neighbour,
                                int g, int grid_n, float* Hf, float* Hi, float& qx, float& qy) {
  const int i = g / grid_n, j = g - i * grid_n;
  const float ns = neighbour * J.grid_scale;
  float cx = cur_x, cy = cur_y;
  condition(cx, cy, J.cond);
  const float halfWidth = ns / 2.f;
  const float stepSize = ns / (float)(grid_n - 1);
  qx = cx - halfWidth + (float)i * stepSize;
  qy = cy - halfWidth + (float)j * stepSize;
  condition(qx, qy, J.cond_inv);
  float xc, yc;
  apply_h(xc, yc, J.t_can, qx, qy);
  const float Q11 = J.Q11, Q22 = J.Q22, Q33 = J.Q33;
  float H[9];
  H[0] = Q33;              H[1] = Q22 * xc * yc;        H[2] = -Q33 * xc;
  H[3] = 0.f;              H[4] = -Q11 * xc * xc - Q33; H[5] = -Q33 * yc;
  H[6] = -Q11 * xc;        H[7] = Q22 * yc;             H[8] = -Q33;
  const float d0 = sqrt_rn((Q22 * Q33 / Q11 * (Q11 * xc * xc + Q22 * yc * yc + Q33)));
  const float d1 = Q33;
  const float d2 = sqrt_rn(-Q22 * (Q11 * xc * xc + Q33));
  for (int r = 0; r < 3; ++r) {
    H[r * 3 + 0] *= d0;
    H[r * 3 + 1] *= d1;
    H[r * 3 + 2] *= d2;
  }
  mul3(J.t_inv_can, H, Hf);
  inverse3(Hf, Hi);
}

// extractSignalUsingHomography (Identification.cpp:216-280), first part,
// one thread per cut: the stop back-projected by Hi, the sample positions
// from begin*stop to end*stop in the marker plane as the reference's running
// sums, kept at the start of every kRefineChunk-sample chunk (the chunks'
// threads then resume the sums from there).
constexpr int kMaxChunks = (kMaxSamples + kRefineChunk - 1) / kRefineChunk;
struct CutWalk {
  float2 start[kMaxChunks];
  float2 step;
};

__device__ void cut_walk_starts(const float* Hi, float2 stop, float begin_sig, float end_sig,
                                int n_samples, CutWalk& w) {
  float bx, by;
  apply_h(bx, by, Hi, stop.x, stop.y);
  float xs, ys, xe, ye;
  if (begin_sig != 0.f) {
    xs = bx * begin_sig;
    ys = by * begin_sig;
  } else {
    xs = 0;
    ys = 0;
  }
  if (end_sig != 1.f) {
    xe = bx * end_sig;
    ye = by * end_sig;
  } else {
    xe = bx;
    ye = by;
  }
  const float stepX = (xe - xs) / ((float)n_samples - 1.f), stepY = (ye - ys) / ((float)n_samples - 1.f);
  float x = xs, y = ys;
  const int nchunk = (n_samples + kRefineChunk - 1) / kRefineChunk;
  for (int c = 0; c < nchunk; ++c) {
    w.start[c] = make_float2(x, y);
    for (int t = 0; t < kRefineChunk; ++t) {
      x += stepX;
      y += stepY;
    }
  }
  w.step = make_float2(stepX, stepY);
}

// Whether every sample of a cut surely falls inside the image at the grid
// point whose homography is H (inverse Hi).  The samples lie, up to the
// rounding of their running sums and of the mapping (far below a pixel), on
// the image of the marker-plane segment t * b, t in [begin, end], b = the
// stop back-projected (homogeneous (b0, b1, b2) = Hi (stop, 1)).  When the
// projective denominator keeps one sign along t in [0, 1], H maps that
// segment to a straight one ending at the stop; H maps the unit disk onto
// the outer ellipse's interior, so with begin * b inside the disk the inner
// end lies in the ellipse.  Stop and ellipse box 2 px inside the image
// (box_inside) then keep every sample inside [0, w-1) x [0, h-1).
__device__ __forceinline__ bool cut_surely_inside(const float* H, const float* Hi, float2 stop,
                                                  float begin_sig, float end_sig, int w, int h) {
  if (!(end_sig <= 1.f && stop.x >= 2.f && stop.x <= (float)w - 3.f && stop.y >= 2.f &&
        stop.y <= (float)h - 3.f))
    return false;
  const float b0 = Hi[0] * stop.x + Hi[1] * stop.y + Hi[2];
  const float b1 = Hi[3] * stop.x + Hi[4] * stop.y + Hi[5];
  const float b2 = Hi[6] * stop.x + Hi[7] * stop.y + Hi[8];
  const float e0 = H[8] * b2, e1 = H[6] * b0 + H[7] * b1 + H[8] * b2;
  const float scale = fabsf(H[6] * b0) + fabsf(H[7] * b1) + fabsf(H[8] * b2);
  const bool keeps_sign = (e0 > 0.f) == (e1 > 0.f) && fabsf(e0) > 1e-3f * scale &&
                          fabsf(e1) > 1e-3f * scale;
  return keeps_sign && begin_sig * begin_sig * (b0 * b0 + b1 * b1) <= 0.98f * (b2 * b2);
}

// Second part for samples [c*kRefineChunk, ...) of a cut: mapped by H, half
// bilinear value; 0 and the flag when out of the image.
__device__ bool sample_chunk(const Img& im, const float* H, const CutWalk& w, int n_samples, int c,
                             float* out) {
  float x = w.start[c].x, y = w.start[c].y;
  const float stepX = w.step.x, stepY = w.step.y;
  const int i0 = c * kRefineChunk;
  const float xmax = (float)(im.w - 1), ymax = (float)(im.h - 1);
  bool oob = false;
  // per group of kLoadBatch samples: positions first, then every corner load
  // (an outside sample reads pixel (0, 0), so the loads need no branch and
  // are in flight together), then the values
#pragma unroll
  for (int t0 = 0; t0 < kRefineChunk; t0 += kLoadBatch) {
    float xr[kLoadBatch], yr[kLoadBatch];
    bool in[kLoadBatch];
#pragma unroll
    for (int t = 0; t < kLoadBatch; ++t) {
      apply_h(xr[t], yr[t], H, x, y);
      const bool live = i0 + t0 + t < n_samples;
      in[t] = live && xr[t] >= 0.f && xr[t] < xmax && yr[t] >= 0.f && yr[t] < ymax;
      oob |= live && !in[t];
      x += stepX;
      y += stepY;
    }
    uint32_t qd[kLoadBatch];
#pragma unroll
    for (int t = 0; t < kLoadBatch; ++t) qd[t] = in[t] ? im.quad((int)xr[t], (int)yr[t]) : im.quad(0, 0);
#pragma unroll
    for (int t = 0; t < kLoadBatch; ++t) {
      if (i0 + t0 + t >= n_samples) continue;
      out[i0 + t0 + t] = in[t] ? bilinear_of(xr[t], yr[t], qd[t]) : 0.f;
    }
  }
  return oob;
}

__device__ __forceinline__ int pair_offset(int k, int nc) { return k * nc - k * (k - 1) / 2; }

// Row k's share of sum_{k<l valid} sum_i (s_ki - s_li)^2 from the Gram
// matrix C (packed upper triangle): (m-1) C_kk - 2 sum_{l>k valid} C_kl,
// valid cuts being the set bits of `valid`.  The row's entries are read
// kRowBatch at a time (indices clamped into the row, so the reads need no
// branch and are in flight together) and subtracted in column order,
// skipping invalid columns as the plain loop does.
constexpr int kRowBatch = 8;
__device__ __forceinline__ double gram_row(const double* C, uint32_t valid, int k, int nc, int m) {
  if (!((valid >> k) & 1u)) return 0.0;
  const double* row = C + pair_offset(k, nc);
  const int last = nc - 1 - k;  // row[0 .. last]
  double acc = (double)(m - 1) * row[0];
  for (int d0 = 1; d0 <= last; d0 += kRowBatch) {
    double v[kRowBatch];
#pragma unroll
    for (int j = 0; j < kRowBatch; ++j) v[j] = row[min(d0 + j, last)];
#pragma unroll
    for (int j = 0; j < kRowBatch; ++j)
      if (d0 + j <= last && ((valid >> (k + d0 + j)) & 1u)) acc -= 2.0 * v[j];
  }
  return acc;
}

// Sum over the 64 lanes of a wavefront in a fixed tree order (lane 0 holds
// the result): identical wherever it is evaluated.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t cut_bits(int nc) { return nc >= 32 ? ~0u : (1u << nc) - 1u; }

// Selection step of an iteration (what the reference's loop does after
// evaluating its grid), by wavefront 0 with lane g on grid point g: the
// sticky out-of-bounds flags in grid-point order (a prefix OR over the
// lanes), each point's residual (its block's own when the valid cuts are the
// ones it assumed, otherwise recomputed from its Gram matrix with the same
// row sums and reduction order), the first minimum, the next centre.  Called
// by a whole block (one barrier); every block that calls it with the same
// inputs gets the same state.
__device__ void pick_state(const RefineJobDev& J, int grid_n, const RefineStateDev& prev,
                           const RefineGridDev* GM, const double* gram_m, RefineStateDev& S) {
  const int tid = threadIdx.x, nc = J.n_cuts, G2 = grid_n * grid_n;
  if (tid < 64) {
    const uint32_t cuts = cut_bits(nc);
    const bool on = tid < G2;
    const uint32_t sp = prev.sticky;
    uint32_t own = 0;
    double acc = 0.0;
    int ml = 0, hg = 0;
    float gx = 0.f, gy = 0.f;
    if (on) {
      const RefineGridDev& G = GM[tid];
      own = G.oob;
      acc = G.cost_local;
      ml = G.m_local;
      gx = G.gx;
      gy = G.gy;
      hg = G.has_gram;
    }
    uint32_t run = own;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(run, o, 64);
      if (tid >= o) run |= v;
    }
    run |= sp;
    const uint32_t valid = ~run & cuts;
    const int mv = __popc(valid);
    const bool same = valid == (~(sp | own) & cuts) && mv == ml;
    // points whose valid cuts differ from the ones their block assumed, one
    // at a time over the whole wavefront, as the block itself sums.  Such a
    // point's block found a flag possible at an earlier point and stored its
    // Gram matrix; were one missing, the marker fails rather than keep a
    // residual over the wrong cuts (cut_surely_inside rules that out).
    const bool lost = __ballot(on && !same && !hg) != 0;
    for (unsigned long long redo = __ballot(on && !same); redo; redo &= redo - 1) {
      const int gg = __builtin_ctzll(redo);
      const uint32_t vg = __shfl(valid, gg, 64);
      const int mg = __shfl(mv, gg, 64);
      const double part = tid < nc ? gram_row(gram_m + (size_t)gg * kMaxPairs, vg, tid, nc, mg) : 0.0;
      const double sum = __shfl(wave_sum(part), 0, 64);
      if (tid == gg) acc = sum;
    }
    const float res = mv < 2 ? FLT_MAX : (float)acc / (float)(mv * (mv - 1) / 2);
    // the first minimum among the points with two valid cuts and a residual
    // below FLT_MAX: lexicographic (residual, index) minimum over the lanes
    const bool pickable = on && mv >= 2 && res < FLT_MAX;
    float key = pickable ? res : INFINITY;
    int idx = pickable ? tid : 64;
    for (int o = 32; o > 0; o >>= 1) {
      const float k2 = __shfl_xor(key, o, 64);
      const int i2 = __shfl_xor(idx, o, 64);
      if (k2 < key || (k2 == key && i2 < idx)) {
        key = k2;
        idx = i2;
      }
    }
    const uint32_t sticky = __shfl(run, G2 - 1, 64);
    const int b = idx < 64 && !lost ? idx : -1;
    const float bx = __shfl(gx, b < 0 ? 0 : b, 64), by = __shfl(gy, b < 0 ? 0 : b, 64);
    if (tid == 0) {
      S.sticky = sticky;
      S.neighbour = prev.neighbour;
      if (b < 0) {
        S.ok = 0;
        S.best_g = prev.best_g;
        S.cx = prev.cx;
        S.cy = prev.cy;
        S.min_res = prev.min_res;
      } else {
        S.ok = 1;
        S.best_g = b;
        S.cx = bx;
        S.cy = by;
        S.neighbour /= (float)((grid_n - 1) / 2);
        S.min_res = key;
      }
    }
  }
  __syncthreads();
}

// The state a marker starts iteration `it` with: the job's start for the
// first one, otherwise the selection over the previous iteration's grid
// (buffers of parity (it - 1) & 1).
__device__ void state_at(const RefineJobDev& J, int it, int m, int n_jobs, int grid_n,
                         const RefineWork& W, RefineStateDev& S) {
  const int tid = threadIdx.x;
  if (it == 0) {
    if (tid == 0) {
      S.cx = J.c0x;
      S.cy = J.c0y;
      S.neighbour = J.neighbour0;
      S.min_res = FLT_MAX;
      S.ok = 1;
      S.best_g = 0;
      S.sticky = 0;
    }
    __syncthreads();
    return;
  }
  const int pb = (it - 1) & 1;
  const RefineStateDev& prev = W.state[(size_t)pb * n_jobs + m];
  if (!prev.ok) {
    if (tid == 0) S = prev;
    __syncthreads();
    return;
  }
  pick_state(J, grid_n, prev, W.grid + ((size_t)pb * n_jobs + m) * G2MAX,
             W.gram + ((size_t)pb * n_jobs + m) * G2MAX * kMaxPairs, S);
}

// Iteration `it` of the grid search: block (g, m) takes marker m's state --
// with `fused`, derived from the previous iteration by the block itself
// (every block of the marker derives the same one; block 0 stores it),
// otherwise as k_refine_pick left it -- then evaluates grid point g:
// homography, cut signals in LDS, out-of-bounds flags, Gram matrix and the
// residual under its own flags.  Buffers alternate with the iteration's
// parity.  Fusing saves a launch per iteration when few markers leave the
// chip idle; with many markers the repeated selection costs more than the
// launch.
__global__ __launch_bounds__(kRefineBlock) void k_refine_iter(
    const RefineJobDev* __restrict__ jobs, int it, int n_samples, int grid_n, int fused,
    RefineWork W) {
  __shared__ float gH[G2MAX][9], gHi[G2MAX][9];
  __shared__ float sig[kMaxCuts][kMaxSamples + 1];
  __shared__ int oobk[kMaxCuts];
  __shared__ double sC[kMaxPairs];  // the Gram matrix, or the per-sample terms
  __shared__ CutWalk walk[kMaxCuts];
  __shared__ RefineStateDev S;
  __shared__ int maybe_out;
  __shared__ uint32_t own_oob;
  static_assert(kMaxPairs >= kMaxSamples, "sC holds the per-sample terms");

  // Workgroups are dealt to the 8 XCDs in turn, each XCD with its own L2:
  // renumber them so that an XCD runs consecutive (grid point, marker)
  // pairs, i.e. whole markers, and each marker's image area is fetched into
  // one L2 only.
  const int n_jobs = gridDim.y, tid = threadIdx.x;
  int g, m;
  {
    const int n = gridDim.x * gridDim.y, hw = blockIdx.y * gridDim.x + blockIdx.x;
    const int x = hw % kXcds, k = hw / kXcds, q = n / kXcds, r = n % kXcds;
    const int logical = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
    g = logical % gridDim.x;
    m = logical / gridDim.x;
  }
  const RefineJobDev& J = jobs[m];
  if (it >= J.n_iter) return;
  const int cb = it & 1;
  if (fused) {
    state_at(J, it, m, n_jobs, grid_n, W, S);
    if (g == 0 && tid == 0) W.state[(size_t)cb * n_jobs + m] = S;
  } else {
    if (tid == 0) S = W.state[(size_t)cb * n_jobs + m];
    __syncthreads();
  }
  if (!S.ok) return;
  const int nc = J.n_cuts;
  const Img im{J.quad, J.iw, J.ih};
  RefineGridDev& GD = W.grid[((size_t)cb * n_jobs + m) * G2MAX + g];
  // the homographies of grid points 0 .. g, lane p on point p (the block
  // stores its own for the selection and the read-out)
  if (tid <= g) {
    float Hf[9], Hi[9], qx, qy;
    grid_homography(J, S.cx, S.cy, S.neighbour, tid, grid_n, Hf, Hi, qx, qy);
    for (int r = 0; r < 9; ++r) {
      gH[tid][r] = Hf[r];
      gHi[tid][r] = Hi[r];
    }
    if (tid == g) {
      for (int r = 0; r < 9; ++r) {
        GD.H[r] = Hf[r];
        GD.Hi[r] = Hi[r];
      }
      GD.gx = qx;
      GD.gy = qy;
    }
  }
  if (tid < kMaxCuts) oobk[tid] = 0;
  if (tid == 0) maybe_out = g > 0 && !J.box_inside;
  __syncthreads();
  const float* sH = gH[g];
  if (tid < nc) cut_walk_starts(gHi[g], J.stops[tid], J.begin_sig, J.end_sig, n_samples, walk[tid]);
  // may a valid cut leave the image at a grid point before this one?
  if (J.box_inside) {
    for (int p = tid; p < g * nc; p += kRefineBlock) {
      const int gp = p / nc, k = p - gp * nc;
      if (!((S.sticky >> k) & 1u) &&
          !cut_surely_inside(gH[gp], gHi[gp], J.stops[k], J.begin_sig, J.end_sig, J.iw, J.ih))
        maybe_out = 1;
    }
  }
  __syncthreads();
  const int nchunk = (n_samples + kRefineChunk - 1) / kRefineChunk;
  for (int w = tid; w < nc * nchunk; w += kRefineBlock) {
    const int k = w / nchunk, c = w - k * nchunk;
    if (sample_chunk(im, sH, walk[k], n_samples, c, sig[k])) atomicOr(&oobk[k], 1);
  }
  __syncthreads();
  if (tid < 64) {
    const uint32_t o = (uint32_t)__ballot(tid < nc && oobk[tid] != 0);
    if (tid == 0) own_oob = o;
  }
  __syncthreads();
  const uint32_t oob = own_oob;
  const uint32_t valid = ~(S.sticky | oob) & cut_bits(nc);
  const int mv = __popc(valid);
  if (!maybe_out) {
    // the valid cuts are known: per sample, sum and sum of squares over them
    // in cut order (read kRowBatch cuts at a time, indices clamped, so the
    // reads are in flight together), then the terms summed in a fixed order
    for (int i = tid; i < n_samples; i += kRefineBlock) {
      double s1 = 0.0, s2 = 0.0;
      for (int k0 = 0; k0 < nc; k0 += kRowBatch) {
        float v[kRowBatch];
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) v[j] = sig[min(k0 + j, nc - 1)][i];
#pragma unroll
        for (int j = 0; j < kRowBatch; ++j) {
          if (k0 + j < nc && ((valid >> (k0 + j)) & 1u)) {
            const double d = (double)v[j];
            s1 += d;
            s2 += d * d;
          }
        }
      }
      sC[i] = (double)mv * s2 - s1 * s1;
    }
    __syncthreads();
    if (tid < 64) {
      double part = 0.0;
      for (int i = tid; i < n_samples; i += 64) part += sC[i];
      const double sum = wave_sum(part);
      if (tid == 0) {
        GD.cost_local = sum;
        GD.m_local = (int16_t)mv;
        GD.has_gram = 0;
        GD.oob = oob;
      }
    }
    return;
  }
  double* C = W.gram + (((size_t)cb * n_jobs + m) * G2MAX + g) * kMaxPairs;
  const int npairs = nc * (nc + 1) / 2;
  for (int q = tid; q < npairs; q += kRefineBlock) {
    int k = 0, r = q;
    while (r >= nc - k) {
      r -= nc - k;
      ++k;
    }
    const int l = k + r;
    double acc = 0.0;
    for (int i = 0; i < n_samples; ++i) acc += (double)sig[k][i] * (double)sig[l][i];
    C[q] = acc;
    sC[q] = acc;
  }
  __syncthreads();
  // residual of this grid point if no earlier grid point of the iteration
  // flags another cut (the selection checks that and recomputes otherwise)
  if (tid < 64) {
    const double part = tid < nc ? gram_row(sC, valid, tid, nc, mv) : 0.0;
    const double sum = wave_sum(part);
    if (tid == 0) {
      GD.cost_local = sum;
      GD.m_local = (int16_t)mv;
      GD.has_gram = 1;
      GD.oob = oob;
    }
  }
}

// The state a marker starts iteration it + 1 with (after k_refine_iter of
// iteration it), one block per marker; k_refine_init gives iteration 0's.
__global__ __launch_bounds__(kRefineBlock) void k_refine_pick(const RefineJobDev* __restrict__ jobs,
                                                              int it, int grid_n, RefineWork W) {
  __shared__ RefineStateDev S;
  const int m = blockIdx.x, n_jobs = gridDim.x;
  const RefineJobDev& J = jobs[m];
  if (it + 1 > J.n_iter) return;
  state_at(J, it + 1, m, n_jobs, grid_n, W, S);
  if (threadIdx.x == 0) W.state[(size_t)((it + 1) & 1) * n_jobs + m] = S;
}

__global__ __launch_bounds__(kRefineBlock) void k_refine_init(const RefineJobDev* __restrict__ jobs,
                                                              int grid_n, RefineWork W) {
  __shared__ RefineStateDev S;
  const int m = blockIdx.x;
  state_at(jobs[m], 0, m, gridDim.x, grid_n, W, S);
  if (threadIdx.x == 0) W.state[m] = S;
}

// Signals at the last iteration's best grid point, then the per-cut
// statistics and distances of orazioDistanceRobust (Identification.cpp:
// 88-160).
// The identification read-out at one homography per marker: the grid
// search's choice after its last iteration (given == null), or a given
// homography with its centre and score (given: 9 floats per job, centres 2
// per job, scores), whose out-of-bounds cuts are the ones leaving the image
// there.  Signals, median profile, orazio statistics, distance to every
// bank profile.
__global__ __launch_bounds__(kFinalBlock) void k_refine_final(
    const RefineJobDev* __restrict__ jobs, int n_samples, int grid_n,
    const float* __restrict__ inv_ratio, int bank_rows, int bank_cols, RefineWork W,
    const float* __restrict__ given_H, const float* __restrict__ given_c,
    const double* __restrict__ given_cost, RefineOutDev* __restrict__ out,
    float* __restrict__ dist) {
  __shared__ float sig[kMaxCuts][kMaxSamples + 1];
  __shared__ float ordered[kMaxSamples][kMaxCuts];  // per sample, the valid cuts' values sorted
  __shared__ float bar[kMaxSamples];
  __shared__ int valid_cut[kMaxCuts];
  __shared__ int n_valid;
  __shared__ int oob_at[kMaxCuts];
  __shared__ CutWalk walk[kMaxCuts];
  __shared__ float sH[9], sHi[9];
  __shared__ float stat_mub[kMaxCuts], stat_muw[kMaxCuts], stat_var[kMaxCuts];
  __shared__ float xs[kMaxSamples];
  __shared__ float ratio[kMaxBankRows * kMaxRatios];
  __shared__ uint8_t parity[kMaxSamples][kMaxBankRows];  // [sample][bank row]
  // a sample's term of the distance, for each ring parity
  __shared__ float term[2][kMaxCuts][kMaxSamples + 1];
  __shared__ RefineStateDev S;
  __shared__ double inv[kMaxSamples + 1];
  const int m = blockIdx.x, tid = threadIdx.x, n_jobs = gridDim.x;
  const RefineJobDev& J = jobs[m];
  RefineOutDev& O = out[m];
  const bool given = given_H != nullptr;
  if (!given && J.n_iter == 0) {
    if (tid == 0) O.ok = 0;
    return;
  }
  for (int q = tid; q < bank_rows * kMaxRatios; q += kFinalBlock) ratio[q] = inv_ratio[q];
  fill_inv(inv);
  const int nc = J.n_cuts;
  if (given) {
    if (tid == 0) {
      const double c = given_cost[m];
      S.ok = c < DBL_MAX && isfinite(c);
      S.cx = given_c[2 * m];
      S.cy = given_c[2 * m + 1];
      S.min_res = (float)c;
      for (int r = 0; r < 9; ++r) sH[r] = given_H[9 * m + r];
      inverse3(sH, sHi);
    }
    if (tid < kMaxCuts) oob_at[tid] = 0;
    __syncthreads();
  } else {
    // the selection after the last iteration
    state_at(J, J.n_iter, m, n_jobs, grid_n, W, S);
    if (S.ok && tid < 9) {
      const RefineGridDev& GB =
          W.grid[((size_t)((J.n_iter - 1) & 1) * n_jobs + m) * G2MAX + S.best_g];
      sH[tid] = GB.H[tid];
      sHi[tid] = GB.Hi[tid];
    }
    __syncthreads();
  }
  if (!S.ok) {
    if (tid == 0) O.ok = 0;
    return;
  }
  const Img im{J.quad, J.iw, J.ih};
  if (tid < nc) cut_walk_starts(sHi, J.stops[tid], J.begin_sig, J.end_sig, n_samples, walk[tid]);
  __syncthreads();
  const int nchunk = (n_samples + kRefineChunk - 1) / kRefineChunk;
  for (int w = tid; w < nc * nchunk; w += kFinalBlock) {
    const int k = w / nchunk, c = w - k * nchunk;
    if (sample_chunk(im, sH, walk[k], n_samples, c, sig[k]) && given) atomicOr(&oob_at[k], 1);
  }
  __syncthreads();
  // The outputs live in host memory: they are stored after the last
  // barrier, which would otherwise wait for the stores to reach the host.
  if (given && tid < 64) {
    const uint32_t oob = (uint32_t)__ballot(tid < nc && oob_at[tid] != 0);
    if (tid == 0) S.sticky = oob;
  }
  __syncthreads();
  const uint32_t sticky = S.sticky;
  if (tid < 64) {
    const uint32_t valid = ~sticky & cut_bits(nc);
    if (tid < nc && ((valid >> tid) & 1u)) valid_cut[__popc(valid & ((1u << tid) - 1u))] = tid;
    if (tid == 0) n_valid = __popc(valid);
  }
  // orazioDistanceRobust statistics per valid cut (Identification.cpp:88-121)
  if (tid < nc && !((sticky >> tid) & 1u)) {
    // the signal is read kStatBatch samples at a time (indices clamped,
    // so the reads are in flight together), then consumed in order
    constexpr int kStatBatch = 8;
    const float* sg = sig[tid];
    MeanVar<false> mv(inv);
    for (int i0 = 30; i0 < n_samples; i0 += kStatBatch) {
      float v[kStatBatch];
#pragma unroll
      for (int j = 0; j < kStatBatch; ++j) v[j] = sg[min(i0 + j, n_samples - 1)];
#pragma unroll
      for (int j = 0; j < kStatBatch; ++j)
        if (i0 + j < n_samples) mv.add(v[j]);
    }
    float s_mean = mv.mean, s_var = mv.var;
    if (mv.tiny) {  // again, dividing where a product was subnormal
      MeanVar<true> me(inv);
      for (int i = 30; i < n_samples; ++i) me.add(sg[i]);
      s_mean = me.mean;
      s_var = me.var;
    }
    float sumInf = 0, sumSup = 0;
    unsigned nInf = 0, nSup = 0;
    bool acc = false;
    for (int i0 = 0; i0 < n_samples; i0 += kStatBatch) {
      float v[kStatBatch];
#pragma unroll
      for (int j = 0; j < kStatBatch; ++j) v[j] = sg[min(i0 + j, n_samples - 1)];
      // without branches: the signal is non-negative, so adding +0 to a
      // sum leaves it unchanged
#pragma unroll
      for (int j = 0; j < kStatBatch; ++j) {
        const bool live = i0 + j < n_samples;
        const bool lo = live && v[j] < s_mean;
        acc = acc || lo;
        const bool take_lo = acc && lo, take_hi = live && acc && !lo;
        sumInf += take_lo ? v[j] : 0.f;
        nInf += take_lo;
        sumSup += take_hi ? v[j] : 0.f;
        nSup += take_hi;
      }
    }
    stat_var[tid] = s_var;
    stat_muw[tid] = sumSup / (float)nSup;
    stat_mub[tid] = sumInf / (float)nInf;
  } else if (tid == kFinalBlock - 64) {
    // sample positions x_i along the profile (another wavefront)
    const float stepX = (J.end_sig - J.begin_sig) / ((float)n_samples - 1.f);
    float x = J.begin_sig;
    for (int i = 0; i < n_samples; ++i) {
      xs[i] = x;
      x += stepX;
    }
  }
  __syncthreads();
  // the median profile over the valid cuts (Identification.cpp:920-934):
  // each value lands at its rank (ties broken by cut order), then the middle
  // order statistic, or the two middle ones summed in float and halved in
  // double, as the host's partial_sort_copy median does
  const int mv = n_valid;
  for (int q = tid; q < n_samples * mv; q += kFinalBlock) {
    const int i = q / mv, j = q - i * mv;
    const float v = sig[valid_cut[j]][i];
    int rank = 0;
    for (int l = 0; l < mv; ++l) {
      const float u = sig[valid_cut[l]][i];
      rank += (u < v) || (u == v && l < j);
    }
    ordered[i][rank] = v;
  }
  __syncthreads();
  for (int i = tid; i < n_samples; i += kFinalBlock) {
    float b = 0.f;
    if (mv % 2) {
      b = ordered[i][mv / 2];
    } else if (mv > 0) {
      const float sum = ordered[i][mv / 2] + ordered[i][mv / 2 - 1];
      b = (float)((double)sum / 2.0);
    }
    bar[i] = b;
  }
  __syncthreads();
  float bar_lo = FLT_MAX, bar_hi = -FLT_MAX;  // lane 0's are the block's
  if (tid < 64) {
    for (int i = tid; i < n_samples; i += 64) {
      bar_lo = fminf(bar_lo, bar[i]);
      bar_hi = fmaxf(bar_hi, bar[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
      bar_lo = fminf(bar_lo, __shfl_xor(bar_lo, o, 64));
      bar_hi = fmaxf(bar_hi, __shfl_xor(bar_hi, o, 64));
    }
  }
  // ring parity of every bank profile at every sample position, and each
  // sample's term under either parity (:124-160): the division is done once
  // per sample instead of once per sample and bank row
  for (int q = tid; q < bank_rows * n_samples; q += kFinalBlock) {
    const int i = q / bank_rows, idc = q - i * bank_rows;
    int ldum = 0;
    for (int j = 0; j < bank_cols; ++j)
      if (ratio[idc * kMaxRatios + j] <= xs[i]) ++ldum;
    parity[i][idc] = (uint8_t)(ldum & 1);
  }
  for (int q = tid; q < nc * n_samples; q += kFinalBlock) {
    const int k = q / n_samples, i = q - k * n_samples;
    if ((sticky >> k) & 1u) continue;
    const float v = sig[k][i], v2 = 2.0f * stat_var[k];
    const float tb = fmaxf(v - stat_mub[k], 0.f);
    const float tw = fminf(v - stat_muw[k], 0.f);
    term[1][k][i] = tb * tb / v2;
    term[0][k][i] = tw * tw / v2;
  }
  __syncthreads();
  // distance of each valid cut to each bank profile, summed in sample order
  for (int q = tid; q < nc * bank_rows; q += kFinalBlock) {
    const int k = q / bank_rows, idc = q - k * bank_rows;
    if ((sticky >> k) & 1u) continue;
    const float* t0 = term[0][k];
    const float* t1 = term[1][k];
    float distance = 0;
    constexpr int kSumBatch = 8;  // terms read together, then added in sample order
    for (int i0 = 0; i0 < n_samples; i0 += kSumBatch) {
      float v[kSumBatch];
#pragma unroll
      for (int j = 0; j < kSumBatch; ++j) {
        const int i = min(i0 + j, n_samples - 1);
        v[j] = parity[i][idc] ? t1[i] : t0[i];
      }
#pragma unroll
      for (int j = 0; j < kSumBatch; ++j)
        if (i0 + j < n_samples) distance += v[j];
    }
    dist[J.dist_offset + k * bank_rows + idc] = distance;
  }
  if (tid == 0) {
    O.ok = 1;
    O.cx = S.cx;
    O.cy = S.cy;
    for (int r = 0; r < 9; ++r) O.H[r] = sH[r];
    O.min_res = S.min_res;
    O.bar_min = bar_lo;
    O.bar_max = bar_hi;
  }
  if (tid < nc) O.oob[tid] = (uint8_t)((sticky >> tid) & 1u);
}

// A refinement job takes the stops of the selection it names (one block of
// 64 threads per job): its cut count is the selected count, and a job whose
// selection is empty gets no iteration.
__global__ __launch_bounds__(64) void k_take_stops(const RefineJobDev* __restrict__ src,
                                                   RefineJobDev* __restrict__ jobs,
                                                   const SelectOutDev* __restrict__ sel) {
  static_assert(sizeof(RefineJobDev) % 4 == 0, "jobs copied as words");
  constexpr int kWords = (int)(sizeof(RefineJobDev) / 4);
  const uint32_t* from = reinterpret_cast<const uint32_t*>(src + blockIdx.x);
  uint32_t* to = reinterpret_cast<uint32_t*>(jobs + blockIdx.x);
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) to[i] = from[i];
  __syncthreads();
  RefineJobDev& J = jobs[blockIdx.x];
  const int s = J.sel_index;
  if (s < 0) return;
  const SelectOutDev& o = sel[s];
  const int ns = o.n_collected == 0 ? 0 : min(o.n_selected, kMaxCuts);
  if (threadIdx.x < ns) J.stops[threadIdx.x] = make_float2(o.stops[threadIdx.x].x, o.stops[threadIdx.x].y);
  if (threadIdx.x == 0) {
    J.n_cuts = ns;
    if (ns == 0) J.n_iter = 0;
  }
}

// The centre the grid search moves to after its first iteration (the job's
// start when that iteration found no readable grid point).
__global__ __launch_bounds__(256) void k_refine_coarse_out(const RefineJobDev* __restrict__ jobs,
                                                           int n_jobs, RefineWork W,
                                                           float* __restrict__ centers) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_jobs) return;
  const RefineJobDev& J = jobs[m];
  const RefineStateDev& S = W.state[(size_t)n_jobs + m];  // parity of iteration 1
  const bool moved = J.n_iter >= 1 && S.ok;
  centers[2 * m] = moved ? S.cx : J.c0x;
  centers[2 * m + 1] = moved ? S.cy : J.c0y;
}

}  // namespace

int launch_refine_coarse(const RefineJobDev* jobs, int n_jobs, int n_samples, int grid_n,
                         const RefineWork& work, float* centers, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  const dim3 grid(grid_n * grid_n, n_jobs);
  k_refine_init<<<n_jobs, kRefineBlock, 0, s>>>(jobs, grid_n, work);
  k_refine_iter<<<grid, kRefineBlock, 0, s>>>(jobs, 0, n_samples, grid_n, 0, work);
  k_refine_pick<<<n_jobs, kRefineBlock, 0, s>>>(jobs, 0, grid_n, work);
  k_refine_coarse_out<<<(n_jobs + 255) / 256, 256, 0, s>>>(jobs, n_jobs, work, centers);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_refine_readout(const RefineJobDev* jobs, int n_jobs, int n_samples, const float* H,
                          const float* centers, const double* cost, const float* inv_ratio,
                          int bank_rows, int bank_cols, const RefineWork& work, RefineOutDev* out,
                          float* dist, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  k_refine_final<<<n_jobs, kFinalBlock, 0, s>>>(jobs, n_samples, kMaxGrid, inv_ratio, bank_rows,
                                                 bank_cols, work, H, centers, cost, out, dist);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_take_stops(const RefineJobDev* src, RefineJobDev* jobs, int n_jobs,
                      const SelectOutDev* sel, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  k_take_stops<<<n_jobs, 64, 0, s>>>(src, jobs, sel);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_select_cuts(const SelectJobDev* jobs, int n_jobs, const float4* pts, float start_sig,
                       int n_samples, int n_ref, int select_max, SelectOutDev* out, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  k_select_cuts<<<n_jobs, 256, 0, s>>>(jobs, pts, start_sig, n_samples, n_ref, select_max, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_refine(const RefineJobDev* jobs, int n_jobs, int max_iter, int n_samples, int grid_n,
                  const float* inv_ratio, int bank_rows, int bank_cols, const RefineWork& work,
                  RefineOutDev* out, float* dist, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  const dim3 grid(grid_n * grid_n, n_jobs);
  const int fused = grid.x * grid.y <= kFuseMaxBlocks;
  if (!fused) k_refine_init<<<n_jobs, kRefineBlock, 0, s>>>(jobs, grid_n, work);
  for (int it = 0; it < max_iter; ++it) {
    k_refine_iter<<<grid, kRefineBlock, 0, s>>>(jobs, it, n_samples, grid_n, fused, work);
    if (!fused && it + 1 < max_iter)
      k_refine_pick<<<n_jobs, kRefineBlock, 0, s>>>(jobs, it, grid_n, work);
  }
  k_refine_final<<<n_jobs, kFinalBlock, 0, s>>>(jobs, n_samples, grid_n, inv_ratio, bank_rows,
                                                 bank_cols, work, nullptr, nullptr, nullptr, out,
                                                 dist);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace cctagx
