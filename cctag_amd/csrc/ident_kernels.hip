// gfx950 kernels of CCTag identification.
//
// k_select_cuts  identify_step_1: collectCuts + selectCutCheapUniform
//                (Identification.cpp:396-709, 1172-1301); one block per
//                marker, one thread per outer point.  Bit-exact with the
//                reference's serial float arithmetic.
// k_refine       refineConicFamilyGlob (Identification.cpp:813-1158): every
//                grid-search iteration of a marker inside one persistent
//                1024-thread block, then the final signals and the per-cut
//                distances to every bank profile of orazioDistanceRobust
//                (:50-160).  The pairwise sum of squared differences uses
//                sum_{i<j} (s_i - s_j)^2 = m sum s^2 - (sum s)^2 in double
//                instead of the reference's sequential float sum; everything
//                else (grid, homographies, sampling, the sticky out-of-bounds
//                flag of a cut) follows the reference's float operations.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ident_device.hpp"

namespace cctagx {
namespace {

__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }

struct Img {
  const uint8_t* p;
  int w, h;
  // Identification.hpp:197-222: half of the bilinear value.
  __device__ __forceinline__ float bilinear(float x, float y) const {
    const int px = (int)x, py = (int)y;
    const uint8_t* p0 = p + px + py * w;
    const float fx = x - px, fy = y - py;
    const float fx1 = 1.0f - fx, fy1 = 1.0f - fy;
    const float w1 = fx1 * fy1, w2 = fx * fy1, w3 = fx1 * fy, w4 = fx * fy;
    return ((float)p0[0] * w1 + (float)p0[1] * w2 + (float)p0[w] * w3 + (float)p0[w + 1] * w4) / 2;
  }
};

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
struct MeanVar {
  float mean = 0.f, var = 0.f;
  unsigned cnt = 0;
  __device__ __forceinline__ void add(float v) {
    ++cnt;
    mean = (mean * (float)(cnt - 1) + v) / (float)cnt;
    if (cnt > 1) {
      const float t = v - mean;
      var = var * (float)(cnt - 1) / (float)cnt + t * t / (float)(cnt - 1);
    }
  }
};

// Identification.cpp:396-459 for a cut whose signal is only consumed
// sample by sample; returns false when a sample falls outside the image
// (the caller then discards what f saw).  Sample positions follow the
// reference's running sums; the image reads of kWalkBatch samples are issued
// together so their latencies overlap.
constexpr int kWalkBatch = 8;

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
    float v[kWalkBatch];
#pragma unroll
    for (int t = 0; t < kWalkBatch; ++t) v[t] = i0 + t < n ? im.bilinear(px[t], py[t]) : 0.f;
#pragma unroll
    for (int t = 0; t < kWalkBatch; ++t)
      if (i0 + t < n) f(i0 + t, v[t]);
  }
  return true;
}

__constant__ float kRefineKernels[3][9] = {
    {-0.0000f, -0.0003f, -0.1065f, -0.7863f, .0f, 0.7863f, 0.1065f, 0.0003f, 0.0000f},
    {-0.0044f, -0.0540f, -0.2376f, -0.3450f, .0f, 0.3450f, 0.2376f, 0.0540f, 0.0044f},
    {-0.0366f, -0.1113f, -0.1801f, -0.1594f, .0f, 0.1594f, 0.1801f, 0.1113f, 0.0366f}};

__global__ __launch_bounds__(256) void k_select_cuts(const SelectJobDev* __restrict__ jobs,
                                                      const float4* __restrict__ pts,
                                                      float start_sig, int n_samples, int n_ref,
                                                      int select_max, SelectOutDev* __restrict__ out) {
  __shared__ float var[kMaxOuterPoints];
  __shared__ uint8_t valid[kMaxOuterPoints];
  __shared__ uint8_t cand[kMaxOuterPoints];
  __shared__ float4 stop[kMaxOuterPoints];
  __shared__ float var_max;
  const SelectJobDev J = jobs[blockIdx.x];
  const Img im{J.img, J.iw, J.ih};
  const int n = J.n_outer;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const float4 op = pts[J.pt_offset + t];
    MeanVar mv;
    const bool ok = cut_walk(im, J.cx, J.cy, op.x, op.y, start_sig, 1.f, n_samples,
                             [&](int, float v) { mv.add(v); });
    valid[t] = ok;
    var[t] = mv.var;
    stop[t] = op;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = -FLT_MAX;
    int nc = 0;
    for (int t = 0; t < n; ++t)
      if (valid[t]) {
        m = nc == 0 ? var[t] : fmaxf(m, var[t]);
        ++nc;
      }
    var_max = m;
    out[blockIdx.x].n_collected = nc;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    cand[t] = 0;
    if (!valid[t]) continue;
    float4 s = stop[t];
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
    float sig[32];
    const bool ok = cut_walk(im, psx, psy, pex, pey, 0.f, 1.f, n_ref, [&](int i, float v) { sig[i] = v; });
    if (ok) {
      float key[3], loc[3];
      for (int k = 0; k < 3; ++k) {
        float best = 0.f;
        int bi = 0;
        for (int i = 0; i < n_ref; ++i) {
          float tmp = 0;
          for (int j = 0; j < 9; ++j) {
            const int r = i - 4 + j;
            const float v = r < 0 ? sig[0] : (r >= n_ref ? sig[n_ref - 1] : sig[r]);
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
      cand[t] = var[t] / var_max > 0.5f;
    }
    stop[t] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nc = 0, ni = 0;
    for (int t = 0; t < n; ++t) {
      if (!valid[t]) continue;
      ++nc;
      if (cand[t]) ++ni;
    }
    const int select = min(select_max, nc);
    const float step = fmaxf(1.f, (float)ni / (float)select);
    int nsel = 0;
    for (int k = 0;; ++k) {
      const int at = (int)((float)k * step);
      if (at < ni && nsel < select) {
        // at-th candidate in collected order
        int seen = 0;
        for (int t = 0; t < n; ++t)
          if (valid[t] && cand[t]) {
            if (seen == at) {
              out[blockIdx.x].stops[nsel++] = stop[t];
              break;
            }
            ++seen;
          }
      } else {
        break;
      }
    }
    out[blockIdx.x].n_selected = nsel;
  }
}

// ---------------------------------------------------------------------------
constexpr int kRefineThreads = 1024;
constexpr int kRefineChunk = 10;  // samples per sampling work item

__global__ __launch_bounds__(kRefineThreads) void k_refine(
    const RefineJobDev* __restrict__ jobs, int n_samples, int grid_n,
    const float* __restrict__ inv_ratio, int bank_rows, int bank_cols,
    float* __restrict__ scratch, RefineOutDev* __restrict__ out) {
  constexpr int G2MAX = kMaxGrid * kMaxGrid;
  __shared__ float sH[G2MAX][9], sHi[G2MAX][9];
  __shared__ float gx[G2MAX], gy[G2MAX];
  __shared__ int oob_here[G2MAX][kMaxCuts];
  __shared__ uint8_t valid[G2MAX][kMaxCuts];
  __shared__ uint8_t sticky[kMaxCuts];
  __shared__ int mvalid[G2MAX];
  __shared__ float res[G2MAX];
  __shared__ double contrib[G2MAX][kMaxSamples];
  __shared__ float cur_x, cur_y, neighbour;
  __shared__ int best_g, ok;
  __shared__ float stat_mub[kMaxCuts], stat_muw[kMaxCuts], stat_var[kMaxCuts];

  const RefineJobDev& J = jobs[blockIdx.x];
  const Img im{J.img, J.iw, J.ih};
  const int nc = J.n_cuts, G2 = grid_n * grid_n;
  float* scr = scratch + (size_t)blockIdx.x * G2MAX * kMaxCuts * kMaxSamples;
  const int tid = threadIdx.x;
  if (tid < kMaxCuts) sticky[tid] = 0;
  if (tid == 0) {
    cur_x = J.c0x;
    cur_y = J.c0y;
    neighbour = J.neighbour0;
    ok = 1;
    best_g = 0;
  }
  __syncthreads();
  float min_res = FLT_MAX;
  for (int it = 0; it < J.n_iter; ++it) {
    // getNearbyPoints + computeHomographyFromEllipseAndImagedCenter per grid point
    if (tid < G2) {
      const int i = tid / grid_n, j = tid - i * grid_n;
      const float ns = neighbour * J.grid_scale;
      float cx = cur_x, cy = cur_y;
      condition(cx, cy, J.cond);
      const float halfWidth = ns / 2.f;
      const float stepSize = ns / (float)(grid_n - 1);
      float qx = cx - halfWidth + (float)i * stepSize, qy = cy - halfWidth + (float)j * stepSize;
      condition(qx, qy, J.cond_inv);
      gx[tid] = qx;
      gy[tid] = qy;
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
      float Hf[9], Hi[9];
      mul3(J.t_inv_can, H, Hf);
      inverse3(Hf, Hi);
      for (int r = 0; r < 9; ++r) {
        sH[tid][r] = Hf[r];
        sHi[tid][r] = Hi[r];
      }
    }
    __syncthreads();
    // extractSignalUsingHomography for every (grid point, cut), split into
    // chunks of kRefineChunk samples so that all lanes gather at once; a
    // chunk replays the running sums of the positions before it.
    for (int q = tid; q < G2 * nc; q += kRefineThreads) oob_here[q / nc][q % nc] = 0;
    __syncthreads();
    const int nchunk = (n_samples + kRefineChunk - 1) / kRefineChunk;
    for (int pk = tid; pk < G2 * nc * nchunk; pk += kRefineThreads) {
      const int c = pk % nchunk, gk = pk / nchunk;
      const int g = gk / nc, k = gk - g * nc;
      float bx, by;
      apply_h(bx, by, sHi[g], J.stops[k].x, J.stops[k].y);
      float xs, ys, xe, ye;
      if (J.begin_sig != 0.f) {
        xs = bx * J.begin_sig;
        ys = by * J.begin_sig;
      } else {
        xs = 0;
        ys = 0;
      }
      if (J.end_sig != 1.f) {
        xe = bx * J.end_sig;
        ye = by * J.end_sig;
      } else {
        xe = bx;
        ye = by;
      }
      const float stepX = (xe - xs) / ((float)n_samples - 1.f), stepY = (ye - ys) / ((float)n_samples - 1.f);
      float x = xs, y = ys;
      const int i0 = c * kRefineChunk;
      for (int i = 0; i < i0; ++i) {
        x += stepX;
        y += stepY;
      }
      const float* H = sH[g];
      const float xmax = (float)(im.w - 1), ymax = (float)(im.h - 1);
      float v[kRefineChunk];
      bool in[kRefineChunk];
      bool oob = false;
#pragma unroll
      for (int t = 0; t < kRefineChunk; ++t) {
        float xr, yr;
        apply_h(xr, yr, H, x, y);
        in[t] = i0 + t < n_samples && xr >= 0.f && xr < xmax && yr >= 0.f && yr < ymax;
        oob |= (i0 + t < n_samples) && !in[t];
        v[t] = in[t] ? im.bilinear(xr, yr) : 0.f;
        x += stepX;
        y += stepY;
      }
      float* dst = scr + ((size_t)g * kMaxCuts + k) * kMaxSamples + i0;
#pragma unroll
      for (int t = 0; t < kRefineChunk; ++t)
        if (in[t]) dst[t] = v[t];
      if (oob) atomicOr(&oob_here[g][k], 1);
    }
    __syncthreads();
    // sticky out-of-bounds flags, in grid-point evaluation order
    if (tid < nc) {
      uint8_t run = sticky[tid];
      for (int g = 0; g < G2; ++g) {
        run |= oob_here[g][tid];
        valid[g][tid] = !run;
      }
      sticky[tid] = run;
    }
    __syncthreads();
    if (tid < G2) {
      int m = 0;
      for (int k = 0; k < nc; ++k) m += valid[tid][k];
      mvalid[tid] = m;
    }
    __syncthreads();
    // costFunctionGlob: per sample, m sum s^2 - (sum s)^2 over valid cuts
    __threadfence_block();
    for (int q = tid; q < G2 * n_samples; q += kRefineThreads) {
      const int g = q / n_samples, ii = q - g * n_samples;
      double s1 = 0.0, s2 = 0.0;
      for (int k = 0; k < nc; ++k)
        if (valid[g][k]) {
          const double s = (double)scr[((size_t)g * kMaxCuts + k) * kMaxSamples + ii];
          s1 += s;
          s2 += s * s;
        }
      contrib[g][ii] = (double)mvalid[g] * s2 - s1 * s1;
    }
    __syncthreads();
    if (tid < G2) {
      const int m = mvalid[tid];
      if (m < 2) {
        res[tid] = FLT_MAX;
      } else {
        double acc = 0.0;
        for (int ii = 0; ii < n_samples; ++ii) acc += contrib[tid][ii];
        const int pairs = m * (m - 1) / 2;
        res[tid] = (float)acc / (float)pairs;
      }
    }
    __syncthreads();
    if (tid == 0) {
      int b = -1;
      float mr = FLT_MAX;
      for (int g = 0; g < G2; ++g)
        if (mvalid[g] >= 2 && res[g] < mr) {
          mr = res[g];
          b = g;
        }
      if (b < 0) {
        ok = 0;
      } else {
        best_g = b;
        cur_x = gx[b];
        cur_y = gy[b];
        neighbour /= (float)((grid_n - 1) / 2);
      }
    }
    __syncthreads();
    if (!ok) break;
    min_res = res[best_g];
  }
  RefineOutDev& O = out[blockIdx.x];
  if (!ok || J.n_iter == 0) {
    if (tid == 0) O.ok = 0;
    return;
  }
  const int bg = best_g;
  if (tid == 0) {
    O.ok = 1;
    O.cx = cur_x;
    O.cy = cur_y;
    for (int r = 0; r < 9; ++r) O.H[r] = sH[bg][r];
    O.min_res = min_res;
  }
  if (tid < nc) O.oob[tid] = sticky[tid];
  for (int q = tid; q < nc * n_samples; q += kRefineThreads) {
    const int k = q / n_samples, ii = q - k * n_samples;
    O.signals[k * kMaxSamples + ii] = scr[((size_t)bg * kMaxCuts + k) * kMaxSamples + ii];
  }
  // orazioDistanceRobust statistics per valid cut (Identification.cpp:88-121)
  if (tid < nc && !sticky[tid]) {
    const float* s = scr + ((size_t)bg * kMaxCuts + tid) * kMaxSamples;
    MeanVar mv;
    for (int i = 30; i < n_samples; ++i) mv.add(s[i]);
    float sumInf = 0, sumSup = 0;
    unsigned nInf = 0, nSup = 0;
    bool acc = false;
    for (int i = 0; i < n_samples; ++i) {
      const float v = s[i];
      if (!acc && v < mv.mean) acc = true;
      if (acc) {
        if (v < mv.mean) {
          sumInf += v;
          ++nInf;
        } else {
          sumSup += v;
          ++nSup;
        }
      }
    }
    stat_var[tid] = mv.var;
    stat_muw[tid] = sumSup / (float)nSup;
    stat_mub[tid] = sumInf / (float)nInf;
  }
  __syncthreads();
  // distance of each valid cut to each bank profile (:124-160)
  for (int q = tid; q < nc * bank_rows; q += kRefineThreads) {
    const int k = q / bank_rows, idc = q - k * bank_rows;
    if (sticky[k]) continue;
    const float* s = scr + ((size_t)bg * kMaxCuts + k) * kMaxSamples;
    const float stepX = (J.end_sig - J.begin_sig) / ((float)n_samples - 1.f);
    const float mub = stat_mub[k], muw = stat_muw[k], v2 = 2.0f * stat_var[k];
    float x = J.begin_sig;
    float distance = 0;
    for (int i = 0; i < n_samples; ++i) {
      int ldum = 0;
      for (int j = 0; j < bank_cols; ++j)
        if (inv_ratio[idc * kMaxRatios + j] <= x) ++ldum;
      float dd;
      if ((ldum & 1) != 0) {
        const float t = fmaxf(s[i] - mub, 0.f);
        dd = t * t / v2;
      } else {
        const float t = fminf(s[i] - muw, 0.f);
        dd = t * t / v2;
      }
      distance += dd;
      x += stepX;
    }
    O.id_dist[k * kMaxBankRows + idc] = distance;
  }
}

}  // namespace

int launch_select_cuts(const SelectJobDev* jobs, int n_jobs, const float4* pts, float start_sig,
                       int n_samples, int n_ref, int select_max, SelectOutDev* out, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  k_select_cuts<<<n_jobs, 256, 0, s>>>(jobs, pts, start_sig, n_samples, n_ref, select_max, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_refine(const RefineJobDev* jobs, int n_jobs, int n_samples, int grid_n,
                  const float* inv_ratio, int bank_rows, int bank_cols, float* scratch,
                  RefineOutDev* out, hipStream_t s) {
  if (n_jobs <= 0) return 0;
  k_refine<<<n_jobs, kRefineThreads, 0, s>>>(jobs, n_samples, grid_n, inv_ratio, bank_rows,
                                             bank_cols, scratch, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace cctagx
