// Differentiable identification refinement on gfx950.
//
// The grid search of refineConicFamilyGlob (Identification.cpp:813-1158)
// scores an imaged centre c by sampling every selected cut through the
// homography H(c) of the outer ellipse and the centre
// (computeHomographyFromEllipseAndImagedCenter, :739-798) and summing the
// squared differences between cuts.  Here that score is a differentiable
// function of H:
//
//   q_k  = pi(H^-1 p_k)                 marker-plane direction of stop p_k
//   y_ki = pi(H (t_i q_k, 1))           t_i = begin + i (end - begin)/(n - 1)
//   s_ki = bilinear(I, y_ki) / 2        Identification.hpp:197-222
//   C    = sum_i sum_{k<l valid} (s_ki - s_li)^2 / (m (m-1) / 2)
//
// (pi(w) = (w0/w2, w1/w2); a cut is valid when all its samples are inside
// the image).  k_grad_eval runs the forward pass and the adjoint pass of one
// marker per block: dC/ds_ki = 2 (m s_ki - S_i) / pairs with S_i the sum over
// valid cuts, lambda = Jpi(w)^T dC/dy, and
//
//   dC/dH = sum_k [ sum_i lambda_ki u_ki^T  -  H^-T Jpi(v_k)^T (H_{:,01}^T L_k) v_k^T ]
//
// with v_k = H^-1 p_k and L_k = sum_i t_i lambda_ki (the second term is the
// dependence of the cut direction q_k on H).  The centre gradient chains
// dC/dH with dH/dc of the closed form, evaluated in forward mode with dual
// numbers.  Sample positions and interpolation run in T (float in
// production, double for checking); every sum is in double.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "grad_device.hpp"

namespace cctagx {
namespace {

// Dual number carrying d/dcx and d/dcy (forward mode).
struct D2 {
  double v, dx, dy;
};
__device__ __forceinline__ D2 dc(double v) { return D2{v, 0.0, 0.0}; }
__device__ __forceinline__ D2 operator+(D2 a, D2 b) { return D2{a.v + b.v, a.dx + b.dx, a.dy + b.dy}; }
__device__ __forceinline__ D2 operator-(D2 a, D2 b) { return D2{a.v - b.v, a.dx - b.dx, a.dy - b.dy}; }
__device__ __forceinline__ D2 operator-(D2 a) { return D2{-a.v, -a.dx, -a.dy}; }
__device__ __forceinline__ D2 operator*(D2 a, D2 b) {
  return D2{a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy};
}
__device__ __forceinline__ D2 operator/(D2 a, D2 b) {
  const double inv = 1.0 / b.v, r = a.v * inv;
  return D2{r, (a.dx - r * b.dx) * inv, (a.dy - r * b.dy) * inv};
}
__device__ __forceinline__ D2 dsqrt(D2 a) {
  const double r = sqrt(a.v), h = 0.5 / r;
  return D2{r, a.dx * h, a.dy * h};
}

// Closed-form homography of the ellipse's canonic form and an imaged centre
// (Identification.cpp:739-798), in double with derivatives w.r.t. (cx, cy).
__device__ void closed_form(const GradJobDev& J, double cx, double cy, D2 H[9]) {
  const D2 x{cx, 1.0, 0.0}, y{cy, 0.0, 1.0};
  const float* T = J.t_can;
  const D2 u = dc(T[0]) * x + dc(T[1]) * y + dc(T[2]);
  const D2 v = dc(T[3]) * x + dc(T[4]) * y + dc(T[5]);
  const D2 w = dc(T[6]) * x + dc(T[7]) * y + dc(T[8]);
  const D2 xc = u / w, yc = v / w;
  const D2 Q11 = dc(J.Q11), Q22 = dc(J.Q22), Q33 = dc(J.Q33);
  D2 Hc[9] = {Q33, Q22 * xc * yc, -(Q33 * xc),
              dc(0.0), -(Q11 * xc * xc) - Q33, -(Q33 * yc),
              -(Q11 * xc), Q22 * yc, -Q33};
  const D2 d0 = dsqrt(Q22 * Q33 / Q11 * (Q11 * xc * xc + Q22 * yc * yc + Q33));
  const D2 d1 = Q33;
  const D2 d2 = dsqrt(-(Q22 * (Q11 * xc * xc + Q33)));
  for (int r = 0; r < 3; ++r) {
    Hc[r * 3 + 0] = Hc[r * 3 + 0] * d0;
    Hc[r * 3 + 1] = Hc[r * 3 + 1] * d1;
    Hc[r * 3 + 2] = Hc[r * 3 + 2] * d2;
  }
  const float* Ti = J.t_inv_can;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      H[i * 3 + j] = dc(Ti[i * 3 + 0]) * Hc[0 * 3 + j] + dc(Ti[i * 3 + 1]) * Hc[1 * 3 + j] +
                     dc(Ti[i * 3 + 2]) * Hc[2 * 3 + j];
}

__device__ bool inverse3d(const double* m, double* out) {
  const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8],
               c02 = m[3] * m[7] - m[4] * m[6];
  const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  if (!(det != 0.0) || !isfinite(det)) return false;
  const double id = 1.0 / det;
  out[0] = c00 * id;
  out[1] = (m[2] * m[7] - m[1] * m[8]) * id;
  out[2] = (m[1] * m[5] - m[2] * m[4]) * id;
  out[3] = c01 * id;
  out[4] = (m[0] * m[8] - m[2] * m[6]) * id;
  out[5] = (m[2] * m[3] - m[0] * m[5]) * id;
  out[6] = c02 * id;
  out[7] = (m[1] * m[6] - m[0] * m[7]) * id;
  out[8] = (m[0] * m[4] - m[1] * m[3]) * id;
  return true;
}

constexpr int kGradBlock = 256;
constexpr int kLanesPerCut = 8;  // adjoint lanes per cut (kMaxCuts * 8 = kGradBlock)
static_assert(kMaxCuts * kLanesPerCut == kGradBlock, "one lane group per cut");

template <class T>
struct EvalShared {
  double H[9], Hi[9];
  double v[kMaxCuts][3];       // H^-1 p_k
  T s[kMaxCuts][kMaxSamples];
  int oob[kMaxCuts];
  double S[kMaxSamples];       // sum over valid cuts of s_ki
  double part[kMaxSamples];
  double G[kMaxCuts][9];
  double cost, gH[9];
  int m, ok;
};

// Sample i of cut k at the homography (Hs in T): the image position, its
// half-bilinear value and the value's derivatives along image x and y.
// False when the position is outside [0, w-1) x [0, h-1).
template <class T>
__device__ __forceinline__ bool sample_at(const GradJobDev& J, const T* Hs, T qx, T qy, T t, T& val,
                                          T& dx, T& dy) {
  const T ux = t * qx, uy = t * qy;
  const T w0 = Hs[0] * ux + Hs[1] * uy + Hs[2], w1 = Hs[3] * ux + Hs[4] * uy + Hs[5],
          w2 = Hs[6] * ux + Hs[7] * uy + Hs[8];
  const T x = w0 / w2, y = w1 / w2;
  if (!(x >= (T)0 && x < (T)(J.iw - 1) && y >= (T)0 && y < (T)(J.ih - 1))) return false;
  const int ix = (int)x, iy = (int)y;
  const uint32_t q = J.quad[(size_t)iy * J.iw + ix];
  const T fx = x - (T)ix, fy = y - (T)iy;
  const T a = (T)(q & 255u), b = (T)((q >> 8) & 255u), c = (T)((q >> 16) & 255u), d = (T)(q >> 24);
  val = (a * ((T)1 - fx) * ((T)1 - fy) + b * fx * ((T)1 - fy) + c * ((T)1 - fx) * fy + d * fx * fy) / (T)2;
  dx = ((b - a) * ((T)1 - fy) + (d - c) * fy) / (T)2;
  dy = ((c - a) * ((T)1 - fx) + (d - b) * fx) / (T)2;
  return true;
}

// Forward (and with `grad`, adjoint) pass of one marker at the homography
// in E.H; the whole block takes part.  E.cost is DBL_MAX and E.ok 0 when
// fewer than two cuts are valid or H is singular.
template <class T>
__device__ void eval_marker(const GradJobDev& J, int n_samples, bool grad, EvalShared<T>& E) {
  const int tid = threadIdx.x;
  const int nc = J.n_cuts;
  if (tid == 0) E.ok = inverse3d(E.H, E.Hi) ? 1 : 0;
  if (tid < kMaxCuts) E.oob[tid] = 0;
  __syncthreads();
  if (!E.ok) {
    if (tid == 0) E.cost = DBL_MAX;
    __syncthreads();
    return;
  }
  if (tid < nc) {
    const double px = J.stops[tid].x, py = J.stops[tid].y;
    for (int r = 0; r < 3; ++r) E.v[tid][r] = E.Hi[r * 3 + 0] * px + E.Hi[r * 3 + 1] * py + E.Hi[r * 3 + 2];
  }
  __syncthreads();
  T Hs[9];
  for (int e = 0; e < 9; ++e) Hs[e] = (T)E.H[e];
  const T b = (T)J.begin_sig, step = ((T)J.end_sig - (T)J.begin_sig) / (T)(n_samples - 1);
  for (int q = tid; q < nc * n_samples; q += kGradBlock) {
    const int k = q / n_samples, i = q - k * n_samples;
    const T qx = (T)(E.v[k][0] / E.v[k][2]), qy = (T)(E.v[k][1] / E.v[k][2]);
    T val, dx, dy;
    if (sample_at<T>(J, Hs, qx, qy, b + (T)i * step, val, dx, dy)) {
      E.s[k][i] = val;
    } else {
      E.s[k][i] = (T)0;
      atomicOr(&E.oob[k], 1);
    }
  }
  __syncthreads();
  if (tid == 0) {
    int m = 0;
    for (int k = 0; k < nc; ++k) m += E.oob[k] == 0;
    E.m = m;
  }
  __syncthreads();
  const int m = E.m;
  if (m < 2) {
    if (tid == 0) {
      E.cost = DBL_MAX;
      E.ok = 0;
    }
    __syncthreads();
    return;
  }
  const double pairs = 0.5 * (double)m * (double)(m - 1);
  for (int i = tid; i < n_samples; i += kGradBlock) {
    double s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < nc; ++k)
      if (!E.oob[k]) {
        const double v = (double)E.s[k][i];
        s1 += v;
        s2 += v * v;
      }
    E.S[i] = s1;
    E.part[i] = (double)m * s2 - s1 * s1;
  }
  __syncthreads();
  if (tid == 0) {
    double acc = 0.0;
    for (int i = 0; i < n_samples; ++i) acc += E.part[i];
    E.cost = acc / pairs;
  }
  if (!grad) {
    __syncthreads();
    return;
  }
  // adjoint: lanes k*8 .. k*8+7 sum cut k's samples
  {
    const int k = tid / kLanesPerCut, j = tid - k * kLanesPerCut;
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, L[3] = {0, 0, 0};
    if (k < nc && !E.oob[k]) {
      const double vx = E.v[k][0] / E.v[k][2], vy = E.v[k][1] / E.v[k][2];
      const T qx = (T)vx, qy = (T)vy;
      for (int i = j; i < n_samples; i += kLanesPerCut) {
        const T tt = b + (T)i * step;
        T val, dsx, dsy;
        sample_at<T>(J, Hs, qx, qy, tt, val, dsx, dsy);
        const double a = 2.0 * ((double)m * (double)E.s[k][i] - E.S[i]) / pairs;
        const double gx = a * (double)dsx, gy = a * (double)dsy;
        const double t = (double)tt;
        const double u0 = t * vx, u1 = t * vy;
        const double w0 = E.H[0] * u0 + E.H[1] * u1 + E.H[2];
        const double w1 = E.H[3] * u0 + E.H[4] * u1 + E.H[5];
        const double w2 = E.H[6] * u0 + E.H[7] * u1 + E.H[8];
        const double l0 = gx / w2, l1 = gy / w2, l2 = -(gx * w0 + gy * w1) / (w2 * w2);
        const double lam[3] = {l0, l1, l2}, u[3] = {u0, u1, 1.0};
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) A[r * 3 + c] += lam[r] * u[c];
          L[r] += t * lam[r];
        }
      }
    }
    for (int o = kLanesPerCut / 2; o > 0; o >>= 1) {
      for (int e = 0; e < 9; ++e) A[e] += __shfl_down(A[e], o, kLanesPerCut);
      for (int e = 0; e < 3; ++e) L[e] += __shfl_down(L[e], o, kLanesPerCut);
    }
    if (j == 0 && k < nc) {
      if (E.oob[k]) {
        for (int e = 0; e < 9; ++e) E.G[k][e] = 0.0;
      } else {
        const double* H = E.H;
        const double* v = E.v[k];
        const double z0 = H[0] * L[0] + H[3] * L[1] + H[6] * L[2];
        const double z1 = H[1] * L[0] + H[4] * L[1] + H[7] * L[2];
        const double nu[3] = {z0 / v[2], z1 / v[2], -(z0 * v[0] + z1 * v[1]) / (v[2] * v[2])};
        for (int a = 0; a < 3; ++a) {
          const double rho = E.Hi[0 * 3 + a] * nu[0] + E.Hi[1 * 3 + a] * nu[1] + E.Hi[2 * 3 + a] * nu[2];
          for (int c = 0; c < 3; ++c) E.G[k][a * 3 + c] = A[a * 3 + c] - rho * v[c];
        }
      }
    }
  }
  __syncthreads();
  if (tid < 9) {
    double acc = 0.0;
    for (int k = 0; k < nc; ++k) acc += E.G[k][tid];
    E.gH[tid] = acc;
  }
  __syncthreads();
}

template <class T>
__global__ __launch_bounds__(kGradBlock) void k_grad_eval(const GradJobDev* __restrict__ jobs,
                                                          const float* __restrict__ Hin, int n_samples,
                                                          double* __restrict__ cost,
                                                          double* __restrict__ grad) {
  __shared__ EvalShared<T> E;
  const int m = blockIdx.x;
  const GradJobDev& J = jobs[m];
  if (threadIdx.x < 9) E.H[threadIdx.x] = (double)Hin[m * 9 + threadIdx.x];
  __syncthreads();
  eval_marker<T>(J, n_samples, grad != nullptr, E);
  if (threadIdx.x == 0) {
    cost[m] = E.cost;
    if (grad) {
      if (!E.ok) {
        for (int e = 0; e < 10; ++e) grad[m * 10 + e] = 0.0;
      } else {
        for (int e = 0; e < 8; ++e) grad[m * 10 + e] = E.gH[e];
        // marker-plane translation of the centre: H' = H [1 0 dx; 0 1 dy; 0 0 1]
        grad[m * 10 + 8] = E.gH[2] * E.H[0] + E.gH[5] * E.H[3] + E.gH[8] * E.H[6];
        grad[m * 10 + 9] = E.gH[2] * E.H[1] + E.gH[5] * E.H[4] + E.gH[8] * E.H[7];
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_grad_homography(const GradJobDev* __restrict__ jobs, int n,
                                                         const float* __restrict__ centers,
                                                         float* __restrict__ H) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  D2 h[9];
  closed_form(jobs[m], centers[2 * m], centers[2 * m + 1], h);
  for (int e = 0; e < 9; ++e) H[m * 9 + e] = (float)h[e].v;
}

// `steps` evaluations per marker, each one forward + adjoint pass at H(c):
// a step that lowers the score is kept and the step length grows by 1.5,
// otherwise the centre returns to the best one and the length halves; the
// next centre moves the best one by the length along -dC/dc.
template <class T>
__global__ __launch_bounds__(kGradBlock) void k_grad_descend(const GradJobDev* __restrict__ jobs,
                                                             int steps, int n_samples,
                                                             const float* __restrict__ start,
                                                             float* __restrict__ centers,
                                                             double* __restrict__ cost) {
  __shared__ EvalShared<T> E;
  __shared__ double dH[9][2];
  __shared__ double cx, cy, best_x, best_y, best_cost, gbx, gby, alpha;
  __shared__ int stop;
  const int m = blockIdx.x, tid = threadIdx.x;
  const GradJobDev& J = jobs[m];
  if (tid == 0) {
    cx = start ? start[2 * m] : J.c0x;
    cy = start ? start[2 * m + 1] : J.c0y;
    best_x = cx;
    best_y = cy;
    best_cost = DBL_MAX;
    gbx = gby = 0.0;
    alpha = J.alpha0;
    stop = 0;
  }
  __syncthreads();
  for (int it = 0; it < steps; ++it) {
    if (tid == 0) {
      D2 h[9];
      closed_form(J, cx, cy, h);
      for (int e = 0; e < 9; ++e) {
        E.H[e] = h[e].v;
        dH[e][0] = h[e].dx;
        dH[e][1] = h[e].dy;
      }
    }
    __syncthreads();
    eval_marker<T>(J, n_samples, true, E);
    if (tid == 0) {
      const bool better = E.ok && E.cost < best_cost;
      if (better) {
        double gx = 0.0, gy = 0.0;
        for (int e = 0; e < 9; ++e) {
          gx += E.gH[e] * dH[e][0];
          gy += E.gH[e] * dH[e][1];
        }
        if (it > 0) alpha *= 1.5;
        best_x = cx;
        best_y = cy;
        best_cost = E.cost;
        gbx = gx;
        gby = gy;
      } else {
        alpha *= 0.5;
      }
      const double gn = sqrt(gbx * gbx + gby * gby);
      if (!(gn > 0.0) || !isfinite(gn) || best_cost == DBL_MAX) {
        stop = 1;
      } else {
        cx = best_x - alpha * gbx / gn;
        cy = best_y - alpha * gby / gn;
      }
    }
    __syncthreads();
    if (stop) break;
  }
  if (tid == 0) {
    centers[2 * m] = (float)best_x;
    centers[2 * m + 1] = (float)best_y;
    cost[m] = best_cost;
  }
}

// Per marker the descent's start: its coarse grid centre when it has one,
// else the job's c0.
__global__ __launch_bounds__(256) void k_grad_start(const GradJobDev* __restrict__ jobs, int n,
                                                    const int* __restrict__ coarse_of,
                                                    const float* __restrict__ coarse,
                                                    float* __restrict__ start) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int j = coarse_of[m];
  start[2 * m] = j >= 0 ? coarse[2 * j] : jobs[m].c0x;
  start[2 * m + 1] = j >= 0 ? coarse[2 * j + 1] : jobs[m].c0y;
}

// One word per pixel: its 2x2 neighbourhood, indices clamped to the image.
__global__ __launch_bounds__(256) void k_quads(const uint8_t* __restrict__ img, int w, int h,
                                               uint32_t* __restrict__ quad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int y = i / w, x = i - y * w;
  const uint8_t* r0 = img + i;
  const int dx = x + 1 < w ? 1 : 0, dy = y + 1 < h ? w : 0;
  quad[i] = (uint32_t)r0[0] | (uint32_t)r0[dx] << 8 | (uint32_t)r0[dy] << 16 |
            (uint32_t)r0[dy + dx] << 24;
}

}  // namespace

int launch_quads(const uint8_t* img, int w, int h, uint32_t* quad, hipStream_t s) {
  if (w < 2 || h < 2) return -1;
  k_quads<<<(w * h + 255) / 256, 256, 0, s>>>(img, w, h, quad);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_grad_start(const GradJobDev* jobs, int n, const int* coarse_of, const float* coarse,
                      float* start, hipStream_t s) {
  if (n <= 0) return 0;
  k_grad_start<<<(n + 255) / 256, 256, 0, s>>>(jobs, n, coarse_of, coarse, start);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_grad_eval(const GradJobDev* jobs, int n, const float* H, int n_samples, bool fp64,
                     double* cost, double* grad, hipStream_t s) {
  if (n <= 0) return 0;
  if (fp64) k_grad_eval<double><<<n, kGradBlock, 0, s>>>(jobs, H, n_samples, cost, grad);
  else k_grad_eval<float><<<n, kGradBlock, 0, s>>>(jobs, H, n_samples, cost, grad);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_grad_homography(const GradJobDev* jobs, int n, const float* centers, float* H,
                           hipStream_t s) {
  if (n <= 0) return 0;
  k_grad_homography<<<(n + 255) / 256, 256, 0, s>>>(jobs, n, centers, H);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_grad_descend(const GradJobDev* jobs, int n, int steps, int n_samples, bool fp64,
                        const float* start, float* centers, double* cost, hipStream_t s) {
  if (n <= 0) return 0;
  if (fp64)
    k_grad_descend<double><<<n, kGradBlock, 0, s>>>(jobs, steps, n_samples, start, centers, cost);
  else
    k_grad_descend<float><<<n, kGradBlock, 0, s>>>(jobs, steps, n_samples, start, centers, cost);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace cctagx
