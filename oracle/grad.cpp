// ORACLE -- test infrastructure only.  Float64 restatement of the score the
// differentiable centre refinement evaluates, with its gradient w.r.t. the
// nine homography entries by forward-mode differentiation (one directional
// derivative per entry), independent of the product's adjoint pass.
//
// The score is the grid search's (refineConicFamilyGlob,
// Identification.cpp:813-1158; costFunctionGlob :1122-1158) as a function
// of the homography H (marker plane -> image): for every stop p_k,
// q_k = pi(H^-1 p_k); samples y_ki = pi(H (t_i q_k, 1)) with t_i =
// begin + i (end - begin)/(n - 1); s_ki = half the bilinear value at y_ki
// (Identification.hpp:197-222); a cut is valid when every sample is inside
// [0, w-1) x [0, h-1); C = sum_i sum_{k<l valid} (s_ki - s_li)^2 / pairs.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

bool inv3(const double* m, double* o) {
  const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8],
               c02 = m[3] * m[7] - m[4] * m[6];
  const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  if (det == 0.0 || !std::isfinite(det)) return false;
  o[0] = c00 / det;
  o[1] = (m[2] * m[7] - m[1] * m[8]) / det;
  o[2] = (m[1] * m[5] - m[2] * m[4]) / det;
  o[3] = c01 / det;
  o[4] = (m[0] * m[8] - m[2] * m[6]) / det;
  o[5] = (m[2] * m[3] - m[0] * m[5]) / det;
  o[6] = c02 / det;
  o[7] = (m[1] * m[6] - m[0] * m[7]) / det;
  o[8] = (m[0] * m[4] - m[1] * m[3]) / det;
  return true;
}

void mv3(const double* M, const double* x, double* y) {
  for (int r = 0; r < 3; ++r) y[r] = M[r * 3] * x[0] + M[r * 3 + 1] * x[1] + M[r * 3 + 2] * x[2];
}

}  // namespace

extern "C" {

// Returns the score (DBL_MAX when fewer than two cuts are valid or H is
// singular); grad9 (optional) receives dC/dH, row-major; *n_valid the valid
// cut count.
double oracle_grad_cost(const uint8_t* img, int w, int h, const double* H, const float* stops,
                        int nc, int n_samples, double begin, double end, double* grad9,
                        int* n_valid) {
  double Hi[9];
  if (grad9)
    for (int e = 0; e < 9; ++e) grad9[e] = 0.0;
  if (n_valid) *n_valid = 0;
  if (!inv3(H, Hi)) return DBL_MAX;
  const int ne = grad9 ? 9 : 0;
  std::vector<double> s((size_t)nc * n_samples, 0.0), ds((size_t)nc * n_samples * 9, 0.0);
  std::vector<char> valid(nc, 1);
  for (int k = 0; k < nc; ++k) {
    const double p[3] = {stops[2 * k], stops[2 * k + 1], 1.0};
    double v[3];
    mv3(Hi, p, v);
    const double q[2] = {v[0] / v[2], v[1] / v[2]};
    for (int i = 0; i < n_samples; ++i) {
      const double t = begin + (double)i * (end - begin) / (double)(n_samples - 1);
      const double u[3] = {t * q[0], t * q[1], 1.0};
      double wv[3];
      mv3(H, u, wv);
      const double x = wv[0] / wv[2], y = wv[1] / wv[2];
      if (!(x >= 0.0 && x < (double)(w - 1) && y >= 0.0 && y < (double)(h - 1))) {
        valid[k] = 0;
        continue;
      }
      const int ix = (int)x, iy = (int)y;
      const double fx = x - ix, fy = y - iy;
      const uint8_t* px = img + (size_t)iy * w + ix;
      const double a = px[0], b = px[1], c = px[w], d = px[w + 1];
      s[(size_t)k * n_samples + i] =
          (a * (1 - fx) * (1 - fy) + b * fx * (1 - fy) + c * (1 - fx) * fy + d * fx * fy) / 2.0;
      const double sx = ((b - a) * (1 - fy) + (d - c) * fy) / 2.0;
      const double sy = ((c - a) * (1 - fx) + (d - b) * fx) / 2.0;
      for (int e = 0; e < ne; ++e) {
        // direction dH = E_e (entry e set to 1)
        const int er = e / 3, ec = e % 3;
        // dv = -Hi E_e v ; dq = Jpi(v) dv
        double dv[3];
        for (int r = 0; r < 3; ++r) dv[r] = -Hi[r * 3 + er] * v[ec];
        const double dq0 = (dv[0] * v[2] - v[0] * dv[2]) / (v[2] * v[2]);
        const double dq1 = (dv[1] * v[2] - v[1] * dv[2]) / (v[2] * v[2]);
        const double du[3] = {t * dq0, t * dq1, 0.0};
        // dw = E_e u + H du
        double dw[3];
        mv3(H, du, dw);
        dw[er] += u[ec];
        const double dx = (dw[0] * wv[2] - wv[0] * dw[2]) / (wv[2] * wv[2]);
        const double dy = (dw[1] * wv[2] - wv[1] * dw[2]) / (wv[2] * wv[2]);
        ds[((size_t)k * n_samples + i) * 9 + e] = sx * dx + sy * dy;
      }
    }
  }
  int m = 0;
  for (int k = 0; k < nc; ++k) m += valid[k];
  if (n_valid) *n_valid = m;
  if (m < 2) return DBL_MAX;
  const double pairs = 0.5 * m * (m - 1);
  double cost = 0.0;
  for (int i = 0; i < n_samples; ++i)
    for (int k = 0; k < nc; ++k) {
      if (!valid[k]) continue;
      for (int l = k + 1; l < nc; ++l) {
        if (!valid[l]) continue;
        const double d = s[(size_t)k * n_samples + i] - s[(size_t)l * n_samples + i];
        cost += d * d;
        for (int e = 0; e < ne; ++e)
          grad9[e] += 2.0 * d *
                      (ds[((size_t)k * n_samples + i) * 9 + e] - ds[((size_t)l * n_samples + i) * 9 + e]);
      }
    }
  for (int e = 0; e < ne; ++e) grad9[e] /= pairs;
  return cost / pairs;
}

}  // extern "C"
