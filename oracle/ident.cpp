// ORACLE -- test infrastructure only (see oracle.hpp).
//
// CPU restatement of the identification of a frame's markers, serial and in
// the reference's loop order (Detection.cpp:873-965):
//   identify_step_1 (Identification.cpp:1172-1301): getSortedOuterPoints
//     (geometry/Ellipse.cpp:259-298), collectCuts / cutInterpolated
//     (:396-496), selectCutCheapUniform (:546-601), outerEdgeRefinement /
//     convImageCut (:604-709), pointOnEllipse (EllipseFromPoints.cpp:29-49);
//   identify_step_2 (:1318-1480): refineConicFamilyGlob (:813-965),
//     imageCenterOptimizationGlob (:980-1058), getNearbyPoints
//     (:1071-1109), conditionerFromEllipse (optimization/conditioner.hpp:
//     22-46), computeHomographyFromEllipseAndImagedCenter (:739-798),
//     Ellipse::getCanonicForm (Ellipse.cpp:161-214), costFunctionGlob
//     (:1122-1158), getSignals / extractSignalUsingHomography (:216-280,
//     :718-728), orazioDistanceRobust (:50-193) and the majority vote,
//     score and status (:1408-1480).
// The geometry (ellipse matrix and parameters, 3x3 inverse, canonical form)
// is the oracle's own, written from the reference's formulas; the sums run
// in the reference's order, including costFunctionGlob's float sum of
// double squares and the sticky out-of-bounds flag of an ImageCut.
//
// Assumptions the reference cannot confirm here (no OpenCV / Eigen / Boost
// build): unqualified cos / sin of float arguments in EllipseFromPoints.cpp
// are the C double functions; Eigen's 3x3 products sum their three terms
// left to right and its 3x3 inverse is the cofactor formula; Boost's
// variance accumulator is the immediate (incremental) mean and variance and
// tag::mean the lazy sum / count.
#include "oracle_ident.hpp"

#include "../cctag_amd/csrc/host/pool.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <stdexcept>
#include <utility>
#include <vector>

namespace oracle {

namespace {

struct M3 {
  float m[3][3];
};

M3 zero3() {
  M3 r;
  for (auto& row : r.m)
    for (float& v : row) v = 0.f;
  return r;
}

// Eigen's coefficient-based 3x3 product: three terms summed left to right.
M3 mul(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}

M3 transposed(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

float cof(const M3& a, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}

// Matrix3f::inverse() / computeInverseWithCheck: cofactors of column 0,
// det = their dot product with column 0, entry (i, j) = cofactor(j, i) / det.
bool invert(const M3& a, M3& out, float threshold) {
  const float c0 = cof(a, 0, 0), c1 = cof(a, 1, 0), c2 = cof(a, 2, 0);
  const float det = c0 * a.m[0][0] + c1 * a.m[1][0] + c2 * a.m[2][0];
  if (threshold > 0.f && !(std::fabs(det) > threshold)) return false;
  const float inv = 1.f / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out.m[i][j] = cof(a, j, i) * inv;
  return true;
}

M3 inverse(const M3& a) {
  M3 r;
  invert(a, r, 0.f);
  return r;
}

// applyHomography (Identification.hpp:102-114).
void apply_h(float& xr, float& yr, const M3& H, float x, float y) {
  const float u = H.m[0][0] * x + H.m[0][1] * y + H.m[0][2];
  const float v = H.m[1][0] * x + H.m[1][1] * y + H.m[1][2];
  const float w = H.m[2][0] * x + H.m[2][1] * y + H.m[2][2];
  xr = u / w;
  yr = v / w;
}

// optimization::condition (conditioner.hpp:60-67).
void condition(float& x, float& y, const M3& T) {
  const float u = T.m[0][0] * x + T.m[0][1] * y + T.m[0][2];
  const float v = T.m[1][0] * x + T.m[1][1] * y + T.m[1][2];
  const float w = T.m[2][0] * x + T.m[2][1] * y + T.m[2][2];
  x = u / w;
  y = v / w;
}

// Ellipse::computeMatrix (Ellipse.cpp:216-241).
M3 ellipse_matrix(float cx, float cy, float a, float b, float angle) {
  if (a < 0.f || b < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
  M3 tmp;
  tmp.m[0][0] = std::cos(angle);
  tmp.m[0][1] = -std::sin(angle);
  tmp.m[0][2] = cx;
  tmp.m[1][0] = std::sin(angle);
  tmp.m[1][1] = std::cos(angle);
  tmp.m[1][2] = cy;
  tmp.m[2][0] = 0.f;
  tmp.m[2][1] = 0.f;
  tmp.m[2][2] = 1.f;
  M3 diag = zero3();
  diag.m[0][0] = 1.f / (a * a);
  diag.m[1][1] = 1.f / (b * b);
  diag.m[2][2] = -1.f;
  M3 inv;
  if (!invert(tmp, inv, 1e-5f)) throw std::domain_error("Singular matrix!");
  return mul(transposed(inv), mul(diag, inv));
}

// Ellipse::computeParameters (Ellipse.cpp:103-156): throws when a semi-axis
// is imaginary.
struct EllParams {
  float cx, cy, a, b, angle;
};
EllParams ellipse_params(const M3& M) {
  const float p0 = M.m[0][0], p1 = 2.f * M.m[0][1], p2 = M.m[1][1];
  const float p3 = 2.f * M.m[0][2], p4 = 2.f * M.m[1][2], p5 = M.m[2][2];
  const float theta = 0.5f * std::atan2(p1, p0 - p2);
  const float c = std::cos(theta), s = std::sin(theta);
  const float ss = s * s, cc = c * c, cs = s * c;
  const float Ao = p5;
  const float Au = p3 * c + p4 * s;
  const float Av = -p3 * s + p4 * c;
  const float Auu = p0 * cc + p2 * ss + p1 * cs;
  const float Avv = p0 * ss + p2 * cc - p1 * cs;
  if (Auu == 0 || Avv == 0) return EllParams{0.f, 0.f, 0.f, 0.f, 0.f};
  const float tu = -Au / (2.0f * Auu);
  const float tv = -Av / (2.0f * Avv);
  const float wc = Ao - Auu * tu * tu - Avv * tv * tv;
  const float Ru = -wc / Auu, Rv = -wc / Avv;
  auto sgn = [](float v) { return (float)((v > 0.f) - (v < 0.f)); };
  const float a = std::sqrt(std::fabs(Ru)) * sgn(Ru);
  const float b = std::sqrt(std::fabs(Rv)) * sgn(Rv);
  if (a < 0.f || b < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
  return EllParams{tu * c - tv * s, tu * s + tv * c, a, b, theta};
}

// Ellipse::getCanonicForm (Ellipse.cpp:161-214).
void canonic_form(const M3& M, M3& C, M3& Tp, M3& Td) {
  const float q1 = M.m[0][0], q2 = M.m[0][1], q3 = M.m[0][2];
  const float q4 = M.m[1][1], q5 = M.m[1][2], q6 = M.m[2][2];
  const float par1 = q1, par2 = 2 * q2, par3 = q4, par4 = 2 * q3, par5 = 2 * q5;
  const float th = 0.5f * std::atan2(par2, par1 - par3);
  const float c = std::cos(th), s = std::sin(th);
  const float ss = s * s, cc = c * c, cs = s * c;
  const float Au = par4 * c + par5 * s;
  const float Av = -par4 * s + par5 * c;
  const float Auu = par1 * cc + par3 * ss + par2 * cs;
  const float Avv = par1 * ss + par3 * cc - par2 * cs;
  const float tu = -Au / (2 * Auu);
  const float tv = -Av / (2 * Avv);
  const float u = tu * c - tv * s;
  const float v = tu * s + tv * c;
  const float qt1 = c * (c * q1 + q2 * s) + s * (c * q2 + q4 * s);
  const float qt2 = c * (c * q2 + q4 * s) - s * (c * q1 + q2 * s);
  const float qt3 = c * q3 + q5 * s + u * (c * q1 + q2 * s) + v * (c * q2 + q4 * s);
  const float qt4 = c * (c * q4 - q2 * s) - s * (c * q2 - q1 * s);
  const float qt5 = c * q5 - q3 * s + u * (c * q2 - q1 * s) + v * (c * q4 - q2 * s);
  const float qt6 = q6 + u * (q3 + q1 * u + q2 * v) + v * (q5 + q2 * u + q4 * v) + q3 * u + q5 * v;
  C.m[0][0] = qt1; C.m[0][1] = qt2; C.m[0][2] = qt3;
  C.m[1][0] = qt2; C.m[1][1] = qt4; C.m[1][2] = qt5;
  C.m[2][0] = qt3; C.m[2][1] = qt5; C.m[2][2] = qt6;
  Tp.m[0][0] = c;  Tp.m[0][1] = s; Tp.m[0][2] = -c * u - s * v;
  Tp.m[1][0] = -s; Tp.m[1][1] = c; Tp.m[1][2] = s * u - c * v;
  Tp.m[2][0] = 0;  Tp.m[2][1] = 0; Tp.m[2][2] = c * c + s * s;
  Td.m[0][0] = c; Td.m[0][1] = -s; Td.m[0][2] = u;
  Td.m[1][0] = s; Td.m[1][1] = c;  Td.m[1][2] = v;
  Td.m[2][0] = 0; Td.m[2][1] = 0;  Td.m[2][2] = 1.f;
}

// computeHomographyFromEllipseAndImagedCenter (Identification.cpp:739-798).
M3 homography(const M3& E, float cx, float cy) {
  M3 C, Tc, Tic;
  canonic_form(E, C, Tc, Tic);
  float xc, yc;
  apply_h(xc, yc, Tc, cx, cy);
  const float Q11 = C.m[0][0], Q22 = C.m[1][1], Q33 = C.m[2][2];
  M3 H;
  H.m[0][0] = Q33;
  H.m[1][0] = 0.f;
  H.m[2][0] = -Q11 * xc;
  H.m[0][1] = Q22 * xc * yc;
  H.m[1][1] = -Q11 * xc * xc - Q33;
  H.m[2][1] = Q22 * yc;
  H.m[0][2] = -Q33 * xc;
  H.m[1][2] = -Q33 * yc;
  H.m[2][2] = -Q33;
  const float d[3] = {std::sqrt((Q22 * Q33 / Q11 * (Q11 * xc * xc + Q22 * yc * yc + Q33))), Q33,
                      std::sqrt(-Q22 * (Q11 * xc * xc + Q33))};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) H.m[i][j] *= d[j];
  return mul(Tic, H);
}

// conditionerFromEllipse (conditioner.hpp:22-46) with the cached mean.
M3 conditioner_of(float cx, float cy, float mean_ab) {
  const float sqrt2 = std::sqrt(2.f);
  M3 T = zero3();
  T.m[0][0] = sqrt2 / mean_ab;
  T.m[0][2] = -sqrt2 * cx / mean_ab;
  T.m[1][1] = sqrt2 / mean_ab;
  T.m[1][2] = -sqrt2 * cy / mean_ab;
  T.m[2][2] = 1.f;
  return T;
}

struct Cut {  // ImageCut.hpp:16-99
  DPt start, stop;
  float begin = 0.f, end = 1.f;
  std::vector<float> sig;
  bool oob = false;
};

// Identification.hpp:197-222: half the bilinear value.
float bilinear(const uint8_t* img, int stride, float x, float y) {
  const int px = (int)x, py = (int)y;
  const uint8_t* p0 = img + px + (size_t)py * stride;
  const float fx = x - px, fy = y - py;
  const float fx1 = 1.0f - fx, fy1 = 1.0f - fy;
  const float w1 = fx1 * fy1, w2 = fx * fy1, w3 = fx1 * fy, w4 = fx * fy;
  return (p0[0] * w1 + p0[1] * w2 + p0[stride] * w3 + p0[stride + 1] * w4) / 2;
}

// Identification.cpp:396-459.
void cut_interpolated(Cut& cut, const uint8_t* img, int w, int h) {
  float xs, ys, xe, ye;
  const float diffX = cut.stop.x - cut.start.x, diffY = cut.stop.y - cut.start.y;
  if (cut.begin != 0.f) {
    xs = cut.start.x + diffX * cut.begin;
    ys = cut.start.y + diffY * cut.begin;
  } else {
    xs = cut.start.x;
    ys = cut.start.y;
  }
  if (cut.end != 1.f) {
    xe = cut.start.x + diffX * cut.end;
    ye = cut.start.y + diffY * cut.end;
  } else {
    xe = cut.stop.x;
    ye = cut.stop.y;
  }
  const size_t n = cut.sig.size();
  const float stepX = (xe - xs) / (n - 1.f), stepY = (ye - ys) / (n - 1.f);
  float x = xs, y = ys;
  for (size_t i = 0; i < n; ++i) {
    if (x >= 1.f && x < w - 1 && y >= 1.f && y < h - 1) {
      cut.sig[i] = bilinear(img, w, x, y);
    } else {
      cut.oob = true;
      break;
    }
    x += stepX;
    y += stepY;
  }
}

// boost::accumulators over [b, e): the immediate mean and variance of
// features<tag::variance>.
void mean_var(const float* b, const float* e, float& mean, float& var) {
  mean = 0.f;
  var = 0.f;
  size_t cnt = 0;
  for (const float* p = b; p != e; ++p) {
    ++cnt;
    mean = (mean * (cnt - 1) + *p) / cnt;
    if (cnt > 1) {
      const float tmp = *p - mean;
      var = var * (cnt - 1) / cnt + tmp * tmp / (cnt - 1);
    }
  }
}

// convImageCut (Identification.cpp:659-709): (max response, its index).
std::pair<float, float> conv_cut(const float* kernel, const Cut& cut) {
  const size_t n = cut.sig.size();
  std::vector<float> out(n);
  for (ptrdiff_t i = 0; i < (ptrdiff_t)n; ++i) {
    float tmp = 0;
    for (size_t j = 0; j < 9; ++j) {
      const ptrdiff_t r = i - 4 + (ptrdiff_t)j;
      if (r < 0) tmp += cut.sig[0] * kernel[j];
      else if (r >= (ptrdiff_t)n) tmp += cut.sig[n - 1] * kernel[j];
      else tmp += cut.sig[r] * kernel[j];
    }
    out[i] = tmp;
  }
  const auto it = std::max_element(out.begin(), out.end());
  return {*it, (float)std::distance(out.begin(), it)};
}

// outerEdgeRefinement (Identification.cpp:604-657).
bool outer_edge_refinement(Cut& cut, const uint8_t* img, int w, int h, float scale, size_t nref) {
  const float len = (float)(3.f * std::sqrt(2.0) * scale);
  const float half = len / 2.f;
  const float gn = std::sqrt(cut.stop.dx * cut.stop.dx + cut.stop.dy * cut.stop.dy);
  const float gdx = cut.stop.dx / gn, gdy = cut.stop.dy / gn;
  const float hx = half * gdx, hy = half * gdy;
  Cut c;
  c.start = DPt{cut.stop.x - hx, cut.stop.y - hy, 0.f, 0.f};
  c.stop = DPt{cut.stop.x + half * gdx, cut.stop.y + half * gdy, cut.stop.dx, cut.stop.dy};
  c.sig.assign(nref, 0.f);
  cut_interpolated(c, img, w, h);
  if (c.oob) return false;
  static const float kA[9] = {-0.0000f, -0.0003f, -0.1065f, -0.7863f, .0f, 0.7863f, 0.1065f, 0.0003f, 0.0000f};
  static const float kB[9] = {-0.0044f, -0.0540f, -0.2376f, -0.3450f, .0f, 0.3450f, 0.2376f, 0.0540f, 0.0044f};
  static const float kC[9] = {-0.0366f, -0.1113f, -0.1801f, -0.1594f, .0f, 0.1594f, 0.1801f, 0.1113f, 0.0366f};
  std::map<float, float> res;  // insert keeps the first kernel of equal maxima
  res.insert(conv_cut(kA, c));
  res.insert(conv_cut(kB, c));
  res.insert(conv_cut(kC, c));
  const float loc = res.rbegin()->second;
  const float step = len / ((float)nref - 1.f);
  cut.stop = DPt{c.start.x + step * loc * gdx, c.start.y + step * loc * gdy, cut.stop.dx, cut.stop.dy};
  return true;
}

// pointOnEllipse (EllipseFromPoints.cpp:29-49), with the C double cos / sin.
DPt point_on_ellipse(const EllParams& e, const DPt& p) {
  const double ca = std::cos((double)e.angle), sa = std::sin((double)e.angle);
  const float x = p.x - e.cx, y = p.y - e.cy;
  float tx = (float)((double)x * ca + (double)y * sa);
  float ty = (float)(-(double)x * sa + (double)y * ca);
  const float cs = (float)std::sqrt((double)(tx * tx / (e.a * e.a) + ty * ty / (e.b * e.b)));
  tx /= cs;
  ty /= cs;
  DPt r = p;
  r.x = (float)((double)tx * ca - (double)ty * sa + (double)e.cx);
  r.y = (float)((double)tx * sa + (double)ty * ca + (double)e.cy);
  return r;
}

// extractSignalUsingHomography (Identification.cpp:216-280): the
// out-of-bounds flag is never cleared, and a sample outside the image keeps
// its previous value.
void extract_signal(Cut& cut, const uint8_t* img, int w, int h, const M3& H, const M3& Hinv) {
  float bx, by;
  apply_h(bx, by, Hinv, cut.stop.x, cut.stop.y);
  float xs, ys, xe, ye;
  if (cut.begin != 0.f) {
    xs = bx * cut.begin;
    ys = by * cut.begin;
  } else {
    xs = 0;
    ys = 0;
  }
  if (cut.end != 1.f) {
    xe = bx * cut.end;
    ye = by * cut.end;
  } else {
    xe = bx;
    ye = by;
  }
  const size_t n = cut.sig.size();
  const float stepX = (xe - xs) / (n - 1.f), stepY = (ye - ys) / (n - 1.f);
  float x = xs, y = ys;
  for (size_t i = 0; i < n; ++i) {
    float xr, yr;
    apply_h(xr, yr, H, x, y);
    if (xr >= 0.f && xr < w - 1 && yr >= 0.f && yr < h - 1) cut.sig[i] = bilinear(img, w, xr, yr);
    else cut.oob = true;
    x += stepX;
    y += stepY;
  }
}

// getSignals (Identification.cpp:718-728).
void get_signals(std::vector<Cut>& cuts, const M3& H, const uint8_t* img, int w, int h) {
  const M3 Hinv = inverse(H);
  for (Cut& c : cuts) extract_signal(c, img, w, h, H, Hinv);
}

// costFunctionGlob (Identification.cpp:1122-1158): res is a float, each
// std::pow(float, 2) a double.
float cost_function(const M3& H, std::vector<Cut>& cuts, const uint8_t* img, int w, int h,
                    bool& flag) {
  flag = true;
  get_signals(cuts, H, img, w, h);
  float res = 0;
  size_t resSize = 0;
  for (size_t i = 0; i + 1 < cuts.size(); ++i)
    for (size_t j = i + 1; j < cuts.size(); ++j)
      if (!cuts[i].oob && !cuts[j].oob) {
        const auto &is = cuts[i].sig, &js = cuts[j].sig;
        for (size_t ii = 0; ii < js.size(); ++ii) res += std::pow(is[ii] - js[ii], 2);
        ++resSize;
      }
  if (resSize == 0) {
    flag = false;
    return std::numeric_limits<float>::max();
  }
  return res / resSize;
}

// computeMedian (Identification.hpp:344-371).
float compute_median(const std::vector<float>& vec) {
  const size_t s = vec.size() / 2 + 1;
  std::vector<float> out(s);
  std::partial_sort_copy(vec.begin(), vec.end(), out.begin(), out.end());
  if (vec.size() % 2) return out.back();
  return (float)((out[s - 1] + out[s - 2]) / 2.0);
}

// conditionerFromEllipse's cached scale (conditioner.hpp:27).
float mean_ab_of(float a, float b) { return (a + b) / 2.f; }

// The per-frame state of conditionerFromEllipse's cached (a + b) / 2.
struct ConditionerCache {
  bool set = false;
  float mean_ab = 0.f;
};

// getNearbyPoints (Identification.cpp:1071-1109) on the GRID pattern.
void nearby_points(const M3& E, const EllParams& e, float cx, float cy, float neighbour, size_t G,
                   ConditionerCache& cache, std::vector<std::pair<float, float>>& out) {
  if (!cache.set) {
    cache.set = true;
    cache.mean_ab = mean_ab_of(e.a, e.b);
  }
  const M3 T = conditioner_of(e.cx, e.cy, cache.mean_ab);
  const M3 Tinv = inverse(T);
  const EllParams te = ellipse_params(mul(mul(transposed(Tinv), E), Tinv));  // projectiveTransform
  neighbour *= std::max(te.a, te.b);
  float ccx = cx, ccy = cy;
  condition(ccx, ccy, T);
  const float halfWidth = neighbour / 2.f;
  const float stepSize = neighbour / (G - 1);
  out.clear();
  for (int i = 0; i < (int)G; ++i)
    for (int j = 0; j < (int)G; ++j) {
      float x = ccx - halfWidth + i * stepSize, y = ccy - halfWidth + j * stepSize;
      condition(x, y, Tinv);
      out.emplace_back(x, y);
    }
}

}  // namespace

// orazioDistanceRobust (Identification.cpp:50-160) for one cut: the
// distance of the signal to every bank row's profile.
void cut_distances(const std::vector<float>& s, float begin, float end,
                   const std::vector<std::vector<float>>& bank, std::vector<float>& out) {
  float medianSig, varSig;
  mean_var(s.data() + 30, s.data() + s.size(), medianSig, varSig);
  float sumInf = 0, sumSup = 0;
  size_t nInf = 0, nSup = 0;
  bool acc = false;
  for (float v : s) {
    if (!acc && v < medianSig) acc = true;
    if (acc) {
      if (v < medianSig) {
        sumInf += v;
        ++nInf;
      } else {
        sumSup += v;
        ++nSup;
      }
    }
  }
  const float muw = sumSup / nSup, mub = sumInf / nInf;
  const float stepX = (end - begin) / (s.size() - 1.f);
  std::vector<float> digit(s.size());
  out.assign(bank.size(), 0.f);
  for (size_t idc = 0; idc < bank.size(); ++idc) {
    float x = begin;
    for (float& d : digit) {
      long ldum = 0;
      for (float j : bank[idc])
        if (1.f / j <= x) ++ldum;
      d = (float)(-(ldum % 2) * 2 + 1);
      x += stepX;
    }
    float distance = 0;
    for (size_t i = 0; i < s.size(); ++i) {  // dis, Identification.hpp:373-383
      float dd;
      if (digit[i] == -1) {
        const float t = std::max(s[i] - mub, 0.f);
        dd = t * t / (2.0f * varSig);
      } else {
        const float t = std::min(s[i] - muw, 0.f);
        dd = t * t / (2.0f * varSig);
      }
      distance += dd;
    }
    out[idc] = distance;
  }
}

namespace {

// orazioDistanceRobust (Identification.cpp:160-190) for one in-bounds cut:
// pushes the best row's exp(-distance) onto its list (a std::map keyed by
// the probability: of equal keys the last row stays).
void orazio_vote(const Cut& cut, const std::vector<std::vector<float>>& bank,
                 std::vector<std::vector<float>>& vScore) {
  std::vector<float> d;
  cut_distances(cut.sig, cut.begin, cut.end, bank, d);
  std::map<float, int> sortedId;
  for (size_t idc = 0; idc < bank.size(); ++idc) sortedId[std::exp(-d[idc])] = (int)idc;
  const auto best = sortedId.rbegin();
  vScore[best->second].push_back(best->first);
}

namespace status {
constexpr int id_reliable = 1;
constexpr int too_few_outer_points = -1;
constexpr int no_collected_cuts = -1;
constexpr int no_selected_cuts = -2;
constexpr int opti_has_diverged = -3;
constexpr int id_not_reliable = -4;
constexpr int degenerate = -5;
}  // namespace status

float signal_begin(int n_crowns) {  // Identification.cpp:1221-1235
  if (n_crowns == 3) return 1 - (2 * n_crowns - 1) * 0.15f;
  if (n_crowns == 4) return 0.26f;
  return 0.f;
}

// identify_step_1 (Identification.cpp:1172-1301): the selected cuts.
int step_1(const IdentMarker& m, const uint8_t* img, int w, int h, const IdentParams& p,
           std::vector<Cut>& selected) {
  std::vector<DPt> outer;
  sorted_outer_points(m.cx, m.cy, m.points, outer, (size_t)p.n_samples_outer_ellipse);
  if (outer.size() < 5) return status::too_few_outer_points;
  const float startSig = signal_begin(p.n_crowns);
  // collectCuts (Identification.cpp:471-496)
  std::vector<Cut> cuts;
  for (const DPt& op : outer) {
    Cut c;
    c.start = DPt{m.cx, m.cy, 0.f, 0.f};
    c.stop = op;
    c.begin = startSig;
    c.end = 1.f;
    c.sig.assign((size_t)p.sample_cut_length, 0.f);
    cut_interpolated(c, img, w, h);
    if (!c.oob) cuts.push_back(std::move(c));
  }
  if (cuts.empty()) return status::no_collected_cuts;
  // selectCutCheapUniform (Identification.cpp:546-601)
  const size_t selectSize = std::min((size_t)p.num_cuts_in_ident_step, cuts.size());
  std::vector<float> var;
  for (const Cut& c : cuts) {
    float mean, v;
    mean_var(c.sig.data(), c.sig.data() + c.sig.size(), mean, v);
    var.push_back(v);
  }
  const float varMax = *std::max_element(var.begin(), var.end());
  const EllParams e{m.cx, m.cy, m.a, m.b, m.angle};
  std::vector<size_t> ind;
  for (size_t i = 0; i < cuts.size(); ++i) {
    Cut& c = cuts[i];
    c.stop = point_on_ellipse(e, c.stop);
    if (outer_edge_refinement(c, img, w, h, m.scale,
                              (size_t)p.num_samples_outer_edge_points_refinement))
      if (var[i] / varMax > 0.5f) ind.push_back(i);
  }
  const float step = std::max(1.f, (float)ind.size() / (float)selectSize);
  selected.clear();
  for (size_t k = 0;; ++k) {
    if ((size_t)(k * step) < ind.size() && selected.size() < selectSize)
      selected.push_back(cuts[ind[(size_t)(k * step)]]);
    else
      break;
  }
  if (selected.empty()) return status::no_selected_cuts;
  return status::id_reliable;
}

void store_h(IdentMarker& m, const M3& H) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.H[i * 3 + j] = H.m[i][j];
}

// identify_step_2 (Identification.cpp:1318-1480) with refineConicFamilyGlob
// (:813-965) and imageCenterOptimizationGlob (:980-1058).
int step_2(IdentMarker& m, std::vector<Cut>& cuts, const uint8_t* img, int w, int h,
           const IdentParams& p, const std::vector<std::vector<float>>& bank,
           ConditionerCache& cache) {
  const M3 E = ellipse_matrix(m.cx, m.cy, m.a, m.b, m.angle);
  const EllParams e{m.cx, m.cy, m.a, m.b, m.angle};
  float residual = std::numeric_limits<float>::max();
  const size_t G = (size_t)p.imaged_center_n_grid_sample;
  float neighbour = p.imaged_center_neighbour_size;
  const float maxSemiAxis = std::max(m.a, m.b);
  float cx = m.center_x, cy = m.center_y;
  M3 H = zero3();
  std::vector<std::pair<float, float>> grid;
  bool converged = true;
  while (neighbour * maxSemiAxis > 0.02) {
    nearby_points(E, e, cx, cy, neighbour, G, cache, grid);
    float minRes = std::numeric_limits<float>::max();
    bool has = false;
    float bx = 0.f, by = 0.f;
    M3 bestH = zero3();
    for (const auto& q : grid) {
      const M3 Ht = homography(E, q.first, q.second);
      bool readable = true;
      const float res = cost_function(Ht, cuts, img, w, h, readable);
      if (readable) {
        has = true;
        if (res < minRes) {
          minRes = res;
          bx = q.first;
          by = q.second;
          bestH = Ht;
        }
      }
    }
    residual = minRes;
    cx = bx;
    cy = by;
    H = bestH;
    if (!has) {
      converged = false;
      break;
    }
    neighbour /= float((G - 1) / 2);
  }
  m.center_x = cx;
  m.center_y = cy;
  store_h(m, H);
  if (converged) {
    get_signals(cuts, H, img, w, h);
    std::vector<size_t> correct;
    for (size_t k = 0; k < cuts.size(); ++k)
      if (!cuts[k].oob) correct.push_back(k);
    const size_t n = cuts[0].sig.size();
    std::vector<float> bar(n), along(correct.size());
    for (size_t s = 0; s < n; ++s) {
      for (size_t k = 0; k < correct.size(); ++k) along[k] = cuts[correct[k]].sig[s];
      bar[s] = correct.empty() ? 0.f : compute_median(along);
    }
    const float vMin = *std::min_element(bar.begin(), bar.end());
    const float vMax = *std::max_element(bar.begin(), bar.end());
    residual = std::sqrt(residual) / (vMax - vMin);
    if (residual > 2.7f) converged = false;
  }
  m.residual = residual;
  m.quality = 1.f / residual;
  if (!converged) return status::opti_has_diverged;
  // orazioDistanceRobust over the in-bounds cuts, then the majority vote
  std::vector<std::vector<float>> vScore(bank.size());
  for (const Cut& c : cuts)
    if (!c.oob) orazio_vote(c, bank, vScore);
  size_t maxSize = 0;
  int iMax = 0;
  for (size_t i = 0; i < vScore.size(); ++i)
    if (vScore[i].size() > maxSize) {
      iMax = (int)i;
      maxSize = vScore[i].size();
    }
  float score = 0;
  for (float pr : vScore[iMax]) score += pr;
  score /= vScore[iMax].size();
  m.id = iMax;
  try {  // the rings' ellipses through the homography (:1440-1458)
    const M3 inv = inverse(H);
    for (float rr : bank[iMax]) {
      const M3 circle = ellipse_matrix(0.f, 0.f, 1.f / rr, 1.f / rr, 0.f);
      (void)ellipse_params(mul(mul(transposed(inv), circle), inv));
    }
  } catch (...) {
    return status::degenerate;
  }
  return score > p.min_ident_proba ? status::id_reliable : status::id_not_reliable;
}

}  // namespace

void sorted_outer_points(float cx, float cy, const std::vector<DPt>& points, std::vector<DPt>& out,
                         size_t requested) {
  std::vector<std::pair<float, size_t>> ang;
  ang.reserve(points.size());
  for (size_t i = 0; i < points.size(); ++i)
    ang.emplace_back(std::atan2(points[i].y - cy, points[i].x - cx), i);
  std::sort(ang.begin(), ang.end());
  const size_t nOuter = std::min(requested, points.size());
  const float step = std::max(1.f, (float)points.size() / (float)(nOuter - 1));
  out.clear();
  for (size_t k = 0;; ++k) {
    const float f = k * step;
    if (!(f < (float)ang.size())) break;  // (size_t)f is then out of range too
    const size_t i = (size_t)f;
    if (i >= ang.size()) break;
    out.push_back(points[ang[i].second]);
  }
}

float start_signal(int n_crowns) { return signal_begin(n_crowns); }

bool sample_cut(const uint8_t* img, int w, int h, const float* H, const DPt& stop, float begin,
                float end, int n, bool homography, float* sig) {
  Cut c;
  c.start = DPt{0.f, 0.f, 0.f, 0.f};
  c.stop = stop;
  c.begin = begin;
  c.end = end;
  c.sig.assign((size_t)n, 0.f);
  if (homography) {
    M3 M;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M.m[i][j] = H[i * 3 + j];
    extract_signal(c, img, w, h, M, inverse(M));
  } else {
    c.start = DPt{H[0], H[1], 0.f, 0.f};
    cut_interpolated(c, img, w, h);
  }
  std::copy(c.sig.begin(), c.sig.end(), sig);
  return c.oob;
}

bool refine_stop(const uint8_t* img, int w, int h, DPt& stop, float scale, int n_ref) {
  Cut c;
  c.stop = stop;
  if (!outer_edge_refinement(c, img, w, h, scale, (size_t)n_ref)) return false;
  stop = c.stop;
  return true;
}

void homography_from_center(float cx, float cy, float a, float b, float angle, float x, float y,
                            float Hout[9]) {
  const M3 H = homography(ellipse_matrix(cx, cy, a, b, angle), x, y);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Hout[i * 3 + j] = H.m[i][j];
}

void conditioner(float cx, float cy, float a, float b, float Tout[9]) {
  const M3 T = conditioner_of(cx, cy, mean_ab_of(a, b));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Tout[i * 3 + j] = T.m[i][j];
}

int grid_iterations(float a, float b, float neighbour_size, int grid_n) {
  // refineConicFamilyGlob's loop (Identification.cpp:869-890), counted
  const float maxSemiAxis = std::max(a, b);
  int it = 0;
  while (neighbour_size * maxSemiAxis > 0.02 && it < 1000) {
    ++it;
    neighbour_size /= float((grid_n - 1) / 2);
  }
  return it;
}

void identify(const uint8_t* img, int w, int h, const IdentParams& p,
              const std::vector<std::vector<float>>& bank, std::vector<IdentMarker>& markers,
              cctagx::TaskPool* pool) {
  const int n = (int)markers.size();
  auto for_each = [&](const std::function<void(int)>& fn) {
    if (pool && pool->threads() > 1) pool->run(n, fn);
    else for (int i = 0; i < n; ++i) fn(i);
  };
  std::vector<std::vector<Cut>> selected(n);
  std::vector<int> detected(n, -1);
  for_each([&](int i) {
    detected[i] = step_1(markers[i], img, w, h, p, selected[i]);
    markers[i].stops.clear();
    for (const Cut& c : selected[i]) markers[i].stops.push_back(c.stop);
  });
  // the serial loop's first grid search (the first marker with selected
  // cuts whose neighbourhood loop runs) sets the conditioner cache
  ConditionerCache cache;
  for (int i = 0; i < n && !cache.set; ++i)
    if (detected[i] == status::id_reliable &&
        grid_iterations(markers[i].a, markers[i].b, p.imaged_center_neighbour_size,
                        p.imaged_center_n_grid_sample) > 0) {
      cache.set = true;
      cache.mean_ab = mean_ab_of(markers[i].a, markers[i].b);
    }
  for_each([&](int i) {
    if (detected[i] == status::id_reliable)
      detected[i] = step_2(markers[i], selected[i], img, w, h, p, bank, cache);
    markers[i].status = detected[i];
  });
}

}  // namespace oracle
