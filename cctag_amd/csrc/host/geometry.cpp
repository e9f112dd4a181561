// Host ellipse geometry (see geometry.hpp for the reference files).
//
// Unqualified cos/sin/sqrt calls of the reference's EllipseFromPoints.cpp
// resolve to the C double functions, so those expressions are evaluated in
// double and rounded on assignment, as written there.  Eigen's decompositions
// (EigenSolver<Matrix3f>, JacobiSVD) are restated from their published
// algorithms; bit equality with an Eigen build is not claimed for them.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <utility>

namespace cctagx {

namespace {
constexpr float kPi = 3.14159265358979323846f;

int sign_of(float v) { return (v > 0.f) - (v < 0.f); }
}  // namespace

void Ellipse::init(float cx, float cy, float a, float b, float angle) {
  if (a < 0.f || b < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
  cx_ = cx;
  cy_ = cy;
  a_ = a;
  b_ = b;
  angle_ = angle;
  compute_matrix();
}

void Ellipse::compute_parameters() {
  const float p0 = M_.m[0][0], p1 = 2.f * M_.m[0][1], p2 = M_.m[1][1];
  const float p3 = 2.f * M_.m[0][2], p4 = 2.f * M_.m[1][2], p5 = M_.m[2][2];
  const float thetarad = 0.5f * std::atan2(p1, p0 - p2);
  const float cost = std::cos(thetarad);
  const float sint = std::sin(thetarad);
  const float sin_squared = sint * sint;
  const float cos_squared = cost * cost;
  const float cos_sin = sint * cost;
  const float Ao = p5;
  const float Au = p3 * cost + p4 * sint;
  const float Av = -p3 * sint + p4 * cost;
  const float Auu = p0 * cos_squared + p2 * sin_squared + p1 * cos_sin;
  const float Avv = p0 * sin_squared + p2 * cos_squared - p1 * cos_sin;
  if (Auu == 0 || Avv == 0) {
    cx_ = cy_ = a_ = b_ = angle_ = 0.f;
    return;
  }
  const float tuCentre = -Au / (2.0f * Auu);
  const float tvCentre = -Av / (2.0f * Avv);
  const float wCentre = Ao - Auu * tuCentre * tuCentre - Avv * tvCentre * tvCentre;
  cx_ = tuCentre * cost - tvCentre * sint;
  cy_ = tuCentre * sint + tvCentre * cost;
  const float Ru = -wCentre / Auu;
  const float Rv = -wCentre / Avv;
  const float aAux = std::sqrt(std::fabs(Ru)) * sign_of(Ru);
  const float bAux = std::sqrt(std::fabs(Rv)) * sign_of(Rv);
  if (aAux < 0.f || bAux < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
  a_ = aAux;
  b_ = bAux;
  angle_ = thetarad;
}

void Ellipse::compute_matrix() {
  Mat3 tmp;
  tmp.m[0][0] = std::cos(angle_); tmp.m[0][1] = -std::sin(angle_); tmp.m[0][2] = cx_;
  tmp.m[1][0] = std::sin(angle_); tmp.m[1][1] = std::cos(angle_);  tmp.m[1][2] = cy_;
  tmp.m[2][0] = 0.f;              tmp.m[2][1] = 0.f;               tmp.m[2][2] = 1.f;
  Mat3 diag = Mat3::identity();
  diag.m[0][0] = 1.f / (a_ * a_);
  diag.m[1][1] = 1.f / (b_ * b_);
  diag.m[2][2] = -1.f;
  Mat3 inv;
  if (!inverse3(tmp, inv, 1e-5f)) throw std::domain_error("Singular matrix!");
  M_ = mul(transpose(inv), mul(diag, inv));
}

void Ellipse::canonic_form(Mat3& C, Mat3& Tp, Mat3& Td) const {
  const float q1 = M_.m[0][0], q2 = M_.m[0][1], q3 = M_.m[0][2];
  const float q4 = M_.m[1][1], q5 = M_.m[1][2], q6 = M_.m[2][2];
  const float par1 = q1, par2 = 2 * q2, par3 = q4, par4 = 2 * q3, par5 = 2 * q5;
  const float thetarad = 0.5f * std::atan2(par2, par1 - par3);
  const float cost = std::cos(thetarad);
  const float sint = std::sin(thetarad);
  const float sin_squared = sint * sint;
  const float cos_squared = cost * cost;
  const float cos_sin = sint * cost;
  const float Au = par4 * cost + par5 * sint;
  const float Av = -par4 * sint + par5 * cost;
  const float Auu = par1 * cos_squared + par3 * sin_squared + par2 * cos_sin;
  const float Avv = par1 * sin_squared + par3 * cos_squared - par2 * cos_sin;
  const float tuCentre = -Au / (2 * Auu);
  const float tvCentre = -Av / (2 * Avv);
  const float uCentre = tuCentre * cost - tvCentre * sint;
  const float vCentre = tuCentre * sint + tvCentre * cost;
  const float qt1 = cost * (cost * q1 + q2 * sint) + sint * (cost * q2 + q4 * sint);
  const float qt2 = cost * (cost * q2 + q4 * sint) - sint * (cost * q1 + q2 * sint);
  const float qt3 = cost * q3 + q5 * sint + uCentre * (cost * q1 + q2 * sint) +
                    vCentre * (cost * q2 + q4 * sint);
  const float qt4 = cost * (cost * q4 - q2 * sint) - sint * (cost * q2 - q1 * sint);
  const float qt5 = cost * q5 - q3 * sint + uCentre * (cost * q2 - q1 * sint) +
                    vCentre * (cost * q4 - q2 * sint);
  const float qt6 = q6 + uCentre * (q3 + q1 * uCentre + q2 * vCentre) +
                    vCentre * (q5 + q2 * uCentre + q4 * vCentre) + q3 * uCentre + q5 * vCentre;
  C.m[0][0] = qt1; C.m[0][1] = qt2; C.m[0][2] = qt3;
  C.m[1][0] = qt2; C.m[1][1] = qt4; C.m[1][2] = qt5;
  C.m[2][0] = qt3; C.m[2][1] = qt5; C.m[2][2] = qt6;
  Tp.m[0][0] = cost;  Tp.m[0][1] = sint; Tp.m[0][2] = -cost * uCentre - sint * vCentre;
  Tp.m[1][0] = -sint; Tp.m[1][1] = cost; Tp.m[1][2] = sint * uCentre - cost * vCentre;
  Tp.m[2][0] = 0;     Tp.m[2][1] = 0;    Tp.m[2][2] = cost * cost + sint * sint;
  Td.m[0][0] = cost; Td.m[0][1] = -sint; Td.m[0][2] = uCentre;
  Td.m[1][0] = sint; Td.m[1][1] = cost;  Td.m[1][2] = vCentre;
  Td.m[2][0] = 0;    Td.m[2][1] = 0;     Td.m[2][2] = 1.f;
}

Ellipse scaled(const Ellipse& e, float s) {
  Ellipse r;
  r.set_center(e.cx() * s, e.cy() * s);
  r.set_a(e.a() * s);
  r.set_b(e.b() * s);
  r.set_angle(e.angle());
  return r;
}

bool is_in_ellipse(const Ellipse& e, float x, float y) {
  const Vec3 p{{x, y, 1.f}};
  const float s1 = dot3(p, mul(e.matrix(), p));
  const float s2 = dot3(p, mul(e.matrix(), e.center_h()));
  return s1 * s2 > 0;
}

bool is_overlapping_ellipses(const Ellipse& e1, const Ellipse& e2) {
  return is_in_ellipse(e1, e2.cx(), e2.cy()) || is_in_ellipse(e2, e1.cx(), e1.cy());
}

bool is_in_hull(const Ellipse& qin, const Ellipse& qout, float x, float y) {
  const Vec3 p{{x, y, 1.f}};
  const float s1 = dot3(p, mul(qin.matrix(), p));
  const float s2 = dot3(p, mul(qout.matrix(), p));
  return s1 * s2 < 0;
}

bool is_on_the_same_side(const Pt2& p1, const Pt2& p2, const Vec3& line) {
  const float s1 = dot3(p1.h(), line);
  const float s2 = dot3(p2.h(), line);
  return s1 * s2 > 0;
}

void compute_hull(const Ellipse& e, float delta, Ellipse& qin, Ellipse& qout) {
  qin = Ellipse(e.cx(), e.cy(), std::max(e.a() - delta, 0.001f), std::max(e.b() - delta, 0.001f),
                e.angle());
  qout = Ellipse(e.cx(), e.cy(), e.a() + delta, e.b() + delta, e.angle());
}

float distance_point_ellipse(float x, float y, const Ellipse& q) {
  const Mat3& Q = q.matrix();
  const float a0 = x * x, a1 = 2 * x * y, a2 = 2 * x, a3 = y * y, a4 = 2 * y, a5 = 1;
  const float tmp1 = x * Q.m[0][0] + y * Q.m[0][1] + 1.f * Q.m[0][2];
  const float tmp2 = x * Q.m[0][1] + y * Q.m[1][1] + 1.f * Q.m[1][2];
  const float denom = tmp1 * tmp1 + tmp2 * tmp2;
  const float d = a0 * Q.m[0][0] + a1 * Q.m[0][1] + a2 * Q.m[0][2] + a3 * Q.m[1][1] +
                  a4 * Q.m[1][2] + a5 * Q.m[2][2];
  return (d * d) / denom;
}

void distance_points_ellipse(std::vector<float>& dist, const std::vector<Pt2>& pts,
                             const Ellipse& q) {
  distance_points_conic(dist, pts, q.matrix());
}

void distance_points_conic(std::vector<float>& dist, const std::vector<Pt2>& pts, const Mat3& Q) {
  dist.resize(pts.size());
  for (size_t i = 0; i < pts.size(); ++i) {
    const float x = pts[i].x, y = pts[i].y, w = 1.f;
    const float tmp1 = x * Q.m[0][0] + (y * Q.m[0][1] + w * Q.m[0][2]);
    const float tmp2 = x * Q.m[0][1] + (y * Q.m[1][1] + w * Q.m[1][2]);
    const float denom = tmp1 * tmp1 + tmp2 * tmp2;
    const float twox = x * 2.f, twoy = y * 2.f;
    float d = Q.m[0][0] * (x * x);
    d = d + Q.m[0][1] * (twox * y);
    d = d + Q.m[0][2] * twox;
    d = d + Q.m[1][1] * (y * y);
    d = d + Q.m[1][2] * twoy;
    d = d + Q.m[2][2];
    d = d * d;
    dist[i] = d / denom;
  }
}

// Ellipse::compute_parameters rotates the conic by half the angle of
// (p0 - p2, p1), which makes Auu the larger and Avv the smaller eigenvalue
// of the quadratic part M = [t0 t1; t1 t2]; the centre value is
// w = 1 - u^T M^-1 u with u = (t3, t4), the semi-axes are real unless w has
// the eigenvalues' sign, and (a/b)^2 = Avv/Auu.  Products of two floats are
// exact in double, so these invariants carry only double rounding; the float
// derivation is trusted to within rho, 1000 float ulps scaled by M's
// conditioning s / |Avv|, on Auu, Avv, the ratio and the terms of w (its
// errors, from the rounded rotation angle, the rotated coefficients and
// their squares, stay below about 80 ulps times that conditioning), and
// outcomes closer than that to a boundary are left to it.
int conic_trial_class_fast(const float t[5], float lo, float hi) {
  const double a = t[0], b = t[1], c = t[2], u = t[3], v = t[4];
  const double det = a * c - b * b;
  if (!(det > 0.0)) return -1;
  const double half = 0.5 * (a + c);
  const double rad = std::sqrt(0.25 * (a - c) * (a - c) + b * b);
  const double big = half >= 0.0 ? half + rad : half - rad;  // larger magnitude
  const double small = det / big;
  const double s = std::fabs(a) + std::fabs(c) + 2.0 * std::fabs(b);
  const double lmin = std::fabs(small);
  if (!(s < 1e3 * lmin)) return -1;
  const double rho = 1000.0 * 0x1p-24 * (1.0 + s / lmin);
  // w = 1 - q: the float path's terms Au^2/(4 Auu) and Av^2/(4 Avv) share
  // the sign of q and sum to it, each off by at most rho of q
  const double q = (c * u * u - 2.0 * b * u * v + a * v * v) / det;
  const double w = 1.0 - q;
  if (!(std::fabs(w) > rho * (1.0 + std::fabs(q)))) return -1;
  if ((w > 0.0) == (half > 0.0)) return 0;
  // (a/b)^2: Avv/Auu, at most 1 for a positive definite M, at least 1 else
  const double r = half > 0.0 ? small / big : big / small;
  const double lo2 = (double)lo * (double)lo, hi2 = (double)hi * (double)hi;
  const double m = 2.0 * rho;
  if (r * (1.0 + m) < lo2 || r * (1.0 - m) > hi2) return 1;
  if (r * (1.0 - m) > lo2 && r * (1.0 + m) < hi2) return 2;
  return -1;
}

int conic_trial_class_exact(const Mat3& Q, float lo, float hi) {
  try {
    const Ellipse q(Q);
    const float ratio = q.a() / q.b();
    return (ratio < lo || ratio > hi) ? 1 : 2;
  } catch (const std::invalid_argument&) {
    return 0;
  }
}

int intersect_ellipse_with_line(const Ellipse& e, float y, bool horizontal, float out[2]) {
  const Mat3& C = e.matrix();
  float a, b, c;
  if (horizontal) {
    a = C.m[0][0];
    b = 2 * (y * C.m[0][1] + C.m[0][2]);
    c = C.m[1][1] * (y * y) + 2 * y * C.m[2][1] + C.m[2][2];
  } else {
    a = C.m[1][1];
    b = 2 * (y * C.m[0][1] + C.m[1][2]);
    c = C.m[0][0] * (y * y) + 2 * y * C.m[0][2] + C.m[2][2];
  }
  const float disc = (b * b) / 4.0f - a * c;
  if (disc > 0.f) {
    const float sq = std::sqrt(disc);
    out[0] = (-b / 2.0f - sq) / a;
    out[1] = (-b / 2.0f + sq) / a;
    return 2;
  }
  if (disc == 0.f) {
    out[0] = -b / (2.0f * a);
    return 1;
  }
  return 0;
}

namespace {
// EllipseFromPoints.cpp:67-75, with the double C cos/sin.
void ellipse_point(const Ellipse& e, float theta, float& px, float& py) {
  const float x = (float)((double)e.a() * std::cos((double)theta));
  const float y = (float)((double)e.b() * std::sin((double)theta));
  const double ca = std::cos((double)e.angle()), sa = std::sin((double)e.angle());
  px = (float)((double)x * ca - (double)y * sa + (double)e.cx());
  py = (float)((double)x * sa + (double)y * ca + (double)e.cy());
}
}  // namespace

int rasterize_ellipse_perimeter(const Ellipse& e) {
  // computeIntermediatePoints (EllipseFromPoints.cpp:77-108)
  const float sa = std::sin(e.angle()), ca = std::cos(e.angle());
  float a = -e.b() * sa - e.b() * ca;
  float b = -e.a() * ca + e.a() * sa;
  const float t11 = std::atan2(-a, b);
  const float t12 = t11 + kPi;
  a = -e.b() * sa + e.b() * ca;
  b = -e.a() * ca - e.a() * sa;
  const float t21 = std::atan2(-a, b);
  const float t22 = t21 + kPi;
  float v[4][2];
  ellipse_point(e, t11, v[0][0], v[0][1]);
  ellipse_point(e, t12, v[1][0], v[1][1]);
  ellipse_point(e, t21, v[2][0], v[2][1]);
  ellipse_point(e, t22, v[3][0], v[3][1]);
  int p[4][2];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 2; ++k) p[i][k] = (int)std::round(v[i][k]);
  // pt11 = p[0], pt12 = p[1], pt21 = p[2], pt22 = p[3]
  float diff1, diff2;
  {
    const float mx = (float)std::abs(p[3][0] - p[0][0]);
    const float my = (float)std::abs(p[3][1] - p[0][1]);
    diff1 = mx > my ? mx : my;
  }
  {
    const float mx = (float)std::abs(p[1][0] - p[3][0]);
    const float my = (float)std::abs(p[1][1] - p[3][1]);
    diff2 = mx > my ? mx : my;
  }
  return (int)((diff1 + diff2) * 2.f);
}

Pt2 point_on_ellipse(float cx, float cy, float ea, float eb, double ca, double sa, const Pt2& p) {
  const float x = p.x - cx;
  const float y = p.y - cy;
  float tx = (float)((double)x * ca + (double)y * sa);
  float ty = (float)(-(double)x * sa + (double)y * ca);
  const float cs = (float)std::sqrt((double)(tx * tx / (ea * ea) + ty * ty / (eb * eb)));
  tx /= cs;
  ty /= cs;
  Pt2 r = p;
  r.x = (float)((double)tx * ca - (double)ty * sa + (double)cx);
  r.y = (float)((double)tx * sa + (double)ty * ca + (double)cy);
  return r;
}

Pt2 point_on_ellipse(const Ellipse& e, const Pt2& p) {
  return point_on_ellipse(e.cx(), e.cy(), e.a(), e.b(), std::cos((double)e.angle()),
                          std::sin((double)e.angle()), p);
}

void sorted_outer_points(const Ellipse& e, const std::vector<Pt2>& points, std::vector<Pt2>& out,
                         size_t requested) {
  // The points in the order a sort of (atan2 angle about the centre, index)
  // pairs gives, by a stable least-significant-digit radix sort of the
  // indices on the angles' bit patterns made monotone (-0 folded onto +0,
  // which compares equal to it; atan2 of finite coordinates is never NaN),
  // then every step-th of them.
  const size_t n = points.size();
  out.clear();
  if (n == 0) return;
  std::vector<uint32_t> key(n), idx(n), tmp(n);
  for (size_t i = 0; i < n; ++i) {
    float a = std::atan2(points[i].y - e.cy(), points[i].x - e.cx());
    if (a == 0.f) a = 0.f;
    uint32_t b;
    std::memcpy(&b, &a, 4);
    key[i] = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    idx[i] = (uint32_t)i;
  }
  for (int shift = 0; shift < 32; shift += 8) {
    size_t count[257] = {};
    for (size_t i = 0; i < n; ++i) ++count[((key[i] >> shift) & 255u) + 1];
    if (count[((key[0] >> shift) & 255u) + 1] == n) continue;  // one digit value throughout
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (size_t i = 0; i < n; ++i) tmp[count[(key[idx[i]] >> shift) & 255u]++] = idx[i];
    idx.swap(tmp);
  }
  const size_t n_outer = std::min(requested, n);
  const float step = std::max(1.f, (float)n / (float)(n_outer - 1));
  for (size_t k = 0;; ++k) {
    const float f = k * step;
    if (!(f < (float)n)) break;
    const size_t i = (size_t)f;
    if (i >= n) break;
    out.push_back(points[idx[i]]);
  }
}

float median_ref(std::vector<float>& v) {
  // the element a full sort would put at n/2 (leaves v partially ordered)
  std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
  return v[v.size() / 2];
}

int median_below(const std::vector<float>& v, float bound) {
  // sorted[n/2] < bound exactly when more than n/2 elements are below it
  size_t below = 0;
  bool nan = false;
  for (const float x : v) {
    below += x < bound;
    nan |= x != x;
  }
  if (nan) return -1;
  return below > v.size() / 2 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Eigen-restatements used by the fits.
namespace {

// Real eigenpairs of a 3x3 nonsymmetric matrix (double): eigenvalues from the
// characteristic cubic, eigenvectors as the largest cross product of two rows
// of (M - lambda I), normalised.  Complex pairs are skipped.
int eigen3_real(const double M[3][3], double vals[3], double vecs[3][3]) {
  const double a = -(M[0][0] + M[1][1] + M[2][2]);
  const double b = M[0][0] * M[1][1] - M[0][1] * M[1][0] + M[0][0] * M[2][2] -
                   M[0][2] * M[2][0] + M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                     M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  const double c = -det;
  // depressed cubic t^3 + p t + q with x = t - a/3
  const double p = b - a * a / 3.0;
  const double q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c;
  int n = 0;
  double roots[3];
  const double disc = q * q / 4.0 + p * p * p / 27.0;
  if (p < 0 && disc <= 0) {
    const double r = std::sqrt(-p / 3.0);
    double arg = (3.0 * q) / (2.0 * p) * std::sqrt(-3.0 / p);
    arg = std::max(-1.0, std::min(1.0, arg));
    const double phi = std::acos(arg) / 3.0;
    for (int k = 0; k < 3; ++k) roots[n++] = 2.0 * r * std::cos(phi - 2.0 * M_PI * k / 3.0) - a / 3.0;
  } else {
    const double sq = std::sqrt(std::max(disc, 0.0));
    roots[n++] = std::cbrt(-q / 2.0 + sq) + std::cbrt(-q / 2.0 - sq) - a / 3.0;
  }
  int out = 0;
  for (int k = 0; k < n; ++k) {
    const double l = roots[k];
    // one Newton polish of the root on the characteristic polynomial
    double x = l;
    for (int it = 0; it < 2; ++it) {
      const double f = ((x + a) * x + b) * x + c;
      const double df = (3 * x + 2 * a) * x + b;
      if (df != 0) x -= f / df;
    }
    double A[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = M[i][j] - (i == j ? x : 0.0);
    double best[3] = {0, 0, 0}, bn = -1;
    for (int r0 = 0; r0 < 3; ++r0)
      for (int r1 = r0 + 1; r1 < 3; ++r1) {
        const double cx = A[r0][1] * A[r1][2] - A[r0][2] * A[r1][1];
        const double cy = A[r0][2] * A[r1][0] - A[r0][0] * A[r1][2];
        const double cz = A[r0][0] * A[r1][1] - A[r0][1] * A[r1][0];
        const double nn = cx * cx + cy * cy + cz * cz;
        if (nn > bn) {
          bn = nn;
          best[0] = cx; best[1] = cy; best[2] = cz;
        }
      }
    if (bn <= 0) continue;
    const double inv = 1.0 / std::sqrt(bn);
    vals[out] = x;
    for (int i = 0; i < 3; ++i) vecs[out][i] = best[i] * inv;
    ++out;
  }
  return out;
}

// Jacobi rotation of a symmetric 2x2 block (Golub & Van Loan, sym.schur2), as
// JacobiRotation::makeJacobi.
void make_jacobi(float x, float y, float z, float& c, float& s) {
  const float deno = 2.f * std::fabs(y);
  if (deno < FLT_MIN) {
    c = 1.f;
    s = 0.f;
    return;
  }
  const float tau = (x - z) / deno;
  const float w = std::sqrt(tau * tau + 1.f);
  const float t = tau > 0.f ? 1.f / (tau + w) : 1.f / (tau - w);
  const float sign_t = t > 0.f ? 1.f : -1.f;
  const float n = 1.f / std::sqrt(t * t + 1.f);
  s = -sign_t * (y / std::fabs(y)) * std::fabs(t) * n;
  c = n;
}

// x_i = c x_i + s y_i ; y_i = -s x_i + c y_i
void rotate(float* x, float* y, int stride, float c, float s) {
  for (int i = 0; i < 2; ++i) {
    const float xi = x[i * stride], yi = y[i * stride];
    x[i * stride] = c * xi + s * yi;
    y[i * stride] = -s * xi + c * yi;
  }
}

// Two-sided Jacobi SVD of a 2x2 matrix (JacobiSVD<Matrix2f> with full U):
// singular values descending, U columns signed so the values are positive.
void svd2(const float S[2][2], float vals[2], float U[2][2]) {
  float scale = std::max(std::max(std::fabs(S[0][0]), std::fabs(S[0][1])),
                         std::max(std::fabs(S[1][0]), std::fabs(S[1][1])));
  if (scale == 0.f) scale = 1.f;
  float W[2][2] = {{S[0][0] / scale, S[0][1] / scale}, {S[1][0] / scale, S[1][1] / scale}};
  U[0][0] = 1; U[0][1] = 0; U[1][0] = 0; U[1][1] = 1;
  const float precision = 2.f * FLT_EPSILON;
  float max_diag = std::max(std::fabs(W[0][0]), std::fabs(W[1][1]));
  for (int sweep = 0; sweep < 16; ++sweep) {
    const float threshold = std::max(FLT_MIN, precision * max_diag);
    if (!(std::fabs(W[1][0]) > threshold || std::fabs(W[0][1]) > threshold)) break;
    // real_2x2_jacobi_svd(W, p=1, q=0)
    float m[2][2] = {{W[1][1], W[1][0]}, {W[0][1], W[0][0]}};
    float c1 = 1.f, s1 = 0.f;
    const float t = m[0][0] + m[1][1], d = m[1][0] - m[0][1];
    if (std::fabs(d) >= FLT_MIN) {
      const float u = t / d;
      const float tmp = std::sqrt(1.f + u * u);
      s1 = 1.f / tmp;
      c1 = u / tmp;
    }
    rotate(&m[0][0], &m[1][0], 1, c1, s1);  // m.applyOnTheLeft(0, 1, rot1)
    float cr, sr;
    make_jacobi(m[0][0], m[0][1], m[1][1], cr, sr);
    // j_left = rot1 * j_right^T
    const float cl = c1 * cr - s1 * (-sr);
    const float sl = c1 * (-sr) + s1 * cr;
    rotate(&W[1][0], &W[0][0], 1, cl, sl);   // W.applyOnTheLeft(1, 0, j_left)
    rotate(&U[0][1], &U[0][0], 2, cl, sl);   // U.applyOnTheRight(1, 0, j_left^T)
    rotate(&W[0][1], &W[0][0], 2, cr, -sr);  // W.applyOnTheRight(1, 0, j_right)
    max_diag = std::max(max_diag, std::max(std::fabs(W[0][0]), std::fabs(W[1][1])));
  }
  for (int i = 0; i < 2; ++i) {
    const float v = W[i][i];
    vals[i] = std::fabs(v) * scale;
    if (v < 0) {
      U[0][i] = -U[0][i];
      U[1][i] = -U[1][i];
    }
  }
  if (vals[1] > vals[0]) {
    std::swap(vals[0], vals[1]);
    std::swap(U[0][0], U[0][1]);
    std::swap(U[1][0], U[1][1]);
  }
}

// Smallest-eigenvalue eigenvector of a symmetric 4x4 matrix (cyclic Jacobi,
// double precision).
void min_eigvec4(double A[4][4], double out[4]) {
  double V[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0;
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j) off += A[i][j] * A[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) {
        if (A[p][q] == 0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  int m = 0;
  for (int i = 1; i < 4; ++i)
    if (A[i][i] < A[m][m]) m = i;
  for (int k = 0; k < 4; ++k) out[k] = V[k][m];
}

}  // namespace

void fit_ellipse(const std::vector<Pt2>& pts, Ellipse& e) {
  const size_t n = pts.size();
  if (n < 5) throw std::domain_error("fitEllipse: at least 5 points are needed");
  // get_offset (Fitting.cpp:44-52)
  float sx = 0, sy = 0;
  for (const Pt2& p : pts) {
    sx += p.x;
    sy += p.y;
  }
  const float ox = sx / (float)n, oy = sy / (float)n;
  // scatter matrices (Fitting.cpp:54-86), accumulated in double
  double S1[3][3] = {}, S2[3][3] = {}, S3[3][3] = {};
  for (const Pt2& p : pts) {
    const double x = (double)(p.x - ox), y = (double)(p.y - oy);
    const double q[3] = {x * x, x * y, y * y};
    const double l[3] = {x, y, 1.0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        S1[i][j] += q[i] * q[j];
        S2[i][j] += q[i] * l[j];
        S3[i][j] += l[i] * l[j];
      }
  }
  // fit_solver (Fitting.cpp:88-154)
  const double d3 = S3[0][0] * (S3[1][1] * S3[2][2] - S3[1][2] * S3[2][1]) -
                    S3[0][1] * (S3[1][0] * S3[2][2] - S3[1][2] * S3[2][0]) +
                    S3[0][2] * (S3[1][0] * S3[2][1] - S3[1][1] * S3[2][0]);
  if (!(std::fabs(d3) > 1e-5))
    throw std::domain_error("fit_solver: the input points appear to be linearly dependent");
  double S3i[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
      S3i[i][j] = (S3[i1][j1] * S3[i2][j2] - S3[i1][j2] * S3[i2][j1]) / d3;
    }
  double T[3][3], R[3][3], M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += S3i[i][k] * S2[j][k];  // S3^-1 S2^T
      T[i][j] = -s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = S1[i][j];
      for (int k = 0; k < 3; ++k) s += S2[i][k] * T[k][j];
      R[i][j] = s;
    }
  // C1^-1 = [[0,0,1/2],[0,-1,0],[1/2,0,0]]
  for (int j = 0; j < 3; ++j) {
    M[0][j] = 0.5 * R[2][j];
    M[1][j] = -R[1][j];
    M[2][j] = 0.5 * R[0][j];
  }
  double vals[3], vecs[3][3];
  const int ne = eigen3_real(M, vals, vecs);
  const double eps = std::numeric_limits<float>::epsilon();
  int imin = -1;
  double minv = std::numeric_limits<float>::max();
  for (int i = 0; i < ne; ++i) {
    const double cond = 4 * vecs[i][0] * vecs[i][2] - vecs[i][1] * vecs[i][1];
    if (cond > eps && cond < minv) {
      imin = i;
      minv = cond;
    }
  }
  if (imin == -1) throw std::domain_error("fit_solver: degeneracy");
  float coef[6];
  for (int i = 0; i < 3; ++i) coef[i] = (float)vecs[imin][i];
  for (int i = 0; i < 3; ++i)
    coef[3 + i] = (float)(T[i][0] * coef[0] + T[i][1] * coef[1] + T[i][2] * coef[2]);

  // to_ellipse (Fitting.cpp:158-205)
  const float feps = std::numeric_limits<float>::epsilon();
  float idet = coef[0] * coef[2] - coef[1] * coef[1] / 4;
  idet = idet > feps ? 1.f / idet : 0;
  const float scale = std::sqrt(idet / 4);
  if (scale < feps) throw std::domain_error("to_ellipse_2: singularity 1");
  for (float& c : coef) c *= scale;
  const float aa = coef[0], bb = coef[1], cc = coef[2], dd = coef[3], ee = coef[4];
  float ff = coef[5];
  const float c0 = (-dd * cc + ee * bb / 2) * 2, c1 = (-aa * ee + dd * bb / 2) * 2;
  ff += aa * c0 * c0 + bb * c0 * c1 + cc * c1 * c1 + dd * c0 + ee * c1;
  if (std::fabs(ff) < feps) throw std::domain_error("to_ellipse_2: singularity 2");
  float S[2][2] = {{aa, bb / 2}, {bb / 2, cc}};
  for (auto& row : S)
    for (float& v : row) v /= -ff;
  float sv[2], U[2][2];
  svd2(S, sv, U);
  const float cx = c0 + ox, cy = c1 + oy;
  const float r0 = std::sqrt(1.f / sv[0]), r1 = std::sqrt(1.f / sv[1]);
  const float angle = kPi - std::atan2(U[0][1], U[1][1]);
  if (!(r0 > 0) || !(r1 > 0))
    throw std::domain_error("Degenerate ellipse after fitEllipse => line or point.");
  e.set_parameters(cx, cy, r0, r1, angle);
}

void fit_circle(const std::vector<Pt2>& pts, Ellipse& e) {
  double A[4][4] = {};
  for (const Pt2& p : pts) {
    const float r[4] = {p.x, p.y, 1.f, p.x * p.x + p.y * p.y};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) A[i][j] += (double)r[i] * (double)r[j];
  }
  double v[4];
  min_eigvec4(A, v);
  const float V0 = (float)v[0], V1 = (float)v[1], V2 = (float)v[2], V3 = (float)v[3];
  const float xC = -0.5f * V0 / V3;
  const float yC = -0.5f * V1 / V3;
  const float radius = std::sqrt(xC * xC + yC * yC - V2 / V3);
  if (!(radius > 0.f)) throw std::domain_error("Degenerate circle in circleFitting");
  e.set_parameters(xC, yC, radius, radius, 0);
}

}  // namespace cctagx
