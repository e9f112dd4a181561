// Small fixed-size float linear algebra for the host geometry.  Sums are
// evaluated left to right; no FMA contraction (the library and the oracle
// are both built with -ffp-contract=off), so host results are identical in
// every binary that compiles these headers.
#pragma once

#include <cmath>
#include <stdexcept>

namespace cctagx {

struct Vec3 {
  float v[3];
  float& operator()(int i) { return v[i]; }
  float operator()(int i) const { return v[i]; }
};

struct Mat3 {
  float m[3][3];
  float& operator()(int i, int j) { return m[i][j]; }
  float operator()(int i, int j) const { return m[i][j]; }
  static Mat3 zero() {
    Mat3 r;
    for (auto& row : r.m)
      for (float& x : row) x = 0.0f;
    return r;
  }
  static Mat3 identity() {
    Mat3 r = zero();
    r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0f;
    return r;
  }
};

inline Mat3 mul(const Mat3& a, const Mat3& b) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}

inline Vec3 mul(const Mat3& a, const Vec3& x) {
  Vec3 r;
  for (int i = 0; i < 3; ++i) r.v[i] = a.m[i][0] * x.v[0] + a.m[i][1] * x.v[1] + a.m[i][2] * x.v[2];
  return r;
}

inline Mat3 transpose(const Mat3& a) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

inline float dot3(const Vec3& a, const Vec3& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }

// Cofactor of (i, j) with cyclic indices: m(i1,j1) m(i2,j2) - m(i1,j2) m(i2,j1).
inline float cofactor3(const Mat3& m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m.m[i1][j1] * m.m[i2][j2] - m.m[i1][j2] * m.m[i2][j1];
}

inline float det3(const Mat3& m) {
  return cofactor3(m, 0, 0) * m.m[0][0] + cofactor3(m, 1, 0) * m.m[1][0] + cofactor3(m, 2, 0) * m.m[2][0];
}

// Adjugate inverse (the cofactor formula Eigen uses for 3x3 matrices).
// Returns false when |det| <= threshold.
inline bool inverse3(const Mat3& m, Mat3& out, float threshold = 0.0f) {
  const float c0 = cofactor3(m, 0, 0), c1 = cofactor3(m, 1, 0), c2 = cofactor3(m, 2, 0);
  const float det = c0 * m.m[0][0] + c1 * m.m[1][0] + c2 * m.m[2][0];
  if (!(std::fabs(det) > threshold)) return false;
  const float invdet = 1.0f / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out.m[i][j] = cofactor3(m, j, i) * invdet;
  return true;
}

inline Mat3 inverse3(const Mat3& m) {
  Mat3 r;
  inverse3(m, r, 0.0f);
  return r;
}

// 2-D point with the reference's Point2d / DirectedPoint2d payload.
struct Pt2 {
  float x = 0, y = 0;
  float dx = 0, dy = 0;  // gradient (DirectedPoint2d)
  Pt2() = default;
  Pt2(float x_, float y_) : x(x_), y(y_) {}
  Pt2(float x_, float y_, float dx_, float dy_) : x(x_), y(y_), dx(dx_), dy(dy_) {}
  Vec3 h() const { return Vec3{{x, y, 1.0f}}; }
};

// Point2d(mT * p) normalised (optimization/conditioner.hpp:66-72).
inline void condition(Pt2& p, const Mat3& T) {
  const Vec3 c = mul(T, p.h());
  p.x = c.v[0] / c.v[2];
  p.y = c.v[1] / c.v[2];
}

// applyHomography (Identification.hpp:102-114).
inline void apply_homography(float& xr, float& yr, const Mat3& H, float x, float y) {
  const float u = H.m[0][0] * x + H.m[0][1] * y + H.m[0][2];
  const float v = H.m[1][0] * x + H.m[1][1] * y + H.m[1][2];
  const float w = H.m[2][0] * x + H.m[2][1] * y + H.m[2][2];
  xr = u / w;
  yr = v / w;
}

}  // namespace cctagx
