// Ellipse geometry of the candidate stage and of identification (host code
// shared by the library and by the oracle's CPU pipeline).
// Reference: src/cctag/geometry/{Ellipse,EllipseFromPoints,Distance,Circle}.*,
// src/cctag/Fitting.cpp, src/cctag/EllipseGrowing.hpp.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "linalg.hpp"

namespace cctagx {

// Ellipse.hpp:24-186: conic matrix plus (centre, a, b, angle); a and b are
// the semi-axes along the ellipse's x and y axes (y-major convention).
class Ellipse {
 public:
  Ellipse() : M_(Mat3::zero()) {}
  explicit Ellipse(const Mat3& m) : M_(m) { compute_parameters(); }
  Ellipse(float cx, float cy, float a, float b, float angle) { init(cx, cy, a, b, angle); }

  const Mat3& matrix() const { return M_; }
  float cx() const { return cx_; }
  float cy() const { return cy_; }
  float a() const { return a_; }
  float b() const { return b_; }
  float angle() const { return angle_; }
  Vec3 center_h() const { return Vec3{{cx_, cy_, 1.0f}}; }

  void set_center(float cx, float cy) { cx_ = cx; cy_ = cy; compute_matrix(); }
  void set_a(float a) {
    if (a < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
    a_ = a;
    compute_matrix();
  }
  void set_b(float b) {
    if (b < 0.f) throw std::invalid_argument("Semi axes must be real positive!");
    b_ = b;
    compute_matrix();
  }
  void set_angle(float angle) { angle_ = angle; compute_matrix(); }
  void set_matrix(const Mat3& m) { M_ = m; compute_parameters(); }
  void set_parameters(float cx, float cy, float a, float b, float angle) { init(cx, cy, a, b, angle); }

  // Ellipse.cpp:94-101: the ellipse of (T^T M) T.
  Ellipse transform(const Mat3& T) const { return Ellipse(mul(mul(transpose(T), M_), T)); }
  // Ellipse.cpp:161-214.
  void canonic_form(Mat3& canonic, Mat3& t_primal, Mat3& t_dual) const;

 private:
  void init(float cx, float cy, float a, float b, float angle);
  void compute_parameters();  // Ellipse.cpp:103-156
  void compute_matrix();      // Ellipse.cpp:216-241
  Mat3 M_;
  float cx_ = 0, cy_ = 0, a_ = 0, b_ = 0, angle_ = 0;
};

// Ellipse.cpp:243-249.
Ellipse scaled(const Ellipse& e, float s);

// EllipseGrowing.hpp:24-98.
bool is_in_ellipse(const Ellipse& e, float x, float y);
bool is_overlapping_ellipses(const Ellipse& e1, const Ellipse& e2);
bool is_in_hull(const Ellipse& qin, const Ellipse& qout, float x, float y);
bool is_on_the_same_side(const Pt2& p1, const Pt2& p2, const Vec3& line);
// EllipseGrowing.cpp:370-384.
void compute_hull(const Ellipse& e, float delta, Ellipse& qin, Ellipse& qout);

// Distance.cpp:24-48 (single point).
float distance_point_ellipse(float x, float y, const Ellipse& q);
// Distance.cpp:323-417: the batched form (AVX2 operation order).
void distance_points_ellipse(std::vector<float>& dist, const std::vector<Pt2>& pts, const Ellipse& q);
// The same distances to the conic x^T Q x = 0 (the ellipse's matrix).
void distance_points_conic(std::vector<float>& dist, const std::vector<Pt2>& pts, const Mat3& Q);

// Outcome of a RANSAC trial's classification (Vote.cpp:486-503) for the
// conic t0 x^2 + 2 t1 xy + t2 y^2 + 2 t3 x + 2 t4 y + 1 with t0 t2 - t1^2 > 0:
// 0 when Ellipse(Q) throws (imaginary semi-axes), 1 when the semi-axis
// ratio a/b lies outside [lo, hi], 2 otherwise.  conic_trial_class_fast
// decides it from the conic's invariants in double and returns -1 where
// the float derivation's rounding could decide otherwise;
// conic_trial_class_exact runs the float derivation itself.
int conic_trial_class_fast(const float t[5], float lo, float hi);
int conic_trial_class_exact(const Mat3& Q, float lo, float hi);

// EllipseFromPoints.cpp:157-197.
int intersect_ellipse_with_line(const Ellipse& e, float y, bool horizontal, float out[2]);
// EllipseFromPoints.cpp:233-269.
int rasterize_ellipse_perimeter(const Ellipse& e);
// EllipseFromPoints.cpp:29-49 (keeps p's gradient); the second form takes the
// ellipse as centre, semi-axes and the double cos / sin of its angle.
Pt2 point_on_ellipse(const Ellipse& e, const Pt2& p);
Pt2 point_on_ellipse(float cx, float cy, float a, float b, double ca, double sa, const Pt2& p);
// Ellipse.cpp:259-298.
void sorted_outer_points(const Ellipse& e, const std::vector<Pt2>& points, std::vector<Pt2>& out,
                         size_t requested);

// Fitting.cpp:44-220 (Halir & Flusser direct least squares); throws
// std::domain_error on degenerate input like the reference.
void fit_ellipse(const std::vector<Pt2>& pts, Ellipse& e);
// Fitting.cpp:337-364.
void fit_circle(const std::vector<Pt2>& pts, Ellipse& e);

// Statistic.hpp:140-145: v[n/2] after sort.
float median_ref(std::vector<float>& v);
// Whether median_ref(v) < bound, decided by one counting pass without
// reordering v: 1 or 0, or -1 when v holds a NaN (no strict weak order;
// the caller then takes median_ref itself).
int median_below(const std::vector<float>& v, float bound);

// PCG32 XSH-RR 64/32 (O'Neill 2014), seeded like pcg32(271828) of
// Statistic.cpp:19, with pcg_extras::bounded_rand.
class Pcg32 {
 public:
  explicit Pcg32(uint64_t seed = 271828u) { reseed(seed); }
  void reseed(uint64_t seed) { state_ = bump(seed + kInc); }
  uint32_t next() {
    const uint64_t old = state_;
    state_ = bump(state_);
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
  }
  uint32_t bounded(uint32_t n) {
    const uint32_t threshold = (uint32_t)(0u - n) % n;
    for (;;) {
      const uint32_t r = next();
      if (r >= threshold) return r % n;
    }
  }
  // Statistic.cpp:17-29: 5 distinct values in [0, n).
  // With n < 5 the reference loops forever; this throws instead.
  void rand_5_k(int perm[5], size_t n) {
    if (n < 5) throw std::domain_error("rand_5_k: fewer than 5 values to draw from");
    for (int i = 0; i < 5; ++i) {
      int r;
      bool dup;
      do {
        r = (int)bounded((uint32_t)n);
        dup = false;
        for (int k = 0; k < i; ++k) dup = dup || perm[k] == r;
      } while (dup);
      perm[i] = r;
    }
  }

 private:
  static constexpr uint64_t kMul = 6364136223846793005ull;
  static constexpr uint64_t kInc = 1442695040888963407ull;
  static uint64_t bump(uint64_t s) { return s * kMul + kInc; }
  uint64_t state_ = 0;
};

}  // namespace cctagx
