// Candidate stage (see candidate.hpp).  The reference runs its three loops
// with tbb::parallel_for (Detection.cpp:616,644,680) over one shared
// EdgePointCollection and a thread_local rand_5_k stream (Statistic.cpp:18),
// so its result depends on the schedule.  Here loop one keeps the serial
// seed order (a seed already linked into an earlier segment is skipped), and
// every loop-two / loop-three task owns its random stream and scratch flags,
// so tasks are independent and the result is the same on any thread count.
#include "candidate.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <deque>
#include <exception>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#ifdef CCTAGX_TRACE
#include <cstdio>
#define TRACE(...) std::fprintf(stderr, __VA_ARGS__)
#else
#define TRACE(...) ((void)0)
#endif

namespace cctagx {

namespace {
constexpr float kPi = 3.14159265358979323846f;
constexpr int kEdgeNotFound = -1, kConvexityLost = -2;
constexpr int kNoWeight = 0, kInvGradWeight = 1;

inline uint32_t packxy(int x, int y) { return ((uint32_t)x << 16) | ((uint32_t)y & 0xFFFFu); }

inline float dist2d(float ax, float ay, float bx, float by) {
  const float dx = bx - ax, dy = by - ay;
  return std::sqrt(dx * dx + dy * dy);
}

inline Pt2 pt_of(const EdgePoint& e) { return Pt2((float)e.x, (float)e.y, e.gx, e.gy); }

// Per-candidate bits of EdgePoint::processed (EllipseGrowing.cpp:332): tasks
// of different candidates touch different bits of the same words.
inline void bits_set(uint64_t& w, uint64_t m) { __atomic_fetch_or(&w, m, __ATOMIC_RELAXED); }
inline void bits_clear(uint64_t& w, uint64_t m) { __atomic_fetch_and(&w, ~m, __ATOMIC_RELAXED); }
inline bool bits_test(const uint64_t& w, uint64_t m) {
  return (__atomic_load_n(&w, __ATOMIC_RELAXED) & m) != 0;
}

// The auxiliary "processed" flags of one addCandidateFlow call
// (Types.hpp:142-155), private to the calling thread and cleared on every
// exit.  Testing a null point throws std::out_of_range like the reference.
class AuxFlags {
 public:
  // n: the point count; expect: about how many points will be set
  AuxFlags(size_t n, size_t expect) : f_(storage()) {
    if (f_.size() < n) f_.resize(n, 0);
    touched_.reserve(expect);
  }
  ~AuxFlags() {
    for (int q : touched_) f_[q] = 0;
  }
  bool test(int i) const {
    if (i < 0) throw std::out_of_range("EdgePointCollection: null point");
    return f_[i] != 0;
  }
  void set(int i) {
    f_[i] = 1;
    touched_.push_back(i);
  }

 private:
  static std::vector<uint8_t>& storage() {
    static thread_local std::vector<uint8_t> v;
    return v;
  }
  std::vector<uint8_t>& f_;
  std::vector<int> touched_;
};
}  // namespace

void EdgeSet::prepare(const LevelEdges& L) {
  w_ = L.w;
  h_ = L.h;
  shape_w_ = L.shape_w;
  shape_h_ = L.shape_h;
  bits_ = L.bits;
  const size_t nw = ((size_t)L.w * L.h + 31) / 32;
  rank_.resize(nw);
  before_ = L.before;
  after_ = L.after;
  vindex_ = L.vindex;
  vlist_ = L.vlist;
  pts.resize(L.n);
  pin_.assign(L.n, 0);
  seeds.assign(L.seeds, L.seeds + L.n_seeds);
}

int32_t EdgeSet::word_edges(size_t w0, size_t w1) const {
  int32_t n = 0;
  for (size_t j = w0; j < w1; ++j) n += __builtin_popcount(bits_[j]);
  return n;
}

void EdgeSet::rank_words(size_t w0, size_t w1, int32_t before) {
  for (size_t j = w0; j < w1; ++j) {
    rank_[j] = before;
    before += __builtin_popcount(bits_[j]);
  }
}


void EdgeSet::fill(const LevelEdges& L, int begin, int end, int min_votes) {
  for (int i = begin; i < end; ++i) {
    EdgePoint& e = pts[i];
    e.x = L.x[i];
    e.y = L.y[i];
    e.gx = (float)L.gx[i];
    e.gy = (float)L.gy[i];
    e.norm = std::sqrt(e.gx * e.gx + e.gy * e.gy);
    e.flow_length = L.flow[i];
    const int nv = L.vindex[i + 1] - L.vindex[i];
    e.is_max = nv >= min_votes ? nv : -1;
    e.processed = 0;
    e.n_segment_out = -1;
  }
}

Marker Marker::make(const Ellipse& outer, std::vector<std::vector<Pt2>> pts, int level,
                    float scale, float quality) {
  Marker m;
  m.id = -1;
  m.center_img = Pt2(outer.cx(), outer.cy());
  m.outer_ellipse = outer;
  m.points = std::move(pts);
  m.homography = Mat3::zero();
  m.quality = quality;
  m.pyramid_level = level;
  m.scale = scale;
  m.outer_ellipse.set_center(m.outer_ellipse.cx() + 0.5f, m.outer_ellipse.cy() + 0.5f);
  m.rescaled_outer_ellipse = scaled(m.outer_ellipse, scale);
  m.status = 0;
  return m;
}

Ellipse Marker::half_circle() const {
  Ellipse a = rescaled_outer_ellipse;
  a.set_a(a.b() * 0.5f);
  a.set_b(a.b() * 0.5f);
  return a;
}

bool Marker::is_equal(const Marker& other) const {
  return is_overlapping_ellipses(half_circle(), other.half_circle());
}

// ---------------------------------------------------------------------------
// Points already visited by one linking (a flat open-addressing set of
// packed coordinates; a linking visits at most a few hundred points).
class VisitedSet {
 public:
  VisitedSet() { slots_.fill(kEmpty); }
  bool contains(uint32_t k) const {
    for (uint32_t i = hash(k);; i = (i + 1) & (kSize - 1)) {
      if (slots_[i] == k) return true;
      if (slots_[i] == kEmpty) return false;
    }
  }
  void insert(uint32_t k) {
    if (count_ * 2 >= kSize) throw std::length_error("VisitedSet full");
    for (uint32_t i = hash(k);; i = (i + 1) & (kSize - 1)) {
      if (slots_[i] == k) return;
      if (slots_[i] == kEmpty) {
        slots_[i] = k;
        ++count_;
        return;
      }
    }
  }

 private:
  static constexpr uint32_t kSize = 2048, kEmpty = 0xFFFFFFFFu;
  static uint32_t hash(uint32_t k) { return (k * 2654435761u) >> (32 - 11); }
  std::array<uint32_t, kSize> slots_;
  size_t count_ = 0;
};

// The segment of one linking: the seed, at most 100 points appended on one
// side and 100 prepended on the other (edge_linking_dir's step bound), in a
// fixed buffer; and the points it flags (the seed, then each direction's
// share of the segment).
class SegBuf {
 public:
  void push_back(int v) {
    if (tail_ == kCap) throw std::length_error("edge linking segment");
    buf_[tail_++] = v;
  }
  void push_front(int v) {
    if (head_ == 0) throw std::length_error("edge linking segment");
    buf_[--head_] = v;
  }
  size_t size() const { return (size_t)(tail_ - head_); }
  const int* begin() const { return buf_ + head_; }
  const int* end() const { return buf_ + tail_; }

 private:
  static constexpr int kCap = 256;
  int buf_[kCap];
  int head_ = kCap / 2, tail_ = kCap / 2;
};

class MarkBuf {
 public:
  void push_back(int v) {
    if (n_ == kCap) throw std::length_error("edge linking marks");
    buf_[n_++] = v;
  }
  const int* begin() const { return buf_; }
  const int* end() const { return buf_ + n_; }

 private:
  static constexpr int kCap = 1 + 2 * 256;
  int buf_[kCap];
  int n_ = 0;
};

// The last `window` points of a linking direction (a ring), whose gradient
// angles the reference keeps.
class PointWindow {
 public:
  explicit PointWindow(size_t window) : w_(window) {
    thread_local std::vector<int> storage;
    if (storage.size() < w_) storage.resize(w_);
    v_ = storage.data();
  }
  void push(int q) {
    size_t at = first_ + n_;
    if (at >= w_) at -= w_;
    v_[at] = q;
    last_ = at;
    if (n_ < w_) ++n_;
    else if (++first_ == w_) first_ = 0;
  }
  size_t size() const { return n_; }
  int front() const { return v_[first_]; }
  int back() const { return v_[last_]; }

 private:
  size_t w_, first_ = 0, n_ = 0, last_ = 0;
  int* v_;
};

// The linking's two decisions on gradient angles.  Both only ask on which
// side of a few directions an angle or an angle difference lies, so away
// from those directions they follow from the gradient vectors themselves:
// the sine and cosine of a difference are the normalised cross and dot
// products (exact integers in double), an angle's sector is its nearest
// multiple of pi/4.  The float angles differ from the exact ones by less
// than 1e-6 rad; within kAngleMargin = 1e-4 of a boundary the float
// expressions are evaluated instead, so every decision is the one the float
// angles give.
constexpr double kAngleMargin = 1e-4;

float gradient_angle_of(float gx, float gy) {
  return std::fmod(std::atan2(gy, gx) + 2.0f * kPi, 2.0f * kPi);
}

int linking_shift(float gxf, float gyf) {
  const double gx = gxf, gy = gyf, n = std::sqrt(gx * gx + gy * gy);
  if (n > 0.0) {
    static const double r = 0.70710678118654752;
    static const double cx[8] = {1, r, 0, -r, -1, -r, 0, r}, cy[8] = {0, r, 1, r, 0, -r, -1, -r};
    int best = 0;
    double d1 = -2.0 * n, d2 = -2.0 * n;  // largest and second largest projections
    for (int k = 0; k < 8; ++k) {
      const double d = gx * cx[k] + gy * cy[k];
      if (d > d1) {
        d2 = d1;
        d1 = d;
        best = k;
      } else if (d > d2) {
        d2 = d;
      }
    }
    // next to a sector boundary the two nearest directions project alike:
    // d1 - d2 = 2 sin(pi/8) n sin(distance to the boundary)
    if (d1 - d2 > 0.7653668647301796 * n * kAngleMargin) return best;
  }
  const float angle = gradient_angle_of(gxf, gyf);
  const int shifting = (int)std::round(((angle + kPi / 4.0f) / (2.0f * kPi)) * 8.0f) - 1;
  return ((shifting % 8) + 8) % 8;
}

bool convexity_lost(float fxf, float fyf, float bxf, float byf, int dir, bool full) {
  const double fx = fxf, fy = fyf, bx = bxf, by = byf;
  const double nn = std::sqrt((fx * fx + fy * fy) * (bx * bx + by * by));
  if (nn > 0.0) {
    const double sn = dir * (by * fx - bx * fy) / nn, cs = (bx * fx + by * fy) / nn;
    const double s_thr = (double)-0.707f;
    if (full) {
      if (std::fabs(sn) > kAngleMargin) return sn < 0.0;
    } else if (std::fabs(sn) > kAngleMargin && std::fabs(cs) > kAngleMargin &&
               std::fabs(sn - s_thr) > kAngleMargin) {
      return ((sn < s_thr) && (cs > 0.0)) || ((sn < 0.0) && (cs < 0.0));
    }
  }
  const float d = gradient_angle_of(bxf, byf) - gradient_angle_of(fxf, fyf);
  if (full) return dir * std::sin(d) < 0.0f;
  const float s = dir * std::sin(d), c = std::cos(d);
  return ((s < -0.707f) && (c > 0.f)) || ((s < 0.f) && (c < 0.f));
}

// Vote.cpp:255-394: convex edge linking from a seed in both directions.
// Points the reference flags processed_in are appended to `marks` (the
// flags themselves are set by the caller, in seed order).
static void edge_linking_dir(const EdgeSet& es, VisitedSet& processed, int p, int dir,
                             SegBuf& seg, size_t window, float average_vote_min,
                             MarkBuf& marks, uint8_t* visiting) {
  PointWindow phi(window);
  size_t i = 0;
  bool found = true;
  int stop = 0;
  const size_t maxLength = 100;
  float averageVote = (float)es.voters_size(p);
  static const int xoff[8] = {1, 1, 0, -1, -1, -1, 0, 1};
  const int yoff[8] = {0, dir * -1, dir * -1, dir * -1, 0, dir * 1, dir * 1, dir * 1};
  while ((i < maxLength) && found && (averageVote >= average_vote_min)) {
    phi.push(p);
    const int shifting = linking_shift(es.pts[p].gx, es.pts[p].gy);
    int j = 0;
    stop = 0;
    while (!stop) {
      if (j >= 8) {
        stop = kEdgeNotFound;
      } else {
        const int k = dir == 1 ? (8 - shifting + j) % 8 : (shifting + j) % 8;
        const int sx = es.pts[p].x + xoff[k];
        const int sy = es.pts[p].y + yoff[k];
        if (sx >= 0 && sx < es.shape_w() && sy >= 0 && sy < es.shape_h() && es.at(sx, sy) >= 0 &&
            !processed.contains(packxy(sx, sy))) {
          const EdgePoint &f = es.pts[phi.front()], &b = es.pts[phi.back()];
          if (convexity_lost(f.gx, f.gy, b.gx, b.gy, dir, phi.size() == window))
            stop = kConvexityLost;
          if (stop == 0) {
            processed.insert(packxy(es.pts[p].x, es.pts[p].y));
            p = es.at(sx, sy);
            if (visiting) __atomic_store_n(&visiting[p], (uint8_t)1, __ATOMIC_RELAXED);
            if (dir > 0) seg.push_back(p);
            else seg.push_front(p);
            averageVote = (averageVote * seg.size() + es.voters_size(p)) / (seg.size() + 1.f);
            stop = 1;
          }
          processed.insert(packxy(es.pts[p].x, es.pts[p].y));
        }
      }
      ++j;
    }
    found = stop == 1;
    ++i;
  }
  if ((i == maxLength) || (stop == kConvexityLost)) {
    if (seg.size() > window) {
      size_t n = 0;
      for (int q : seg) {
        if (n == seg.size() - window) break;
        marks.push_back(q);
        ++n;
      }
    }
  } else if (stop == kEdgeNotFound) {
    for (int q : seg) marks.push_back(q);
  }
}

// visiting (optional): a byte per point, set for every point the linking
// walks onto as it goes.
static void edge_linking(const EdgeSet& es, SegBuf& seg, int pmax, size_t window,
                         float average_vote_min, MarkBuf& marks, uint8_t* visiting) {
  VisitedSet processed;
  if (pmax >= 0) {
    seg.push_back(pmax);
    marks.push_back(pmax);
    processed.insert(packxy(es.pts[pmax].x, es.pts[pmax].y));
    edge_linking_dir(es, processed, pmax, 1, seg, window, average_vote_min, marks, visiting);
    edge_linking_dir(es, processed, pmax, -1, seg, window, average_vote_min, marks, visiting);
  }
}

// Vote.cpp:396-412.
static void children_of(const EdgeSet& es, const std::vector<int>& edges, std::vector<int>& children) {
  size_t voteMax = 1;
  for (int e : edges) voteMax = std::max((int)voteMax, es.voters_size(e));
  for (int e : edges) {
    const int n = es.voters_size(e);
    if (n > 0 && (size_t)n >= voteMax / 14)
      for (const int32_t* v = es.voters_begin(e); v != es.voters_end(e); ++v) children.push_back(*v);
  }
}

// 5x5 partial-pivot LU solve (PartialPivLU: unblocked elimination, first
// largest pivot, column-oriented triangular solves).  Returns false when a
// pivot is zero (determinant zero).
static bool lu_solve5(float A[5][5], float b[5], float x[5]) {
  int perm[5] = {0, 1, 2, 3, 4};
  int ntrans = 0;
  bool singular = false;
  for (int k = 0; k < 5; ++k) {
    int piv = k;
    float best = std::fabs(A[k][k]);
    for (int i = k + 1; i < 5; ++i)
      if (std::fabs(A[i][k]) > best) {
        best = std::fabs(A[i][k]);
        piv = i;
      }
    if (best != 0.f) {
      if (piv != k) {
        for (int j = 0; j < 5; ++j) std::swap(A[k][j], A[piv][j]);
        std::swap(perm[k], perm[piv]);
        ++ntrans;
      }
      for (int i = k + 1; i < 5; ++i) A[i][k] /= A[k][k];
    } else {
      singular = true;
    }
    for (int i = k + 1; i < 5; ++i)
      for (int j = k + 1; j < 5; ++j) A[i][j] -= A[i][k] * A[k][j];
  }
  const float det = (A[0][0] * A[1][1]) * (A[2][2] * (A[3][3] * A[4][4])) * (ntrans % 2 ? -1.f : 1.f);
  if (singular || det == 0.f) return false;
  for (int i = 0; i < 5; ++i) x[i] = b[perm[i]];
  for (int i = 0; i < 5; ++i)
    for (int s = i + 1; s < 5; ++s) x[s] -= x[i] * A[s][i];
  for (int i = 4; i >= 0; --i) {
    x[i] /= A[i][i];
    for (int s = 0; s < i; ++s) x[s] -= x[i] * A[s][i];
  }
  return true;
}

// Vote.cpp:414-576: RANSAC-like ellipse fit over a subsample, then filter.
static void outlier_removal(const EdgeSet& es, const std::vector<int>& children,
                            std::vector<int>& filtered, float& SmFinal, float threshold,
                            int weightedType, size_t maxSize, Pcg32& rng) {
  filtered.reserve(children.size());
  const size_t nSub = std::min(children.size(), maxSize);
  const float step = (float)children.size() / (float)nSub;
  if (nSub < 5) return;
  Ellipse qm;
  float Sm = 10000000.0f;
  const float f = 1.f;
  std::vector<Pt2> pts;
  std::vector<float> weights;
  pts.reserve(nSub + 1);
  if (weightedType == kInvGradWeight) weights.reserve(nSub + 1);
  size_t k = 0, iEdge = 0;
  for (int c : children) {
    if (iEdge == (size_t)(k * step)) {
      ++k;
      const EdgePoint& e = es.pts[c];
      pts.emplace_back((float)e.x, (float)e.y);
      if (weightedType == kInvGradWeight) weights.push_back(255 / e.norm);
    }
    ++iEdge;
  }
  size_t counter = 0, guard = 0;
  std::vector<float> dist;
  while (counter < 70 && guard++ < 100000) {
    int perm[5];
    rng.rand_5_k(perm, pts.size());
    float A[5][5], b[5], t[5];
    for (int i = 0; i < 5; ++i) {
      const float x = pts[perm[i]].x, y = pts[perm[i]].y;
      A[i][0] = x * x;
      A[i][1] = 2.0f * x * y;
      A[i][2] = y * y;
      A[i][3] = 2.0f * f * x;
      A[i][4] = 2.0f * f * y;
      b[i] = -f * f;
    }
    if (!lu_solve5(A, b, t)) {
      ++counter;
      continue;
    }
    if (!(t[0] * t[2] - t[1] * t[1] > 0)) {
      ++counter;
      continue;
    }
    Mat3 Q = Mat3::zero();
    Q.m[0][0] = t[0];
    Q.m[0][1] = Q.m[1][0] = t[1];
    Q.m[0][2] = Q.m[2][0] = t[3];
    Q.m[1][1] = t[2];
    Q.m[1][2] = Q.m[2][1] = t[4];
    Q.m[2][2] = 1.f;
    // Ellipse(Q) is built only for the ratio test and the throw on
    // imaginary semi-axes (a dropped, uncounted trial); the invariants
    // decide both without it except next to a boundary
    int cls = conic_trial_class_fast(t, 0.04f, 25.f);
    if (cls < 0) cls = conic_trial_class_exact(Q, 0.04f, 25.f);
    if (cls == 0) continue;
    if (cls == 1) {
      ++counter;
      continue;
    }
    distance_points_conic(dist, pts, Q);
    if (weightedType != kNoWeight)
      for (size_t i = 0; i < dist.size(); ++i) dist[i] = dist[i] * weights[i];
    // most fits lose to the best so far: a counting pass rejects them
    // with the result the median comparison would give
    const float S = median_below(dist, Sm) != 0 ? median_ref(dist) : Sm;
    if (S < Sm) {
      counter = 0;
      qm = Ellipse(Q);
      Sm = S;
    } else {
      ++counter;
    }
  }
  std::vector<float> vDistFinal;
  vDistFinal.reserve(children.size());
  for (int c : children) {
    const EdgePoint& e = es.pts[c];
    float d = std::numeric_limits<float>::max();
    if (weightedType == kNoWeight) d = distance_point_ellipse((float)e.x, (float)e.y, qm);
    else if (weightedType == kInvGradWeight)
      d = distance_point_ellipse((float)e.x, (float)e.y, qm) * 255 / e.norm;
    if (d < threshold * Sm) {
      filtered.push_back(c);
      vDistFinal.push_back(d);
    }
  }
  if (vDistFinal.empty()) return;
  SmFinal = median_ref(vDistFinal);
}

// Fitting.cpp:228-316.
static float inner_prod_min(const EdgeSet& es, const std::vector<int>& fc, float thr, int& p1,
                            int& p2) {
  int pAngle1 = -1;
  float mn = 1.1f, distMax = 0.f;
  const EdgePoint& p0 = es.pts[fc.front()];
  float normGrad = std::sqrt(p0.gx * p0.gx + p0.gy * p0.gy);
  const float gx0 = p0.gx / normGrad, gy0 = p0.gy / normGrad;
  for (size_t i = 1; i < fc.size(); ++i) {
    const EdgePoint& c = es.pts[fc[i]];
    normGrad = std::sqrt(c.gx * c.gx + c.gy * c.gy);
    const float gx = c.gx / normGrad, gy = c.gy / normGrad;
    const float ip = gx0 * gx + gy0 * gy;
    if (ip <= thr) return ip;
    if (ip < mn) {
      mn = ip;
      pAngle1 = fc[i];
    }
    const float d = dist2d((float)p0.x, (float)p0.y, (float)c.x, (float)c.y);
    if (d > distMax) {
      distMax = d;
      p1 = fc[i];
    }
  }
  const EdgePoint& pa = es.pts[pAngle1];
  normGrad = std::sqrt(pa.gx * pa.gx + pa.gy * pa.gy);
  const float gxmin = pa.gx / normGrad, gymin = pa.gy / normGrad;
  mn = 1.f;
  distMax = 0.f;
  for (size_t i = 0; i < fc.size(); ++i) {
    const EdgePoint& c = es.pts[fc[i]];
    normGrad = std::sqrt(c.gx * c.gx + c.gy * c.gy);
    const float chgx = c.gx / normGrad, chgy = c.gy / normGrad;
    const float ip = gxmin * chgx + gymin * chgy;
    if (ip <= thr) return ip;
    if (ip < mn) mn = ip;
    const EdgePoint& q1 = es.pts[p1];
    const float d = dist2d((float)q1.x, (float)q1.y, (float)c.x, (float)c.y);
    if (d > distMax) {
      distMax = d;
      p2 = fc[i];
    }
  }
  return mn;
}

static std::vector<Pt2> as_points(const EdgeSet& es, const std::vector<int>& idx) {
  std::vector<Pt2> v;
  v.reserve(idx.size());
  for (int i : idx) v.push_back(pt_of(es.pts[i]));
  return v;
}

// EllipseGrowing.cpp:302-325.
static bool ellipse_growing_init(const EdgeSet& es, const std::vector<int>& fc, Ellipse& e) {
  int p1 = fc.front(), p2 = fc.front();
  const float mn = inner_prod_min(es, fc, 0.25f, p1, p2);
  if (mn <= 0.25f) {
    fit_ellipse(as_points(es, fc), e);
    return true;
  }
  fit_circle(as_points(es, fc), e);
  return false;
}

// EllipseGrowing.cpp:327-368, with an explicit stack in the recursion's order
// (the caller's, reused from call to call).
struct HullFrame {
  int x, y, i;
};
static void connected_point(std::vector<int>& pts, uint64_t mask, EdgeSet& es, const Ellipse& qin,
                            const Ellipse& qout, int x0, int y0, std::vector<HullFrame>& stack) {
  static const int xoff[8] = {1, 1, 0, -1, -1, -1, 0, 1};
  static const int yoff[8] = {0, -1, -1, -1, 0, 1, 1, 1};
  using Frame = HullFrame;
  stack.clear();
  bits_set(es.pts[es.at(x0, y0)].processed, mask);
  stack.push_back({x0, y0, 0});
  while (!stack.empty()) {
    Frame& f = stack.back();
    if (f.i >= 8) {
      stack.pop_back();
      continue;
    }
    const int sx = f.x + xoff[f.i], sy = f.y + yoff[f.i];
    ++f.i;
    if (!(sx >= 0 && sx < es.shape_w() && sy >= 0 && sy < es.shape_h())) continue;
    const int ie = es.at(sx, sy);
    if (ie < 0) continue;
    EdgePoint& e = es.pts[ie];
    if (!is_in_hull(qin, qout, (float)e.x, (float)e.y) || bits_test(e.processed, mask)) continue;
    const float ex = qin.cx() - e.x, ey = qin.cy() - e.y;
    if (e.gx * ex + e.gy * ey < 0) {
      pts.push_back(ie);
      bits_set(e.processed, mask);
      stack.push_back({sx, sy, 0});
    }
  }
}

// EllipseGrowing.cpp:386-402.
static void ellipse_hull(EdgeSet& es, std::vector<int>& pts, const Ellipse& e, float delta,
                         uint64_t mask) {
  Ellipse qin, qout;
  compute_hull(e, delta, qin, qout);
  const size_t n0 = pts.size();
  std::vector<HullFrame> stack;
  stack.reserve(64);
  for (size_t i = 0; i < n0; ++i) {
    const EdgePoint& p = es.pts[pts[i]];
    connected_point(pts, mask, es, qin, qout, p.x, p.y, stack);
  }
}

static int count_in_hull(const EdgeSet& es, const std::vector<int>& pts, const Ellipse& e,
                         float delta) {
  Ellipse qin, qout;
  compute_hull(e, delta, qin, qout);
  int n = 0;
  for (int i : pts)
    if (is_in_hull(qin, qout, (float)es.pts[i].x, (float)es.pts[i].y)) ++n;
  return n;
}

// EllipseGrowing.cpp:404-510.
static void ellipse_growing2(EdgeSet& es, const std::vector<int>& fc, std::vector<int>& outer,
                             Ellipse& ellipse, float width, size_t run_id, bool good_init) {
  const uint64_t mask = (uint64_t)1 << run_id;
  outer.reserve(fc.size() * 3);
  for (int c : fc) {
    outer.push_back(c);
    bits_set(es.pts[c].processed, mask);
  }
  int lastSize = 0;
  int nIter = 0;
  if (!good_init) {
    int newSize = (int)outer.size();
    int maxNb = newSize;
    int nIterMax = 0;
    std::vector<std::vector<int>> sets;
    std::vector<Ellipse> ellipses;
    sets.push_back(outer);
    ellipses.push_back(ellipse);
    ++nIter;
    while (newSize - lastSize > 0) {
      lastSize = count_in_hull(es, outer, ellipse, width);
      ellipse_hull(es, outer, ellipse, width, mask);
      sets.push_back(outer);
      ellipses.push_back(ellipse);
      fit_circle(as_points(es, outer), ellipse);
      newSize = count_in_hull(es, outer, ellipse, width);
      if (newSize > maxNb) {
        maxNb = newSize;
        nIterMax = nIter;
      }
      ++nIter;
    }
    outer = sets[nIterMax];
    ellipse = ellipses[nIterMax];
    for (auto& s : sets)
      for (int i : s) bits_clear(es.pts[i].processed, mask);
    for (int i : outer) bits_set(es.pts[i].processed, mask);
  }
  lastSize = 0;
  fit_ellipse(as_points(es, outer), ellipse);
  while ((int)outer.size() - lastSize > 0) {
    lastSize = (int)outer.size();
    ellipse_hull(es, outer, ellipse, width, mask);
    fit_ellipse(as_points(es, outer), ellipse);
  }
}

// EllipseGrowing.cpp:80-203.
static bool add_candidate_flow(const EdgeSet& es, const std::vector<int>& fc,
                               const std::vector<int>& outer_pts, const Ellipse& outer,
                               std::vector<std::vector<Pt2>>& cctagPoints, size_t numCircles) {
  cctagPoints.resize(numCircles);
  cctagPoints.back().reserve(cctagPoints.back().size() + outer_pts.size());
  for (int e : outer_pts) cctagPoints.back().push_back(pt_of(es.pts[e]));
  for (size_t c = 0; c + 1 < numCircles; ++c) cctagPoints[c].reserve(cctagPoints[c].size() + fc.size());
  AuxFlags aux(es.pts.size(), fc.size() * numCircles);
  size_t nGradientOut = 0, nAddedPoint = 0;
  for (int p0 : fc) {
    int dir = -1;
    int p = p0;
    const Pt2 outerPoint((float)es.pts[p0].x, (float)es.pts[p0].y);
    const float a = outerPoint.x - outer.cx();
    const float b = outerPoint.y - outer.cy();
    const Vec3 line{{a, b, -a * outer.cx() - b * outer.cy()}};
    for (size_t j = 1; j < numCircles; ++j) {
      p = dir == -1 ? es.before(p) : es.after(p);
      // a missing link throws out of here (Types.hpp:150-155)
      if (!aux.test(p)) {
        aux.set(p);
        const EdgePoint& ep = es.pts[p];
        const float normGrad = std::sqrt(ep.gx * ep.gx + ep.gy * ep.gy);
        const float ge0 = ep.gx / normGrad, ge1 = ep.gy / normGrad;
        float t0 = outer.cx() - ep.x, t1 = outer.cy() - ep.y;
        const float dc = std::sqrt(t0 * t0 + t1 * t1);
        t0 /= dc;
        t1 /= dc;
        const Pt2 add = pt_of(ep);
        if (is_in_ellipse(outer, add.x, add.y) && is_on_the_same_side(outerPoint, add, line)) {
          if ((float(-dir) * (ge0 * t0 + ge1 * t1) < -0.5f) && (j >= numCircles - 2)) ++nGradientOut;
          cctagPoints[numCircles - j - 1].push_back(add);
          if (j >= numCircles - 2) ++nAddedPoint;
        } else {
          cctagPoints.clear();
          return false;
        }
      }
      dir = -dir;
    }
  }
  if (float(nGradientOut) / float(nAddedPoint) > 0.5f) {
    cctagPoints.clear();
    return false;
  }
  return true;
}

namespace {
struct Candidate {
  int seed = -1;
  std::vector<int> convex_segment;
  std::vector<int> outer_points;
  Ellipse outer;
  std::vector<int> filtered_children;
  int score = 0;
  size_t label = 0;
  float average_received_vote = 0.f;
};
}  // namespace

// Vote.cpp:578-742.
static bool is_another_segment(const EdgeSet& es, Ellipse& outerEllipse, std::vector<int>& outerPts,
                               const Candidate& another,
                               std::vector<std::vector<Pt2>>& cctagPoints, size_t numCircles,
                               float thrMedian, Pcg32& rng) {
  const std::vector<int>& anotherOuter = another.outer_points;
  Ellipse qm;
  float Sm = std::numeric_limits<float>::max();
  const std::vector<Pt2> pts = as_points(es, outerPts);
  const std::vector<Pt2> anotherPts = as_points(es, anotherOuter);
  std::vector<float> distRef;
  distance_points_ellipse(distRef, pts, outerEllipse);
  const float SRef = median_ref(distRef);
  size_t cnt = 0;
  std::vector<float> dist, anotherDist;
  std::vector<Pt2> points;
  points.reserve(8);
  while (cnt < 100) {
    points.clear();
    int perm[5];
    rng.rand_5_k(perm, outerPts.size());
    for (int i = 0; i < 4; ++i) points.emplace_back(pts[perm[i]].x, pts[perm[i]].y);
    rng.rand_5_k(perm, anotherOuter.size());
    for (int i = 0; i < 4; ++i)
      points.emplace_back((float)es.pts[anotherOuter[perm[i]]].x, (float)es.pts[anotherOuter[perm[i]]].y);
    // the fit is outside the try (Vote.cpp:646-647): a degenerate fit ends
    // the search and, through the caller, the candidate
    Ellipse eToto;
    fit_ellipse(points, eToto);
    try {
      Ellipse q(eToto.matrix());
      const float ratio = q.a() / q.b();
      if ((ratio < 0.12) || (ratio > 8)) {
        ++cnt;
        continue;
      }
      // distances are >= 0, so S1 + S2 >= max(S1, S2) after rounding too:
      // either median at or above Sm rejects the fit, decided by counting
      distance_points_ellipse(dist, pts, q);
      if (median_below(dist, Sm) == 0) {
        ++cnt;
        continue;
      }
      const float S1 = median_ref(dist);
      distance_points_ellipse(anotherDist, anotherPts, q);
      if (median_below(anotherDist, Sm) == 0) {
        ++cnt;
        continue;
      }
      const float S2 = median_ref(anotherDist);
      const float S = S1 + S2;
      if (S < Sm) {
        cnt = 0;
        qm = q;
        Sm = S;
      } else {
        ++cnt;
      }
    } catch (...) {
      ++cnt;
    }
  }
  const float thr = 6;
  if (!(Sm < thr * SRef)) return false;
  std::vector<int> tmpPts = outerPts;
  tmpPts.insert(tmpPts.end(), anotherOuter.begin(), anotherOuter.end());
  Ellipse tmpEllipse;
  fit_ellipse(as_points(es, tmpPts), tmpEllipse);
  const float quality = (float)tmpPts.size() / (float)rasterize_ellipse_perimeter(tmpEllipse);
  if (!(quality < 1.1)) return false;
  std::vector<float> vDistFinal;
  for (int p : tmpPts) vDistFinal.push_back(distance_point_ellipse((float)es.pts[p].x, (float)es.pts[p].y, tmpEllipse));
  const float SmFinal = median_ref(vDistFinal);
  if (SmFinal < thrMedian) {
    if (add_candidate_flow(es, another.filtered_children, anotherOuter, tmpEllipse, cctagPoints,
                           numCircles)) {
      outerPts = tmpPts;
      outerEllipse = tmpEllipse;
      return true;
    }
    return false;
  }
  return false;
}

// Detection.cpp:69-113, in two steps.  The linking from a seed depends only
// on the edge points, so it is computed for every seed up front (in
// parallel); the reference's skip of a seed already flagged by an earlier
// seed's linking is then applied in seed order.
namespace {
struct SeedLink {
  SegBuf seg;
  MarkBuf marks;  // points the linking flags processed_in
  std::exception_ptr error;
};
// A seed's linking in loop one.
enum SeedState : int { kFree = 0, kTaken = 1, kLinked = 2 };
}  // namespace

static void link_seed(int seed, const EdgeSet& es, const HostParams& p, SeedLink& out,
                      uint8_t* visiting) {
  edge_linking(es, out.seg, seed, p.window_size_on_inner_elliptic_segment, p.average_vote_min,
               out.marks, visiting);
}

static void construct_flow_component(int seed, EdgeSet& es, SeedLink& link,
                                     std::vector<std::unique_ptr<Candidate>>& loopOne) {
  if (es.test_processed_in(seed)) return;
  for (int q : link.marks) es.set_processed_in(q, true);
  auto cand = std::make_unique<Candidate>();
  cand->seed = seed;
  cand->convex_segment.assign(link.seg.begin(), link.seg.end());
  int nReceivedVote = 0, nVotedPoints = 0;
  for (int q : cand->convex_segment) {
    const int vs = es.voters_size(q);
    nReceivedVote += vs;
    if (vs > 0) ++nVotedPoints;
  }
  cand->average_received_vote = (float)(nReceivedVote * nReceivedVote) / (float)nVotedPoints;
  auto it = std::lower_bound(loopOne.begin(), loopOne.end(), cand,
                             [](const std::unique_ptr<Candidate>& a, const std::unique_ptr<Candidate>& b) {
                               return a->average_received_vote > b->average_received_vote;
                             });
  loopOne.insert(it, std::move(cand));
}

// Detection.cpp:115-267, up to the segment label: what one loop-two task
// computes.  `labeled` is set when the reference would have reached the
// label assignment (enough filtered points, no exception before it).
namespace {
struct LoopTwoTask {
  Candidate cand;
  bool labeled = false;
  bool accepted = false;
};
}  // namespace

static void complete_flow_component(LoopTwoTask& t, EdgeSet& es, size_t run_id,
                                    const HostParams& p) {
  Candidate& cand = t.cand;
  Pcg32 rng(271828u);
  try {
    std::vector<int> children;
    children_of(es, cand.convex_segment, children);
    TRACE("  loop-one seed (%d,%d) seg %zu avg %.1f children %zu\n", es.pts[cand.seed].x, es.pts[cand.seed].y,
          cand.convex_segment.size(), cand.average_received_vote, children.size());
    if (children.size() < (size_t)p.min_points_segment_candidate) return;
    cand.score = (int)children.size();
    float SmFinal = 1e+10f;
    std::vector<int>& filtered = cand.filtered_children;
    outlier_removal(es, children, filtered, SmFinal, p.thresh_robust_estimation_of_outer_ellipse,
                    kInvGradWeight, 60, rng);
    TRACE("    filtered %zu SmFinal %g\n", filtered.size(), SmFinal);
    if (filtered.size() < 5) return;
    t.labeled = true;
    std::vector<int>& outer = cand.outer_points;
    Ellipse& outerEllipse = cand.outer;
    const bool goodInit = ellipse_growing_init(es, filtered, outerEllipse);
    ellipse_growing2(es, filtered, outer, outerEllipse, p.ellipse_growing_elliptic_hull_width,
                     run_id, goodInit);
    std::vector<float> vDist;
    vDist.reserve(outer.size());
    for (int q : outer) vDist.push_back(distance_point_ellipse((float)es.pts[q].x, (float)es.pts[q].y, outerEllipse));
    SmFinal = median_ref(vDist);
    if (SmFinal > p.thr_median_distance_ellipse) return;
    const float quality = (float)outer.size() / (float)rasterize_ellipse_perimeter(outerEllipse);
    if (quality > 1.1) return;
    const float ratio = outerEllipse.a() / outerEllipse.b();
    if ((ratio < 0.05) || (ratio > 20)) return;
    t.accepted = true;
  } catch (const std::exception& ex) {
    TRACE("    exception %s\n", ex.what());
  }
}

// Detection.cpp:164-180: the segment label of each labelled candidate, in
// candidate order -- the first label already carried by one of its
// filtered points, else a new one.
static void assign_labels(EdgeSet& es, std::vector<LoopTwoTask>& tasks,
                          std::vector<Candidate>& loopTwo) {
  size_t nSegmentOut = 0;
  for (LoopTwoTask& t : tasks) {
    if (!t.labeled) continue;
    const std::vector<int>& filtered = t.cand.filtered_children;
    long nSegmentCommon = -1;
    for (int q : filtered)
      if (es.pts[q].n_segment_out != -1) {
        nSegmentCommon = es.pts[q].n_segment_out;
        break;
      }
    const size_t nLabel = nSegmentCommon == -1 ? nSegmentOut++ : (size_t)nSegmentCommon;
    for (int q : filtered) es.pts[q].n_segment_out = (int32_t)nLabel;
    t.cand.label = nLabel;
    if (t.accepted) loopTwo.push_back(std::move(t.cand));
  }
}

// Detection.cpp:278-377.
static void flow_component_assembling(const EdgeSet& es, float& quality, size_t ic,
                                      const std::vector<Candidate>& loopTwo, Ellipse& outerEllipse,
                                      std::vector<int>& outerPts,
                                      std::vector<std::vector<Pt2>>& cctagPoints,
                                      const HostParams& p, Pcg32& rng) {
  const Candidate& cand = loopTwo[ic];
  int score = -1;
  size_t iMax = 0;
  const float ratioExpension = 2.5f;
  const EdgePoint& s = es.pts[cand.seed];
  const Ellipse circle((float)s.x, (float)s.y, s.flow_length * ratioExpension,
                       s.flow_length * ratioExpension, 0.f);
  for (size_t i = 0; i < loopTwo.size(); ++i) {
    if (i == ic) continue;
    const Candidate& another = loopTwo[i];
    if (cand.label == another.label) continue;
    const float r = es.pts[another.seed].flow_length / s.flow_length;
    if (!((r > 0.666) && (r < 1.5))) continue;
    if (is_in_ellipse(circle, (float)es.pts[another.seed].x, (float)es.pts[another.seed].y) &&
        another.score > score) {
      score = another.score;
      iMax = i;
    }
  }
  if (score > 0) {
    const Candidate& sel = loopTwo[iMax];
    if (is_another_segment(es, outerEllipse, outerPts, sel, cctagPoints, p.n_crowns * 2,
                           p.thr_median_distance_ellipse, rng))
      quality = (float)outerPts.size() / (float)rasterize_ellipse_perimeter(outerEllipse);
  }
}

// Detection.cpp:379-566.
static void loop_two_iteration(std::vector<Marker>& markers, const EdgeSet& es,
                               const std::vector<Candidate>& loopTwo, size_t ic, int level,
                               float scale, const HostParams& p, Pcg32& rng) {
  const Candidate& cand = loopTwo[ic];
  std::vector<int> outerPts = cand.outer_points;
  Ellipse outerEllipse = cand.outer;
  std::vector<std::vector<Pt2>> cctagPoints;
  try {
    float quality = (float)outerPts.size() / (float)rasterize_ellipse_perimeter(outerEllipse);
    if (p.search_for_another_segment && (quality > 0.25) && (quality < 0.7))
      flow_component_assembling(es, quality, ic, loopTwo, outerEllipse, outerPts, cctagPoints, p, rng);
    if (!add_candidate_flow(es, cand.filtered_children, cand.outer_points, outerEllipse,
                            cctagPoints, p.n_crowns * 2))
      return;
    const Ellipse rescaleEllipse(outerEllipse.cx(), outerEllipse.cy(), outerEllipse.a() * scale,
                                 outerEllipse.b() * scale, outerEllipse.angle());
    const int realPixelPerimeter = rasterize_ellipse_perimeter(rescaleEllipse);
    const float rs = quality * realPixelPerimeter;
    if (((quality <= 0.35) && (rs >= 300.0)) ||
        ((quality <= 0.45f) && (rs >= 200.0) && (rs < 300.0)) ||
        ((quality <= 0.5f) && (rs >= 100.0) && (rs < 200.0)) ||
        ((quality <= 0.5f) && (rs >= 70.0) && (rs < 100.0)) ||
        ((quality <= 0.96f) && (rs >= 50.0) && (rs < 70.0)) || (rs < 50.0))
      return;
    const float ratio = outerEllipse.a() / outerEllipse.b();
    if (ratio > 8.0 || ratio < 0.125) return;
    Ellipse qin, qout;
    compute_hull(outerEllipse, 3.6f, qin, qout);
    for (int q : outerPts)
      if (!is_in_hull(qin, qout, (float)es.pts[q].x, (float)es.pts[q].y)) return;
    float quality2 = 0;
    for (int q : outerPts) quality2 += es.pts[q].norm;
    quality2 *= scale;
    markers.push_back(Marker::make(outerEllipse, cctagPoints, level, scale, quality2));
  } catch (...) {
  }
}

static void rescale_marker(Marker& m, const EdgeSet& level0);

void detect_candidates(std::vector<EdgeSet>& levels, const std::vector<int>& level_h,
                       const HostParams& p, TaskPool* pool, std::vector<Marker>& markers,
                       HostTimes* times) {
  using Clock = std::chrono::steady_clock;
  auto t0 = Clock::now();
  auto lap = [&](double HostTimes::*field) {
    const auto t1 = Clock::now();
    if (times) times->*field += std::chrono::duration<double, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  const int nl = (int)levels.size();
  // Loop one (Detection.cpp:603-627): seeds by received votes
  // (Multiresolution.cpp:239-242, std::sort: its order among equal votes is
  // the library's; the input order is the order in which the points reached
  // min votes, as on the CPU), then flow components in seed order.  A seed
  // flagged by an earlier seed's linking is skipped, so the flow components
  // are committed in seed order by one task while the others link seeds
  // ahead of it (a linking reads only the edge data).  A linker passes over
  // a seed that the commits so far would flag, or that another linking
  // (finished or under way) has walked onto -- most likely flagged by a
  // commit to come; the committer links any seed nobody has taken.
  std::vector<std::vector<std::unique_ptr<Candidate>>> loopOne(nl);
  std::vector<size_t> n_seeds(nl, 0);
  for (int l = 0; l < nl; ++l) {
    EdgeSet& es = levels[l];
    const size_t nMaxSeeds = (size_t)std::max(level_h[l] / 2, p.maximum_nb_seeds);
    n_seeds[l] = std::min(es.seeds.size(), nMaxSeeds);
  }
  // the sorts on (votes, seed) pairs: introsort's moves depend only on the
  // comparisons' outcomes, so the order (ties included) is the one sorting
  // the seeds through the vote lookup gives
  parallel_for(pool, nl, [&](int l) {
    EdgeSet& es = levels[l];
    if (es.seeds.size() < 2) return;
    std::vector<std::pair<int32_t, int32_t>> kv(es.seeds.size());
    for (size_t i = 0; i < kv.size(); ++i) kv[i] = {es.pts[es.seeds[i]].is_max, es.seeds[i]};
    std::sort(kv.begin(), kv.end(), [](const std::pair<int32_t, int32_t>& a,
                                       const std::pair<int32_t, int32_t>& b) { return a.first > b.first; });
    for (size_t i = 0; i < kv.size(); ++i) es.seeds[i] = kv[i].second;
  });
  {
    std::vector<std::pair<int, int>> order;  // (level, seed rank), level by level
    for (int l = 0; l < nl; ++l)
      for (size_t i = 0; i < n_seeds[l]; ++i) order.emplace_back(l, (int)i);
    const int n = (int)order.size();
    // a linking's result is allocated by the thread that links the seed
    std::vector<std::unique_ptr<SeedLink>> links(n);
    std::unique_ptr<std::atomic<int>[]> state(new std::atomic<int>[n]);  // kFree .. kLinked
    for (int k = 0; k < n; ++k) state[k].store(kFree, std::memory_order_relaxed);
    // the commits' processed_in flags as the linkers see them, and the
    // points some linking has walked onto or would flag if its seed is
    // accepted (a hint only; plain bytes accessed atomically)
    std::vector<std::vector<uint8_t>> flagged(nl), tentative(nl);
    for (int l = 0; l < nl; ++l) {
      flagged[l].assign(levels[l].pts.size(), 0);
      tentative[l].assign(levels[l].pts.size(), 0);
    }
    std::atomic<int> next{0}, committed{0};
    std::atomic<bool> done{false};
    const int tasks = pool ? pool->threads() : 1;
    const int lookahead = 64 * tasks;
    auto link_one = [&](int k) {
      const int l = order[k].first;
      links[k].reset(new SeedLink());
      try {
        link_seed(levels[l].seeds[order[k].second], levels[l], p, *links[k], tentative[l].data());
      } catch (...) {
        links[k]->error = std::current_exception();
      }
      for (const int q : links[k]->marks) __atomic_store_n(&tentative[l][q], (uint8_t)1, __ATOMIC_RELAXED);
      state[k].store(kLinked, std::memory_order_release);
    };
    auto linker = [&] {
      for (;;) {
        if (done.load(std::memory_order_acquire)) return;
        const int k = next.fetch_add(1, std::memory_order_relaxed);
        if (k >= n) return;
        while (k - committed.load(std::memory_order_acquire) > lookahead) {
          if (done.load(std::memory_order_acquire)) return;
          std::this_thread::yield();
        }
        int expect = kFree;
        if (!state[k].compare_exchange_strong(expect, kTaken)) continue;
        // a seed an earlier linking covers is most likely skipped: leave it
        // to the committer, which links it itself if it is not
        const int l = order[k].first, seed = levels[l].seeds[order[k].second];
        if (__atomic_load_n(&flagged[l][seed], __ATOMIC_RELAXED) ||
            __atomic_load_n(&tentative[l][seed], __ATOMIC_RELAXED)) {
          state[k].store(kFree, std::memory_order_release);
          continue;
        }
        link_one(k);
      }
    };
    auto committer = [&] {
      struct Done {
        std::atomic<bool>& d;
        ~Done() { d.store(true, std::memory_order_release); }
      } guard{done};
      for (int k = 0; k < n; ++k) {
        const int l = order[k].first;
        EdgeSet& es = levels[l];
        const int seed = es.seeds[order[k].second];
        if (!es.test_processed_in(seed)) {
          // link it here unless a linker holds it; a linker that passes it
          // over hands it back free
          for (;;) {
            int expect = kFree;
            if (state[k].compare_exchange_strong(expect, kTaken)) {
              link_one(k);
              break;
            }
            if (expect == kLinked) break;
            std::this_thread::yield();
          }
          std::atomic_thread_fence(std::memory_order_acquire);
          SeedLink& link = *links[k];
          if (link.error) std::rethrow_exception(link.error);
          construct_flow_component(seed, es, link, loopOne[l]);
          for (const int q : link.marks) __atomic_store_n(&flagged[l][q], (uint8_t)1, __ATOMIC_RELAXED);
        }
        committed.store(k + 1, std::memory_order_release);
      }
    };
    parallel_for(pool, tasks, [&](int t) {
      if (t == 0) committer();
      else linker();
    });
  }
  lap(&HostTimes::loop_one);
  // Loops two and three and the level-0 re-fit as one flow of tasks
  // (Detection.cpp:629-688, Multiresolution.cpp:319-401): one loop-two task
  // per (level, candidate); a level's segment labels are assigned by the
  // task that ends its loop two, which then queues the level's loop-three
  // tasks; each marker a loop-three task finds above level 0 is queued for
  // its re-fit.  A thread takes loop-two tasks first, then queued ones.  The
  // loop-two tasks of a level flip their own bits of its points' candidate
  // masks, the labels write its points' segment numbers, loop three and the
  // re-fit only read the edge sets, and every task writes only its own
  // candidate or marker, so the results (and their order) are those of the
  // three loops run one after another.
  std::vector<std::vector<LoopTwoTask>> two(nl);
  std::vector<std::pair<int, int>> jobs2;
  for (int l = 0; l < nl; ++l) {
    const size_t n = std::min(loopOne[l].size(), (size_t)p.maximum_nb_candidates_loop_two);
    two[l].resize(n);
    for (size_t i = 0; i < n; ++i) {
      two[l][i].cand = std::move(*loopOne[l][i]);
      jobs2.emplace_back(l, (int)i);
    }
  }
  std::vector<std::vector<Candidate>> loopTwo(nl);
  std::vector<std::vector<std::vector<Marker>>> found(nl);  // [level][candidate]
  struct Queued {
    int l, i, k;  // k < 0: loop three of candidate i of level l; else re-fit of its marker k
  };
  std::mutex qm;
  std::deque<Queued> queue;
  std::exception_ptr error;  // the first exception a task raised (under qm)
  const int n2 = (int)jobs2.size();
  std::atomic<int> next2{0}, pending{n2};  // pending: tasks queued, running or to hand out
  std::unique_ptr<std::atomic<int>[]> left2(new std::atomic<int>[nl]);
  for (int l = 0; l < nl; ++l) left2[l].store((int)two[l].size(), std::memory_order_relaxed);
  using Tick = Clock::duration::rep;
  std::atomic<Tick> end2{0}, end3{0}, labels_ticks{0};
  auto stamp = [](std::atomic<Tick>& at) {
    const Tick now = Clock::now().time_since_epoch().count();
    Tick cur = at.load(std::memory_order_relaxed);
    while (now > cur && !at.compare_exchange_weak(cur, now, std::memory_order_relaxed)) {
    }
  };
  auto guarded = [&](auto&& fn) {
    try {
      fn();
    } catch (...) {
      std::lock_guard<std::mutex> lk(qm);
      if (!error) error = std::current_exception();
    }
  };
  // labels of level l, then its loop-three tasks
  auto finish_level = [&](int l) {
    const auto a = Clock::now();
    guarded([&] { assign_labels(levels[l], two[l], loopTwo[l]); });
    labels_ticks.fetch_add((Clock::now() - a).count(), std::memory_order_relaxed);
    found[l].resize(loopTwo[l].size());
    const int n = (int)loopTwo[l].size();
    if (n == 0) return;
    // the candidates that will search for another segment (the long tasks)
    // go to the front of the queue, as do the re-fits a finished task asks
    // for (the last step of its chain)
    std::vector<uint8_t> long_task(n, 0);
    if (p.search_for_another_segment)
      for (int i = 0; i < n; ++i) {
        const Candidate& c = loopTwo[l][i];
        try {  // an order hint only: a failure leaves the task at the back
          const float quality =
              (float)c.outer_points.size() / (float)rasterize_ellipse_perimeter(c.outer);
          long_task[i] = quality > 0.25 && quality < 0.7;
        } catch (...) {
        }
      }
    pending.fetch_add(n, std::memory_order_relaxed);
    std::lock_guard<std::mutex> lk(qm);
    for (int i = n - 1; i >= 0; --i)
      if (long_task[i]) queue.push_front(Queued{l, i, -1});
    for (int i = 0; i < n; ++i)
      if (!long_task[i]) queue.push_back(Queued{l, i, -1});
  };
  for (int l = 0; l < nl; ++l)
    if (two[l].empty()) finish_level(l);
  auto work = [&](int) {
    for (int spins = 0;;) {
      const int k = next2.fetch_add(1, std::memory_order_relaxed);
      if (k < n2) {
        const int l = jobs2[k].first, i = jobs2[k].second;
        guarded([&] { complete_flow_component(two[l][i], levels[l], (size_t)i, p); });
        if (left2[l].fetch_sub(1, std::memory_order_acq_rel) == 1) finish_level(l);
        stamp(end2);
        pending.fetch_sub(1, std::memory_order_release);
        continue;
      }
      Queued t{-1, -1, -1};
      {
        std::lock_guard<std::mutex> lk(qm);
        if (!queue.empty()) {
          t = queue.front();
          queue.pop_front();
        }
      }
      if (t.l >= 0 && t.k < 0) {
        std::vector<Marker>& out = found[t.l][t.i];
        guarded([&] {
          Pcg32 rng(271828u);
          loop_two_iteration(out, levels[t.l], loopTwo[t.l], (size_t)t.i, t.l,
                             (float)std::pow(2.0, t.l), p, rng);
        });
        int refits = 0;
        for (Marker& m : out) {
          if (m.pyramid_level > 0) ++refits;
          else m.rescaled_outer_ellipse_points = m.points.back();
        }
        if (refits > 0) {
          pending.fetch_add(refits, std::memory_order_relaxed);
          std::lock_guard<std::mutex> lk(qm);
          for (int m = (int)out.size() - 1; m >= 0; --m)
            if (out[m].pyramid_level > 0) queue.push_front(Queued{t.l, t.i, m});
        }
        stamp(end3);
        pending.fetch_sub(1, std::memory_order_release);
        continue;
      }
      if (t.l >= 0) {
        guarded([&] { rescale_marker(found[t.l][t.i][t.k], levels[0]); });
        pending.fetch_sub(1, std::memory_order_release);
        continue;
      }
      if (pending.load(std::memory_order_acquire) == 0) return;
      if (++spins % 64 == 0) std::this_thread::yield();
    }
  };
  const auto t_flow = Clock::now();
  parallel_for(pool, pool ? pool->threads() : 1, work);
  if (error) std::rethrow_exception(error);
  for (int l = nl - 1; l >= 0; --l)
    for (std::vector<Marker>& f : found[l])
      for (Marker& m : f) markers.push_back(std::move(m));
  if (times) {
    // loop two: until its last task ended; loop three: from there until its
    // last task ended; re-fit: the rest.  The labels' own time (inside the
    // other two) is reported beside them.
    using Ms = std::chrono::duration<double, std::milli>;
    const Clock::time_point e2(Clock::duration(end2.load())), e3(Clock::duration(end3.load()));
    const Clock::time_point t_end = Clock::now();
    const Clock::time_point a2 = std::max(e2, t_flow), a3 = std::max(e3, a2);
    times->loop_two += Ms(a2 - t_flow).count();
    times->loop_three += Ms(a3 - a2).count();
    times->rescale += Ms(t_end - a3).count();
    times->labels += Ms(Clock::duration(labels_ticks.load())).count();
  }
}

// Multiresolution.cpp:43-170.
static bool intersect_line_two_ellipses(int y, const Ellipse& qIn, const Ellipse& qOut,
                                        const EdgeSet& es, std::vector<int>& hull) {
  float io[2], ii[2];
  const int no = intersect_ellipse_with_line(qOut, (float)y, true, io);
  const int ni = intersect_ellipse_with_line(qIn, (float)y, true, ii);
  auto take = [&](int x) {
    const int ie = es.at(x, y);
    if (ie >= 0) {
      const EdgePoint& e = es.pts[ie];
      const float c0 = qIn.cx() - e.x, c1 = qIn.cy() - e.y;
      if (e.gx * c0 + e.gy * c1 < 0) hull.push_back(ie);
    }
  };
  const int W = es.shape_w();
  if (no == 2 && ni == 2) {
    const int b1 = std::max(0, (int)io[0]), e1 = std::min(W - 1, (int)ii[0]);
    const int b2 = std::max(0, (int)ii[1]), e2 = std::min(W - 1, (int)io[1]);
    for (int x = b1; x <= e1; ++x) take(x);
    for (int x = b2; x <= e2; ++x) take(x);
  } else if (no == 2 && ni <= 1) {
    const int b = std::max(0, (int)io[0]), e = std::min(W - 1, (int)io[1]);
    for (int x = b; x <= e; ++x) take(x);
  } else if (no == 1 && ni == 0) {
    if (io[0] >= 0 && io[0] < W) take((int)io[0]);
  } else {
    return false;
  }
  return true;
}

// Multiresolution.cpp:319-401: the level-0 re-fit of one marker found above
// level 0 (own rand_5_k stream seeded 271828).
static void rescale_marker(Marker& m, const EdgeSet& level0) {
  Pcg32 rng(271828u);
  const float scale = m.scale;
  Ellipse rescaled = m.rescaled_outer_ellipse;
  std::vector<int> hull;
  Ellipse qIn, qOut;
  compute_hull(rescaled, scale, qIn, qOut);
  const float yCenter = rescaled.cy();
  const int maxY = std::max((int)yCenter, 0);
  const int minY = std::min((int)yCenter, level0.shape_h() - 1);
  for (int y = maxY; y < level0.shape_h(); ++y)
    if (!intersect_line_two_ellipses(y, qIn, qOut, level0, hull)) break;
  for (int y = minY; y >= 0; --y)
    if (!intersect_line_two_ellipses(y, qIn, qOut, level0, hull)) break;
  if (hull.size() < 5) return;
  std::vector<int> refit;
  float SmFinal = 1e+10f;
  outlier_removal(level0, hull, refit, SmFinal, 20.0f, kNoWeight, 60, rng);
  if (refit.size() < 5) return;
  try {
    fit_ellipse(as_points(level0, refit), rescaled);
    std::vector<Pt2> pts;
    for (int q : refit) pts.push_back(pt_of(level0.pts[q]));
    m.center_img = Pt2(m.center_img.x * scale, m.center_img.y * scale);
    m.rescaled_outer_ellipse = rescaled;
    m.rescaled_outer_ellipse_points = pts;
  } catch (...) {
  }
}

std::vector<int> update_markers(const std::vector<Marker>& markers,
                                const std::vector<Ellipse>& half, const std::vector<int>& order) {
  std::vector<int> kept;
  for (const int add : order) {
    const Marker& a = markers[add];
    bool flag = false;
    for (int& cur : kept) {
      const Marker& c = markers[cur];
      if (c.status > 0 && a.status > 0 && is_overlapping_ellipses(half[cur], half[add])) {
        if (a.quality > c.quality) cur = add;
        flag = true;
      }
    }
    if (!flag) kept.push_back(add);
  }
  return kept;
}

}  // namespace cctagx
