// ORACLE -- test infrastructure only (see oracle.hpp and oracle_cand.hpp).
//
// The CCTag candidate stage restated from the reference, one function per
// reference function, in the reference's data layout (an edge-point
// collection with per-point flags and voter lists, std::list segments,
// a recursive hull walk, the linking's float angles), serial within each
// task.  Reference: alicevision/CCTag v1.0.3, src/cctag/.
#include "oracle_cand.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <deque>
#include <limits>
#include <list>
#include <memory>
#include <set>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../cctag_amd/csrc/host/geometry.hpp"
#include "../cctag_amd/csrc/host/pool.hpp"

namespace oracle {
namespace {

using cctagx::Ellipse;
using cctagx::Mat3;
using cctagx::Pcg32;
using cctagx::Pt2;
using cctagx::Vec3;
using Marker = cctagx::Marker;

constexpr float kPi = 3.14159265358979323846f;  // boost::math::constants::pi<float>()

// ---------------------------------------------------------------------------
// EdgePointCollection (Types.hpp:19-158) of one level and the mutable fields
// of its EdgePoints (EdgePoint.hpp:24-79).  The collection's shape is the
// level-0 size (Multiresolution.cpp:286); only the level's own pixels hold
// points.  Point index -1 is the null pointer: the flag accessors throw
// std::out_of_range on it (Types.hpp:51-63, 142-155).
struct Collection {
  const LevelResult* L = nullptr;
  int shape_w = 0, shape_h = 0;
  std::vector<float> norm;             // EdgePoint::_normGrad
  std::vector<int32_t> n_segment_out;  // EdgePoint::_nSegmentOut
  std::vector<uint64_t> processed;     // EdgePoint::_processed (one bit per loop-two run)
  std::vector<uint8_t> processed_in;   // EdgePointCollection::_processedIn

  explicit Collection(const LevelResult& lv) : L(&lv), shape_w(lv.W0), shape_h(lv.H0) {
    const size_t n = lv.ex.size();
    norm.resize(n);
    for (size_t i = 0; i < n; ++i) {
      const float dx = (float)lv.edx[i], dy = (float)lv.edy[i];
      norm[i] = std::sqrt(dx * dx + dy * dy);
    }
    n_segment_out.assign(n, -1);
    processed.assign(n, 0);
    processed_in.assign(n, 0);
  }
  int n() const { return (int)L->ex.size(); }
  int x(int i) const { return L->ex[i]; }
  int y(int i) const { return L->ey[i]; }
  float dX(int i) const { return (float)L->edx[i]; }
  float dY(int i) const { return (float)L->edy[i]; }
  // operator()(x, y): the point at a pixel inside the shape, or -1
  int at(int px, int py) const { return L->edge_at(px, py); }
  int voters_size(int i) const { return L->voters_index[i + 1] - L->voters_index[i]; }
  int before(int i) const { return L->before[check(i)]; }
  int after(int i) const { return L->after[check(i)]; }
  void set_processed_in(int i, bool f) { processed_in[check(i)] = f; }
  bool test_processed_in(int i) const { return processed_in[check(i)] != 0; }
  Pt2 directed(int i) const { return Pt2((float)x(i), (float)y(i), dX(i), dY(i)); }
  static int check(int i) {
    if (i < 0) throw std::out_of_range("EdgePointCollection: null point");
    return i;
  }
  // _processed bit of a loop-two run; runs of one level may be in flight on
  // several threads, each on its own bit of the shared words
  void set_bit(int i, uint64_t m) { __atomic_fetch_or(&processed[i], m, __ATOMIC_RELAXED); }
  void clear_bit(int i, uint64_t m) { __atomic_fetch_and(&processed[i], ~m, __ATOMIC_RELAXED); }
  bool test_bit(int i, uint64_t m) const {
    return (__atomic_load_n(&processed[i], __ATOMIC_RELAXED) & m) != 0;
  }
};

// Candidate.hpp:19-56.
struct Candidate {
  int seed = -1;
  std::list<int> convex_edge_segment;
  std::vector<int> outer_ellipse_points;
  Ellipse outer_ellipse;
  std::vector<int> filtered_children;
  int score = 0;
  size_t n_label = 0;
  float average_received_vote = 0.f;
};

std::vector<Pt2> positions(const Collection& c, const std::vector<int>& idx) {
  std::vector<Pt2> v;
  v.reserve(idx.size());
  for (int i : idx) v.emplace_back((float)c.x(i), (float)c.y(i));
  return v;
}

// ---------------------------------------------------------------------------
// Vote.cpp:249-253.
uint32_t packxy(int x, int y) {
  const unsigned ux = x, uy = y;
  return (ux << 16) | (uy & 0xFFFF);
}

constexpr int kEdgeNotFound = -1;  // Vote.cpp:40
constexpr int kConvexityLost = -2;  // Vote.cpp:41

// edgeLinkingDir (Vote.cpp:272-392).
void edge_linking_dir(Collection& c, std::set<uint32_t>& processed, int p, int dir,
                      std::list<int>& seg, size_t window, float averageVoteMin) {
  std::deque<float> phi;
  size_t i = 0;
  bool found = true;
  int stop = 0;
  const size_t maxLength = 100;
  float averageVote = (float)c.voters_size(p);
  static const int xoff[] = {1, 1, 0, -1, -1, -1, 0, 1};
  while ((i < maxLength) && found && (averageVote >= averageVoteMin)) {
    const float angle =
        std::fmod(float(std::atan2(c.dY(p), c.dX(p))) + 2.0f * kPi, 2.0f * kPi);
    phi.push_back(angle);
    if (phi.size() > window) phi.pop_front();
    const int shifting =
        (int)std::round(((angle + kPi / 4.0f) / (2.0f * kPi)) * 8.0f) - 1;
    const int yoff[] = {0, dir * -1, dir * -1, dir * -1, 0, dir * 1, dir * 1, dir * 1};
    int j = 0;
    stop = 0;
    while (!stop) {
      if (j >= 8) {
        stop = kEdgeNotFound;
      } else {
        int sx, sy;
        if (dir == 1) {
          sx = c.x(p) + xoff[(8 - shifting + j) % 8];
          sy = c.y(p) + yoff[(8 - shifting + j) % 8];
        } else {
          sx = c.x(p) + xoff[(shifting + j) % 8];
          sy = c.y(p) + yoff[(shifting + j) % 8];
        }
        if (sx >= 0 && sx < c.shape_w && sy >= 0 && sy < c.shape_h && c.at(sx, sy) >= 0 &&
            processed.find(packxy(sx, sy)) == processed.end()) {
          if (phi.size() == window) {
            if (dir * std::sin(phi.back() - phi.front()) < 0.0f) stop = kConvexityLost;
          } else {
            const float s = dir * std::sin(phi.back() - phi.front());
            const float co = std::cos(phi.back() - phi.front());
            if (((s < -0.707f) && (co > 0.f)) || ((s < 0.f) && (co < 0.f))) stop = kConvexityLost;
          }
          if (stop == 0) {
            processed.insert(packxy(c.x(p), c.y(p)));
            p = c.at(sx, sy);
            if (dir > 0) seg.push_back(p);
            else seg.push_front(p);
            averageVote = (averageVote * seg.size() + c.voters_size(p)) / (seg.size() + 1.f);
            stop = 1;
          }
          processed.insert(packxy(c.x(p), c.y(p)));
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
        c.set_processed_in(q, true);
        ++n;
      }
    }
  } else if (stop == kEdgeNotFound) {
    for (int q : seg) c.set_processed_in(q, true);
  }
}

// edgeLinking (Vote.cpp:255-270).
void edge_linking(Collection& c, std::list<int>& seg, int pmax, size_t window,
                  float averageVoteMin) {
  std::set<uint32_t> processed;
  if (pmax < 0) return;
  seg.push_back(pmax);
  c.set_processed_in(pmax, true);
  processed.insert(packxy(c.x(pmax), c.y(pmax)));
  edge_linking_dir(c, processed, pmax, 1, seg, window, averageVoteMin);
  edge_linking_dir(c, processed, pmax, -1, seg, window, averageVoteMin);
}

// constructFlowComponentFromSeed (Detection.cpp:69-113).
void construct_flow_component(int seed, Collection& c,
                              std::vector<std::unique_ptr<Candidate>>& loopOne,
                              const CandidateParams& p) {
  if (c.test_processed_in(seed)) return;
  std::unique_ptr<Candidate> cand(new Candidate);
  cand->seed = seed;
  edge_linking(c, cand->convex_edge_segment, seed, (size_t)p.window_size_on_inner_elliptic_segment,
               p.average_vote_min);
  int nReceivedVote = 0, nVotedPoints = 0;
  for (int q : cand->convex_edge_segment) {
    const int vs = c.voters_size(q);
    nReceivedVote += vs;
    if (vs > 0) ++nVotedPoints;
  }
  cand->average_received_vote = (float)(nReceivedVote * nReceivedVote) / (float)nVotedPoints;
  auto it = std::lower_bound(loopOne.begin(), loopOne.end(), cand,
                             [](const std::unique_ptr<Candidate>& c1, const std::unique_ptr<Candidate>& c2) {
                               return c1->average_received_vote > c2->average_received_vote;
                             });
  loopOne.insert(it, std::move(cand));
}

// childrenOf (Vote.cpp:394-412).
void children_of(const Collection& c, const std::list<int>& edges, std::list<int>& children) {
  size_t voteMax = 1;
  for (int e : edges) voteMax = (size_t)std::max((int)voteMax, c.voters_size(e));
  for (int e : edges) {
    const int b = c.L->voters_index[e], en = c.L->voters_index[e + 1];
    if (en > b && (size_t)(en - b) >= voteMax / 14)
      for (int v = b; v < en; ++v) children.push_back(c.L->voters_list[v]);
  }
}

// A.determinant() != 0 and A.lu().solve(b) of Vote.cpp:485-489 for the 5x5
// float system: Eigen's PartialPivLU (unblocked for this size: per column the
// first largest magnitude is the pivot, the rows are swapped, the column
// below it divided by it and the rank-one update subtracted; solving applies
// the row swaps, then unit-lower and upper column-oriented substitutions).
// The determinant is the transpositions' sign times the pivots' product,
// (d0 d1)(d2 (d3 d4)); false when it is zero.
bool solve_5x5(std::array<std::array<float, 5>, 5> A, const std::array<float, 5>& b,
               std::array<float, 5>& x) {
  std::array<int, 5> transp{};
  int n_transp = 0;
  for (int k = 0; k < 5; ++k) {
    int big = k;
    float bigv = std::fabs(A[k][k]);
    for (int i = k + 1; i < 5; ++i)
      if (std::fabs(A[i][k]) > bigv) {
        bigv = std::fabs(A[i][k]);
        big = i;
      }
    transp[k] = big;
    if (bigv != 0.f) {
      if (big != k) {
        std::swap(A[k], A[big]);
        ++n_transp;
      }
      for (int i = k + 1; i < 5; ++i) A[i][k] = A[i][k] / A[k][k];
    }
    for (int i = k + 1; i < 5; ++i)
      for (int j = k + 1; j < 5; ++j) A[i][j] = A[i][j] - A[i][k] * A[k][j];
  }
  const float det =
      (n_transp % 2 ? -1.f : 1.f) * ((A[0][0] * A[1][1]) * (A[2][2] * (A[3][3] * A[4][4])));
  if (!(det != 0.f)) return false;
  x = b;
  for (int k = 0; k < 5; ++k) std::swap(x[k], x[transp[k]]);
  for (int i = 0; i < 5; ++i)
    for (int s = i + 1; s < 5; ++s) x[s] = x[s] - x[i] * A[s][i];
  for (int i = 4; i >= 0; --i) {
    x[i] = x[i] / A[i][i];
    for (int s = 0; s < i; ++s) x[s] = x[s] - x[i] * A[s][i];
  }
  return true;
}

constexpr int kNoWeight = 0;       // Params.hpp:20
constexpr int kInvGradWeight = 1;  // Params.hpp:21

// outlierRemoval (Vote.cpp:414-576); the subsample's distances use the
// batched point-to-ellipse form, the final filter the single-point form, as
// there.  A trial whose conic is not an ellipse (Ellipse(Q) throws) draws
// again without counting; the reference would loop forever if every trial
// threw, so the trials are bounded as in the library (100000).
void outlier_removal(const Collection& c, const std::list<int>& children,
                     std::vector<int>& filtered, float& SmFinal, float threshold,
                     int weightedType, size_t maxSize, Pcg32& rng) {
  filtered.reserve(children.size());
  const size_t nSubsampleSize = std::min(children.size(), maxSize);
  const float step = (float)children.size() / (float)nSubsampleSize;
  if (nSubsampleSize < 5) return;
  Ellipse qm;
  float Sm = 10000000.0f;
  const float f = 1.f;
  std::vector<Pt2> pts;
  std::vector<float> weights;
  size_t k = 0, iEdgePoint = 0;
  for (int e : children) {
    if (iEdgePoint == (size_t)(k * step)) {
      ++k;
      pts.emplace_back((float)c.x(e), (float)c.y(e));
      if (weightedType == kInvGradWeight) weights.push_back(255 / c.norm[e]);
    }
    ++iEdgePoint;
  }
  size_t counter = 0, trials = 0;
  while (counter < 70 && trials++ < 100000) {
    int perm[5];
    rng.rand_5_k(perm, pts.size());
    std::array<std::array<float, 5>, 5> A{};
    std::array<float, 5> b{}, temp{};
    for (int i = 0; i < 5; ++i) {
      const float px = pts[perm[i]].x, py = pts[perm[i]].y;
      A[i][0] = px * px;
      A[i][1] = 2.0f * px * py;
      A[i][2] = py * py;
      A[i][3] = 2.0f * f * px;
      A[i][4] = 2.0f * f * py;
      b[i] = -f * f;
    }
    if (!solve_5x5(A, b, temp)) {
      ++counter;
      continue;
    }
    if (!(temp[0] * temp[2] - temp[1] * temp[1] > 0)) {
      ++counter;
      continue;
    }
    Mat3 Q = Mat3::zero();
    Q.m[0][0] = temp[0];
    Q.m[0][1] = Q.m[1][0] = temp[1];
    Q.m[0][2] = Q.m[2][0] = temp[3];
    Q.m[1][1] = temp[2];
    Q.m[1][2] = Q.m[2][1] = temp[4];
    Q.m[2][2] = 1.f;
    try {
      const Ellipse q(Q);
      const float ratioSemiAxes = q.a() / q.b();
      if ((ratioSemiAxes < 0.04f) || (ratioSemiAxes > 25)) {
        ++counter;
        continue;
      }
      std::vector<float> dist;
      cctagx::distance_points_ellipse(dist, pts, q);
      if (weightedType != kNoWeight)
        for (size_t d = 0; d < dist.size(); ++d) dist[d] = dist[d] * weights[d];
      const float S = cctagx::median_ref(dist);
      if (S < Sm) {
        counter = 0;
        qm = Ellipse(Q);
        Sm = S;
      } else {
        ++counter;
      }
    } catch (...) {
    }
  }
  std::vector<float> vDistFinal;
  for (int e : children) {
    float distFinal = std::numeric_limits<float>::max();
    if (weightedType == kNoWeight)
      distFinal = cctagx::distance_point_ellipse((float)c.x(e), (float)c.y(e), qm);
    else if (weightedType == kInvGradWeight)
      distFinal = cctagx::distance_point_ellipse((float)c.x(e), (float)c.y(e), qm) * 255 / c.norm[e];
    if (distFinal < threshold * Sm) {
      filtered.push_back(e);
      vDistFinal.push_back(distFinal);
    }
  }
  if (vDistFinal.empty()) return;
  SmFinal = cctagx::median_ref(vDistFinal);
}

// distancePoints2D (geometry/Distance.hpp:19-23) on integer pixel positions.
float distance_points_2d(int x1, int y1, int x2, int y2) {
  return std::sqrt((float)((x2 - x1) * (x2 - x1)) + (float)((y2 - y1) * (y2 - y1)));
}

// innerProdMin (Fitting.cpp:228-316); p1, p2 as there (the caller ignores them).
float inner_prod_min(const Collection& c, const std::vector<int>& fc, float thrCosDiffMax,
                     int& p1, int& p2) {
  int pAngle1 = -1;
  float min = 1.1f, distMax = 0.f;
  const int p0 = fc.front();
  float normGrad = std::sqrt(c.dX(p0) * c.dX(p0) + c.dY(p0) * c.dY(p0));
  const float gx0 = c.dX(p0) / normGrad, gy0 = c.dY(p0) / normGrad;
  for (size_t k = 1; k < fc.size(); ++k) {
    const int q = fc[k];
    normGrad = std::sqrt(c.dX(q) * c.dX(q) + c.dY(q) * c.dY(q));
    const float gx = c.dX(q) / normGrad, gy = c.dY(q) / normGrad;
    const float innerProd = gx0 * gx + gy0 * gy;
    if (innerProd <= thrCosDiffMax) return innerProd;
    if (innerProd < min) {
      min = innerProd;
      pAngle1 = q;
    }
    const float dist = distance_points_2d(c.x(p0), c.y(p0), c.x(q), c.y(q));
    if (dist > distMax) {
      distMax = dist;
      p1 = q;
    }
  }
  normGrad = std::sqrt(c.dX(pAngle1) * c.dX(pAngle1) + c.dY(pAngle1) * c.dY(pAngle1));
  const float gxmin = c.dX(pAngle1) / normGrad, gymin = c.dY(pAngle1) / normGrad;
  min = 1.f;
  distMax = 0.f;
  for (int q : fc) {
    normGrad = std::sqrt(c.dX(q) * c.dX(q) + c.dY(q) * c.dY(q));
    const float chgx = c.dX(q) / normGrad, chgy = c.dY(q) / normGrad;
    const float innerProd = gxmin * chgx + gymin * chgy;
    if (innerProd <= thrCosDiffMax) return innerProd;
    if (innerProd < min) min = innerProd;
    const float dist = distance_points_2d(c.x(p1), c.y(p1), c.x(q), c.y(q));
    if (dist > distMax) {
      distMax = dist;
      p2 = q;
    }
  }
  return min;
}

// ellipseGrowingInit (EllipseGrowing.cpp:302-325) with isGoodEGPoints (:226-234).
bool ellipse_growing_init(const Collection& c, const std::vector<int>& fc, Ellipse& e) {
  int p1 = fc.front(), p2 = fc.front();
  const float thrCosDiffMax = 0.25f;
  if (inner_prod_min(c, fc, thrCosDiffMax, p1, p2) <= thrCosDiffMax) {
    cctagx::fit_ellipse(positions(c, fc), e);
    return true;
  }
  cctagx::fit_circle(positions(c, fc), e);
  return false;
}

// connectedPoint (EllipseGrowing.cpp:327-368), recursive as there.
void connected_point(std::vector<int>& pts, uint64_t mask, Collection& c, const Ellipse& qIn,
                     const Ellipse& qOut, int x, int y) {
  c.set_bit(c.at(x, y), mask);
  static const int xoff[] = {1, 1, 0, -1, -1, -1, 0, 1};
  static const int yoff[] = {0, -1, -1, -1, 0, 1, 1, 1};
  for (int i = 0; i < 8; ++i) {
    const int sx = x + xoff[i], sy = y + yoff[i];
    if (sx >= 0 && sx < c.shape_w && sy >= 0 && sy < c.shape_h) {
      const int e = c.at(sx, sy);
      if (e >= 0 && cctagx::is_in_hull(qIn, qOut, (float)c.x(e), (float)c.y(e)) &&
          !c.test_bit(e, mask)) {
        const float gx = c.dX(e), gy = c.dY(e);
        const float ox = qIn.cx() - c.x(e), oy = qIn.cy() - c.y(e);
        if (gx * ox + gy * oy < 0) {
          pts.push_back(e);
          c.set_bit(e, mask);
          connected_point(pts, mask, c, qIn, qOut, sx, sy);
        }
      }
    }
  }
}

// ellipseHull (EllipseGrowing.cpp:386-402).
void ellipse_hull(Collection& c, std::vector<int>& pts, const Ellipse& ellipse, float delta,
                  uint64_t mask) {
  Ellipse qIn, qOut;
  cctagx::compute_hull(ellipse, delta, qIn, qOut);
  const size_t initSize = pts.size();
  for (size_t i = 0; i < initSize; ++i) {
    const int e = pts[i];
    connected_point(pts, mask, c, qIn, qOut, c.x(e), c.y(e));
  }
}

int in_hull_count(const Collection& c, const std::vector<int>& pts, const Ellipse& ellipse,
                  float delta) {
  Ellipse qIn, qOut;
  cctagx::compute_hull(ellipse, delta, qIn, qOut);
  int n = 0;
  for (int q : pts)
    if (cctagx::is_in_hull(qIn, qOut, (float)c.x(q), (float)c.y(q))) ++n;
  return n;
}

// ellipseGrowing2 (EllipseGrowing.cpp:404-510).
void ellipse_growing2(Collection& c, const std::vector<int>& filteredChildren,
                      std::vector<int>& outerEllipsePoints, Ellipse& ellipse, float width,
                      size_t runId, bool goodInit) {
  const uint64_t threadMask = (uint64_t)1 << runId;
  outerEllipsePoints.reserve(filteredChildren.size() * 3);
  for (int q : filteredChildren) {
    outerEllipsePoints.push_back(q);
    c.set_bit(q, threadMask);
  }
  int lastSizePoints = 0;
  int nIter = 0;
  if (!goodInit) {
    int newSizePoints = (int)outerEllipsePoints.size();
    int maxNbPoints = newSizePoints;
    int nIterMax = 0;
    std::vector<std::vector<int>> edgePointsSets;
    edgePointsSets.push_back(outerEllipsePoints);
    std::vector<Ellipse> ellipsesSets;
    ellipsesSets.push_back(ellipse);
    ++nIter;
    while (newSizePoints - lastSizePoints > 0) {
      lastSizePoints = in_hull_count(c, outerEllipsePoints, ellipse, width);
      ellipse_hull(c, outerEllipsePoints, ellipse, width, threadMask);
      edgePointsSets.push_back(outerEllipsePoints);
      ellipsesSets.push_back(ellipse);
      cctagx::fit_circle(positions(c, outerEllipsePoints), ellipse);
      newSizePoints = in_hull_count(c, outerEllipsePoints, ellipse, width);
      if (newSizePoints > maxNbPoints) {
        maxNbPoints = newSizePoints;
        nIterMax = nIter;
      }
      ++nIter;
    }
    outerEllipsePoints = edgePointsSets[nIterMax];
    ellipse = ellipsesSets[nIterMax];
    for (const auto& set : edgePointsSets)
      for (int q : set) c.clear_bit(q, threadMask);
    for (int q : outerEllipsePoints) c.set_bit(q, threadMask);
  }
  lastSizePoints = 0;
  nIter = 0;
  cctagx::fit_ellipse(positions(c, outerEllipsePoints), ellipse);
  while ((int)outerEllipsePoints.size() - lastSizePoints > 0) {
    lastSizePoints = (int)outerEllipsePoints.size();
    ellipse_hull(c, outerEllipsePoints, ellipse, width, threadMask);
    cctagx::fit_ellipse(positions(c, outerEllipsePoints), ellipse);
    ++nIter;
  }
}

// completeFlowComponent (Detection.cpp:115-267) up to the segment label,
// which the caller assigns in candidate order.  `labeled`: the reference
// would reach the label (enough filtered points, nothing thrown before);
// `accepted`: it would append the candidate to vCandidateLoopTwo.
struct LoopTwo {
  Candidate cand;
  bool labeled = false;
  bool accepted = false;
};

void complete_flow_component(LoopTwo& t, Collection& c, size_t runId, const CandidateParams& p) {
  Candidate& cand = t.cand;
  Pcg32 rng(271828u);
  try {
    std::list<int> children;
    children_of(c, cand.convex_edge_segment, children);
    if (children.size() < (size_t)p.min_points_segment_candidate) return;
    cand.score = (int)children.size();
    float SmFinal = 1e+10f;
    outlier_removal(c, children, cand.filtered_children, SmFinal,
                    p.thresh_robust_estimation_of_outer_ellipse, kInvGradWeight, 60, rng);
    if (cand.filtered_children.size() < 5) return;
    t.labeled = true;
    const bool goodInit = ellipse_growing_init(c, cand.filtered_children, cand.outer_ellipse);
    ellipse_growing2(c, cand.filtered_children, cand.outer_ellipse_points, cand.outer_ellipse,
                     p.ellipse_growing_elliptic_hull_width, runId, goodInit);
    std::vector<float> vDistFinal;
    for (int q : cand.outer_ellipse_points)
      vDistFinal.push_back(cctagx::distance_point_ellipse((float)c.x(q), (float)c.y(q), cand.outer_ellipse));
    SmFinal = cctagx::median_ref(vDistFinal);
    if (SmFinal > p.thr_median_distance_ellipse) return;
    const float quality = (float)cand.outer_ellipse_points.size() /
                          (float)cctagx::rasterize_ellipse_perimeter(cand.outer_ellipse);
    if (quality > 1.1) return;
    const float ratioSemiAxes = cand.outer_ellipse.a() / cand.outer_ellipse.b();
    if ((ratioSemiAxes < 0.05) || (ratioSemiAxes > 20)) return;
    t.accepted = true;
  } catch (...) {
  }
}

// ---------------------------------------------------------------------------
// addCandidateFlowtoCCTag (EllipseGrowing.cpp:80-203).  The processed_aux
// flags are the call's own (the reference sets and clears them all within
// the call, so no other call can observe them).
bool add_candidate_flow(const Collection& c, const std::vector<int>& filteredChildren,
                        const std::vector<int>& outerEllipsePoints, const Ellipse& outerEllipse,
                        std::vector<std::vector<Pt2>>& cctagPoints, size_t numCircles) {
  cctagPoints.resize(numCircles);
  for (int e : outerEllipsePoints) cctagPoints.back().push_back(c.directed(e));
  std::set<int> aux;  // processed_aux
  size_t nGradientOut = 0, nAddedPoint = 0;
  for (int p0 : filteredChildren) {
    int dir = -1;
    int p = p0;
    const Pt2 outerPoint((float)c.x(p0), (float)c.y(p0));
    const float a = outerPoint.x - outerEllipse.cx();
    const float b = outerPoint.y - outerEllipse.cy();
    const Vec3 lineThroughCenter{{a, b, -a * outerEllipse.cx() - b * outerEllipse.cy()}};
    for (size_t j = 1; j < numCircles; ++j) {
      p = dir == -1 ? c.before(p) : c.after(p);
      Collection::check(p);  // test_processed_aux of the null point throws
      if (!aux.count(p)) {
        aux.insert(p);
        const float normGrad = std::sqrt(c.dX(p) * c.dX(p) + c.dY(p) * c.dY(p));
        const float gradE0 = c.dX(p) / normGrad, gradE1 = c.dY(p) / normGrad;
        float toto0 = outerEllipse.cx() - c.x(p), toto1 = outerEllipse.cy() - c.y(p);
        const float distancePointToCenter = std::sqrt(toto0 * toto0 + toto1 * toto1);
        toto0 /= distancePointToCenter;
        toto1 /= distancePointToCenter;
        const Pt2 pointToAdd = c.directed(p);
        if (cctagx::is_in_ellipse(outerEllipse, pointToAdd.x, pointToAdd.y) &&
            cctagx::is_on_the_same_side(outerPoint, pointToAdd, lineThroughCenter)) {
          if ((float(-dir) * (gradE0 * toto0 + gradE1 * toto1) < -0.5f) && (j >= numCircles - 2))
            ++nGradientOut;
          cctagPoints[numCircles - j - 1].push_back(pointToAdd);
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

// isAnotherSegment (Vote.cpp:578-742).
bool is_another_segment(const Collection& c, Ellipse& outerEllipse,
                        std::vector<int>& outerEllipsePoints, const Candidate& another,
                        std::vector<std::vector<Pt2>>& cctagPoints, size_t numCircles,
                        float thrMedianDistanceEllipse, Pcg32& rng) {
  const std::vector<int>& anotherOuter = another.outer_ellipse_points;
  Ellipse qm;
  float Sm = std::numeric_limits<float>::max();
  const std::vector<Pt2> pts = positions(c, outerEllipsePoints);
  const std::vector<Pt2> anotherPts = positions(c, anotherOuter);
  std::vector<float> distRef;
  cctagx::distance_points_ellipse(distRef, pts, outerEllipse);
  const float SRef = cctagx::median_ref(distRef);
  size_t cnt = 0;
  std::vector<Pt2> points;
  while (cnt < 100) {
    points.clear();
    int perm[5];
    rng.rand_5_k(perm, outerEllipsePoints.size());
    for (int i = 0; i < 4; ++i) points.emplace_back(pts[perm[i]].x, pts[perm[i]].y);
    rng.rand_5_k(perm, anotherOuter.size());
    for (int i = 0; i < 4; ++i)
      points.emplace_back((float)c.x(anotherOuter[perm[i]]), (float)c.y(anotherOuter[perm[i]]));
    Ellipse eToto;
    cctagx::fit_ellipse(points, eToto);  // a throw here leaves the loop (and the call)
    try {
      const Ellipse q(eToto.matrix());
      const float ratioSemiAxes = q.a() / q.b();
      if ((ratioSemiAxes < 0.12) || (ratioSemiAxes > 8)) {
        ++cnt;
        continue;
      }
      std::vector<float> dist;
      cctagx::distance_points_ellipse(dist, pts, q);
      const float S1 = cctagx::median_ref(dist);
      std::vector<float> anotherDist;
      cctagx::distance_points_ellipse(anotherDist, anotherPts, q);
      const float S2 = cctagx::median_ref(anotherDist);
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
  if (Sm < thr * SRef) {
    std::vector<int> outerEllipsePointsTemp = outerEllipsePoints;
    outerEllipsePointsTemp.insert(outerEllipsePointsTemp.end(), anotherOuter.begin(), anotherOuter.end());
    Ellipse outerEllipseTemp;
    cctagx::fit_ellipse(positions(c, outerEllipsePointsTemp), outerEllipseTemp);
    const float quality = (float)outerEllipsePointsTemp.size() /
                          (float)cctagx::rasterize_ellipse_perimeter(outerEllipseTemp);
    if (quality < 1.1) {
      std::vector<float> vDistFinal;
      for (int q : outerEllipsePointsTemp)
        vDistFinal.push_back(cctagx::distance_point_ellipse((float)c.x(q), (float)c.y(q), outerEllipseTemp));
      const float SmFinal = cctagx::median_ref(vDistFinal);
      if (SmFinal < thrMedianDistanceEllipse) {
        if (add_candidate_flow(c, another.filtered_children, anotherOuter, outerEllipseTemp,
                               cctagPoints, numCircles)) {
          outerEllipsePoints = outerEllipsePointsTemp;
          outerEllipse = outerEllipseTemp;
          return true;
        }
        return false;
      }
    } else {
      return false;
    }
  } else {
    return false;
  }
  return false;
}

// flowComponentAssembling (Detection.cpp:278-377).
void flow_component_assembling(const Collection& c, float& quality, size_t iCandidate,
                               const std::vector<Candidate>& loopTwo, Ellipse& outerEllipse,
                               std::vector<int>& outerEllipsePoints,
                               std::vector<std::vector<Pt2>>& cctagPoints,
                               const CandidateParams& p, Pcg32& rng) {
  const Candidate& candidate = loopTwo[iCandidate];
  int score = -1;
  size_t iMax = 0;
  const float ratioExpension = 2.5f;
  const float seedFlow = c.L->flow_length[candidate.seed];
  const Ellipse circularResearchArea((float)c.x(candidate.seed), (float)c.y(candidate.seed),
                                     seedFlow * ratioExpension, seedFlow * ratioExpension, 0.f);
  for (size_t i = 0; i < loopTwo.size(); ++i) {
    if (i == iCandidate) continue;
    const Candidate& another = loopTwo[i];
    if (candidate.n_label == another.n_label) continue;
    const float anotherFlow = c.L->flow_length[another.seed];
    if ((anotherFlow / seedFlow > 0.666) && (anotherFlow / seedFlow < 1.5)) {
      if (cctagx::is_in_ellipse(circularResearchArea, float(c.x(another.seed)), float(c.y(another.seed))) &&
          another.score > score) {
        score = another.score;
        iMax = i;
      }
    }
  }
  if (score > 0) {
    const Candidate& selected = loopTwo[iMax];
    if (is_another_segment(c, outerEllipse, outerEllipsePoints, selected, cctagPoints,
                           (size_t)p.n_crowns * 2, p.thr_median_distance_ellipse, rng))
      quality = (float)outerEllipsePoints.size() / (float)cctagx::rasterize_ellipse_perimeter(outerEllipse);
  }
}

// CCTag's constructor (CCTag.hpp:68-93): the outer ellipse's centre shifted by
// half a pixel, the rescaled ellipse = it scaled by the level's scale.
Marker make_marker(const Ellipse& outerEllipse, std::vector<std::vector<Pt2>> points,
                   int pyramidLevel, float scale, float quality) {
  Marker m;
  m.id = -1;
  m.center_img = Pt2(outerEllipse.cx(), outerEllipse.cy());
  m.points = std::move(points);
  m.outer_ellipse = outerEllipse;
  m.homography = Mat3::zero();
  m.quality = quality;
  m.pyramid_level = pyramidLevel;
  m.scale = scale;
  m.outer_ellipse.set_center(m.outer_ellipse.cx() + 0.5f, m.outer_ellipse.cy() + 0.5f);
  m.rescaled_outer_ellipse = cctagx::scaled(m.outer_ellipse, scale);
  m.status = 0;
  return m;
}

// cctagDetectionFromEdgesLoopTwoIteration (Detection.cpp:379-566).
void loop_two_iteration(std::vector<Marker>& markers, const Collection& c,
                        const std::vector<Candidate>& loopTwo, size_t iCandidate, int pyramidLevel,
                        float scale, const CandidateParams& p) {
  Pcg32 rng(271828u);
  const Candidate& candidate = loopTwo[iCandidate];
  std::vector<int> outerEllipsePoints = candidate.outer_ellipse_points;
  Ellipse outerEllipse = candidate.outer_ellipse;
  std::vector<std::vector<Pt2>> cctagPoints;
  try {
    float quality = (float)outerEllipsePoints.size() / (float)cctagx::rasterize_ellipse_perimeter(outerEllipse);
    if (p.search_for_another_segment && (quality > 0.25) && (quality < 0.7))
      flow_component_assembling(c, quality, iCandidate, loopTwo, outerEllipse, outerEllipsePoints,
                                cctagPoints, p, rng);
    if (!add_candidate_flow(c, candidate.filtered_children, candidate.outer_ellipse_points,
                            outerEllipse, cctagPoints, (size_t)p.n_crowns * 2))
      return;
    const Ellipse rescaleEllipse(outerEllipse.cx(), outerEllipse.cy(), outerEllipse.a() * scale,
                                 outerEllipse.b() * scale, outerEllipse.angle());
    const int realPixelPerimeter = cctagx::rasterize_ellipse_perimeter(rescaleEllipse);
    const float realSizeOuterEllipsePoints = quality * realPixelPerimeter;
    if (((quality <= 0.35) && (realSizeOuterEllipsePoints >= 300.0)) ||
        ((quality <= 0.45f) && (realSizeOuterEllipsePoints >= 200.0) && (realSizeOuterEllipsePoints < 300.0)) ||
        ((quality <= 0.5f) && (realSizeOuterEllipsePoints >= 100.0) && (realSizeOuterEllipsePoints < 200.0)) ||
        ((quality <= 0.5f) && (realSizeOuterEllipsePoints >= 70.0) && (realSizeOuterEllipsePoints < 100.0)) ||
        ((quality <= 0.96f) && (realSizeOuterEllipsePoints >= 50.0) && (realSizeOuterEllipsePoints < 70.0)) ||
        (realSizeOuterEllipsePoints < 50.0))
      return;
    const float ratioSemiAxes = outerEllipse.a() / outerEllipse.b();
    if (ratioSemiAxes > 8.0 || ratioSemiAxes < 0.125) return;
    Ellipse qIn, qOut;
    cctagx::compute_hull(outerEllipse, 3.6f, qIn, qOut);
    for (int q : outerEllipsePoints)
      if (!cctagx::is_in_hull(qIn, qOut, (float)c.x(q), (float)c.y(q))) return;
    float quality2 = 0;
    for (int q : outerEllipsePoints) quality2 += c.norm[q];
    quality2 *= scale;
    markers.push_back(make_marker(outerEllipse, cctagPoints, pyramidLevel, scale, quality2));
  } catch (...) {
  }
}

template <class Fn>
void for_each_task(cctagx::TaskPool* pool, int n, Fn&& fn) {
  if (pool && pool->threads() > 1 && n > 1) pool->run(n, fn);
  else for (int i = 0; i < n; ++i) fn(i);
}

// cctagDetectionFromEdges (Detection.cpp:568-693) on one level: loop one in
// seed order; loop two's tasks (independent: each writes its own candidate
// and its own bit of the points' _processed words), then the segment labels
// in candidate order (Detection.cpp:158-188); loop three's tasks.
void detection_from_edges(std::vector<Marker>& markers, Collection& c, int pyramidLevel,
                          float scale, const CandidateParams& p, cctagx::TaskPool* pool) {
  const LevelResult& L = *c.L;
  std::vector<int> seeds(L.seeds.begin(), L.seeds.end());
  if (seeds.size() > 1)  // Multiresolution.cpp:240-243, receivedMoreVoteThan
    std::sort(seeds.begin(), seeds.end(),
              [&](int a, int b) { return L.is_max[a] > L.is_max[b]; });
  if (seeds.empty()) return;
  const size_t nMaximumNbSeeds = (size_t)std::max(L.h / 2, p.maximum_nb_seeds);
  const size_t nSeedsToProcess = std::min(seeds.size(), nMaximumNbSeeds);
  std::vector<std::unique_ptr<Candidate>> loopOne;
  for (size_t i = 0; i < nSeedsToProcess; ++i) construct_flow_component(seeds[i], c, loopOne, p);
  const size_t nLoopTwo = std::min(loopOne.size(), (size_t)p.maximum_nb_candidates_loop_two);
  std::vector<LoopTwo> tasks(nLoopTwo);
  for (size_t i = 0; i < nLoopTwo; ++i) tasks[i].cand = *loopOne[i];
  for_each_task(pool, (int)nLoopTwo, [&](int i) { complete_flow_component(tasks[i], c, (size_t)i, p); });
  std::vector<Candidate> loopTwo;
  size_t nSegmentOut = 0;
  for (LoopTwo& t : tasks) {
    if (!t.labeled) continue;
    long nSegmentCommon = -1;
    for (int q : t.cand.filtered_children)
      if (c.n_segment_out[q] != -1) {
        nSegmentCommon = c.n_segment_out[q];
        break;
      }
    const size_t nLabel = nSegmentCommon == -1 ? nSegmentOut++ : (size_t)nSegmentCommon;
    for (int q : t.cand.filtered_children) c.n_segment_out[q] = (int32_t)nLabel;
    t.cand.n_label = nLabel;
    if (t.accepted) loopTwo.push_back(std::move(t.cand));
  }
  std::vector<std::vector<Marker>> found(loopTwo.size());
  for_each_task(pool, (int)loopTwo.size(), [&](int i) {
    loop_two_iteration(found[i], c, loopTwo, (size_t)i, pyramidLevel, scale, p);
  });
  for (auto& f : found)
    for (Marker& m : f) markers.push_back(std::move(m));
}

// intersectLineToTwoEllipses (Multiresolution.cpp:43-147).
bool intersect_line_two_ellipses(int y, const Ellipse& qIn, const Ellipse& qOut,
                                 const Collection& c, std::vector<int>& pointsInHull) {
  float io[2], ii[2];
  const int nOut = cctagx::intersect_ellipse_with_line(qOut, (float)y, true, io);
  const int nIn = cctagx::intersect_ellipse_with_line(qIn, (float)y, true, ii);
  auto consider = [&](int x) {
    const int e = c.at(x, y);
    if (e < 0) return;
    const float cx = qIn.cx() - c.x(e), cy = qIn.cy() - c.y(e);
    if (c.dX(e) * cx + c.dY(e) * cy < 0) pointsInHull.push_back(e);
  };
  if (nOut == 2 && nIn == 2) {
    const int begin1 = std::max(0, (int)io[0]);
    const int end1 = std::min(c.shape_w - 1, (int)ii[0]);
    const int begin2 = std::max(0, (int)ii[1]);
    const int end2 = std::min(c.shape_w - 1, (int)io[1]);
    for (int x = begin1; x <= end1; ++x) consider(x);
    for (int x = begin2; x <= end2; ++x) consider(x);
  } else if (nOut == 2 && nIn <= 1) {
    const int begin = std::max(0, (int)io[0]);
    const int end = std::min(c.shape_w - 1, (int)io[1]);
    for (int x = begin; x <= end; ++x) consider(x);
  } else if (nOut == 1 && nIn == 0) {
    if ((io[0] >= 0) && (io[0] < c.shape_w)) consider((int)io[0]);
  } else {
    return false;
  }
  return true;
}

// selectEdgePointInEllipticHull (Multiresolution.cpp:149-170).
void select_edge_points_in_hull(const Collection& c, const Ellipse& outerEllipse, float scale,
                                std::vector<int>& pointsInHull) {
  Ellipse qIn, qOut;
  cctagx::compute_hull(outerEllipse, scale, qIn, qOut);
  const float yCenter = outerEllipse.cy();
  const int maxY = std::max(int(yCenter), 0);
  const int minY = std::min(int(yCenter), c.shape_h - 1);
  for (int y = maxY; y < c.shape_h; ++y)
    if (!intersect_line_two_ellipses(y, qIn, qOut, c, pointsInHull)) break;
  for (int y = minY; y >= 0; --y)
    if (!intersect_line_two_ellipses(y, qIn, qOut, c, pointsInHull)) break;
}

// The level-0 re-fit of a marker found above level 0
// (Multiresolution.cpp:338-397); its outlier removal draws from its own
// pcg32(271828).
void rescale_marker(Marker& marker, const Collection& level0) {
  Pcg32 rng(271828u);
  const float scale = marker.scale;
  Ellipse rescaledOuterEllipse = marker.rescaled_outer_ellipse;
  std::vector<int> pointsInHull;
  select_edge_points_in_hull(level0, rescaledOuterEllipse, scale, pointsInHull);
  if (pointsInHull.size() < 5) return;
  std::vector<int> rescaledOuterEllipsePoints;
  float SmFinal = 1e+10f;
  const std::list<int> hullList(pointsInHull.begin(), pointsInHull.end());
  outlier_removal(level0, hullList, rescaledOuterEllipsePoints, SmFinal, 20.0f, kNoWeight, 60, rng);
  if (rescaledOuterEllipsePoints.size() < 5) return;
  try {
    cctagx::fit_ellipse(positions(level0, rescaledOuterEllipsePoints), rescaledOuterEllipse);
    std::vector<Pt2> directed;
    for (int e : rescaledOuterEllipsePoints) directed.push_back(level0.directed(e));
    marker.center_img = Pt2(marker.center_img.x * scale, marker.center_img.y * scale);
    marker.rescaled_outer_ellipse = rescaledOuterEllipse;
    marker.rescaled_outer_ellipse_points = directed;
  } catch (...) {
  }
}

// update (Multiresolution.cpp:172-193).
void update(std::vector<Marker>& markers, const Marker& markerToAdd) {
  auto half = [](const Marker& m) {  // CCTag::isEqual's comparison ellipse (CCTag.cpp:42-59)
    Ellipse e = m.rescaled_outer_ellipse;
    e.set_a(e.b() * 0.5f);
    e.set_b(e.b() * 0.5f);
    return e;
  };
  bool flag = false;
  for (Marker& current : markers) {
    if ((current.status > 0) && (markerToAdd.status > 0) &&
        cctagx::is_overlapping_ellipses(half(current), half(markerToAdd))) {
      if (markerToAdd.quality > current.quality) current = markerToAdd;
      flag = true;
    }
  }
  if (!flag) markers.push_back(markerToAdd);
}

}  // namespace

// cctagMultiresDetection (Multiresolution.cpp:268-413): levels from the
// coarsest, then every marker above level 0 re-fitted to level 0.
void candidates(const std::vector<LevelResult>& levels, const CandidateParams& p,
                std::vector<Marker>& markers, cctagx::TaskPool* pool) {
  markers.clear();
  const int nl = (int)levels.size();
  std::vector<std::unique_ptr<Collection>> cols(nl);
  for (int l = 0; l < nl; ++l) cols[l].reset(new Collection(levels[l]));
  for (int l = nl - 1; l >= 0; --l)
    detection_from_edges(markers, *cols[l], l, (float)std::pow(2.0, l), p, pool);
  for_each_task(pool, (int)markers.size(), [&](int i) {
    Marker& m = markers[i];
    if (m.pyramid_level > 0) rescale_marker(m, *cols[0]);
    else m.rescaled_outer_ellipse_points = m.points.back();
  });
}

void deduplicate(const std::vector<Marker>& markers, std::vector<Marker>& out) {
  std::vector<Marker> prelim, fin;
  for (const Marker& m : markers) update(prelim, m);
  for (const Marker& m : prelim) update(fin, m);
  std::stable_sort(fin.begin(), fin.end(), [](const Marker& a, const Marker& b) { return a.id < b.id; });
  out = std::move(fin);
}

}  // namespace oracle
