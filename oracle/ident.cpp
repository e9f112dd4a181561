// ORACLE -- test infrastructure only (see oracle.hpp).
//
// CPU restatement of the identification hot path: collectCuts /
// cutInterpolated / selectCutCheapUniform / outerEdgeRefinement
// (Identification.cpp:396-709), the imaged-centre grid search
// refineConicFamilyGlob / imageCenterOptimizationGlob / getNearbyPoints /
// computeHomographyFromEllipseAndImagedCenter / costFunctionGlob /
// extractSignalUsingHomography (:216-280, :739-1158) and the per-cut distances
// of orazioDistanceRobust (:50-193).  Serial, in the reference's loop order,
// including its sequential float sum of squared differences and the sticky
// out-of-bounds flag of an ImageCut.
#include "oracle_ident.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <map>
#include <vector>

namespace oracle {

using cctagx::Mat3;
using cctagx::Pt2;

namespace {

struct Cut {  // ImageCut.hpp:16-99
  Pt2 start, stop;
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

// boost::accumulators immediate mean + variance over [b, e).
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

// Identification.cpp:659-709.
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

// Identification.cpp:604-657.
bool outer_edge_refinement(Cut& cut, const uint8_t* img, int w, int h, float scale, size_t nref) {
  const float len = (float)(3.f * std::sqrt(2.0) * scale);
  const float half = len / 2.f;
  const float gn = std::sqrt(cut.stop.dx * cut.stop.dx + cut.stop.dy * cut.stop.dy);
  const float gdx = cut.stop.dx / gn, gdy = cut.stop.dy / gn;
  const float hx = half * gdx, hy = half * gdy;
  const Pt2 pStart(cut.stop.x - hx, cut.stop.y - hy);
  const Pt2 pStop(cut.stop.x + half * gdx, cut.stop.y + half * gdy, cut.stop.dx, cut.stop.dy);
  Cut c;
  c.start = pStart;
  c.stop = pStop;
  c.sig.assign(nref, 0.f);
  cut_interpolated(c, img, w, h);
  if (c.oob) return false;
  static const float kA[9] = {-0.0000f, -0.0003f, -0.1065f, -0.7863f, .0f, 0.7863f, 0.1065f, 0.0003f, 0.0000f};
  static const float kB[9] = {-0.0044f, -0.0540f, -0.2376f, -0.3450f, .0f, 0.3450f, 0.2376f, 0.0540f, 0.0044f};
  static const float kC[9] = {-0.0366f, -0.1113f, -0.1801f, -0.1594f, .0f, 0.1594f, 0.1801f, 0.1113f, 0.0366f};
  std::map<float, float> res;
  res.insert(conv_cut(kA, c));
  res.insert(conv_cut(kB, c));
  res.insert(conv_cut(kC, c));
  const float maxLocation = res.rbegin()->second;
  const float step = len / ((float)nref - 1.f);
  cut.stop = Pt2(pStart.x + step * maxLocation * gdx, pStart.y + step * maxLocation * gdy,
                 cut.stop.dx, cut.stop.dy);
  return true;
}

// Identification.cpp:216-280 (the out-of-bounds flag is never cleared).
void extract_signal(Cut& cut, const uint8_t* img, int w, int h, const Mat3& H, const Mat3& Hinv) {
  float bx, by;
  cctagx::apply_homography(bx, by, Hinv, cut.stop.x, cut.stop.y);
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
    cctagx::apply_homography(xr, yr, H, x, y);
    if (xr >= 0.f && xr < w - 1 && yr >= 0.f && yr < h - 1) cut.sig[i] = bilinear(img, w, xr, yr);
    else cut.oob = true;
    x += stepX;
    y += stepY;
  }
}

// Identification.cpp:739-798 with the per-ellipse canonical form precomputed.
Mat3 homography_from_center(const cctagx::RefineJob& r, const Pt2& c) {
  float xc, yc;
  cctagx::apply_homography(xc, yc, r.t_can, c.x, c.y);
  const float Q11 = r.canonic.m[0][0], Q22 = r.canonic.m[1][1], Q33 = r.canonic.m[2][2];
  Mat3 H;
  H.m[0][0] = Q33;
  H.m[1][0] = 0.f;
  H.m[2][0] = -Q11 * xc;
  H.m[0][1] = Q22 * xc * yc;
  H.m[1][1] = -Q11 * xc * xc - Q33;
  H.m[2][1] = Q22 * yc;
  H.m[0][2] = -Q33 * xc;
  H.m[1][2] = -Q33 * yc;
  H.m[2][2] = -Q33;
  const float d0 = std::sqrt((Q22 * Q33 / Q11 * (Q11 * xc * xc + Q22 * yc * yc + Q33)));
  const float d1 = Q33;
  const float d2 = std::sqrt(-Q22 * (Q11 * xc * xc + Q33));
  for (int i = 0; i < 3; ++i) {
    H.m[i][0] *= d0;
    H.m[i][1] *= d1;
    H.m[i][2] *= d2;
  }
  return cctagx::mul(r.t_inv_can, H);
}

// Identification.cpp:1122-1158.
float cost(const Mat3& H, std::vector<Cut>& cuts, const uint8_t* img, int w, int h, bool& flag) {
  flag = true;
  const Mat3 Hinv = cctagx::inverse3(H);
  for (Cut& c : cuts) extract_signal(c, img, w, h, H, Hinv);
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

// Identification.cpp:50-160, for one cut: the distance of each bank row.
void orazio_distances(const Cut& cut, const cctagx::IdentConstants& k, float* out) {
  const std::vector<float>& s = cut.sig;
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
  const float stepX = (cut.end - cut.begin) / (s.size() - 1.f);
  for (int idc = 0; idc < k.bank_rows; ++idc) {
    float x = cut.begin;
    float distance = 0;
    std::vector<float> digit(s.size());
    for (float& d : digit) {
      long ldum = 0;
      for (int j = 0; j < k.bank_cols; ++j)
        if (k.inv_ratio[idc][j] <= x) ++ldum;
      d = (float)(-(ldum % 2) * 2 + 1);
      x += stepX;
    }
    for (size_t i = 0; i < s.size(); ++i) {
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

}  // namespace

void CpuIdentBackend::select_cuts(const std::vector<cctagx::CutJob>& jobs,
                                  const cctagx::HostParams& p,
                                  std::vector<cctagx::CutSelection>& out) {
  const cctagx::IdentConstants k = cctagx::ident_constants(p, bank_);
  out.assign(jobs.size(), cctagx::CutSelection());
  cctagx::parallel_for(pool_, (int)jobs.size(), [&](int jb) {
    const cctagx::CutJob& job = jobs[jb];
    std::vector<Cut> cuts;
    for (const Pt2& op : job.outer_points) {  // collectCuts, Identification.cpp:471-496
      Cut c;
      c.start = Pt2(job.cx, job.cy);
      c.stop = op;
      c.begin = k.start_sig;
      c.end = 1.f;
      c.sig.assign(k.n_samples, 0.f);
      cut_interpolated(c, img_, w_, h_);
      if (!c.oob) cuts.push_back(c);
    }
    out[jb].n_collected = (int)cuts.size();
    if (cuts.empty()) return;
    // selectCutCheapUniform, Identification.cpp:546-601
    const size_t selectSize = std::min((size_t)p.num_cuts_in_ident_step, cuts.size());
    std::vector<float> var;
    for (const Cut& c : cuts) {
      float m, v;
      mean_var(c.sig.data(), c.sig.data() + c.sig.size(), m, v);
      var.push_back(v);
    }
    const float varMax = *std::max_element(var.begin(), var.end());
    std::vector<size_t> ind;
    for (size_t i = 0; i < cuts.size(); ++i) {
      Cut& c = cuts[i];
      c.stop = cctagx::point_on_ellipse(job.cx, job.cy, job.ea, job.eb, job.ca, job.sa, c.stop);
      if (outer_edge_refinement(c, img_, w_, h_, job.scale, (size_t)k.n_refine_samples))
        if (var[i] / varMax > 0.5f) ind.push_back(i);
    }
    const float step = std::max(1.f, (float)ind.size() / (float)selectSize);
    for (size_t kk = 0;; ++kk) {
      const size_t at = (size_t)(kk * step);
      if (at < ind.size() && out[jb].stops.size() < selectSize) out[jb].stops.push_back(cuts[ind[at]].stop);
      else break;
    }
  });
}

void CpuIdentBackend::refine(const std::vector<cctagx::RefineJob>& jobs, const cctagx::HostParams& p,
                             std::vector<cctagx::RefineResult>& out) {
  const cctagx::IdentConstants k = cctagx::ident_constants(p, bank_);
  out.assign(jobs.size(), cctagx::RefineResult());
  const size_t G = (size_t)p.imaged_center_n_grid_sample;
  cctagx::parallel_for(pool_, (int)jobs.size(), [&](int jb) {
    const cctagx::RefineJob& r = jobs[jb];
    cctagx::RefineResult& res = out[jb];
    std::vector<Cut> cuts(r.stops.size());
    for (size_t i = 0; i < cuts.size(); ++i) {
      cuts[i].stop = r.stops[i];
      cuts[i].begin = r.begin_sig;
      cuts[i].end = r.end_sig;
      cuts[i].sig.assign(k.n_samples, 0.f);
    }
    Pt2 center = r.center0;
    Mat3 H = Mat3::zero();
    float minRes = std::numeric_limits<float>::max();
    float neighbour = r.neighbour0;
    bool ok = true;
    for (int it = 0; it < r.n_iter; ++it) {
      // getNearbyPoints (Identification.cpp:1071-1109)
      const float ns = neighbour * r.grid_scale;
      Pt2 cc = center;
      cctagx::condition(cc, r.cond);
      const float halfWidth = ns / 2.f;
      const float stepSize = ns / (G - 1);
      std::vector<Pt2> grid;
      for (int i = 0; i < (int)G; ++i)
        for (int j = 0; j < (int)G; ++j) {
          Pt2 q(cc.x - halfWidth + i * stepSize, cc.y - halfWidth + j * stepSize);
          cctagx::condition(q, r.cond_inv);
          grid.push_back(q);
        }
      // imageCenterOptimizationGlob (Identification.cpp:980-1058)
      minRes = std::numeric_limits<float>::max();
      bool has = false;
      Pt2 best;
      Mat3 bestH = Mat3::zero();
      for (const Pt2& q : grid) {
        const Mat3 Ht = homography_from_center(r, q);
        bool readable = true;
        const float v = cost(Ht, cuts, img_, w_, h_, readable);
        if (readable) {
          has = true;
          if (v < minRes) {
            minRes = v;
            best = q;
            bestH = Ht;
          }
        }
      }
      center = best;
      H = bestH;
      if (!has) {
        ok = false;
        break;
      }
      neighbour /= (float)((G - 1) / 2);
    }
    res.ok = ok;
    if (!ok) return;
    res.center = center;
    res.H = H;
    res.min_res = minRes;
    const Mat3 Hinv = cctagx::inverse3(H);  // getSignals at the optimum
    for (Cut& c : cuts) extract_signal(c, img_, w_, h_, H, Hinv);
    res.oob.resize(cuts.size());
    std::vector<float> signals(cuts.size() * k.n_samples);
    res.id_dist.assign(cuts.size() * k.bank_rows, 0.f);
    for (size_t i = 0; i < cuts.size(); ++i) {
      res.oob[i] = cuts[i].oob;
      std::copy(cuts[i].sig.begin(), cuts[i].sig.end(), signals.begin() + i * k.n_samples);
      if (!cuts[i].oob) orazio_distances(cuts[i], k, &res.id_dist[i * k.bank_rows]);
    }
    cctagx::median_profile_range(signals, res.oob, (size_t)k.n_samples, res.bar_min, res.bar_max);
  });
}

}  // namespace oracle
