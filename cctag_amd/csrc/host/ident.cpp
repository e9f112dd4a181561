// Identification host control (see ident.hpp).
#include "ident.hpp"

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <list>
#include <string>

#include "../bank_radii.hpp"

namespace cctagx {

std::vector<std::vector<float>> marker_bank(int n_crowns) {
  const char* text = n_crowns == 4 ? kRadii4 : kRadii3;
  std::vector<std::vector<float>> bank;
  std::vector<int> radii;
  std::string cur;
  auto flush_row = [&]() {
    if (radii.empty()) return;
    std::vector<float> row;
    for (auto it = radii.rbegin(); it != radii.rend(); ++it) {
      char buf[32];
      std::snprintf(buf, sizeof(buf), "%.6f", 100.0 / *it);
      row.push_back(std::strtof(buf, nullptr));
    }
    bank.push_back(row);
    radii.clear();
  };
  for (const char* c = text;; ++c) {
    if (*c == ' ' || *c == ';' || *c == 0) {
      if (!cur.empty()) radii.push_back(std::atoi(cur.c_str()));
      cur.clear();
      if (*c == ';' || *c == 0) flush_row();
      if (*c == 0) break;
    } else {
      cur.push_back(*c);
    }
  }
  return bank;
}

IdentConstants ident_constants(const HostParams& p, const std::vector<std::vector<float>>& bank) {
  IdentConstants k{};
  // Identification.cpp:1221-1235
  if (p.n_crowns == 3) k.start_sig = 1 - (2 * (size_t)p.n_crowns - 1) * 0.15f;
  else if (p.n_crowns == 4) k.start_sig = 0.26f;
  else k.start_sig = 0.f;
  k.n_samples = p.sample_cut_length;
  k.n_refine_samples = p.num_samples_outer_edge_points_refinement;
  k.bank_rows = (int)bank.size();
  k.bank_cols = bank.empty() ? 0 : (int)bank[0].size();
  for (int i = 0; i < k.bank_rows && i < kMaxBank; ++i)
    for (int j = 0; j < k.bank_cols && j < kMaxRatios; ++j) k.inv_ratio[i][j] = 1.f / bank[i][j];
  return k;
}

// Identification.hpp:344-371.
static float compute_median(const std::vector<float>& vec) {
  const size_t s = vec.size() / 2 + 1;
  std::vector<float> out(s);
  std::partial_sort_copy(vec.begin(), vec.end(), out.begin(), out.end());
  if (vec.size() % 2) return out.back();
  return (float)((out[s - 1] + out[s - 2]) / 2.0);
}

void median_profile_range(const std::vector<float>& signals, const std::vector<uint8_t>& oob,
                          size_t n_samples, float& bar_min, float& bar_max) {
  // Identification.cpp:920-934
  std::vector<size_t> correct;
  for (size_t k = 0; k < oob.size(); ++k)
    if (!oob[k]) correct.push_back(k);
  std::vector<float> bar(n_samples), along(correct.size());
  for (size_t s = 0; s < n_samples; ++s) {
    for (size_t k = 0; k < correct.size(); ++k) along[k] = signals[correct[k] * n_samples + s];
    bar[s] = correct.empty() ? 0.f : compute_median(along);
  }
  bar_min = *std::min_element(bar.begin(), bar.end());
  bar_max = *std::max_element(bar.begin(), bar.end());
}

bool make_refine_job(const Ellipse& e, const Pt2& center0, const std::vector<Pt2>& stops,
                     const HostParams& p, const IdentConstants& kc, float mean_ab, RefineJob& r) {
  const float sqrt2 = std::sqrt(2.f);
  e.canonic_form(r.canonic, r.t_can, r.t_inv_can);
  r.cond = Mat3::zero();
  r.cond.m[0][0] = sqrt2 / mean_ab;
  r.cond.m[0][2] = -sqrt2 * e.cx() / mean_ab;
  r.cond.m[1][1] = sqrt2 / mean_ab;
  r.cond.m[1][2] = -sqrt2 * e.cy() / mean_ab;
  r.cond.m[2][2] = 1.f;
  r.cond_inv = inverse3(r.cond);
  float gs = 0.f;
  try {
    const Ellipse te(mul(mul(transpose(r.cond_inv), e.matrix()), r.cond_inv));
    gs = std::max(te.a(), te.b());
  } catch (...) {
    return false;
  }
  r.grid_scale = gs;
  r.max_semi_axis = std::max(e.a(), e.b());
  r.neighbour0 = p.imaged_center_neighbour_size;
  float ns = p.imaged_center_neighbour_size;
  int it = 0;
  const float div = (float)((p.imaged_center_n_grid_sample - 1) / 2);
  // (bounded: a non-finite semi-axis would never end the reference's loop)
  while (ns * r.max_semi_axis > 0.02 && it < 64) {
    ++it;
    ns /= div;
  }
  r.n_iter = it;
  r.center0 = center0;
  r.stops = stops;
  r.begin_sig = kc.start_sig;
  r.end_sig = 1.f;
  {
    const double c = std::cos((double)e.angle()), s = std::sin((double)e.angle());
    const double a = e.a(), b = e.b();
    const double hx = std::sqrt(a * a * c * c + b * b * s * s), hy = std::sqrt(a * a * s * s + b * b * c * c);
    r.box[0] = (float)(e.cx() - hx);
    r.box[1] = (float)(e.cx() + hx);
    r.box[2] = (float)(e.cy() - hy);
    r.box[3] = (float)(e.cy() + hy);
  }
  return true;
}

bool box_inside(const RefineJob& r, int w, int h) {
  return r.box[0] >= 2.f && r.box[1] <= (float)w - 3.f && r.box[2] >= 2.f &&
         r.box[3] <= (float)h - 3.f;
}

void IdentBackend::select_and_refine(const std::vector<CutJob>& cuts,
                                     std::vector<RefineJob>& jobs, const std::vector<int>& sel_of,
                                     const HostParams& p, std::vector<CutSelection>& sel,
                                     std::vector<RefineResult>& out) {
  select_cuts(cuts, p, sel);
  std::vector<RefineJob> run;
  std::vector<size_t> at;
  for (size_t r = 0; r < jobs.size(); ++r) {
    jobs[r].stops = sel[sel_of[r]].stops;
    if (jobs[r].stops.empty()) continue;
    run.push_back(jobs[r]);
    at.push_back(r);
  }
  std::vector<RefineResult> res;
  refine(run, p, res);
  out.assign(jobs.size(), RefineResult());
  for (size_t k = 0; k < at.size(); ++k) out[at[k]] = std::move(res[k]);
}

namespace {
// Working state of one frame between the backend calls.
struct FrameWork {
  std::vector<Marker>* markers;
  int image;
  std::vector<int> detected;
  std::vector<size_t> job_marker;   // select job -> marker
  std::vector<std::vector<Pt2>> stops;
};
}  // namespace

void identify_frames(std::vector<IdentFrame>& frames, const HostParams& p,
                     const std::vector<std::vector<float>>& bank, IdentBackend& backend,
                     TaskPool* pool, HostTimes* times) {
  using Clock = std::chrono::steady_clock;
  auto t0 = Clock::now();
  auto lap = [&](double HostTimes::*field) {
    const auto t1 = Clock::now();
    if (times) times->*field += std::chrono::duration<double, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  std::vector<FrameWork> fw;
  std::vector<std::pair<int, int>> all;  // (frame, marker)
  for (size_t f = 0; f < frames.size(); ++f) {
    IdentFrame& fr = frames[f];
    fw.push_back(FrameWork{fr.markers, fr.image, {}, {}, {}});
    const size_t n = fr.markers->size();
    fw.back().detected.assign(n, -1);
    fw.back().stops.assign(n, {});
    for (size_t i = 0; i < n; ++i) all.emplace_back((int)f, (int)i);
  }
  // --- identify_step_1 (Identification.cpp:1172-1301) ----------------------
  std::vector<CutJob> slot(all.size());
  std::vector<uint8_t> have(all.size(), 0);
  parallel_for(pool, (int)all.size(), [&](int q) {
    FrameWork& F = fw[all[q].first];
    const Marker& m = (*F.markers)[all[q].second];
    const Ellipse& e = m.rescaled_outer_ellipse;
    CutJob& job = slot[q];
    job.image = F.image;
    sorted_outer_points(e, m.rescaled_outer_ellipse_points, job.outer_points,
                        (size_t)p.n_samples_outer_ellipse);
    if (job.outer_points.size() < 5) return;
    job.cx = e.cx();
    job.cy = e.cy();
    job.ca = std::cos((double)e.angle());
    job.sa = std::sin((double)e.angle());
    job.ea = e.a();
    job.eb = e.b();
    job.scale = m.scale;
    have[q] = 1;
  });
  std::vector<CutJob> jobs;
  std::vector<std::pair<int, int>> job_at;  // select job -> (frame, marker)
  for (size_t q = 0; q < all.size(); ++q) {
    FrameWork& F = fw[all[q].first];
    if (!have[q]) {
      F.detected[all[q].second] = status::too_few_outer_points;
      continue;
    }
    jobs.push_back(std::move(slot[q]));
    F.job_marker.push_back(all[q].second);
    job_at.push_back(all[q]);
  }
  // --- identify_step_2 (Identification.cpp:1318-1480) -----------------------
  // prepared for every marker with cuts to select, so that the backend can
  // run both steps at once.  conditionerFromEllipse (conditioner.hpp:22-46)
  // caches (a+b)/2 of the first ellipse it sees: per frame, that of the
  // first marker refined -- guessed as the frame's first marker with cuts,
  // and the frame refined again when its selection turns out empty.
  const IdentConstants kc = ident_constants(p, bank);
  std::vector<int> guess(fw.size(), -1);
  for (size_t f = 0; f < fw.size(); ++f)
    if (!fw[f].job_marker.empty()) guess[f] = (int)fw[f].job_marker[0];
  auto mean_of = [&](size_t f, int i) {
    const Ellipse& e = (*fw[f].markers)[i].rescaled_outer_ellipse;
    return (e.a() + e.b()) / 2.f;
  };
  const std::vector<Pt2> no_stops;
  std::vector<RefineJob> rslot(jobs.size());
  std::vector<uint8_t> made(jobs.size(), 0);  // 1: job made, 2: its conditioned ellipse failed
  parallel_for(pool, (int)jobs.size(), [&](int q) {
    const int f = job_at[q].first, i = job_at[q].second;
    const Marker& m = (*fw[f].markers)[i];
    rslot[q].image = fw[f].image;
    made[q] = make_refine_job(m.rescaled_outer_ellipse, m.center_img, no_stops, p, kc,
                              mean_of(f, guess[f]), rslot[q]) ? 1 : 2;
  });
  std::vector<RefineJob> rjobs;
  std::vector<int> sel_of, rjob_of(jobs.size(), -1);
  for (size_t q = 0; q < jobs.size(); ++q) {
    if (made[q] != 1) continue;
    rjob_of[q] = (int)rjobs.size();
    rjobs.push_back(std::move(rslot[q]));
    sel_of.push_back((int)q);
  }
  lap(&HostTimes::ident_prep);
  std::vector<CutSelection> sel;
  std::vector<RefineResult> res;
  backend.select_and_refine(jobs, rjobs, sel_of, p, sel, res);
  lap(&HostTimes::ident_refine);
  for (size_t q = 0; q < jobs.size(); ++q) {
    FrameWork& F = fw[job_at[q].first];
    const int i = job_at[q].second;
    CutSelection& c = sel[q];
    if (c.n_collected == 0) F.detected[i] = status::no_collected_cuts;
    else if (c.stops.empty()) F.detected[i] = status::no_selected_cuts;
    else {
      F.detected[i] = status::id_reliable;
      F.stops[i] = std::move(c.stops);
      (*F.markers)[i].cut_stops = F.stops[i];
    }
  }
  // frames whose first refined marker is not the guessed one: their jobs
  // again with that marker's (a+b)/2
  std::vector<int> first(fw.size(), -1);
  std::vector<uint8_t> redo(fw.size(), 0);
  for (size_t f = 0; f < fw.size(); ++f) {
    for (size_t i = 0; i < fw[f].markers->size(); ++i)
      if (fw[f].detected[i] == status::id_reliable) {
        first[f] = (int)i;
        break;
      }
    redo[f] = first[f] >= 0 && first[f] != guess[f];
  }
  std::vector<RefineJob> rjobs2;
  std::vector<RefineResult> res2;
  std::vector<int> rjob2_of(jobs.size(), -1);
  {
    std::vector<int> qs;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const int f = job_at[q].first;
      if (redo[f] && fw[f].detected[job_at[q].second] == status::id_reliable) qs.push_back((int)q);
    }
    if (!qs.empty()) {
      std::vector<RefineJob> slot2(qs.size());
      parallel_for(pool, (int)qs.size(), [&](int k) {
        const int q = qs[k], f = job_at[q].first, i = job_at[q].second;
        const Marker& m = (*fw[f].markers)[i];
        slot2[k].image = fw[f].image;
        made[q] = make_refine_job(m.rescaled_outer_ellipse, m.center_img, fw[f].stops[i], p, kc,
                                  mean_of(f, first[f]), slot2[k]) ? 1 : 2;
      });
      for (size_t k = 0; k < qs.size(); ++k)
        if (made[qs[k]] == 1) {
          rjob2_of[qs[k]] = (int)rjobs2.size();
          rjobs2.push_back(std::move(slot2[k]));
        }
      backend.refine(rjobs2, p, res2);
    }
  }
  lap(&HostTimes::ident_refine);
  // the refined markers in frame and marker order, with their job and result
  struct Refined {
    int f;
    size_t i;
    const RefineJob* job;
    const RefineResult* res;
  };
  std::vector<Refined> rj_at;
  for (size_t q = 0; q < jobs.size(); ++q) {
    const int f = job_at[q].first, i = job_at[q].second;
    if (fw[f].detected[i] != status::id_reliable) continue;
    if (made[q] == 2) {
      fw[f].detected[i] = status::opti_has_diverged;
      continue;
    }
    if (redo[f]) rj_at.push_back(Refined{f, (size_t)i, &rjobs2[rjob2_of[q]], &res2[rjob2_of[q]]});
    else rj_at.push_back(Refined{f, (size_t)i, &rjobs[rjob_of[q]], &res[rjob_of[q]]});
  }
  const int nb = (int)bank.size();
  parallel_for(pool, (int)rj_at.size(), [&](int jj) {
    FrameWork& F = fw[rj_at[jj].f];
    const size_t i = rj_at[jj].i;
    std::vector<Marker>& markers = *F.markers;
    std::vector<int>& detected = F.detected;
    const RefineJob& rj = *rj_at[jj].job;
    const RefineResult& r = *rj_at[jj].res;
    Marker& m = markers[i];
    if (!r.ok) {
      m.quality = 1.f / FLT_MAX;
      m.center_img = Pt2(0.f, 0.f);
      detected[i] = status::opti_has_diverged;
      return;
    }
    m.center_img = r.center;
    m.homography = r.H;
    const size_t ncut = rj.stops.size();
    const float magnitude = r.bar_max - r.bar_min;
    const float residual = (float)(std::sqrt((double)r.min_res) / (double)magnitude);
    m.quality = 1.f / residual;
    if (residual > 2.7f) {
      detected[i] = status::opti_has_diverged;
      return;
    }
    // orazioDistanceRobust (Identification.cpp:50-193), serial over cuts.
    std::vector<std::list<float>> vScore(nb);
    std::vector<float> v(nb);
    for (size_t k = 0; k < ncut; ++k) {
      if (r.oob[k]) continue;  // Identification.cpp:78: only in-bounds cuts vote
      // the largest exp(-d) keyed in a std::map<float, int>: the last row
      // among equal keys; a scan gives the same when no key is NaN
      bool nan = false;
      for (int idc = 0; idc < nb; ++idc) {
        v[idc] = std::exp(-r.id_dist[k * nb + idc]);
        nan |= v[idc] != v[idc];
      }
      if (!nan) {
        int best = 0;
        for (int idc = 1; idc < nb; ++idc)
          if (v[idc] >= v[best]) best = idc;
        vScore[best].push_back(v[best]);
      } else {
        std::map<float, int> sortedId;
        for (int idc = 0; idc < nb; ++idc) sortedId[v[idc]] = idc;
        const auto best = sortedId.rbegin();
        vScore[best->second].push_back(best->first);
      }
    }
    size_t maxSize = 0;
    int iMax = 0;
    for (int c = 0; c < nb; ++c)
      if (vScore[c].size() > maxSize) {
        iMax = c;
        maxSize = vScore[c].size();
      }
    float score = 0;
    for (float v : vScore[iMax]) score += v;
    score /= vScore[iMax].size();
    m.id = iMax;
    m.radius_ratios = bank[iMax];
    try {
      const Mat3 inv = inverse3(m.homography);
      for (float rr : m.radius_ratios) {
        const Ellipse circle(0.f, 0.f, 1.f / rr, 1.f / rr, 0.f);
        const Ellipse ring(mul(mul(transpose(inv), circle.matrix()), inv));
        (void)ring;
      }
    } catch (...) {
      detected[i] = status::degenerate;
      return;
    }
    detected[i] = (score > p.min_ident_proba) ? status::id_reliable : status::id_not_reliable;
  });
  for (FrameWork& F : fw)
    for (size_t i = 0; i < F.markers->size(); ++i) (*F.markers)[i].status = F.detected[i];
  lap(&HostTimes::ident_post);
}

void identify_markers(std::vector<Marker>& markers, const HostParams& p,
                      const std::vector<std::vector<float>>& bank, IdentBackend& backend,
                      TaskPool* pool, HostTimes* times) {
  std::vector<IdentFrame> f(1);
  f[0].markers = &markers;
  f[0].image = 0;
  identify_frames(f, p, bank, backend, pool, times);
}

}  // namespace cctagx
