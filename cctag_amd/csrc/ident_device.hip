// GPU identification backend (host side): packs the per-marker jobs the host
// prepared, launches the identification kernels once per step for all
// markers and unpacks the results.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "ident_device.hpp"

namespace cctagx {

long long refine_flops(long long n_iter, long long grid_points, long long cuts, long long samples) {
  // per grid point: 25 FLOP per bilinear cut sample (homography, sampling,
  // step) and 3 per sample of every cut pair's difference (sub, mul, add)
  return n_iter * grid_points * (25 * cuts * samples + 3 * samples * cuts * (cuts - 1) / 2);
}

namespace {
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void copy9(float* dst, const Mat3& m) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dst[i * 3 + j] = m.m[i][j];
}
}  // namespace

int GpuIdentBackend::init(const std::vector<std::vector<float>>& bank, const HostParams& p,
                          hipStream_t s) {
  stream_ = s;
  k_ = ident_constants(p, bank);
  if (k_.bank_rows > kMaxBankRows || k_.bank_rows > kMaxBank || k_.bank_cols > kMaxRatios)
    throw std::runtime_error("marker bank larger than the identification kernels accept");
  for (hipEvent_t& e : ev_)
    if (!e) CCTAGX_HIP(hipEventCreate(&e));
  if (!jobs_copied_) CCTAGX_HIP(hipEventCreate(&jobs_copied_));
  ensure(64, 64 * 160);
  return 0;
}

void GpuIdentBackend::set_bank(const std::vector<std::vector<float>>& bank, const HostParams& p) {
  const IdentConstants k = ident_constants(p, bank);
  if (k.bank_rows > kMaxBankRows || k.bank_rows > kMaxBank || k.bank_cols > kMaxRatios)
    throw std::runtime_error("marker bank larger than the identification kernels accept");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  k_ = k;
  if (d_inv_ratio_)
    hip_check(hipMemcpy(d_inv_ratio_, &k_.inv_ratio[0][0], sizeof(float) * kMaxBank * kMaxRatios,
                        hipMemcpyHostToDevice),
              "hipMemcpy bank");
}

void GpuIdentBackend::ensure(int n_markers, int n_points) {
  if (n_markers <= cap_markers_ && n_points <= cap_points_) return;
  const int nm = std::max(n_markers, cap_markers_), np = std::max(n_points, cap_points_);
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  dev_.release();
  dev_.clear_plan();

  const size_t o_rj = dev_.reserve(sizeof(RefineJobDev) * nm);
  const size_t o_st = dev_.reserve(sizeof(RefineStateDev) * 2 * nm);
  const size_t o_gr = dev_.reserve(sizeof(RefineGridDev) * 2 * (size_t)nm * kMaxGrid * kMaxGrid);
  const size_t o_gm =
      dev_.reserve(sizeof(double) * 2 * (size_t)nm * kMaxGrid * kMaxGrid * kMaxPairs);
  const size_t o_inv = dev_.reserve(sizeof(float) * kMaxBank * kMaxRatios);
  if (dev_.allocate() != 0) throw std::runtime_error(std::string("hipMalloc: ") + last_error());
  d_ref_jobs_ = dev_.at<RefineJobDev>(o_rj);
  d_work_.state = dev_.at<RefineStateDev>(o_st);
  d_work_.grid = dev_.at<RefineGridDev>(o_gr);
  d_work_.gram = dev_.at<double>(o_gm);
  d_inv_ratio_ = dev_.at<float>(o_inv);
  hip_check(hipMemcpy(d_inv_ratio_, &k_.inv_ratio[0][0], sizeof(float) * kMaxBank * kMaxRatios,
                      hipMemcpyHostToDevice),
            "hipMemcpy bank");
  // the select kernel reads its jobs from and writes its results to mapped
  // host memory, as the refinement's last kernel does its results
  if (h_sel_jobs_.ensure(sizeof(SelectJobDev) * nm) || h_pts_.ensure(sizeof(float4) * np) ||
      h_sel_out_.ensure(sizeof(SelectOutDev) * nm, hipHostMallocCoherent) ||
      h_ref_jobs_.ensure(sizeof(RefineJobDev) * nm) ||
      h_ref_out_.ensure(sizeof(RefineOutDev) * nm, hipHostMallocCoherent) ||
      h_ref_dist_.ensure(sizeof(float) * (size_t)nm * kMaxCuts * kMaxBankRows, hipHostMallocCoherent))
    throw std::runtime_error(std::string("hipHostMalloc: ") + last_error());
  cap_markers_ = nm;
  cap_points_ = np;
}

const DeviceImage& GpuIdentBackend::image(int i) const {
  if (i < 0 || i >= (int)images_.size()) throw std::runtime_error("identification job names no image");
  return images_[i];
}

// The selection jobs and their outer points into mapped host memory (the
// caller has called ensure for them).
void GpuIdentBackend::pack_select(const std::vector<CutJob>& jobs) {
  const int n = (int)jobs.size();
  int npts = 0;
  for (const CutJob& j : jobs) npts += (int)j.outer_points.size();
  select_jobs_ = n;
  select_samples_ = (long long)npts * (k_.n_samples + k_.n_refine_samples);
  SelectJobDev* hj = h_sel_jobs_.as<SelectJobDev>();
  float4* hp = h_pts_.as<float4>();
  int off = 0;
  for (int i = 0; i < n; ++i) {
    const CutJob& j = jobs[i];
    if ((int)j.outer_points.size() > kMaxOuterPoints)
      throw std::runtime_error("more outer points than the identification kernels accept");
    const DeviceImage& im = image(j.image);
    hj[i] = SelectJobDev{im.quad, im.w, im.h, j.cx, j.cy, j.ca, j.sa, j.ea, j.eb, j.scale,
                         (int)j.outer_points.size(), off};
    for (const Pt2& q : j.outer_points) hp[off++] = make_float4(q.x, q.y, q.dx, q.dy);
  }
}

void GpuIdentBackend::unpack_select(int n, std::vector<CutSelection>& out) {
  host_sel_.resize(n);
  std::memcpy(host_sel_.data(), h_sel_out_.as<SelectOutDev>(), sizeof(SelectOutDev) * n);
  const SelectOutDev* ho = host_sel_.data();
  out.assign(n, CutSelection());
  for (int i = 0; i < n; ++i) {
    out[i].n_collected = ho[i].n_collected;
    out[i].stops.reserve(ho[i].n_selected);
    for (int k = 0; k < ho[i].n_selected; ++k) {
      const float4 v = ho[i].stops[k];
      out[i].stops.emplace_back(v.x, v.y, v.z, v.w);
    }
  }
}

void GpuIdentBackend::select_cuts(const std::vector<CutJob>& jobs, const HostParams& p,
                                  std::vector<CutSelection>& out) {
  const int n = (int)jobs.size();
  if (n == 0) {
    out.clear();
    return;
  }
  int npts = 0;
  for (const CutJob& j : jobs) npts += (int)j.outer_points.size();
  ensure(n, npts);
  pack_select(jobs);
  hipStream_t s = stream_;
  hipEventRecord(ev_[0], s);
  if (launch_select_cuts(h_sel_jobs_.dev<SelectJobDev>(), n, h_pts_.dev<float4>(), k_.start_sig,
                         k_.n_samples, k_.n_refine_samples, p.num_cuts_in_ident_step,
                         h_sel_out_.dev<SelectOutDev>(), s) != 0)
    throw std::runtime_error("k_select_cuts launch failed");
  hipEventRecord(ev_[1], s);
  hip_check(stream_wait(s), "sync select");
  hipEventElapsedTime(&select_us_, ev_[0], ev_[1]);
  select_us_ *= 1000.f;
  unpack_select(n, out);
}

// Fills the pinned job array for `jobs` (distances packed job after job);
// returns the number of distance floats.
int GpuIdentBackend::pack_refine_jobs(const std::vector<RefineJob>& jobs,
                                      const std::vector<int>* sel_of) {
  // the pinned array may still be the source of a queued copy
  if (jobs_in_flight_) hip_check(hipEventSynchronize(jobs_copied_), "job copy");
  jobs_in_flight_ = false;
  RefineJobDev* hj = h_ref_jobs_.as<RefineJobDev>();
  const int nb = k_.bank_rows;
  int n_dist = 0;
  dist_offset_.resize(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) {
    const RefineJob& r = jobs[i];
    RefineJobDev& d = hj[i];
    const DeviceImage& im = image(r.image);
    d.quad = im.quad;
    d.iw = im.w;
    d.ih = im.h;
    d.Q11 = r.canonic.m[0][0];
    d.Q22 = r.canonic.m[1][1];
    d.Q33 = r.canonic.m[2][2];
    copy9(d.t_can, r.t_can);
    copy9(d.t_inv_can, r.t_inv_can);
    copy9(d.cond, r.cond);
    copy9(d.cond_inv, r.cond_inv);
    d.grid_scale = r.grid_scale;
    d.neighbour0 = r.neighbour0;
    d.begin_sig = r.begin_sig;
    d.end_sig = r.end_sig;
    d.n_iter = r.n_iter;
    d.box_inside = box_inside(r, im.w, im.h) ? 1 : 0;
    d.c0x = r.center0.x;
    d.c0y = r.center0.y;
    if (sel_of) {
      d.sel_index = (*sel_of)[i];
      d.n_cuts = 0;
      d.dist_offset = (int)i * kMaxCuts * nb;
      n_dist = ((int)i + 1) * kMaxCuts * nb;
    } else {
      d.sel_index = -1;
      d.n_cuts = (int)r.stops.size();
      d.dist_offset = n_dist;
      n_dist += d.n_cuts * nb;
      for (size_t k = 0; k < r.stops.size(); ++k) d.stops[k] = make_float2(r.stops[k].x, r.stops[k].y);
    }
    dist_offset_[i] = d.dist_offset;
  }
  dist_total_ = n_dist;
  return n_dist;
}

void GpuIdentBackend::coarse_centers(const std::vector<RefineJob>& jobs, const HostParams& p,
                                     float* d_centers) {
  const int n = (int)jobs.size();
  if (n == 0) return;
  ensure(n, 0);
  pack_refine_jobs(jobs);
  hip_check(hipMemcpyAsync(d_ref_jobs_, h_ref_jobs_.as<RefineJobDev>(), sizeof(RefineJobDev) * n,
                           hipMemcpyHostToDevice, stream_),
            "H2D");
  hip_check(hipEventRecord(jobs_copied_, stream_), "event");
  jobs_in_flight_ = true;
  if (launch_refine_coarse(d_ref_jobs_, n, k_.n_samples, p.imaged_center_n_grid_sample, d_work_,
                           d_centers, stream_) != 0)
    throw std::runtime_error("k_refine coarse launch failed");
}

// Gradient refinement of the jobs already on the device: the grid search's
// first iteration, grad_steps_ descent steps from the centre it picked, the
// closed-form homography there and the read-out (results into the same
// mapped buffers as the grid search's).
void GpuIdentBackend::refine_gradient(const std::vector<RefineJob>& jobs, const HostParams& p) {
  const int n = (int)jobs.size();
  hipStream_t s = stream_;
  if (h_grad_jobs_.ensure(sizeof(GradJobDev) * n) || d_grad_jobs_.ensure(sizeof(GradJobDev) * n) ||
      d_coarse_.ensure(sizeof(float) * 2 * n) || d_gcenters_.ensure(sizeof(float) * 2 * n) ||
      d_gcost_.ensure(sizeof(double) * n) || d_gH_.ensure(sizeof(float) * 9 * n))
    throw std::runtime_error(std::string("gradient refinement buffers: ") + last_error());
  GradJobDev* g = h_grad_jobs_.as<GradJobDev>();
  for (int i = 0; i < n; ++i) {
    const RefineJob& r = jobs[i];
    GradJobDev& d = g[i];
    const DeviceImage& im = image(r.image);
    d.quad = im.quad;
    d.iw = im.w;
    d.ih = im.h;
    d.Q11 = r.canonic.m[0][0];
    d.Q22 = r.canonic.m[1][1];
    d.Q33 = r.canonic.m[2][2];
    copy9(d.t_can, r.t_can);
    copy9(d.t_inv_can, r.t_inv_can);
    d.begin_sig = r.begin_sig;
    d.end_sig = r.end_sig;
    // a quarter of the grid search's first neighbourhood
    d.alpha0 = 0.25f * r.neighbour0 * r.max_semi_axis;
    d.n_cuts = (int)r.stops.size();
    d.c0x = r.center0.x;
    d.c0y = r.center0.y;
    for (size_t k = 0; k < r.stops.size(); ++k) d.stops[k] = make_float2(r.stops[k].x, r.stops[k].y);
  }
  hip_check(hipMemcpyAsync(d_grad_jobs_.as<void>(), g, sizeof(GradJobDev) * n,
                           hipMemcpyHostToDevice, s),
            "H2D");
  const int ns = k_.n_samples;
  if (launch_refine_coarse(d_ref_jobs_, n, ns, p.imaged_center_n_grid_sample, d_work_,
                           d_coarse_.as<float>(), s) != 0 ||
      launch_grad_descend(d_grad_jobs_.as<GradJobDev>(), n, grad_steps_, ns, false,
                          d_coarse_.as<float>(), d_gcenters_.as<float>(), d_gcost_.as<double>(),
                          s) != 0 ||
      launch_grad_homography(d_grad_jobs_.as<GradJobDev>(), n, d_gcenters_.as<float>(),
                             d_gH_.as<float>(), s) != 0 ||
      launch_refine_readout(d_ref_jobs_, n, ns, d_gH_.as<float>(), d_gcenters_.as<float>(),
                            d_gcost_.as<double>(), d_inv_ratio_, k_.bank_rows, k_.bank_cols,
                            d_work_, h_ref_out_.dev<RefineOutDev>(), h_ref_dist_.dev<float>(),
                            s) != 0)
    throw std::runtime_error("gradient refinement launch failed");
}

void GpuIdentBackend::unpack_refine(const std::vector<RefineJob>& jobs,
                                    std::vector<RefineResult>& out) {
  // the headers and the distances leave the mapped buffers in one bulk copy
  // each (field-by-field reads from them took most of the unpacking time)
  const size_t nj = jobs.size();
  host_out_.resize(nj);
  std::memcpy(host_out_.data(), h_ref_out_.as<RefineOutDev>(), sizeof(RefineOutDev) * nj);
  host_dist_.resize((size_t)dist_total_);
  if (dist_total_ > 0)
    std::memcpy(host_dist_.data(), h_ref_dist_.as<float>(), sizeof(float) * (size_t)dist_total_);
  const RefineOutDev* ho = host_out_.data();
  const float* hd = host_dist_.data();
  const int nb = k_.bank_rows;
  out.assign(jobs.size(), RefineResult());
  for (size_t i = 0; i < jobs.size(); ++i) {
    const RefineOutDev& o = ho[i];
    RefineResult& r = out[i];
    const int nc = (int)jobs[i].stops.size();
    r.ok = o.ok != 0 && nc > 0;
    if (!r.ok) continue;
    r.center = Pt2(o.cx, o.cy);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) r.H.m[a][b] = o.H[a * 3 + b];
    r.min_res = o.min_res;
    r.bar_min = o.bar_min;
    r.bar_max = o.bar_max;
    r.oob.assign(o.oob, o.oob + nc);
    r.id_dist.assign((size_t)nc * nb, 0.f);
    const float* src = hd + dist_offset_[i];
    for (int k = 0; k < nc; ++k)
      if (!r.oob[k]) std::copy(src + (size_t)k * nb, src + (size_t)(k + 1) * nb, &r.id_dist[(size_t)k * nb]);
  }
}

void GpuIdentBackend::refine(const std::vector<RefineJob>& jobs, const HostParams& p,
                             std::vector<RefineResult>& out) {
  const int n = (int)jobs.size();
  if (n == 0) {
    out.clear();
    return;
  }
  ensure(n, 0);
  refine_jobs_ = n;
  refine_samples_ = 0;
  refine_flops_ = 0;
  const long long g2 = (long long)p.imaged_center_n_grid_sample * p.imaged_center_n_grid_sample;
  for (const RefineJob& r : jobs) {
    refine_samples_ += (long long)r.n_iter * g2 * (long long)r.stops.size() * k_.n_samples;
    refine_flops_ += refine_flops(r.n_iter, g2, (long long)r.stops.size(), k_.n_samples);
  }
  pack_refine_jobs(jobs);
  hipStream_t s = stream_;
  hip_check(hipMemcpyAsync(d_ref_jobs_, h_ref_jobs_.as<RefineJobDev>(), sizeof(RefineJobDev) * n,
                           hipMemcpyHostToDevice, s),
            "H2D");
  hip_check(hipEventRecord(jobs_copied_, s), "event");
  jobs_in_flight_ = true;
  hipEventRecord(ev_[2], s);
  if (grad_steps_ > 0) {
    refine_gradient(jobs, p);
  } else {
    int max_iter = 0;
    for (const RefineJob& r : jobs) max_iter = std::max(max_iter, r.n_iter);
    if (launch_refine(d_ref_jobs_, n, max_iter, k_.n_samples, p.imaged_center_n_grid_sample,
                      d_inv_ratio_, k_.bank_rows, k_.bank_cols, d_work_,
                      h_ref_out_.dev<RefineOutDev>(), h_ref_dist_.dev<float>(), s) != 0)
      throw std::runtime_error("k_refine launch failed");
  }
  hipEventRecord(ev_[3], s);
  hip_check(stream_wait(s), "sync refine");
  hipEventElapsedTime(&refine_us_, ev_[2], ev_[3]);
  refine_us_ *= 1000.f;
  unpack_refine(jobs, out);
}

void GpuIdentBackend::select_and_refine(const std::vector<CutJob>& cuts,
                                        std::vector<RefineJob>& jobs, const std::vector<int>& sel_of,
                                        const HostParams& p, std::vector<CutSelection>& sel,
                                        std::vector<RefineResult>& out) {
  const int nc = (int)cuts.size(), n = (int)jobs.size();
  if (grad_steps_ > 0 || nc == 0 || n == 0) {
    IdentBackend::select_and_refine(cuts, jobs, sel_of, p, sel, out);
    return;
  }
  if ((int)sel_of.size() != n) throw std::runtime_error("select_and_refine: one selection per job");
  for (const int q : sel_of)
    if (q < 0 || q >= nc) throw std::runtime_error("select_and_refine: selection index out of range");
  int npts = 0;
  for (const CutJob& j : cuts) npts += (int)j.outer_points.size();
  ensure(std::max(nc, n), npts);
  pack_select(cuts);
  pack_refine_jobs(jobs, &sel_of);
  hipStream_t s = stream_;
  hipEventRecord(ev_[0], s);
  if (launch_select_cuts(h_sel_jobs_.dev<SelectJobDev>(), nc, h_pts_.dev<float4>(), k_.start_sig,
                         k_.n_samples, k_.n_refine_samples, p.num_cuts_in_ident_step,
                         h_sel_out_.dev<SelectOutDev>(), s) != 0)
    throw std::runtime_error("k_select_cuts launch failed");
  hipEventRecord(ev_[1], s);
  // k_take_stops brings the refinement jobs over from mapped memory itself
  // (no copy-engine transfer between the selection and the refinement)
  if (launch_take_stops(h_ref_jobs_.dev<RefineJobDev>(), d_ref_jobs_, n,
                        h_sel_out_.dev<SelectOutDev>(), s) != 0)
    throw std::runtime_error("k_take_stops launch failed");
  hip_check(hipEventRecord(jobs_copied_, s), "event");
  jobs_in_flight_ = true;
  hipEventRecord(ev_[2], s);
  int max_iter = 0;
  for (const RefineJob& r : jobs) max_iter = std::max(max_iter, r.n_iter);
  if (launch_refine(d_ref_jobs_, n, max_iter, k_.n_samples, p.imaged_center_n_grid_sample,
                    d_inv_ratio_, k_.bank_rows, k_.bank_cols, d_work_,
                    h_ref_out_.dev<RefineOutDev>(), h_ref_dist_.dev<float>(), s) != 0)
    throw std::runtime_error("k_refine launch failed");
  hipEventRecord(ev_[3], s);
  hip_check(stream_wait(s), "sync select and refine");
  hipEventElapsedTime(&select_us_, ev_[0], ev_[1]);
  select_us_ *= 1000.f;
  hipEventElapsedTime(&refine_us_, ev_[2], ev_[3]);
  refine_us_ *= 1000.f;
  unpack_select(nc, sel);
  refine_jobs_ = n;
  refine_samples_ = 0;
  refine_flops_ = 0;
  const long long g2 = (long long)p.imaged_center_n_grid_sample * p.imaged_center_n_grid_sample;
  for (int r = 0; r < n; ++r) {
    jobs[r].stops = sel[sel_of[r]].stops;
    if (!jobs[r].stops.empty())
    {
      refine_samples_ += (long long)jobs[r].n_iter * g2 * (long long)jobs[r].stops.size() * k_.n_samples;
      refine_flops_ += refine_flops(jobs[r].n_iter, g2, (long long)jobs[r].stops.size(), k_.n_samples);
    }
  }
  unpack_refine(jobs, out);
}

}  // namespace cctagx
