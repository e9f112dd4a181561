// Device-side identification: job / result layouts of the identification
// kernels and the GPU implementation of the host's IdentBackend.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"
#include "grad_device.hpp"
#include "host/ident.hpp"

namespace cctagx {

constexpr int kMaxGrid = 5;        // imagedCenterNGridSample upper bound accepted
constexpr int kMaxBankRows = 128;  // rows of the 4-crown bank
constexpr int kMaxRefineSamples = 32;  // numSamplesOuterEdgePointsRefinement bound

struct SelectJobDev {
  const uint32_t* quad;  // the marker's frame (level 0) as quads, row-major, width iw
  int iw, ih;
  float cx, cy;
  double ca, sa;
  float ea, eb, scale;
  int n_outer, pt_offset;
};

struct SelectOutDev {
  int n_collected, n_selected;
  float4 stops[kMaxCuts];  // x, y, dx, dy
};

struct RefineJobDev {
  const uint32_t* quad;  // the marker's frame (level 0) as quads, row-major, width iw
  int iw, ih;
  float Q11, Q22, Q33;
  float t_can[9], t_inv_can[9], cond[9], cond_inv[9];
  float grid_scale, neighbour0, begin_sig, end_sig;
  int n_iter, n_cuts;
  int box_inside;  // the outer ellipse's bounding box lies 2 px inside the image
  float c0x, c0y;
  int dist_offset;  // first float of its distances in the packed output
  int sel_index;    // >= 0: the selection whose stops the job takes on the device
  float2 stops[kMaxCuts];
};

struct RefineOutDev {
  int ok;
  float cx, cy;
  float H[9];
  float min_res;
  float bar_min, bar_max;  // range of the per-sample median over the valid cuts
  uint8_t oob[kMaxCuts];
};

int launch_select_cuts(const SelectJobDev* jobs, int n_jobs, const float4* pts, float start_sig,
                       int n_samples, int n_ref, int select_max, SelectOutDev* out, hipStream_t s);
// Copies the jobs from src (host memory the device maps) to jobs; those
// with sel_index >= 0 take the stops of that selection (n_cuts = the
// selected count; no iteration when it is zero).
int launch_take_stops(const RefineJobDev* src, RefineJobDev* jobs, int n_jobs,
                      const SelectOutDev* sel, hipStream_t s);
// Grid-search state of one marker between iteration launches.
static_assert(kMaxCuts <= 32, "one bit per cut in the flag words");
struct RefineStateDev {
  float cx, cy, neighbour, min_res;
  int ok, best_g;
  uint32_t sticky;     // bit k: cut k left the image at an earlier grid point
  int pad;
};
// One grid point of the current iteration.
struct RefineGridDev {
  float H[9], Hi[9];
  float gx, gy;
  double cost_local;   // residual sum assuming only this point's own flags
  int16_t m_local;     // valid cuts under that assumption
  int16_t has_gram;    // its Gram matrix was stored (the selection may recompute from it)
  uint32_t oob;        // bit k: a sample of cut k is outside the image
};
static_assert(sizeof(RefineGridDev) == 96, "grid record size");
constexpr int kMaxPairs = kMaxCuts * (kMaxCuts + 1) / 2;
// Two sets (by iteration parity): an iteration reads the previous set and
// writes its own.
struct RefineWork {
  RefineStateDev* state;  // 2 x n_jobs
  RefineGridDev* grid;    // 2 x n_jobs x kMaxGrid^2
  double* gram;           // 2 x n_jobs x kMaxGrid^2 x kMaxPairs
};

// out: one header per job; dist: per job n_cuts x bank_rows distances from
// its dist_offset on (invalid cuts' rows are not written).
// The grid search's first iteration only: per job the centre it moves to
// (centers: 2 per job, device memory).
int launch_refine_coarse(const RefineJobDev* jobs, int n_jobs, int n_samples, int grid_n,
                         const RefineWork& work, float* centers, hipStream_t s);
// The read-out of launch_refine's last kernel at given homographies (9 per
// job), centres (2 per job) and scores instead of the grid search's choice.
int launch_refine_readout(const RefineJobDev* jobs, int n_jobs, int n_samples, const float* H,
                          const float* centers, const double* cost, const float* inv_ratio,
                          int bank_rows, int bank_cols, const RefineWork& work, RefineOutDev* out,
                          float* dist, hipStream_t s);
int launch_refine(const RefineJobDev* jobs, int n_jobs, int max_iter, int n_samples, int grid_n,
                  const float* inv_ratio, int bank_rows, int bank_cols, const RefineWork& work,
                  RefineOutDev* out, float* dist, hipStream_t s);

// Algorithmic FLOPs of the grid search on one marker: n_iter iterations of
// grid_points homographies, each sampling `cuts` cuts of `samples` samples
// and differencing every cut pair (3.1 MFLOP per iteration for 22 cuts of
// 100 samples on a 5x5 grid).
long long refine_flops(long long n_iter, long long grid_points, long long cuts, long long samples);

// A device image the identification jobs sample, as quads: word (x, y)
// holds pixels (x, y), (x+1, y), (x, y+1), (x+1, y+1) in bytes 0..3.
struct DeviceImage {
  const uint32_t* quad;
  int w, h;
};

// IdentBackend on the GPU: jobs packed into pinned memory, one launch per
// step for all markers of the frame.  Errors are thrown as std::runtime_error.
class GpuIdentBackend : public IdentBackend {
 public:
  int init(const std::vector<std::vector<float>>& bank, const HostParams& p, hipStream_t s);
  // Replaces the marker bank the read-out scores (after init).
  void set_bank(const std::vector<std::vector<float>>& bank, const HostParams& p);
  // Images that CutJob::image / RefineJob::image index.
  void set_images(const std::vector<DeviceImage>& images) { images_ = images; }
  void set_image(const uint32_t* dev_quad, int w, int h) { images_.assign(1, DeviceImage{dev_quad, w, h}); }
  void select_cuts(const std::vector<CutJob>& jobs, const HostParams& p,
                   std::vector<CutSelection>& out) override;
  void refine(const std::vector<RefineJob>& jobs, const HostParams& p,
              std::vector<RefineResult>& out) override;
  // Both steps in one submission (the refinement jobs take their stops from
  // the selection on the device) and one wait; the gradient refinement
  // takes the two calls.
  void select_and_refine(const std::vector<CutJob>& cuts, std::vector<RefineJob>& jobs,
                         const std::vector<int>& sel_of, const HostParams& p,
                         std::vector<CutSelection>& sel, std::vector<RefineResult>& out) override;
  // 0: the reference grid search (the default); steps > 0: its first
  // iteration, then `steps` descent steps on the differentiable score from
  // the centre it picked, and the read-out there.
  void set_gradient_steps(int steps) { grad_steps_ = steps; }
  int gradient_steps() const { return grad_steps_; }
  // The centres after the grid search's first iteration, into device memory
  // (2 floats per job), on the backend's stream without synchronizing.
  void coarse_centers(const std::vector<RefineJob>& jobs, const HostParams& p, float* d_centers);
  // Device time of the last select / refine launches, microseconds.
  float last_select_us() const { return select_us_; }
  float last_refine_us() const { return refine_us_; }
  // Work of the last launches: markers and bilinear image samples.
  int last_select_jobs() const { return select_jobs_; }
  int last_refine_jobs() const { return refine_jobs_; }
  long long last_select_samples() const { return select_samples_; }
  long long last_refine_samples() const { return refine_samples_; }
  // Algorithmic FLOPs of the last grid search (refine_flops over its jobs).
  long long last_refine_flops() const { return refine_flops_; }
  void clear_stats() {
    select_us_ = refine_us_ = 0;
    select_jobs_ = refine_jobs_ = 0;
    select_samples_ = refine_samples_ = refine_flops_ = 0;
  }

 private:
  void ensure(int n_markers, int n_points);
  // sel_of: null, or per job the selection it takes its stops from on the
  // device (its distances then at a fixed stride of kMaxCuts rows)
  int pack_refine_jobs(const std::vector<RefineJob>& jobs, const std::vector<int>* sel_of = nullptr);
  void pack_select(const std::vector<CutJob>& jobs);
  void unpack_select(int n, std::vector<CutSelection>& out);
  void unpack_refine(const std::vector<RefineJob>& jobs, std::vector<RefineResult>& out);
  // host copies of the mapped result buffers, and where each job's distances
  // start in them (set by pack_refine_jobs)
  std::vector<SelectOutDev> host_sel_;
  std::vector<RefineOutDev> host_out_;
  std::vector<float> host_dist_;
  std::vector<int> dist_offset_;
  int dist_total_ = 0;
  hipStream_t stream_ = nullptr;
  std::vector<DeviceImage> images_;
  const DeviceImage& image(int i) const;
  IdentConstants k_{};
  int cap_markers_ = 0, cap_points_ = 0;
  DeviceArena dev_;
  PinnedBuffer h_sel_jobs_, h_pts_, h_sel_out_, h_ref_jobs_, h_ref_out_, h_ref_dist_;
  RefineJobDev* d_ref_jobs_ = nullptr;
  RefineWork d_work_{};
  float* d_inv_ratio_ = nullptr;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t jobs_copied_ = nullptr;  // the last copy out of h_ref_jobs_
  bool jobs_in_flight_ = false;
  int grad_steps_ = 0;
  void refine_gradient(const std::vector<RefineJob>& jobs, const HostParams& p);
  PinnedBuffer h_grad_jobs_;
  DeviceBuffer d_grad_jobs_, d_coarse_, d_gcenters_, d_gcost_, d_gH_;
  float select_us_ = 0, refine_us_ = 0;
  int select_jobs_ = 0, refine_jobs_ = 0;
  long long select_samples_ = 0, refine_samples_ = 0, refine_flops_ = 0;
};

}  // namespace cctagx
