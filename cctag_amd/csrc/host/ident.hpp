// Identification: host-side control (shared) around a backend that runs the
// data-parallel parts -- cut collection and selection (identify_step_1),
// imaged-centre grid search (refineConicFamilyGlob) and per-cut ID
// distances (orazioDistanceRobust).  The library's backend is the GPU; the
// oracle's is a CPU restatement.
//
// Reference: Identification.cpp:50-193,396-1480, Identification.hpp,
// ImageCut.hpp, optimization/conditioner.hpp, CCTagMarkersBank.cpp.
#pragma once

#include <cstdint>
#include <vector>

#include "candidate.hpp"
#include "geometry.hpp"

namespace cctagx {

constexpr int kMaxCuts = 32;         // numCutsInIdentStep upper bound accepted
constexpr int kMaxSamples = 128;     // sampleCutLength upper bound accepted
constexpr int kMaxOuterPoints = 256; // nSamplesOuterEllipse upper bound accepted
constexpr int kMaxBank = 128;        // bank rows
constexpr int kMaxRatios = 8;        // radius ratios per row

namespace status {
constexpr int id_reliable = 1;
constexpr int too_few_outer_points = -1;
constexpr int no_collected_cuts = -1;
constexpr int no_selected_cuts = -2;
constexpr int opti_has_diverged = -3;
constexpr int id_not_reliable = -4;
constexpr int degenerate = -5;
}  // namespace status

// Input of identify_step_1 for one marker after the host's getSortedOuterPoints.
struct CutJob {
  int image = 0;                     // index of the frame's image in the backend
  float cx = 0, cy = 0;              // outer ellipse centre (cut start)
  std::vector<Pt2> outer_points;     // sorted, subsampled outer points (cut stops, with gradients)
  double ca = 1, sa = 0;             // cos / sin of the ellipse angle (double C functions)
  float ea = 1, eb = 1;              // ellipse semi-axes
  float scale = 1;                   // marker pyramid scale
};

// Selected cuts of one marker (ImageCut start/stop; start is the centre).
struct CutSelection {
  int n_collected = 0;
  std::vector<Pt2> stops;            // <= numCutsInIdentStep, with gradients
};

// Per-marker constants of the grid search, precomputed on the host (the
// only transcendental functions of the search live here).
struct RefineJob {
  int image = 0;                     // index of the frame's image in the backend
  Mat3 canonic, t_can, t_inv_can;    // Ellipse::getCanonicForm
  Mat3 cond, cond_inv;               // conditionerFromEllipse and its inverse
  float grid_scale = 0;              // max(a, b) of the conditioned ellipse
  float max_semi_axis = 0;           // max(a, b) of the ellipse
  int n_iter = 0;                    // neighbourhood halvings until size*maxSemiAxis <= 0.02
  float neighbour0 = 0.2f;           // imagedCenterNeighbourSize
  Pt2 center0;                       // initial imaged centre (CCTag::centerImg)
  std::vector<Pt2> stops;            // selected cut stops
  float begin_sig = 0.25f, end_sig = 1.f;
  float box[4] = {0, 0, 0, 0};       // the ellipse's bounding box: x min, x max, y min, y max
};

// Whether a job's ellipse box lies at least 2 px inside a w x h image.
bool box_inside(const RefineJob& r, int w, int h);

struct RefineResult {
  bool ok = false;                   // every grid search found a readable point
  Pt2 center;
  Mat3 H = Mat3::zero();
  float min_res = 0;                 // residual of the last grid search
  std::vector<uint8_t> oob;          // per cut, sticky out-of-bounds flag
  // range of the median profile of the valid cuts' signals at the optimum
  // (Identification.cpp:920-934)
  float bar_min = 0, bar_max = 0;
  std::vector<float> id_dist;        // n_cuts x bank rows: orazio distance, valid cuts only
};

// bar_min / bar_max of a RefineResult from the n_cuts x n_samples signals at
// the optimum and the cuts' out-of-bounds flags.
void median_profile_range(const std::vector<float>& signals, const std::vector<uint8_t>& oob,
                          size_t n_samples, float& bar_min, float& bar_max);

class IdentBackend {
 public:
  virtual ~IdentBackend() = default;
  // collectCuts + selectCutCheapUniform for every job.
  virtual void select_cuts(const std::vector<CutJob>& jobs, const HostParams& p,
                           std::vector<CutSelection>& out) = 0;
  // refineConicFamilyGlob's search + final signals + orazio distances.
  virtual void refine(const std::vector<RefineJob>& jobs, const HostParams& p,
                      std::vector<RefineResult>& out) = 0;
  // Both steps at once: select_cuts over `cuts`, then refine over `jobs`,
  // job r taking the stops selected for cut job sel_of[r] (written into its
  // `stops`); a job whose selection is empty is not refined and its result
  // is not ok.  The results are those of the two calls; this default makes
  // them one after the other.
  virtual void select_and_refine(const std::vector<CutJob>& cuts, std::vector<RefineJob>& jobs,
                                 const std::vector<int>& sel_of, const HostParams& p,
                                 std::vector<CutSelection>& sel, std::vector<RefineResult>& out);
};

// Bank rows (CCTagMarkersBank.cpp:22-52), innermost ring first.
std::vector<std::vector<float>> marker_bank(int n_crowns);

// The float-exact parameters of the identification that backends need.
struct IdentConstants {
  float start_sig = 0.25f;           // Identification.cpp:1221-1235
  int n_samples = 100;
  int n_refine_samples = 20;
  int bank_rows = 0, bank_cols = 0;
  float inv_ratio[kMaxBank][kMaxRatios];  // 1.f / rr_j (Identification.cpp:144)
};
IdentConstants ident_constants(const HostParams& p, const std::vector<std::vector<float>>& bank);

// The grid search's per-marker constants (Identification.cpp:1318-1370):
// canonic form of the outer ellipse, conditioner with the cached (a+b)/2
// `mean_ab`, grid scale and iteration count.  False when the conditioned
// ellipse is degenerate.
bool make_refine_job(const Ellipse& e, const Pt2& center0, const std::vector<Pt2>& stops,
                     const HostParams& p, const IdentConstants& kc, float mean_ab, RefineJob& r);

// Runs identification on every marker (Detection.cpp:873-965) and sets
// status, id, centre, homography, quality.  The (a+b)/2 that
// conditionerFromEllipse caches on its first call is taken from the first
// marker refined in this call.
void identify_markers(std::vector<Marker>& markers, const HostParams& p,
                      const std::vector<std::vector<float>>& bank, IdentBackend& backend,
                      TaskPool* pool = nullptr, HostTimes* times = nullptr);

// The same for several frames at once: one select_and_refine call of the
// backend for all their markers (and one refine call for the frames whose
// first marker with cuts gets none selected); each frame's markers come out
// as identify_markers would leave them (the conditioner cache is per frame).
struct IdentFrame {
  std::vector<Marker>* markers = nullptr;
  int image = 0;                     // backend image index of the frame
};
void identify_frames(std::vector<IdentFrame>& frames, const HostParams& p,
                     const std::vector<std::vector<float>>& bank, IdentBackend& backend,
                     TaskPool* pool = nullptr, HostTimes* times = nullptr);

}  // namespace cctagx
