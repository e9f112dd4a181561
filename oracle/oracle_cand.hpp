// ORACLE -- test infrastructure only (see oracle.hpp).
//
// Serial restatement of the CCTag candidate stage on the CPU: from the front
// end's levels (edge points, links, voter lists, seeds) to the marker
// candidates of every level, their level-0 re-fit, and -- after
// identification -- the de-duplication of the final list.  Reference:
// Detection.cpp:69-693, Vote.cpp:249-742, EllipseGrowing.cpp,
// Fitting.cpp:228-364, Multiresolution.cpp:43-413, CCTag.{hpp,cpp}.
//
// It is written from the reference independently of the library's host
// candidate stage (cctag_amd/csrc/host/candidate.cpp, pipeline.cpp), which
// it checks; what the two share are the restated numerical primitives of
// host/geometry.cpp (the Ellipse class and its parameters, the
// Halir-Flusser and circle fits, point-to-ellipse distances, the perimeter
// rasterization, hull tests, medians and the pcg32 stream), as the
// reference's candidate stage uses its own geometry library, Eigen and
// pcg-cpp.  The reference runs its candidate loops on TBB threads, so its
// own random stream and label order depend on the schedule; this oracle
// fixes them as documented for the library (every loop-two task, loop-three
// task and level-0 re-fit draws from its own pcg32(271828); segment labels
// in candidate order), runs the loops in order and, given a pool, runs the
// independent tasks of loops two and three and of the re-fit on it.
#pragma once

#include <vector>

#include "oracle.hpp"
#include "../cctag_amd/csrc/host/candidate.hpp"

namespace oracle {

struct CandidateParams {
  float average_vote_min = 0.f;                        // Params.hpp:33
  float thr_median_distance_ellipse = 3.f;             // Params.hpp:34
  int maximum_nb_seeds = 500;                          // Params.hpp:35
  int maximum_nb_candidates_loop_two = 40;             // Params.hpp:36
  int n_crowns = 3;                                    // Params.hpp:28
  int min_points_segment_candidate = 10;               // Params.hpp:39
  float thresh_robust_estimation_of_outer_ellipse = 30.f;  // Params.hpp:41
  float ellipse_growing_elliptic_hull_width = 2.3f;    // Params.hpp:42
  int window_size_on_inner_elliptic_segment = 20;      // Params.hpp:43
  bool search_for_another_segment = true;              // Params.hpp:56
};

// cctagMultiresDetection (Multiresolution.cpp:268-413) after the front end:
// the markers of every level (coarsest first), re-fitted to level 0.
void candidates(const std::vector<LevelResult>& levels, const CandidateParams& p,
                std::vector<cctagx::Marker>& markers, cctagx::TaskPool* pool = nullptr);

// Detection.cpp:974-989: update() twice over the identified markers, then
// the list sorted by id (a stable sort).
void deduplicate(const std::vector<cctagx::Marker>& markers, std::vector<cctagx::Marker>& out);

}  // namespace oracle
