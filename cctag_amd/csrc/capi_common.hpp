// Helpers shared by the C-ABI entry points of the plain and batched paths.
#pragma once

#include <algorithm>
#include <cmath>

#include "../../include/cctag_amd.h"
#include "frontend.hpp"
#include "host/ident.hpp"
#include "ident_device.hpp"

namespace cctagx {

inline int validate_params(const cctagx_params* p) {
  if (!p) return CCTAGX_ERR_ARG;
  if (p->n_crowns != 3 && p->n_crowns != 4) return CCTAGX_ERR_UNSUPPORTED;
  if (p->angle_voting != 0.0f) return CCTAGX_ERR_ARG;  // Vote.cpp:95-98 throws
  if (p->number_of_processed_multires_layers < 1 || p->number_of_processed_multires_layers > 4)
    return CCTAGX_ERR_UNSUPPORTED;
  if (p->dist_search < 0 || p->min_votes_to_select_candidate < 1) return CCTAGX_ERR_ARG;
  // the linking's convexity test reads the first and last of the window
  if (p->window_size_on_inner_elliptic_segment < 1) return CCTAGX_ERR_ARG;
  // fixed bounds of the identification kernels and of the candidate stage's
  // 64-bit per-candidate masks
  if (p->imaged_center_n_grid_sample < 3 || p->imaged_center_n_grid_sample > kMaxGrid ||
      p->num_cuts_in_ident_step < 1 || p->num_cuts_in_ident_step > kMaxCuts ||
      p->sample_cut_length < 31 || p->sample_cut_length > kMaxSamples ||
      p->num_samples_outer_edge_points_refinement < 9 ||
      p->num_samples_outer_edge_points_refinement > kMaxRefineSamples ||
      p->n_samples_outer_ellipse < 5 || p->n_samples_outer_ellipse > kMaxOuterPoints ||
      p->maximum_nb_candidates_loop_two < 0 || p->maximum_nb_candidates_loop_two > 64)
    return CCTAGX_ERR_UNSUPPORTED;
  return CCTAGX_OK;
}

inline FrontEndParams front_end_params(const cctagx_params& p) {
  // Level.cpp:206-207 passes thr * 256; cvRecode.cpp:67,107-108 orders and floors them.
  float lo = p.canny_thr_low * 256, hi = p.canny_thr_high * 256;
  if (lo > hi) std::swap(lo, hi);
  FrontEndParams f{};
  f.low = (int)std::floor(lo);
  f.high = (int)std::floor(hi);
  f.dist_search = p.dist_search;
  f.thr_grad = p.thr_gradient_mag_in_vote;
  f.ratio_voting = p.ratio_voting;
  f.angle_voting = p.angle_voting;
  f.n_crowns = p.n_crowns;
  f.min_votes = p.min_votes_to_select_candidate;
  return f;
}

}  // namespace cctagx
