// Conversions between the C ABI structs and the host types.
#pragma once

#include "../../../include/cctag_amd.h"
#include "candidate.hpp"

namespace cctagx {

inline HostParams host_params(const cctagx_params& c) {
  HostParams p;
  p.average_vote_min = c.average_vote_min;
  p.thr_median_distance_ellipse = c.thr_median_distance_ellipse;
  p.maximum_nb_seeds = c.maximum_nb_seeds;
  p.maximum_nb_candidates_loop_two = c.maximum_nb_candidates_loop_two;
  p.n_crowns = c.n_crowns;
  p.min_points_segment_candidate = c.min_points_segment_candidate;
  p.min_votes_to_select_candidate = c.min_votes_to_select_candidate;
  p.thresh_robust_estimation_of_outer_ellipse = c.thresh_robust_estimation_of_outer_ellipse;
  p.ellipse_growing_elliptic_hull_width = c.ellipse_growing_elliptic_hull_width;
  p.window_size_on_inner_elliptic_segment = c.window_size_on_inner_elliptic_segment;
  p.n_samples_outer_ellipse = c.n_samples_outer_ellipse;
  p.num_cuts_in_ident_step = c.num_cuts_in_ident_step;
  p.num_samples_outer_edge_points_refinement = c.num_samples_outer_edge_points_refinement;
  p.sample_cut_length = c.sample_cut_length;
  p.imaged_center_n_grid_sample = c.imaged_center_n_grid_sample;
  p.imaged_center_neighbour_size = c.imaged_center_neighbour_size;
  p.min_ident_proba = c.min_ident_proba;
  p.search_for_another_segment = c.search_for_another_segment != 0;
  p.do_identification = c.do_identification != 0;
  return p;
}

inline void to_c_marker(const Marker& m, cctagx_marker& o) {
  o.x = m.center_img.x;
  o.y = m.center_img.y;
  o.id = m.id;
  o.status = m.status;
  o.quality = m.quality;
  o.pyramid_level = m.pyramid_level;
  o.scale = m.scale;
  const Ellipse& e = m.rescaled_outer_ellipse;
  o.ellipse[0] = e.cx();
  o.ellipse[1] = e.cy();
  o.ellipse[2] = e.a();
  o.ellipse[3] = e.b();
  o.ellipse[4] = e.angle();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.homography[i * 3 + j] = m.homography.m[i][j];
}

}  // namespace cctagx
