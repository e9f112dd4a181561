// Device-side job layout and launchers of the differentiable refinement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "host/ident.hpp"

namespace cctagx {

struct GradJobDev {
  const uint32_t* quad;  // the marker's image (level 0) as quads (bytes (x, y), (x+1, y),
                         // (x, y+1), (x+1, y+1) of word (x, y)), row-major, width iw
  int iw, ih;
  float Q11, Q22, Q33;             // canonic form of the outer ellipse
  float t_can[9], t_inv_can[9];    // image <-> canonic frame
  float begin_sig, end_sig;
  float alpha0;                    // first step of the centre descent, px
  int n_cuts;
  float c0x, c0y;                  // initial imaged centre
  float2 stops[kMaxCuts];
};

// The quads of a w x h image (w, h >= 2) into quad (w * h words).
int launch_quads(const uint8_t* img, int w, int h, uint32_t* quad, hipStream_t s);
int launch_grad_eval(const GradJobDev* jobs, int n, const float* H, int n_samples, bool fp64,
                     double* cost, double* grad, hipStream_t s);
int launch_grad_homography(const GradJobDev* jobs, int n, const float* centers, float* H,
                           hipStream_t s);
// start (2 per marker) = coarse[coarse_of[m]] when coarse_of[m] >= 0, else
// the job's c0.
int launch_grad_start(const GradJobDev* jobs, int n, const int* coarse_of, const float* coarse,
                      float* start, hipStream_t s);
// `steps` descent steps per marker from start (2 per marker) or, when start
// is null, from the job's c0.
int launch_grad_descend(const GradJobDev* jobs, int n, int steps, int n_samples, bool fp64,
                        const float* start, float* centers, double* cost, hipStream_t s);

}  // namespace cctagx
