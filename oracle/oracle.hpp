// ORACLE -- test infrastructure only.  Nothing in the product links, calls or
// loads this code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker and as the CPU baseline.
//
// A plain C++ restatement of the reference CCTag CPU detection path
// (alicevision/CCTag v1.0.3, src/cctag/...).  Every function cites the
// reference file:line it follows.  Scalar loops, no SIMD, no threads:
// the accumulation order of every float reduction is spelled out because the
// HIP kernels must reproduce it bit for bit.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

struct Params {
  float canny_low = 0.01f;        // Params.hpp:37 kDefaultCannyThrLow
  float canny_high = 0.04f;       // Params.hpp:38 kDefaultCannyThrHigh
  int dist_search = 30;           // Params.hpp:27 kDefaultDistSearch
  int thr_grad_vote = 2500;       // Params.hpp:30 kDefaultThrGradientMagInVote
  float angle_voting = 0.0f;      // Params.hpp:31
  float ratio_voting = 4.0f;      // Params.hpp:32
  int n_crowns = 3;               // Params.hpp:28
  int min_votes = 3;              // Params.hpp:40 kDefaultMinVotesToSelectCandidate
  int n_levels = 4;               // Params.hpp:45 kDefaultNumberOfProcessedMultiresLayers
};

// One pyramid level after the front end (what the reference keeps in a Level
// plus an EdgePointCollection, Level.hpp:51-56, Types.hpp:19-158).
struct LevelResult {
  int w = 0, h = 0;      // level size
  int W0 = 0, H0 = 0;    // EdgePointCollection shape = level-0 size (Multiresolution.cpp:286)
  std::vector<uint8_t> img;
  std::vector<int16_t> dx, dy;
  std::vector<uint8_t> edges;       // thinned Canny map, {0,255}
  std::vector<int32_t> edge_map;    // w*h, point index or -1
  std::vector<uint32_t> edge_bits;  // ceil(w*h/32): bit i of word j = pixel 32j+i is an edge
  // edge points in raster order (Canny.cpp:24-33)
  std::vector<int16_t> ex, ey, edx, edy;
  std::vector<int32_t> before, after;        // Vote.cpp:79-87
  std::vector<int32_t> chosen;               // Vote.cpp:227-229, -1 = no vote
  std::vector<float> total_dist;             // Vote.cpp:117-183
  std::vector<int32_t> voters_index;         // CSR, n+1 (Types.cpp:66-84)
  std::vector<int32_t> voters_list;
  std::vector<float> flow_length;            // Vote.cpp:232
  std::vector<int32_t> is_max;               // Vote.cpp:236-240, -1 = not a seed
  std::vector<int32_t> seeds;                // push order, Vote.cpp:237-239
  int edge_at(int x, int y) const {
    return (x < w && y < h) ? edge_map[(size_t)y * w + x] : -1;
  }
};

// cv::cvtColor(COLOR_BGR2GRAY) on an 8-bit BGR / BGRA frame.
void bgr_to_gray(const uint8_t* src, int w, int h, size_t stride, int channels, uint8_t* dst);
// Level.cpp:67 -- cv::resize(src, dst, Size(dw, dh)) on 8U, dw <= sw, dh <= sh.
void downsample(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh);
// cvRecode.cpp:74-92 -- 9x9 derivative filters, BORDER_REPLICATE, int16.
void derivatives(const uint8_t* img, int w, int h, int16_t* dx, int16_t* dy);
// The same filters returning the float value before rounding.
void derivatives_float(const uint8_t* img, int w, int h, float* dx, float* dy);
// cvRecode.cpp:143-414 -- L2 magnitude, NMS, stack hysteresis, final map.
void canny(const int16_t* dx, const int16_t* dy, int w, int h, float low, float high,
           uint8_t* edges);
// thinning.cpp:15-51 -- two LUT passes.
void thin(uint8_t* edges, int w, int h);
// Canny.cpp + Bresenham.cpp + Vote.cpp:58-247 on one level.
void edges_and_vote(LevelResult& L, const Params& p);
// ImagePyramid.cpp:35-52 + the above for every level.
void front_end(const uint8_t* img, int w, int h, const Params& p,
               std::vector<LevelResult>& levels);

// 512-entry thinning tables (thinning.cpp:12-13), value 0 or 255.
int thin_lut(int pass, int index);

}  // namespace oracle
