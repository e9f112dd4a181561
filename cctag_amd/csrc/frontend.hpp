// Device-side data layout of the CCTag front end (pyramid .. seeds).
//
// A "plane" is one pyramid level of one frame.  Every front-end kernel takes
// the whole plane set and maps its blocks over the planes' tiles, so one
// launch covers all levels of a frame -- and, for a batch, all frames.
//
// Pixel-indexed arrays (img, dx, dy, cls, lab, ...) hold every plane back to
// back; plane p starts at PlaneDesc::pix.  Edge-indexed arrays (edge list,
// links, votes, voter lists) reuse the same offsets: a plane of N pixels can
// never hold more than N edge points, so plane p's points live at
// [pix, pix + n_edges).  CSR voter offsets need n+1 entries per plane and
// start at pix + p.
#pragma once

#include <cstdint>

namespace cctagx {

struct PlaneDesc {
  int32_t w, h;             // level size
  int32_t shape_w, shape_h; // EdgePointCollection shape = level-0 size of the frame
  int32_t pix;              // first pixel / edge slot of this plane
  int32_t parent;           // plane of the next-finer level, -1 for level 0
  int32_t level;            // pyramid level
  int32_t frame;            // frame index within a batch
  int32_t tile_f;           // first block of the default filter grid (64x32 tiles)
  int32_t tile_s;           // first block of the stencil grids (64x16 tiles)
  int32_t tile_c;           // first block of the compaction grid (4096-px chunks)
  int32_t tile_p;           // first block of the pyramid / quad grid (level-0 planes only)
  int32_t tile_r;           // first block of its level's resampling grid (levels >= 1)
};

// Per-plane counters written by the GPU.
struct PlaneCounts {
  int32_t n_edges;
  int32_t n_seeds;
  int32_t n_voters;
  int32_t pad;
};

// Blocks per plane of the two-pass scans over edge points (voter offsets,
// seed compaction); each block scans one contiguous range.
constexpr int kScanBlocks = 64;

// Tile geometry shared by host launch code and kernels.
constexpr int kFilterTileW = 64, kFilterTileH = 32;
// Filter tile heights the launch can choose from (64 wide, 256 threads: a
// thread owns one column of TH/4 rows); variant 1 is the default 64x32.
constexpr int kFilterVariants = 3;
constexpr int kFilterTileHs[kFilterVariants] = {16, 32, 64};
constexpr int kStencilTileW = 64, kStencilTileH = 16;
constexpr int kChunk = 4096;        // compaction chunk, pixels
constexpr int kPyrTile = 64;        // pyramid tile at level 0 (-> 32, 16, 8)
constexpr int kResTileW = 64, kResTileH = 16;  // resampling tile (one level from its parent)

struct FrontEndParams {
  int32_t low, high;        // integer Canny thresholds (floor of thr * 256)
  int32_t dist_search;
  int32_t thr_grad;
  float ratio_voting;
  float angle_voting;
  int32_t n_crowns;
  int32_t min_votes;
};

// Device buffers of one plane set.
struct FrontEndBuffers {
  uint8_t* img;        // Npix
  uint32_t* quad;      // Npix, level-0 planes only: pixel (x, y)'s 2x2 neighbourhood, bytes
                       // (x, y), (x+1, y), (x, y+1), (x+1, y+1), indices clamped to the plane
  int16_t* dx;         // Npix
  int16_t* dy;         // Npix
  uint8_t* cls;        // Npix: bit0 candidate (passes NMS and low), bit1 strong
  int32_t* lab;        // Npix: union-find parent of candidate pixels
  uint8_t* root_strong;// Npix: component root holds a strong pixel
  uint8_t* edges;      // Npix: Canny map, then thinned in place
  uint32_t* ebits;     // Npix/32: edge bitmask, bit i of word j = pixel 32j+i (plane starts are 256-aligned)
  int32_t* erank;      // Npix/32: index (raster rank in the plane) of word j's first edge pixel
  int32_t* tcount;     // compaction chunks: count, then exclusive offset
  int16_t* ex;         // Npix (edge slots)
  int16_t* ey;
  int16_t* egx;
  int16_t* egy;
  int32_t* before;
  int32_t* after;
  int32_t* chosen;
  float* tdist;
  int32_t* vcount;     // votes per point
  int32_t* vindex;     // CSR offsets, Npix + nplanes
  int32_t* vlist;      // voter indices
  float* flow;
  int32_t* ismax;
  int32_t* seedmark;   // voter index -> point that reached min votes with it
  int32_t* seeds;
  int32_t* scan_part;  // nplanes x kScanBlocks partial sums of the two-pass scans
  PlaneCounts* counts; // nplanes
};

}  // namespace cctagx
