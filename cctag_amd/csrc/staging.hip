// LevelStager and its gather kernel (see staging.hpp).
#include "staging.hpp"

#include <algorithm>
#include <cstring>

namespace cctagx {
namespace {

constexpr uint64_t kSegPiece = 64 * 1024;  // bytes per gather block

// Copies [src, src + n) to d, 16-byte moves when both allow, else 4-byte,
// else bytes (the block's threads share the work).
__device__ __forceinline__ void block_copy(uint8_t* d, const uint8_t* src, uint64_t n) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(d);
  uint64_t done = 0;
  if ((a & 15) == 0) {
    const uint64_t n16 = n / 16;
    for (uint64_t i = threadIdx.x; i < n16; i += blockDim.x)
      reinterpret_cast<uint4*>(d)[i] = reinterpret_cast<const uint4*>(src)[i];
    done = n16 * 16;
  } else if ((a & 3) == 0) {
    const uint64_t n4 = n / 4;
    for (uint64_t i = threadIdx.x; i < n4; i += blockDim.x)
      reinterpret_cast<uint32_t*>(d)[i] = reinterpret_cast<const uint32_t*>(src)[i];
    done = n4 * 4;
  }
  for (uint64_t i = done + threadIdx.x; i < n; i += blockDim.x) d[i] = src[i];
}

// One block per piece.  Reads exactly [src, src + bytes).
__global__ __launch_bounds__(256) void k_pack(const CopySeg* __restrict__ segs, int n_segs,
                                              uint8_t* __restrict__ dst) {
  const CopySeg sg = segs[blockIdx.x];
  block_copy(dst + sg.dst, static_cast<const uint8_t*>(sg.src), sg.bytes);
}

__host__ __device__ inline uint64_t align16(uint64_t v) { return (v + 15) & ~uint64_t(15); }

// The planned gather: every block lays the np planes' arrays out in LDS
// from the device counts (the order and alignment LevelStager::stage uses),
// block 0 writes the header, and the blocks take the kSegPiece pieces of
// the arrays in turn.
constexpr int kPlannedSegs = kMaxPlannedPlanes * kStageArrays;
__global__ __launch_bounds__(256) void k_pack_planned(const PlaneDesc* __restrict__ P, int np,
                                                      FrontEndBuffers b, uint8_t* __restrict__ dst,
                                                      uint64_t capacity) {
  __shared__ const uint8_t* src[kPlannedSegs];
  __shared__ uint64_t bytes[kPlannedSegs], start[kPlannedSegs], first_piece[kPlannedSegs + 1];
  __shared__ uint64_t total;
  const int nseg = np * kStageArrays;
  if (threadIdx.x == 0) {
    uint64_t at = 0, pieces = 0;
    int k = 0;
    for (int p = 0; p < np; ++p) {
      const PlaneDesc d = P[p];
      const PlaneCounts c = b.counts[p];
      const uint64_t N = (uint64_t)d.w * d.h, n = (uint64_t)c.n_edges, nv = (uint64_t)c.n_voters,
                     nsd = (uint64_t)c.n_seeds;
      const void* s_[kStageArrays] = {b.ebits + (d.pix >> 5), b.ex + d.pix,     b.ey + d.pix,
                                      b.egx + d.pix,          b.egy + d.pix,    b.before + d.pix,
                                      b.after + d.pix,        b.vindex + d.pix + p,
                                      b.vlist + d.pix,        b.flow + d.pix,   b.seeds + d.pix};
      const uint64_t z[kStageArrays] = {4 * ((N + 31) / 32), 2 * n, 2 * n, 2 * n, 2 * n,
                                        4 * n, 4 * n, 4 * (n + 1), 4 * nv, 4 * n, 4 * nsd};
      for (int a = 0; a < kStageArrays; ++a, ++k) {
        src[k] = static_cast<const uint8_t*>(s_[a]);
        bytes[k] = z[a];
        start[k] = at;
        first_piece[k] = pieces;
        at = align16(at + z[a]);
        pieces += (z[a] + kSegPiece - 1) / kSegPiece;
      }
    }
    first_piece[nseg] = pieces;
    total = at;
  }
  __syncthreads();
  const bool overflow = kStageDataStart + total > capacity;
  if (blockIdx.x == 0) {
    StageHeader* h = reinterpret_cast<StageHeader*>(dst);
    for (int k = threadIdx.x; k < nseg; k += blockDim.x)
      h->planes[k / kStageArrays].off[k % kStageArrays] = start[k];
    if ((int)threadIdx.x < np) h->planes[threadIdx.x].counts = b.counts[threadIdx.x];
    if (threadIdx.x == 0) {
      h->total = total;
      h->overflow = overflow ? 1u : 0u;
      h->n_planes = (uint32_t)np;
    }
  }
  if (overflow) return;
  uint8_t* data = dst + kStageDataStart;
  for (uint64_t q = blockIdx.x; q < first_piece[nseg]; q += gridDim.x) {
    int k = 0;
    while (first_piece[k + 1] <= q) ++k;
    const uint64_t off = (q - first_piece[k]) * kSegPiece;
    block_copy(data + start[k] + off, src[k] + off, min(kSegPiece, bytes[k] - off));
  }
}

}  // namespace

int LevelStager::grow_planned(size_t data_bytes) {
  return h_planned_.ensure(kStageDataStart + data_bytes, hipHostMallocCoherent) ? -2 : 0;
}

int LevelStager::reserve_planned(const PlaneSet& ps) {
  if (h_planned_.capacity() != 0) return 0;
  // first guess: 4 bytes per pixel of the plane set (a 1080p frame needs
  // about 1); finish_planned reports a frame that needs more
  size_t px = 0;
  for (const PlaneDesc& d : ps.planes) px += (size_t)d.w * d.h;
  return grow_planned(4 * px);
}

int LevelStager::enqueue_planned(const PlaneSet& ps, const PlaneDesc* dev_planes,
                                 const FrontEndBuffers& b, hipStream_t s) {
  const int np = (int)ps.planes.size();
  if (np < 1 || np > kMaxPlannedPlanes) return -1;
  if (h_planned_.capacity() == 0) return -2;  // reserve_planned first
  k_pack_planned<<<128, 256, 0, s>>>(dev_planes, np, b, h_planned_.dev<uint8_t>(),
                                     h_planned_.capacity());
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int LevelStager::finish_planned(const PlaneSet& ps, hipStream_t s, std::vector<LevelEdges>& views) {
  CCTAGX_HIP(stream_wait(s));
  const int np = (int)ps.planes.size();
  StageHeader h;
  std::memcpy(&h, h_planned_.as<uint8_t>(), sizeof(StageHeader));
  if ((int)h.n_planes != np) return -2;
  if (h.overflow) {
    if (grow_planned(h.total + h.total / 4)) return -2;
    return 1;
  }
  last_bytes_ = h.total;
  counts_.resize(np);
  const uint8_t* base = h_planned_.as<uint8_t>() + kStageDataStart;
  views.resize(np);
  for (int i = 0; i < np; ++i) {
    const PlaneDesc& d = ps.planes[i];
    const StagePlane& sp = h.planes[i];
    counts_[i] = sp.counts;
    LevelEdges& v = views[i];
    v.w = d.w;
    v.h = d.h;
    v.shape_w = d.shape_w;
    v.shape_h = d.shape_h;
    v.bits = reinterpret_cast<const uint32_t*>(base + sp.off[0]);
    v.n = sp.counts.n_edges;
    v.x = reinterpret_cast<const int16_t*>(base + sp.off[1]);
    v.y = reinterpret_cast<const int16_t*>(base + sp.off[2]);
    v.gx = reinterpret_cast<const int16_t*>(base + sp.off[3]);
    v.gy = reinterpret_cast<const int16_t*>(base + sp.off[4]);
    v.before = reinterpret_cast<const int32_t*>(base + sp.off[5]);
    v.after = reinterpret_cast<const int32_t*>(base + sp.off[6]);
    v.vindex = reinterpret_cast<const int32_t*>(base + sp.off[7]);
    v.vlist = reinterpret_cast<const int32_t*>(base + sp.off[8]);
    v.flow = reinterpret_cast<const float*>(base + sp.off[9]);
    v.n_seeds = sp.counts.n_seeds;
    v.seeds = reinterpret_cast<const int32_t*>(base + sp.off[10]);
  }
  return 0;
}

int launch_pack(const CopySeg* segs, int n_segs, uint8_t* dst, hipStream_t s) {
  if (n_segs <= 0) return 0;
  k_pack<<<n_segs, 256, 0, s>>>(segs, n_segs, dst);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int LevelStager::stage(const PlaneSet& ps, const FrontEndBuffers& b, hipStream_t s,
                       std::vector<LevelEdges>& views, int first, int count) {
  const int np_all = (int)ps.planes.size();
  if (count < 0) count = np_all - first;
  if (first < 0 || first + count > np_all) return -1;
  if (h_counts_.ensure(sizeof(PlaneCounts) * std::max(count, 1))) return -2;
  PlaneCounts* cnt = h_counts_.as<PlaneCounts>();
  if (count > 0)
    CCTAGX_HIP(hipMemcpyAsync(cnt, b.counts + first, sizeof(PlaneCounts) * count,
                              hipMemcpyDeviceToHost, s));
  CCTAGX_HIP(stream_wait(s));
  counts_.assign(cnt, cnt + count);

  // Segment list: per plane, the arrays in a fixed order.
  std::vector<CopySeg> segs;
  segs.reserve((size_t)count * 16);
  struct Off {
    uint64_t bits, ex, ey, egx, egy, before, after, vindex, vlist, flow, seeds;
  };
  std::vector<Off> offs(count);
  uint64_t at = 0;
  auto add = [&](const void* src, uint64_t bytes) -> uint64_t {
    const uint64_t o = at;
    for (uint64_t done = 0; done < bytes; done += kSegPiece)
      segs.push_back(CopySeg{static_cast<const uint8_t*>(src) + done, o + done,
                             std::min(kSegPiece, bytes - done)});
    at = align16(at + bytes);
    return o;
  };
  for (int i = 0; i < count; ++i) {
    const int l = first + i;
    const PlaneDesc& d = ps.planes[l];
    const uint64_t N = (uint64_t)d.w * d.h, n = (uint64_t)cnt[i].n_edges,
                   nv = (uint64_t)cnt[i].n_voters, nsd = (uint64_t)cnt[i].n_seeds;
    Off& o = offs[i];
    o.bits = add(b.ebits + (d.pix >> 5), 4 * ((N + 31) / 32));
    o.ex = add(b.ex + d.pix, 2 * n);
    o.ey = add(b.ey + d.pix, 2 * n);
    o.egx = add(b.egx + d.pix, 2 * n);
    o.egy = add(b.egy + d.pix, 2 * n);
    o.before = add(b.before + d.pix, 4 * n);
    o.after = add(b.after + d.pix, 4 * n);
    o.vindex = add(b.vindex + d.pix + l, 4 * (n + 1));
    o.vlist = add(b.vlist + d.pix, 4 * nv);
    o.flow = add(b.flow + d.pix, 4 * n);
    o.seeds = add(b.seeds + d.pix, 4 * nsd);
  }
  const uint64_t total = std::max<uint64_t>(at, 16);
  last_bytes_ = at;
  // the gather writes straight into the mapped host buffer: no device
  // staging buffer and no separate copy
  if (h_segs_.ensure(sizeof(CopySeg) * std::max<size_t>(segs.size(), 1)) ||
      d_segs_.ensure(sizeof(CopySeg) * std::max<size_t>(segs.size(), 1)) ||
      h_packed_.ensure(total + total / 4, hipHostMallocCoherent))
    return -2;
  if (!segs.empty()) {
    std::memcpy(h_segs_.as<CopySeg>(), segs.data(), sizeof(CopySeg) * segs.size());
    CCTAGX_HIP(hipMemcpyAsync(d_segs_.as<CopySeg>(), h_segs_.as<CopySeg>(),
                              sizeof(CopySeg) * segs.size(), hipMemcpyHostToDevice, s));
    if (launch_pack(d_segs_.as<CopySeg>(), (int)segs.size(), h_packed_.dev<uint8_t>(), s))
      return -2;
  }
  CCTAGX_HIP(stream_wait(s));

  const uint8_t* base = h_packed_.as<uint8_t>();
  views.resize(count);
  for (int i = 0; i < count; ++i) {
    const PlaneDesc& d = ps.planes[first + i];
    const Off& o = offs[i];
    LevelEdges& v = views[i];
    v.w = d.w;
    v.h = d.h;
    v.shape_w = d.shape_w;
    v.shape_h = d.shape_h;
    v.bits = reinterpret_cast<const uint32_t*>(base + o.bits);
    v.n = cnt[i].n_edges;
    v.x = reinterpret_cast<const int16_t*>(base + o.ex);
    v.y = reinterpret_cast<const int16_t*>(base + o.ey);
    v.gx = reinterpret_cast<const int16_t*>(base + o.egx);
    v.gy = reinterpret_cast<const int16_t*>(base + o.egy);
    v.before = reinterpret_cast<const int32_t*>(base + o.before);
    v.after = reinterpret_cast<const int32_t*>(base + o.after);
    v.vindex = reinterpret_cast<const int32_t*>(base + o.vindex);
    v.vlist = reinterpret_cast<const int32_t*>(base + o.vlist);
    v.flow = reinterpret_cast<const float*>(base + o.flow);
    v.n_seeds = cnt[i].n_seeds;
    v.seeds = reinterpret_cast<const int32_t*>(base + o.seeds);
  }
  return 0;
}

}  // namespace cctagx
