// LevelStager and its gather kernel (see staging.hpp).
#include "staging.hpp"

#include <algorithm>
#include <cstring>

namespace cctagx {
namespace {

constexpr uint64_t kSegPiece = 64 * 1024;  // bytes per gather block

// One block per piece; 16-byte moves when source and destination allow,
// else 4-byte, else bytes.  Reads exactly [src, src + bytes).
__global__ __launch_bounds__(256) void k_pack(const CopySeg* __restrict__ segs, int n_segs,
                                              uint8_t* __restrict__ dst) {
  const CopySeg sg = segs[blockIdx.x];
  const uint8_t* src = static_cast<const uint8_t*>(sg.src);
  uint8_t* d = dst + sg.dst;
  const uint64_t n = sg.bytes;
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

inline uint64_t align16(uint64_t v) { return (v + 15) & ~uint64_t(15); }

}  // namespace

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
    uint64_t bits, ex, ey, egx, egy, before, after, vindex, vlist, flow, ismax, seeds;
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
    o.ismax = add(b.ismax + d.pix, 4 * n);
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
    v.ismax = reinterpret_cast<const int32_t*>(base + o.ismax);
    v.n_seeds = cnt[i].n_seeds;
    v.seeds = reinterpret_cast<const int32_t*>(base + o.seeds);
  }
  return 0;
}

}  // namespace cctagx
