// Host side of the front end: plane planning and device buffer layout.
#include "device.hpp"
#include "frontend_launch.hpp"
#include "frontend_host.hpp"

#include <algorithm>
#include <cstring>

namespace cctagx {

namespace {
inline int cdiv(int a, int b) { return (a + b - 1) / b; }
inline int align256(int v) { return (v + 255) & ~255; }
thread_local std::string g_error;
}  // namespace

void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

bool plan_planes(const std::vector<std::pair<int, int>>& frame_sizes, int n_levels,
                 PlaneSet& out) {
  out = PlaneSet();
  out.n_levels = n_levels;
  out.level_planes.assign(n_levels, {});
  out.n_tiles_r.assign(n_levels, 0);
  int pix = 0, tp = 0, tf = 0, ts = 0, tc = 0;
  for (size_t f = 0; f < frame_sizes.size(); ++f) {
    const int W = frame_sizes[f].first, H = frame_sizes[f].second;
    int lw = W, lh = H;
    for (int l = 0; l < n_levels; ++l) {
      if (l > 0) {
        const int pw = lw, ph = lh;
        lw = pw / 2;
        lh = ph / 2;
        if (lw < 1 || lh < 1) return false;
        if (pw != 2 * lw || ph != 2 * lh) out.exact_halves = false;
      }
      PlaneDesc d{};
      d.w = lw;
      d.h = lh;
      d.shape_w = W;
      d.shape_h = H;
      d.pix = pix;
      d.parent = l == 0 ? -1 : (int)out.planes.size() - 1;
      d.level = l;
      d.frame = (int)f;
      d.tile_f = tf;
      tf += cdiv(lw, kFilterTileW) * cdiv(lh, kFilterTileH);
      for (int v = 0; v < kFilterVariants; ++v) {
        out.filter_starts[v].push_back(out.n_tiles_fv[v]);
        out.n_tiles_fv[v] += cdiv(lw, kFilterTileW) * cdiv(lh, kFilterTileHs[v]);
      }
      d.tile_s = ts;
      ts += cdiv(lw, kStencilTileW) * cdiv(lh, kStencilTileH);
      d.tile_c = tc;
      tc += cdiv(lw * lh, kChunk);
      d.tile_p = tp;
      if (l == 0) tp += cdiv(lw, kPyrTile) * cdiv(lh, kPyrTile);
      d.tile_r = out.n_tiles_r[l];
      if (l > 0) out.n_tiles_r[l] += cdiv(lw, kResTileW) * cdiv(lh, kResTileH);
      out.level_planes[l].push_back((int32_t)out.planes.size());
      out.planes.push_back(d);
      pix += align256(lw * lh);
    }
  }
  out.total_pix = pix;
  out.n_tiles_p = tp;
  out.n_tiles_f = tf;
  out.n_tiles_s = ts;
  out.n_tiles_c = tc;
  return true;
}

int FrontEndDevice::init(const PlaneSet& ps, const FrontEndParams& prm) {
  ps_ = ps;
  arena_.release();
  arena_.clear_plan();
  const size_t N = (size_t)ps.total_pix;
  const size_t NP = ps.planes.size();
  const size_t o_img = arena_.reserve(N);
  const size_t o_quad = arena_.reserve(4 * N);
  const size_t o_dx = arena_.reserve(2 * N);
  const size_t o_dy = arena_.reserve(2 * N);
  const size_t o_cls = arena_.reserve(N);
  const size_t o_lab = arena_.reserve(4 * N);
  const size_t o_rs = arena_.reserve(N);
  const size_t o_edges = arena_.reserve(N);
  const size_t o_ebits = arena_.reserve(N / 8 + 4);
  const size_t o_erank = arena_.reserve(N / 8 + 4);
  const size_t o_tcount = arena_.reserve(4 * (size_t)std::max(ps.n_tiles_c, 1));
  const size_t o_ex = arena_.reserve(2 * N);
  const size_t o_ey = arena_.reserve(2 * N);
  const size_t o_egx = arena_.reserve(2 * N);
  const size_t o_egy = arena_.reserve(2 * N);
  const size_t o_before = arena_.reserve(4 * N);
  const size_t o_after = arena_.reserve(4 * N);
  const size_t o_chosen = arena_.reserve(4 * N);
  const size_t o_tdist = arena_.reserve(4 * N);
  const size_t o_vcount = arena_.reserve(4 * N);
  const size_t o_vindex = arena_.reserve(4 * (N + NP));
  const size_t o_vlist = arena_.reserve(4 * N);
  const size_t o_flow = arena_.reserve(4 * N);
  const size_t o_ismax = arena_.reserve(4 * N);
  const size_t o_seedmark = arena_.reserve(4 * N);
  const size_t o_seeds = arena_.reserve(4 * N);
  const size_t o_part = arena_.reserve(sizeof(int32_t) * NP * kScanBlocks);
  const size_t o_counts = arena_.reserve(sizeof(PlaneCounts) * NP);
  const size_t o_planes = arena_.reserve(sizeof(PlaneDesc) * NP);
  const size_t o_lplanes = arena_.reserve(sizeof(int32_t) * NP);
  size_t o_fstart[kFilterVariants];
  for (int v = 0; v < kFilterVariants; ++v) o_fstart[v] = arena_.reserve(sizeof(int32_t) * NP);
  if (int rc = arena_.allocate()) return rc;
  FrontEndBuffers& b = launch_.buf;
  b.img = arena_.at<uint8_t>(o_img);
  b.quad = arena_.at<uint32_t>(o_quad);
  b.dx = arena_.at<int16_t>(o_dx);
  b.dy = arena_.at<int16_t>(o_dy);
  b.cls = arena_.at<uint8_t>(o_cls);
  b.lab = arena_.at<int32_t>(o_lab);
  b.root_strong = arena_.at<uint8_t>(o_rs);
  b.edges = arena_.at<uint8_t>(o_edges);
  b.ebits = arena_.at<uint32_t>(o_ebits);
  b.erank = arena_.at<int32_t>(o_erank);
  b.tcount = arena_.at<int32_t>(o_tcount);
  b.ex = arena_.at<int16_t>(o_ex);
  b.ey = arena_.at<int16_t>(o_ey);
  b.egx = arena_.at<int16_t>(o_egx);
  b.egy = arena_.at<int16_t>(o_egy);
  b.before = arena_.at<int32_t>(o_before);
  b.after = arena_.at<int32_t>(o_after);
  b.chosen = arena_.at<int32_t>(o_chosen);
  b.tdist = arena_.at<float>(o_tdist);
  b.vcount = arena_.at<int32_t>(o_vcount);
  b.vindex = arena_.at<int32_t>(o_vindex);
  b.vlist = arena_.at<int32_t>(o_vlist);
  b.flow = arena_.at<float>(o_flow);
  b.ismax = arena_.at<int32_t>(o_ismax);
  b.seedmark = arena_.at<int32_t>(o_seedmark);
  b.seeds = arena_.at<int32_t>(o_seeds);
  b.scan_part = arena_.at<int32_t>(o_part);
  b.counts = arena_.at<PlaneCounts>(o_counts);
  PlaneDesc* dp = arena_.at<PlaneDesc>(o_planes);
  CCTAGX_HIP(hipMemcpy(dp, ps.planes.data(), sizeof(PlaneDesc) * NP, hipMemcpyHostToDevice));
  launch_.dev_planes = dp;
  launch_.n_planes = (int)NP;
  {
    std::vector<int32_t> lp;
    launch_.level_plane_start.clear();
    launch_.level_plane_count.clear();
    for (const auto& v : ps.level_planes) {
      launch_.level_plane_start.push_back((int)lp.size());
      launch_.level_plane_count.push_back((int)v.size());
      lp.insert(lp.end(), v.begin(), v.end());
    }
    int32_t* dl = arena_.at<int32_t>(o_lplanes);
    CCTAGX_HIP(hipMemcpy(dl, lp.data(), sizeof(int32_t) * lp.size(), hipMemcpyHostToDevice));
    launch_.dev_level_planes = dl;
    launch_.resample = !ps.exact_halves;
    launch_.n_tiles_r = ps.n_tiles_r;
  }
  launch_.n_levels = ps.n_levels;
  launch_.n_tiles_p = ps.n_tiles_p;
  launch_.n_tiles_f = ps.n_tiles_f;
  for (int v = 0; v < kFilterVariants; ++v) {
    int32_t* fs = arena_.at<int32_t>(o_fstart[v]);
    CCTAGX_HIP(hipMemcpy(fs, ps.filter_starts[v].data(), sizeof(int32_t) * NP, hipMemcpyHostToDevice));
    launch_.dev_filter_starts[v] = fs;
    launch_.n_tiles_fv[v] = ps.n_tiles_fv[v];
  }
  launch_.n_tiles_s = ps.n_tiles_s;
  launch_.n_tiles_c = ps.n_tiles_c;
  // Blocks per plane of the edge-indexed kernels, from the largest plane:
  // their lanes walk rays of dependent mask reads, so a large frame wants
  // more of them in flight (512 up to 3 Mpx, 1024 to 6 Mpx, then 2048:
  // 3840x2160's front end 691 -> 595 us against 512, 1280x720 and 1920x1080
  // best at 512); planes loop over their own edges.
  int max_pix = 0;
  for (const PlaneDesc& d : ps.planes) max_pix = std::max(max_pix, d.w * d.h);
  int eb = 512;
  while (eb < 2048 && (long long)eb * 6144 < (long long)max_pix) eb *= 2;
  launch_.edge_blocks = eb;
  launch_.prm = prm;
  return 0;
}

void copy_rows(TaskPool* pool, uint8_t* dst, const uint8_t* src, size_t row_bytes, int rows,
               size_t stride) {
  const int parts = std::max(1, std::min(rows, 4 * (pool ? pool->threads() : 1)));
  parallel_for(pool, parts, [&](int k) {
    const int r0 = (int)((long long)rows * k / parts), r1 = (int)((long long)rows * (k + 1) / parts);
    if (stride == row_bytes)
      std::memcpy(dst + (size_t)r0 * row_bytes, src + (size_t)r0 * row_bytes,
                  (size_t)(r1 - r0) * row_bytes);
    else
      for (int r = r0; r < r1; ++r) std::memcpy(dst + (size_t)r * row_bytes, src + r * stride, row_bytes);
  });
}

FrontEndDevice::~FrontEndDevice() {
  for (hipEvent_t e : staged_sent_)
    if (e) {
      hipEventSynchronize(e);
      hipEventDestroy(e);
    }
  if (color_sent_) {
    hipEventSynchronize(color_sent_);
    hipEventDestroy(color_sent_);
  }
}

int FrontEndDevice::upload(int frame, const uint8_t* gray, size_t stride, hipStream_t s,
                           TaskPool* pool) {
  int p0 = -1;
  for (size_t i = 0; i < ps_.planes.size(); ++i)
    if (ps_.planes[i].frame == frame && ps_.planes[i].level == 0) {
      p0 = (int)i;
      break;
    }
  if (p0 < 0) return -1;
  const PlaneDesc& d = ps_.planes[p0];
  if (pool) {
    const size_t bytes = (size_t)d.w * d.h;
    if (staged_off_.empty()) {
      // one region per frame, in frame order
      int nf = 0;
      for (const PlaneDesc& q : ps_.planes) nf = std::max(nf, q.frame + 1);
      std::vector<size_t> size(nf, 0);
      for (const PlaneDesc& q : ps_.planes)
        if (q.level == 0) size[q.frame] = (size_t)q.w * q.h;
      staged_off_.assign(nf + 1, 0);
      for (int f = 0; f < nf; ++f) staged_off_[f + 1] = staged_off_[f] + ((size[f] + 255) & ~size_t(255));
      if (staged_.ensure(std::max<size_t>(staged_off_[nf], 1))) {
        staged_off_.clear();
        return -2;
      }
      staged_sent_.assign(nf, nullptr);
      staged_in_flight_.assign(nf, 0);
    }
    if (!staged_sent_[frame])
      CCTAGX_HIP(hipEventCreateWithFlags(&staged_sent_[frame], hipEventDisableTiming));
    if (staged_in_flight_[frame]) CCTAGX_HIP(hipEventSynchronize(staged_sent_[frame]));
    staged_in_flight_[frame] = 0;
    uint8_t* dst = staged_.as<uint8_t>() + staged_off_[frame];
    copy_rows(pool, dst, gray, (size_t)d.w, d.h, stride);
    CCTAGX_HIP(hipMemcpyAsync(launch_.buf.img + d.pix, dst, bytes, hipMemcpyHostToDevice, s));
    CCTAGX_HIP(hipEventRecord(staged_sent_[frame], s));
    staged_in_flight_[frame] = 1;
    return 0;
  }
  // rows packed: one contiguous copy (a 2-D copy from pageable memory goes row by row)
  if (stride == (size_t)d.w)
    CCTAGX_HIP(hipMemcpyAsync(launch_.buf.img + d.pix, gray, (size_t)d.w * d.h,
                              hipMemcpyHostToDevice, s));
  else
    CCTAGX_HIP(hipMemcpy2DAsync(launch_.buf.img + d.pix, d.w, gray, stride, d.w, d.h,
                                hipMemcpyHostToDevice, s));
  return 0;
}

int FrontEndDevice::run(hipStream_t s, StageTimer* timer) {
  return launch_front_end(launch_, s, timer);
}

}  // namespace cctagx
