// strand_oracle.cpp -- CPU restatement of the reference's strand rendering
// hot path, used ONLY as test infrastructure: the correctness reference for
// the HIP kernels and the timed CPU baseline.  Nothing in the shipped library
// links, loads or calls this file.
//
// Every function below restates one WGSL function of the reference and cites
// it (paths relative to the reference root).  Arithmetic follows WGSL rules:
// f32 throughout (no double promotion), no FMA contraction (built with
// -ffp-contract=off), round() is ties-to-even, f32->u32/i32 conversions
// saturate, clamp(e, lo, hi) = min(max(e, lo), hi) literally, including the
// call sites whose arguments are swapped in the reference.
//
// Canonical order (the reference orders several things by atomics, which is
// not reproducible; this restatement fixes one order that the HIP path must
// reproduce bit for bit):
//   * binning tasks are ordered by (frustum, instance, index-slab position);
//   * page ids are ranks of marked page-table entries in page-table order and
//     page reference bases are the exclusive scan of page totals in page order;
//   * references inside a cell are in task order, i.e. sorted by (inst, seg);
//   * raster work bases and tile-run indices are scans over the active-tile
//     queue, which is sorted ascending (as the reference's compaction is);
//   * coarse-interval lists are walked newest-first with ranges inserted in
//     (instance, frustum) slot order;
//   * a shading-atlas texel written by two segments (strand_shading.wgsl:692-694)
//     holds the value of the segment that starts at it.
//
// Parity pins for this file live in tests/test_oracle_pins.py.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../include/strand_rt.h"

namespace {

constexpr uint32_t kExt = SR_COARSE_FINE_TILE_EXTENT;      // 4
constexpr uint32_t kSlices = SR_COARSE_DEPTH_SLICES;       // 16: fine depth slices of a page, default LUT slices
constexpr uint32_t kMaxSlices = SR_MAX_COARSE_DEPTH_SLICES;  // 32
constexpr uint32_t kPage = SR_COARSE_COUNT_PAGE_SIZE;      // 256
constexpr uint32_t kQuant = SR_DEPTH_QUANT_MAX;            // 16777215
constexpr uint32_t kInvalid = SR_INVALID_PTR;
constexpr uint32_t kMarked = 0xFFFFFFFEu;

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

// ---------------------------------------------------------------- WGSL ops
inline float wclamp(float e, float lo, float hi) { return std::fmin(std::fmax(e, lo), hi); }
inline float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline float wfract(float x) { return x - std::floor(x); }
inline uint32_t f2u(float x) {  // saturating f32 -> u32
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
inline int32_t f2i(float x) {  // saturating f32 -> i32
  if (std::isnan(x)) return 0;
  if (x <= -2147483648.0f) return INT32_MIN;
  if (x >= 2147483648.0f) return INT32_MAX;
  return (int32_t)x;
}
inline int32_t iclamp(int32_t e, int32_t lo, int32_t hi) { return std::min(std::max(e, lo), hi); }
inline uint32_t ceil_div(uint32_t x, uint32_t y) { return (x + y - 1u) / y; }

// mat4 (column-major) x vec4, summed as ((c0*x + c1*y) + c2*z) + c3*w.
inline V4 mat_vec(const float* m, V4 v) {
  V4 r;
  r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
  r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
  r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
  r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
  return r;
}

// pack2x16unorm per the WGSL spec: floor(0.5 + 65535 * clamp(e, 0, 1)).
inline uint32_t pack2x16unorm(float a, float b) {
  uint32_t lo = (uint32_t)std::floor(0.5f + 65535.0f * wclamp(a, 0.0f, 1.0f));
  uint32_t hi = (uint32_t)std::floor(0.5f + 65535.0f * wclamp(b, 0.0f, 1.0f));
  return lo | (hi << 16);
}
inline V2 unpack2x16unorm(uint32_t v) {
  return V2{(float)(v & 0xFFFFu) / 65535.0f, (float)(v >> 16) / 65535.0f};
}

// common.wgsl:229-242
inline uint32_t wang_hash(uint32_t seed) {
  uint32_t x = seed;
  x = (x ^ 61u) ^ (x >> 16u);
  x = x + (x << 3u);
  x = x ^ (x >> 4u);
  x = x * 0x27d4eb2du;
  x = x ^ (x >> 15u);
  return x;
}
inline float hash_to_unit_float(uint32_t x) { return (float)x / 4294967296.0f; }

// common.wgsl:206-212
inline float normalize_depth01(float z) {
  if (z >= 0.0f && z <= 1.0f) return z;
  return wclamp(z * 0.5f + 0.5f, 0.0f, 1.0f);
}

// strand_prepass.wgsl:533-539 (round = ties-to-even)
inline uint32_t quantize_depth01(float z) {
  return f2u(std::nearbyint(wclamp(z, 0.0f, 1.0f) * (float)kQuant));
}
inline float dequantize_depth01(uint32_t q) { return (float)q / (float)kQuant; }

// common.wgsl:93-107
inline V4 world_to_screen_raw(V4 pos, const float* clip_from_world, float vw, float vh) {
  V4 clip = mat_vec(clip_from_world, pos);
  if (clip.w <= 0.0f) return V4{-1.0f, -1.0f, -1.0f, -1.0f};
  float nx = clip.x / clip.w, ny = clip.y / clip.w, nz = clip.z / clip.w;
  float sx = 0.0f + (nx * 0.5f + 0.5f) * vw;
  float sy = 0.0f + (ny * -0.5f + 0.5f) * vh;
  return V4{sx, sy, nz, clip.w};
}

// strand_prepass.wgsl:512-516
inline int32_t sgn_i32(float f) {
  if (f > 1e-6f) return 1;
  if (f < -1e-6f) return -1;
  return 0;
}

// common.wgsl:245-262 (ties resolve x, then y, then z)
inline void canonical_min_mask(V3 v, bool m[3]) {
  float mn = std::fmin(std::fmin(v.x, v.y), v.z);
  bool ix = v.x == mn, iy = v.y == mn;
  m[0] = ix;
  m[1] = !ix && iy;
  m[2] = !ix && !iy;
}

struct ClipRange { bool ok; float tmin, tmax; };
// strand_prepass.wgsl:1356-1380
inline ClipRange clip_axis_to_range(float p, float d, float mn, float mx, float tmin, float tmax) {
  if (std::fabs(d) < 1e-6f) {
    if (p < mn || p > mx) return {false, tmin, tmax};
    return {true, tmin, tmax};
  }
  float inv = 1.0f / d;
  float a = (mn - p) * inv, b = (mx - p) * inv;
  if (a > b) std::swap(a, b);
  float omin = std::fmax(tmin, a), omax = std::fmin(tmax, b);
  if (omin > omax) return {false, omin, omax};
  return {true, omin, omax};
}
struct Clipped { bool ok; V3 p0, p1; };
// strand_prepass.wgsl:1382-1410
inline Clipped clip_segment_to_box(V3 p0, V3 p1, V3 bmin, V3 bmax) {
  V3 d{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
  float tmin = 0.0f, tmax = 1.0f;
  ClipRange r = clip_axis_to_range(p0.x, d.x, bmin.x, bmax.x, tmin, tmax);
  if (!r.ok) return {false, p0, p1};
  tmin = r.tmin; tmax = r.tmax;
  r = clip_axis_to_range(p0.y, d.y, bmin.y, bmax.y, tmin, tmax);
  if (!r.ok) return {false, p0, p1};
  tmin = r.tmin; tmax = r.tmax;
  r = clip_axis_to_range(p0.z, d.z, bmin.z, bmax.z, tmin, tmax);
  if (!r.ok) return {false, p0, p1};
  tmin = r.tmin; tmax = r.tmax;
  V3 a{p0.x + d.x * tmin, p0.y + d.y * tmin, p0.z + d.z * tmin};
  V3 b{p0.x + d.x * tmax, p0.y + d.y * tmax, p0.z + d.z * tmax};
  return {true, a, b};
}

void parallel_for(int threads, size_t n, const std::function<void(size_t, size_t)>& fn) {
  if (threads <= 1 || n < 2048) { fn(0, n); return; }
  std::vector<std::thread> pool;
  size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back(fn, b, e);
  }
  for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------ frame
struct Frame {
  // inputs
  const sr_asset_slabs* assets = nullptr;
  uint32_t asset_count = 0;
  std::vector<sr_strand_instance> inst;
  std::vector<sr_frustum_desc> fr;
  std::vector<sr_view> views;
  std::vector<sr_directional_light> lights;
  sr_frame_settings st{};
  int threads = 1;
  uint32_t F = 0, NI = 0;
  uint32_t coarse_total = 0;
  uint32_t ref_capacity = 0;
  // depth-warp LUT slices per coarse tile (sr_frame_settings.coarse_depth_slices)
  uint32_t slices() const { return st.coarse_depth_slices ? st.coarse_depth_slices : kSlices; }
  // the previous frame's shadow history (float2 per atlas texel, .x = the
  // filtered visibility read as shadow_history_prev), or none
  const float* prev_history = nullptr;
  uint64_t prev_history_texels = 0;
  // opaque scene depth per camera (host, pixel-major, opaque_samples each)
  // and its fine-tile reduction (covered, quantised farthest) per fine tile
  const float* opaque_depth = nullptr;
  uint32_t opaque_samples = 0;
  std::vector<uint32_t> opaque_tiles;

  // outputs
  sr_frame_counters ctr{};
  std::vector<sr_broad_instance_meta> broad;
  std::vector<float> keep;
  std::vector<sr_coarse_asset_range> ranges;
  std::vector<sr_binning_task> tasks;
  std::vector<sr_projected_segment> proj;
  std::vector<uint32_t> shading_queue;
  std::vector<sr_coarse_depth_lut_entry> lut;
  std::vector<uint32_t> page_table;
  std::vector<sr_fine_page_meta> page_meta;
  std::vector<uint32_t> cell_counts, cell_offsets;
  std::vector<sr_fine_seg_ref> refs;
  std::vector<uint32_t> active_flags, active_queue;
  std::vector<sr_raster_work_item> work;
  std::vector<sr_raster_tile_run> cam_runs, shadow_runs;
  std::vector<float> dom_depth, dom_opacity, stamp;
  std::vector<uint64_t> dom_depth_base, dom_opacity_base;  // per frustum offsets (lights)
  // virtual DOMs: per light frustum ceil(W/256) x ceil(H/256) pages,
  // concatenated: requested (0/1) and physical page or kInvalid
  std::vector<uint32_t> dom_requests, dom_table;
  std::vector<float> atlas, history;  // atlas float4, history float2 (visibility, current)
  std::vector<uint64_t> image_base;   // per frustum offsets (cameras)
  std::vector<float> image, depth, composite;
  std::vector<uint32_t> atlas_layer_base;  // per frustum: first atlas layer (cameras)
  uint32_t atlas_texels = 1;
  std::vector<double> stage_ms;
  std::vector<std::string> stage_names;
};

// -------------------------------------------------------- asset access
inline const sr_asset_slabs* slab(const Frame& F, uint32_t id) {
  return id < F.asset_count ? &F.assets[id] : nullptr;
}
inline V4 vertex_at(const sr_asset_slabs* a, uint32_t vi) {
  const float* v = a->vertices + 4ull * vi;
  return V4{v[0], v[1], v[2], 1.0f};
}
inline bool light_layer_ok(const Frame& F, uint32_t fi, uint32_t* layer_out) {
  // strand_prepass.wgsl:412-430
  if (fi >= F.F || F.fr[fi].kind != 1u) return false;
  uint32_t layer = 0;
  for (uint32_t i = 0; i < fi; ++i)
    if (F.fr[i].kind == 1u) ++layer;
  if (layer >= F.lights.size()) return false;
  if (layer_out) *layer_out = layer;
  return true;
}

// ---------------------------------------------------- K2 broad_prepass
// strand_prepass.wgsl:295-362 aabb_projected_screen_area and 575-682
// emit_coarse_asset_range_for_aabb share the corner projection.
struct CornerBounds { V2 smin, smax; float zmin, zmax; bool any, near_clip; };
CornerBounds project_aabb(const Frame& F, uint32_t fi, const sr_aabb& bb, const float* w) {
  const sr_frustum_desc& fd = F.fr[fi];
  const float* cfw = F.views[fi].clip_from_world;
  float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
  const float* mn = bb.min; const float* mx = bb.max;
  V4 corners[8] = {
      {mn[0], mn[1], mn[2], 1.0f}, {mn[0], mn[1], mx[2], 1.0f}, {mn[0], mx[1], mn[2], 1.0f},
      {mn[0], mx[1], mx[2], 1.0f}, {mx[0], mn[1], mn[2], 1.0f}, {mx[0], mn[1], mx[2], 1.0f},
      {mx[0], mx[1], mn[2], 1.0f}, {mx[0], mx[1], mx[2], 1.0f}};
  CornerBounds cb{{1e30f, 1e30f}, {-1e30f, -1e30f}, 1.0f, 0.0f, false, false};
  for (int i = 0; i < 8; ++i) {
    V4 wc = mat_vec(w, corners[i]);
    V4 clip = mat_vec(cfw, wc);
    if (clip.w <= 1e-6f) { cb.near_clip = true; continue; }
    float nx = clip.x / clip.w, ny = clip.y / clip.w, nz = clip.z / clip.w;
    if (nz < 0.0f) cb.near_clip = true;
    float rx = 0.0f + (nx * 0.5f + 0.5f) * vw;
    float ry = 0.0f + (ny * -0.5f + 0.5f) * vh;
    cb.any = true;
    cb.smin.x = std::fmin(cb.smin.x, rx); cb.smin.y = std::fmin(cb.smin.y, ry);
    cb.smax.x = std::fmax(cb.smax.x, rx); cb.smax.y = std::fmax(cb.smax.y, ry);
    float z = normalize_depth01(nz);
    cb.zmin = std::fmin(cb.zmin, z); cb.zmax = std::fmax(cb.zmax, z);
  }
  return cb;
}

float aabb_projected_screen_area(const Frame& F, uint32_t fi, const sr_aabb& bb, const float* w) {
  const sr_frustum_desc& fd = F.fr[fi];
  float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
  CornerBounds cb = project_aabb(F, fi, bb, w);
  if (cb.near_clip) return vw * vh;
  if (!cb.any) return -1.0f;
  if (cb.smax.x < 0.0f || cb.smax.y < 0.0f || cb.smin.x >= vw || cb.smin.y >= vh) return -1.0f;
  if (cb.zmax < 0.0f || cb.zmin > 1.0f) return -1.0f;
  V2 cmin{wclamp(cb.smin.x, 0.0f, vw - 1.0f), wclamp(cb.smin.y, 0.0f, vh - 1.0f)};
  V2 cmax{wclamp(cb.smax.x, 0.0f, vw - 1.0f), wclamp(cb.smax.y, 0.0f, vh - 1.0f)};
  float ex = std::fmax(cmax.x - cmin.x, 1.0f), ey = std::fmax(cmax.y - cmin.y, 1.0f);
  return ex * ey;
}

// strand_prepass.wgsl:364-392
float camera_keep(const Frame& F, uint32_t strand_count, float area) {
  if (F.st.stochastic_cull_enabled == 0u) return 1.0f;
  if (area < 0.0f) return 0.0f;
  float target = std::fmax(F.st.target_strands_per_pixel, 1e-5f);
  float min_keep = wclamp(F.st.min_keep_probability, 0.0f, 1.0f);
  return wclamp(area * target / std::fmax((float)strand_count, 1.0f), min_keep, 1.0f);
}
float shadow_keep(const Frame& F, uint32_t strand_count, float area, const sr_frustum_desc& fd) {
  if (F.st.stochastic_cull_enabled == 0u) return 1.0f;
  if (area < 0.0f) return 0.0f;
  float target = std::fmax(F.st.target_strands_per_pixel, 1e-5f);
  float min_keep = wclamp(F.st.min_keep_probability, 0.0f, 1.0f);
  float cap = wclamp(F.st.shadow_keep_probability, 0.0f, 1.0f);
  float area_keep = wclamp(area * target / std::fmax((float)strand_count, 1.0f), min_keep, 1.0f);
  float curved = area_keep * area_keep;
  float cascade_scale = 1.0f / std::exp2((float)fd.cascade_index);
  return wclamp(curved * cascade_scale * cap, min_keep, 1.0f);
}

// strand_prepass.wgsl:575-682
bool emit_coarse_range(Frame& F, uint32_t ii, uint32_t fi, const sr_aabb& bb, const float* w) {
  const sr_frustum_desc& fd = F.fr[fi];
  if (fd.coarse_depth_tile_count == 0u) return false;
  if (fd.kind == 1u && !light_layer_ok(F, fi, nullptr)) return false;
  float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
  CornerBounds cb = project_aabb(F, fi, bb, w);
  if (cb.near_clip) {
    cb.smin = V2{0.0f, 0.0f};
    cb.smax = V2{vw - 1.0f, vh - 1.0f};
    cb.zmax = 1.0f;
    if (!cb.any) cb.zmin = 0.0f;
    cb.any = true;
  }
  if (!cb.any) return false;
  if (cb.smax.x < 0.0f || cb.smax.y < 0.0f || cb.smin.x >= vw || cb.smin.y >= vh) return false;
  if (cb.zmax < 0.0f || cb.zmin > 1.0f) return false;
  uint32_t cpx = std::max(1u, fd.froxel_size_x * kExt), cpy = std::max(1u, fd.froxel_size_y * kExt);
  V2 cmin{wclamp(cb.smin.x, 0.0f, vw - 1.0f), wclamp(cb.smin.y, 0.0f, vh - 1.0f)};
  V2 cmax{wclamp(cb.smax.x, 0.0f, vw - 1.0f), wclamp(cb.smax.y, 0.0f, vh - 1.0f)};
  sr_coarse_asset_range r;
  r.geo_id = ii;
  r.frustum_id = fi;
  r.min_x = std::min(f2u(cmin.x) / cpx, fd.coarse_tiles_x - 1u);
  r.max_x = std::min(f2u(cmax.x) / cpx, fd.coarse_tiles_x - 1u);
  r.min_y = std::min(f2u(cmin.y) / cpy, fd.coarse_tiles_y - 1u);
  r.max_y = std::min(f2u(cmax.y) / cpy, fd.coarse_tiles_y - 1u);
  r.z_min_q = quantize_depth01(cb.zmin);
  r.z_max_q = quantize_depth01(cb.zmax);
  F.ranges[(size_t)ii * F.F + fi] = r;
  return true;
}

// strand_prepass.wgsl:684-760
void broad_prepass(Frame& F) {
  F.broad.assign(F.NI, sr_broad_instance_meta{0, 0, 0.0f, 0.0f});
  F.keep.assign((size_t)F.NI * F.F, 0.0f);
  sr_coarse_asset_range none{};
  none.geo_id = kInvalid;
  F.ranges.assign((size_t)F.NI * F.F, none);
  for (uint32_t ii = 0; ii < F.NI; ++ii) {
    const sr_strand_instance& in = F.inst[ii];
    const sr_asset_slabs* g = slab(F, in.geo_id);
    const sr_asset_slabs* m = slab(F, in.meta_id);
    if (!g || !m || g->strand_count == 0 || m->strand_count == 0) continue;
    const sr_strand_geo& geo = *g->geo;
    uint32_t strand_count = std::min(m->strand_count, geo.strand_count);
    bool visible = false;
    float cam = 0.0f, shd = 0.0f;
    for (uint32_t fi = 0; fi < F.F; ++fi) {
      const sr_frustum_desc& fd = F.fr[fi];
      float k = 0.0f;
      if (fd.kind <= 1u) {
        bool vis = emit_coarse_range(F, ii, fi, geo.aabb, in.world_from_local);
        visible = vis || visible;
        if (vis && fd.kind == 0u) {
          float area = aabb_projected_screen_area(F, fi, geo.aabb, in.world_from_local);
          k = camera_keep(F, strand_count, area);
          cam = std::fmax(cam, k);
        } else if (vis && fd.kind == 1u) {
          float area = aabb_projected_screen_area(F, fi, geo.aabb, in.world_from_local);
          k = shadow_keep(F, strand_count, area, fd);
          shd = std::fmax(shd, k);
        }
      }
      F.keep[(size_t)ii * F.F + fi] = k;
    }
    F.broad[ii] = sr_broad_instance_meta{strand_count, visible ? 1u : 0u, cam, shd};
  }
}

// ------------------------------------------ K3 + K5 broad_strand + fine_prepass
// strand_prepass.wgsl:762-782 (strand keep) and 1016-1140 (projection).
struct StrandRef { uint32_t inst, strand; };

void fine_prepass(Frame& F) {
  // per-frustum task lists built in (instance, seg) order, then concatenated
  std::vector<std::vector<sr_binning_task>> ftasks(F.F);
  std::vector<std::vector<sr_projected_segment>> fproj(F.F);
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    const sr_frustum_desc& fd = F.fr[fi];
    if (fd.kind > 1u) continue;
    if (fd.kind == 1u && !light_layer_ok(F, fi, nullptr)) continue;
    const float* cfw = F.views[fi].clip_from_world;
    float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
    V3 box_min{0.0f, 0.0f, 0.0f};
    V3 box_max{std::fmax(0.0f, vw - 1e-3f), std::fmax(0.0f, vh - 1e-3f), 1.0f};
    float ssx = std::fmax(1.0f, vw), ssy = std::fmax(1.0f, vh);
    uint32_t packed = (0u & 3u) | ((fd.kind == 1u ? 1u : 0u) << 2) | (fi << 3);
    for (uint32_t ii = 0; ii < F.NI; ++ii) {
      const sr_strand_instance& in = F.inst[ii];
      const sr_broad_instance_meta& b = F.broad[ii];
      if (b.visible == 0u) continue;
      const sr_asset_slabs* ms = slab(F, in.meta_id);
      const sr_asset_slabs* is = slab(F, in.index_id);
      const sr_asset_slabs* vs = slab(F, in.vertex_id);
      if (!ms || !is || !vs) continue;
      // K3 is dispatched over (max strands, instances); strand_local must also
      // be inside broad.strand_count.
      uint32_t nstr = b.strand_count;
      float keep_max = std::fmax(b.camera_keep_probability, b.shadow_keep_probability);
      float keep_f = F.keep[(size_t)ii * F.F + fi];
      size_t nseg_guess = is->index_count;
      (void)nseg_guess;
      // per-strand independent work: parallel over strand chunks, each chunk
      // keeps strand order, chunks are concatenated in order
      int nt = std::max(1, F.threads);
      std::vector<std::vector<sr_binning_task>> chunk_t(nt);
      std::vector<std::vector<sr_projected_segment>> chunk_p(nt);
      size_t per = (nstr + nt - 1) / nt;
      auto work = [&](int t) {
        size_t s0 = t * per, s1 = std::min<size_t>(nstr, s0 + per);
        for (size_t s = s0; s < s1; ++s) {
          uint32_t strand = (uint32_t)s;
          uint32_t h = wang_hash((ii * 16777619u) ^ strand);
          float rnd = hash_to_unit_float(h);
          if (rnd > keep_max) continue;  // K3
          if (strand >= ms->strand_count) continue;
          const sr_strand_meta& sm = ms->meta[strand];
          if (sm.count < 2u) continue;
          if (rnd > keep_f) continue;  // K5 per-frustum keep
          uint32_t segs = sm.count - 1u;
          for (uint32_t i = 0; i < segs; ++i) {
            uint32_t seg = sm.offset + i;
            if (seg + 1u >= is->index_count) continue;
            uint32_t vi0 = is->indices[seg], vi1 = is->indices[seg + 1u];
            V4 p0w = mat_vec(in.world_from_local, vertex_at(vs, vi0));
            V4 p1w = mat_vec(in.world_from_local, vertex_at(vs, vi1));
            V4 r0 = world_to_screen_raw(p0w, cfw, vw, vh);
            V4 r1 = world_to_screen_raw(p1w, cfw, vw, vh);
            V3 p0{r0.x, r0.y, normalize_depth01(r0.z)};
            V3 p1{r1.x, r1.y, normalize_depth01(r1.z)};
            Clipped c = clip_segment_to_box(p0, p1, box_min, box_max);
            if (!c.ok) continue;
            chunk_t[t].push_back(sr_binning_task{ii, strand, seg, packed});
            chunk_p[t].push_back(sr_projected_segment{
                pack2x16unorm(wclamp(c.p0.x / ssx, 0.0f, 1.0f), wclamp(c.p0.y / ssy, 0.0f, 1.0f)),
                pack2x16unorm(wclamp(c.p1.x / ssx, 0.0f, 1.0f), wclamp(c.p1.y / ssy, 0.0f, 1.0f)),
                pack2x16unorm(wclamp(c.p0.z, 0.0f, 1.0f), wclamp(c.p1.z, 0.0f, 1.0f))});
          }
        }
      };
      if (nt > 1 && nstr >= 1024) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t) pool.emplace_back(work, t);
        for (auto& th : pool) th.join();
      } else {
        per = nstr;
        work(0);
      }
      // canonical order inside (frustum, instance): index-slab position
      std::vector<std::pair<uint32_t, uint32_t>> order;  // (seg, chunk << 24 | item)
      std::vector<sr_binning_task> lt_all;
      std::vector<sr_projected_segment> lp_all;
      for (int t = 0; t < nt; ++t) {
        lt_all.insert(lt_all.end(), chunk_t[t].begin(), chunk_t[t].end());
        lp_all.insert(lp_all.end(), chunk_p[t].begin(), chunk_p[t].end());
      }
      std::vector<uint32_t> perm(lt_all.size());
      for (uint32_t k = 0; k < perm.size(); ++k) perm[k] = k;
      std::stable_sort(perm.begin(), perm.end(),
                       [&](uint32_t x, uint32_t y) { return lt_all[x].seg_idx < lt_all[y].seg_idx; });
      for (uint32_t k : perm) {
        ftasks[fi].push_back(lt_all[k]);
        fproj[fi].push_back(lp_all[k]);
      }
    }
  }
  F.tasks.clear();
  F.proj.clear();
  F.shading_queue.clear();
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    uint32_t base = (uint32_t)F.tasks.size();
    F.tasks.insert(F.tasks.end(), ftasks[fi].begin(), ftasks[fi].end());
    F.proj.insert(F.proj.end(), fproj[fi].begin(), fproj[fi].end());
    if (F.fr[fi].kind == 0u)
      for (uint32_t t = 0; t < ftasks[fi].size(); ++t) F.shading_queue.push_back(base + t);
  }
}

// ------------------------------------------------ K7 + K8 intervals / LUT
inline uint32_t disjoint_interval_gap(uint32_t amin, uint32_t amax, uint32_t bmin, uint32_t bmax) {
  if (amax < bmin) return bmin - amax;
  if (bmax < amin) return amin - bmax;
  return 0u;
}

// strand_prepass.wgsl:884-1014 over one coarse tile's interval list, given
// newest-first (the linked-list walk order).
void build_lut_for_tile(const std::vector<std::pair<uint32_t, uint32_t>>& ivals, sr_coarse_depth_lut_entry* out,
                        uint32_t S) {
  uint32_t mins[kMaxSlices], maxs[kMaxSlices];
  uint32_t count = 0;
  for (auto& iv : ivals) {
    uint32_t zmin = std::min(iv.first, iv.second), zmax = std::max(iv.first, iv.second);
    bool merged = false;
    for (uint32_t i = 0; i < S; ++i) {
      if (i >= count) break;
      if (zmin <= maxs[i] + 1u && zmax + 1u >= mins[i]) {
        mins[i] = std::min(mins[i], zmin);
        maxs[i] = std::max(maxs[i], zmax);
        merged = true;
      }
    }
    if (!merged) {
      if (count < S) {
        mins[count] = zmin; maxs[count] = zmax; ++count;
      } else {
        uint32_t best = 0, best_gap = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < S; ++i) {
          uint32_t gap = disjoint_interval_gap(mins[i], maxs[i], zmin, zmax);
          if (gap < best_gap) { best_gap = gap; best = i; }
        }
        mins[best] = std::min(mins[best], zmin);
        maxs[best] = std::max(maxs[best], zmax);
      }
    }
  }
  if (count == 0) {
    for (uint32_t i = 0; i < S; ++i) out[i] = sr_coarse_depth_lut_entry{0, 0, 0, 0};
    return;
  }
  for (uint32_t i = 1; i < S; ++i) {  // insertion sort by min
    if (i >= count) break;
    uint32_t j = i, kmin = mins[i], kmax = maxs[i];
    while (!(j == 0u || mins[j - 1u] <= kmin)) {
      mins[j] = mins[j - 1u]; maxs[j] = maxs[j - 1u]; --j;
    }
    mins[j] = kmin; maxs[j] = kmax;
  }
  uint32_t out_slice = 0;
  for (uint32_t iv = 0; iv < S; ++iv) {
    if (iv >= count) break;
    uint32_t base_cnt = std::max(1u, S / count);
    uint32_t rem = S % count;
    uint32_t slice_count = base_cnt + (iv < rem ? 1u : 0u);
    uint32_t zmin = mins[iv], zmax = maxs[iv];
    uint32_t span = std::max(1u, zmax - zmin + 1u);
    for (uint32_t ls = 0; ls < S; ++ls) {
      if (ls >= slice_count || out_slice >= S) break;
      uint32_t start = zmin + (span * ls) / slice_count;
      uint32_t ex_end = zmin + (span * (ls + 1u)) / slice_count;
      uint32_t end = std::min(ex_end > start ? ex_end - 1u : start, zmax);
      out[out_slice] = sr_coarse_depth_lut_entry{start, end, start, std::max(1u, end - start + 1u)};
      ++out_slice;
    }
  }
  while (out_slice < S) out[out_slice++] = sr_coarse_depth_lut_entry{0, 0, 0, 0};
}

// strand_prepass.wgsl:835-882 (intervals) + 884-1014 (LUT)
void build_luts(Frame& F) {
  const uint32_t S = F.slices();
  F.lut.assign((size_t)F.coarse_total * S, sr_coarse_depth_lut_entry{0, 0, 0, 0});
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> lists(F.coarse_total);
  // insertion in slot order (instance, frustum); a list walk is newest-first
  for (uint32_t ii = 0; ii < F.NI; ++ii)
    for (uint32_t fi = 0; fi < F.F; ++fi) {
      const sr_coarse_asset_range& r = F.ranges[(size_t)ii * F.F + fi];
      if (r.geo_id == kInvalid) continue;
      const sr_frustum_desc& fd = F.fr[r.frustum_id];
      if (fd.coarse_depth_tile_count == 0u || r.max_x < r.min_x || r.max_y < r.min_y) continue;
      for (uint32_t cy = r.min_y; cy <= r.max_y; ++cy)
        for (uint32_t cx = r.min_x; cx <= r.max_x; ++cx) {
          uint32_t local = cy * fd.coarse_tiles_x + cx;
          if (local < fd.coarse_depth_tile_count)
            lists[fd.coarse_depth_tile_base + local].push_back({r.z_min_q, r.z_max_q});
        }
    }
  for (uint32_t t = 0; t < F.coarse_total; ++t) {
    std::vector<std::pair<uint32_t, uint32_t>> newest_first(lists[t].rbegin(), lists[t].rend());
    build_lut_for_tile(newest_first, &F.lut[(size_t)t * S], S);
  }
}

// strand_prepass.wgsl:1251-1274
inline uint32_t lut_slice_for_depth(const Frame& F, uint32_t tile, float z) {
  uint32_t zq = quantize_depth01(z);
  const uint32_t S = F.slices();
  const sr_coarse_depth_lut_entry* lut = &F.lut[(size_t)tile * S];
  uint32_t best = kInvalid, best_d = 0xFFFFFFFFu;
  for (uint32_t i = 0; i < S; ++i) {
    const sr_coarse_depth_lut_entry& e = lut[i];
    if (e.virtual_count_q == 0u) continue;
    if (zq >= e.z_min_q && zq <= e.z_max_q) return i;
    uint32_t d = zq > e.z_max_q ? zq - e.z_max_q : e.z_min_q - zq;
    if (d < best_d) { best_d = d; best = i; }
  }
  return best;
}

// ------------------------------------------------- binning traversal
enum class Mode { Mark, Count, Fill };

struct FillCtx {
  std::vector<uint32_t>* cursors;
  sr_fine_seg_ref ref;
};

// fine_tile_occluded_by_opaque (strand_prepass.wgsl:551-573): a camera fine
// tile fully covered by opaque depth hides a slice whose nearest depth
// (z_max, reverse-Z) lies behind the tile's farthest opaque sample.
bool opaque_occluded(const Frame& F, const sr_frustum_desc& fd, uint32_t fine_x, uint32_t fine_y, float nearest) {
  if (F.opaque_tiles.empty() || fd.kind != 0u) return false;
  uint32_t ftx = (fd.screen_width + fd.froxel_size_x - 1u) / fd.froxel_size_x;
  uint32_t fty = (fd.screen_height + fd.froxel_size_y - 1u) / fd.froxel_size_y;
  if (fine_x >= ftx || fine_y >= fty) return false;
  size_t t = (size_t)fd.fine_depth_tile_base + fine_y * ftx + fine_x;
  if (2 * t + 1 >= F.opaque_tiles.size()) return false;
  if (F.opaque_tiles[2 * t] == 0u) return false;
  return nearest <= dequantize_depth01(F.opaque_tiles[2 * t + 1]) - 1e-4f;
}

// reduce_opaque_depth (opaque_depth_reduce.wgsl:35-86): per camera fine
// tile, whether every pixel has opaque depth (min over its samples > 0) and
// the quantised farthest (smallest) such depth.
void reduce_opaque_depth(Frame& F) {
  if (!F.opaque_depth) return;
  size_t tiles = 0;
  for (const auto& fd : F.fr) {
    uint32_t ftx = (fd.screen_width + fd.froxel_size_x - 1u) / fd.froxel_size_x;
    uint32_t fty = (fd.screen_height + fd.froxel_size_y - 1u) / fd.froxel_size_y;
    tiles = std::max<size_t>(tiles, (size_t)fd.fine_depth_tile_base + (size_t)ftx * fty);
  }
  F.opaque_tiles.assign(2 * tiles, 0u);
  uint64_t cam_px = 0;
  const uint32_t S = F.opaque_samples;
  for (const auto& fd : F.fr) {
    if (fd.kind != 0u) continue;
    uint32_t ftx = (fd.screen_width + fd.froxel_size_x - 1u) / fd.froxel_size_x;
    uint32_t fty = (fd.screen_height + fd.froxel_size_y - 1u) / fd.froxel_size_y;
    const float* img = F.opaque_depth + cam_px * S;
    for (uint32_t ty = 0; ty < fty; ++ty)
      for (uint32_t tx = 0; tx < ftx; ++tx) {
        uint32_t x0 = tx * fd.froxel_size_x, y0 = ty * fd.froxel_size_y;
        uint32_t x1 = std::min(x0 + fd.froxel_size_x, fd.screen_width);
        uint32_t y1 = std::min(y0 + fd.froxel_size_y, fd.screen_height);
        uint32_t covered = 0, total = 0;
        float farthest = 1.0f;
        for (uint32_t y = y0; y < y1; ++y)
          for (uint32_t x = x0; x < x1; ++x) {
            ++total;
            const float* d = img + ((size_t)y * fd.screen_width + x) * S;
            float depth = d[0];
            for (uint32_t k = 1; k < S; ++k) depth = std::fmin(depth, d[k]);
            if (depth > 0.0f) {
              ++covered;
              farthest = std::fmin(farthest, depth);
            }
          }
        size_t t = (size_t)fd.fine_depth_tile_base + ty * ftx + tx;
        F.opaque_tiles[2 * t] = (covered == total && total > 0u) ? 1u : 0u;
        F.opaque_tiles[2 * t + 1] = quantize_depth01(farthest);
      }
    cam_px += (uint64_t)fd.screen_width * fd.screen_height;
  }
}

// strand_prepass.wgsl:1412-1575 trace_segment_into_coarse_count_page
void trace_coarse_tile(Frame& F, V3 p0, V3 p1, const sr_frustum_desc& fd, uint32_t tile, uint32_t local,
                       Mode mode, FillCtx* fc, std::atomic<uint32_t>* counts) {
  uint32_t cx = local % fd.coarse_tiles_x, cy = local / fd.coarse_tiles_x;
  uint32_t cpx = std::max(1u, fd.froxel_size_x * kExt), cpy = std::max(1u, fd.froxel_size_y * kExt);
  V2 origin{(float)(cx * cpx), (float)(cy * cpy)};
  V2 tmx{std::fmin(origin.x + (float)cpx, (float)fd.screen_width) - 1e-3f,
         std::fmin(origin.y + (float)cpy, (float)fd.screen_height) - 1e-3f};
  V3 bmax{std::fmax(origin.x, tmx.x), std::fmax(origin.y, tmx.y), 1.0f};
  Clipped c = clip_segment_to_box(p0, p1, V3{origin.x, origin.y, 0.0f}, bmax);
  if (!c.ok) return;
  uint32_t z0s = lut_slice_for_depth(F, tile, c.p0.z), z1s = lut_slice_for_depth(F, tile, c.p1.z);
  if (z0s == kInvalid || z1s == kInvalid) return;
  uint32_t zlo = std::min(z0s, z1s), zhi = std::max(z0s, z1s);
  for (uint32_t cz = zlo; cz <= zhi; ++cz) {
    const sr_coarse_depth_lut_entry& e = F.lut[(size_t)tile * F.slices() + cz];
    if (e.virtual_count_q == 0u) continue;
    float zmin = dequantize_depth01(e.z_min_q), zmax = dequantize_depth01(e.z_max_q);
    Clipped dc = clip_segment_to_box(c.p0, c.p1, V3{origin.x, origin.y, zmin}, V3{bmax.x, bmax.y, zmax});
    if (!dc.ok) continue;
    uint32_t pti = tile * F.slices() + cz;
    if (mode == Mode::Mark) {
      F.page_table[pti] = kMarked;
      continue;
    }
    uint32_t h = F.page_table[pti];
    if (h == 0u || h == kMarked) continue;
    uint32_t page = h - 1u;
    float zspan = std::fmax(zmax - zmin, 1.0f / (float)kQuant);
    float fx = (float)fd.froxel_size_x, fy = (float)fd.froxel_size_y;
    V3 a{(dc.p0.x - origin.x) / fx, (dc.p0.y - origin.y) / fy,
         wclamp((dc.p0.z - zmin) / zspan, 0.0f, 0.999999f) * (float)kSlices};
    V3 b{(dc.p1.x - origin.x) / fx, (dc.p1.y - origin.y) / fy,
         wclamp((dc.p1.z - zmin) / zspan, 0.0f, 0.999999f) * (float)kSlices};
    int32_t f0[3] = {iclamp(f2i(std::floor(a.x)), 0, (int)kExt - 1), iclamp(f2i(std::floor(a.y)), 0, (int)kExt - 1),
                     iclamp(f2i(std::floor(a.z)), 0, (int)kSlices - 1)};
    int32_t f1[3] = {iclamp(f2i(std::floor(b.x)), 0, (int)kExt - 1), iclamp(f2i(std::floor(b.y)), 0, (int)kExt - 1),
                     iclamp(f2i(std::floor(b.z)), 0, (int)kSlices - 1)};
    float dir[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
    float pa[3] = {a.x, a.y, a.z};
    int32_t step[3];
    float safe[3], delta[3], tmax[3];
    for (int k = 0; k < 3; ++k) {
      step[k] = sgn_i32(dir[k]);
      bool tiny = std::fabs(dir[k]) < 1e-6f;
      safe[k] = tiny ? 1e-6f : dir[k];
      delta[k] = std::fabs(1.0f / safe[k]);
      float fl = std::floor(pa[k]);
      tmax[k] = step[k] > 0 ? ((fl + 1.0f) - pa[k]) / safe[k] : (fl - pa[k]) / safe[k];
      if (tiny) tmax[k] = 1e38f;
    }
    int32_t f[3] = {f0[0], f0[1], f0[2]};
    uint32_t safety = 0;
    const sr_fine_page_meta* pm = (mode == Mode::Fill) ? &F.page_meta[page] : nullptr;
    for (;;) {
      if (safety > kExt + kExt + kSlices + 3u) break;
      uint32_t cell = ((uint32_t)f[2] * kExt + (uint32_t)f[1]) * kExt + (uint32_t)f[0];
      if (opaque_occluded(F, fd, cx * kExt + (uint32_t)f[0], cy * kExt + (uint32_t)f[1], zmax)) {
        // skipped cell, the walk goes on (:1541-1555)
      } else if (mode == Mode::Count) {
        counts[(size_t)page * kPage + cell].fetch_add(1u, std::memory_order_relaxed);
      } else if (pm->seg_ref_count != 0u) {  // write_fine_seg_ref, :1329-1354
        size_t gc = (size_t)page * kPage + cell;
        uint32_t w = (*fc->cursors)[gc]++;
        uint64_t dst = (uint64_t)pm->seg_ref_base + F.cell_offsets[gc] + w;
        if (dst < F.refs.size()) F.refs[dst] = fc->ref;
      }
      if (f[0] == f1[0] && f[1] == f1[1] && f[2] == f1[2]) break;
      bool m[3];
      canonical_min_mask(V3{tmax[0], tmax[1], tmax[2]}, m);
      for (int k = 0; k < 3; ++k)
        if (m[k]) { f[k] += step[k]; tmax[k] += delta[k]; }
      if (f[0] < 0 || f[1] < 0 || f[2] < 0 || f[0] >= (int)kExt || f[1] >= (int)kExt || f[2] >= (int)kSlices) break;
      ++safety;
    }
  }
}

// strand_prepass.wgsl:1591-1696 visit_segment_asset_coarse_tiles
void visit_coarse_tiles(Frame& F, V3 p0, V3 p1, uint32_t fi, uint32_t ii, Mode mode, FillCtx* fc,
                        std::atomic<uint32_t>* counts) {
  const sr_coarse_asset_range& r = F.ranges[(size_t)ii * F.F + fi];
  if (r.geo_id != ii || r.frustum_id != fi) return;
  const sr_frustum_desc& fd = F.fr[fi];
  uint32_t cpx = std::max(1u, fd.froxel_size_x * kExt), cpy = std::max(1u, fd.froxel_size_y * kExt);
  V2 smin{std::fmin(p0.x, p1.x), std::fmin(p0.y, p1.y)};
  V2 smax{std::fmax(p0.x, p1.x), std::fmax(p0.y, p1.y)};
  uint32_t min_cx = std::max(r.min_x, std::min(f2u(std::fmax(smin.x, 0.0f)) / cpx, fd.coarse_tiles_x - 1u));
  uint32_t max_cx = std::min(r.max_x, std::min(f2u(std::fmax(smax.x, 0.0f)) / cpx, fd.coarse_tiles_x - 1u));
  uint32_t min_cy = std::max(r.min_y, std::min(f2u(std::fmax(smin.y, 0.0f)) / cpy, fd.coarse_tiles_y - 1u));
  uint32_t max_cy = std::min(r.max_y, std::min(f2u(std::fmax(smax.y, 0.0f)) / cpy, fd.coarse_tiles_y - 1u));
  if (max_cx < min_cx || max_cy < min_cy) return;
  V3 bmin{(float)(min_cx * cpx), (float)(min_cy * cpy), 0.0f};
  V3 bmax{std::fmin((float)((max_cx + 1u) * cpx), (float)fd.screen_width) - 1e-3f,
          std::fmin((float)((max_cy + 1u) * cpy), (float)fd.screen_height) - 1e-3f, 1.0f};
  Clipped c = clip_segment_to_box(p0, p1, bmin, bmax);
  if (!c.ok) return;
  float pa[2] = {c.p0.x / (float)cpx, c.p0.y / (float)cpy};
  float pb[2] = {c.p1.x / (float)cpx, c.p1.y / (float)cpy};
  int32_t f0[2] = {iclamp(f2i(std::floor(pa[0])), (int)min_cx, (int)max_cx),
                   iclamp(f2i(std::floor(pa[1])), (int)min_cy, (int)max_cy)};
  int32_t f1[2] = {iclamp(f2i(std::floor(pb[0])), (int)min_cx, (int)max_cx),
                   iclamp(f2i(std::floor(pb[1])), (int)min_cy, (int)max_cy)};
  int32_t step[2];
  float safe[2], delta[2], tmax[2];
  for (int k = 0; k < 2; ++k) {
    float d = pb[k] - pa[k];
    step[k] = sgn_i32(d);
    bool tiny = std::fabs(d) < 1e-6f;
    safe[k] = tiny ? 1e-6f : d;
    delta[k] = std::fabs(1.0f / safe[k]);
    float fl = std::floor(pa[k]);
    tmax[k] = step[k] > 0 ? ((fl + 1.0f) - pa[k]) / safe[k] : (fl - pa[k]) / safe[k];
    if (tiny) tmax[k] = 1e38f;
  }
  int32_t f[2] = {f0[0], f0[1]};
  uint32_t safety = 0;
  uint32_t max_steps = (max_cx - min_cx + 1u) + (max_cy - min_cy + 1u) + 2u;
  for (;;) {
    if (safety > max_steps) break;
    if (f[0] < (int)min_cx || f[0] > (int)max_cx || f[1] < (int)min_cy || f[1] > (int)max_cy) break;
    uint32_t local = (uint32_t)f[1] * fd.coarse_tiles_x + (uint32_t)f[0];
    if (local < fd.coarse_depth_tile_count)
      trace_coarse_tile(F, p0, p1, fd, fd.coarse_depth_tile_base + local, local, mode, fc, counts);
    if (f[0] == f1[0] && f[1] == f1[1]) break;
    if (tmax[0] <= tmax[1]) { f[0] += step[0]; tmax[0] += delta[0]; }
    else { f[1] += step[1]; tmax[1] += delta[1]; }
    ++safety;
  }
}

inline void task_endpoints(const Frame& F, uint32_t t, uint32_t* fi, V3* p0, V3* p1) {
  const sr_binning_task& tk = F.tasks[t];
  *fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
  const sr_frustum_desc& fd = F.fr[*fi];
  const sr_projected_segment& ps = F.proj[t];
  float W = (float)fd.screen_width, H = (float)fd.screen_height;
  V2 a = unpack2x16unorm(ps.p0_xy), b = unpack2x16unorm(ps.p1_xy), d = unpack2x16unorm(ps.depths);
  *p0 = V3{a.x * W, a.y * H, d.x};
  *p1 = V3{b.x * W, b.y * H, d.y};
}

// K10 mark (:1928-1932) + K11 allocate (:1934-1969) + K13 count (:1986-1990)
void bin_count(Frame& F) {
  const uint32_t S = F.slices();
  size_t ptab = (size_t)F.coarse_total * S;
  F.page_table.assign(ptab, 0u);
  for (uint32_t t = 0; t < F.tasks.size(); ++t) {  // mark: idempotent stores
    uint32_t fi; V3 p0, p1;
    task_endpoints(F, t, &fi, &p0, &p1);
    visit_coarse_tiles(F, p0, p1, fi, F.tasks[t].id_info, Mode::Mark, nullptr, nullptr);
  }
  uint32_t P = 0;
  F.page_meta.clear();
  for (size_t e = 0; e < ptab; ++e) {
    if (F.page_table[e] != kMarked) continue;
    sr_fine_page_meta pm{};
    pm.coarse_tile_id = (uint32_t)(e / S);
    pm.coarse_z = (uint32_t)(e % S);
    F.page_meta.push_back(pm);
    F.page_table[e] = ++P;  // page + 1
  }
  std::vector<std::atomic<uint32_t>> counts((size_t)P * kPage);
  for (auto& c : counts) c.store(0u, std::memory_order_relaxed);
  parallel_for(F.threads, F.tasks.size(), [&](size_t b, size_t e) {
    for (size_t t = b; t < e; ++t) {
      uint32_t fi; V3 p0, p1;
      task_endpoints(F, (uint32_t)t, &fi, &p0, &p1);
      visit_coarse_tiles(F, p0, p1, fi, F.tasks[t].id_info, Mode::Count, nullptr, counts.data());
    }
  });
  F.cell_counts.resize((size_t)P * kPage);
  for (size_t i = 0; i < F.cell_counts.size(); ++i) F.cell_counts[i] = counts[i].load();
  F.ctr.allocated_pages = P;
}

inline uint32_t frustum_for_coarse_tile(const Frame& F, uint32_t tile) {
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    const sr_frustum_desc& fd = F.fr[fi];
    if (tile >= fd.coarse_depth_tile_base && tile < fd.coarse_depth_tile_base + fd.coarse_depth_tile_count) return fi;
  }
  return kInvalid;
}

// K14 prefix_fine_pages (:2051-2143) + append_active_fine_tile (:2207-2238)
void prefix_pages(Frame& F) {
  uint32_t P = F.ctr.allocated_pages;
  F.cell_offsets.assign((size_t)P * kPage, 0u);
  F.active_flags.assign((size_t)F.coarse_total * kExt * kExt, 0u);
  uint64_t base = 0;
  uint32_t overflow = 0;
  for (uint32_t p = 0; p < P; ++p) {
    uint32_t acc = 0;
    for (uint32_t c = 0; c < kPage; ++c) {
      F.cell_offsets[(size_t)p * kPage + c] = acc;
      acc += F.cell_counts[(size_t)p * kPage + c];
    }
    sr_fine_page_meta& pm = F.page_meta[p];
    uint32_t flags = 0;
    uint32_t pbase = 0;
    if (acc > 0u) {
      pbase = (uint32_t)base;
      if (base + acc > F.ref_capacity) flags = 1u;
      base += acc;
    }
    pm.seg_ref_base = pbase;
    pm.seg_ref_count = flags ? 0u : acc;
    pm.work_base = 0; pm.work_count = 0; pm.first_depth_key = 0;
    pm.flags = flags;
    if (flags) ++overflow;
    uint32_t fi = frustum_for_coarse_tile(F, pm.coarse_tile_id);
    if (flags != 0u || fi == kInvalid) continue;
    const sr_frustum_desc& fd = F.fr[fi];
    for (uint32_t xy = 0; xy < kExt * kExt; ++xy) {
      bool has = false;
      for (uint32_t z = 0; z < kSlices; ++z)
        has = has || F.cell_counts[(size_t)p * kPage + z * kExt * kExt + xy] > 0u;
      if (!has) continue;
      uint32_t local_coarse = pm.coarse_tile_id - fd.coarse_depth_tile_base;
      uint32_t cx = local_coarse % fd.coarse_tiles_x, cy = local_coarse / fd.coarse_tiles_x;
      uint32_t fxx = cx * kExt + xy % kExt, fyy = cy * kExt + xy / kExt;
      if (fxx >= ceil_div(fd.screen_width, fd.froxel_size_x) || fyy >= ceil_div(fd.screen_height, fd.froxel_size_y))
        continue;
      F.active_flags[(size_t)pm.coarse_tile_id * kExt * kExt + xy] = 1u;
    }
  }
  F.ctr.overflow_pages = overflow;
  F.ctr.fine_refs = (uint32_t)std::min<uint64_t>(base, 0xFFFFFFFFu);
}

// K15 fill_fine_seg_refs (:2190-2194) + make_fine_seg_ref (:1313-1327)
void bin_fill(Frame& F) {
  F.refs.assign(std::min<uint64_t>(F.ctr.fine_refs, F.ref_capacity), sr_fine_seg_ref{0, 0, 0});
  std::vector<uint32_t> cursors(F.cell_counts.size(), 0u);
  FillCtx fc{&cursors, {}};
  for (uint32_t t = 0; t < F.tasks.size(); ++t) {  // task order == (inst, seg) order per cell
    const sr_binning_task& tk = F.tasks[t];
    uint32_t fi; V3 p0, p1;
    task_endpoints(F, t, &fi, &p0, &p1);
    const sr_strand_instance& in = F.inst[tk.id_info];
    uint32_t mat = 0;
    const sr_asset_slabs* ms = slab(F, in.meta_id);
    if (ms && tk.chunk_id < ms->strand_count) mat = ms->meta[tk.chunk_id].material_idx;
    fc.ref = sr_fine_seg_ref{tk.id_info, tk.seg_idx, mat};
    visit_coarse_tiles(F, p0, p1, fi, tk.id_info, Mode::Fill, &fc, nullptr);
  }
}

// K20 prefix_active_tile_blocks + K21 compact_active_tiles (:2581-2676)
void compact_active(Frame& F) {
  F.active_queue.clear();
  for (uint32_t i = 0; i < F.active_flags.size(); ++i)
    if (F.active_flags[i]) F.active_queue.push_back(i);
  F.ctr.active_tiles = (uint32_t)F.active_queue.size();
}

// strand_prepass.wgsl:2509-2518
inline uint32_t raster_depth_key(const Frame& F, uint32_t pti, uint32_t fine_z) {
  if (pti >= F.lut.size()) return fine_z;
  const sr_coarse_depth_lut_entry& e = F.lut[pti];
  if (e.virtual_count_q == 0u) return fine_z;
  return e.virtual_start_q + (e.virtual_count_q * fine_z) / kSlices;
}

// K23 emit_raster_work (:2678-2834)
void emit_work(Frame& F) {
  F.work.clear(); F.cam_runs.clear(); F.shadow_runs.clear();
  for (uint32_t at : F.active_queue) {
    uint32_t local_fine = at % (kExt * kExt), tile = at / (kExt * kExt);
    uint32_t fi = frustum_for_coarse_tile(F, tile);
    if (fi == kInvalid) continue;
    const sr_frustum_desc& fd = F.fr[fi];
    uint32_t lc = tile - fd.coarse_depth_tile_base;
    uint32_t cx = lc % fd.coarse_tiles_x, cy = lc / fd.coarse_tiles_x;
    uint32_t ftx = ceil_div(fd.screen_width, fd.froxel_size_x);
    uint32_t lx = local_fine % kExt, ly = local_fine / kExt;
    uint32_t fxx = cx * kExt + lx, fyy = cy * kExt + ly;
    uint32_t screen_tile = fyy * ftx + fxx;
    uint32_t base = (uint32_t)F.work.size();
    uint32_t load = 0, count = 0;
    const uint32_t S = F.slices(), lanes = S * kSlices;  // (LUT slice, fine slice) depth lanes
    for (uint32_t lane = 0; lane < lanes; ++lane) {
      uint32_t dl = (lanes - 1u) - lane;  // near (high keys) first
      uint32_t cz = dl / kSlices, fz = dl % kSlices;
      uint32_t pti = tile * S + cz;
      uint32_t h = F.page_table[pti];
      if (h == 0u || h == kMarked) continue;
      uint32_t page = h - 1u;
      const sr_fine_page_meta& pm = F.page_meta[page];
      uint32_t cell = (fz * kExt + ly) * kExt + lx;
      if (!(pm.seg_ref_count > 0u && pm.flags == 0u)) continue;
      uint32_t n = F.cell_counts[(size_t)page * kPage + cell];
      if (n == 0u) continue;
      uint32_t sb = pm.seg_ref_base + F.cell_offsets[(size_t)page * kPage + cell];
      F.work.push_back(sr_raster_work_item{sb, n, fi, screen_tile, raster_depth_key(F, pti, fz), n, page, cell});
      load += n; ++count;
    }
    if (count == 0u) continue;
    sr_raster_tile_run run{base, count, fi, screen_tile, load, 0, 0, 0};
    if (fd.kind == 1u) F.shadow_runs.push_back(run); else F.cam_runs.push_back(run);
  }
  F.ctr.work_items = (uint32_t)F.work.size();
  F.ctr.camera_runs = (uint32_t)F.cam_runs.size();
  F.ctr.shadow_runs = (uint32_t)F.shadow_runs.size();
}

// ------------------------------------------------------ segment fetch
struct SegWorld { V4 v0, v1; };
inline SegWorld seg_world(const Frame& F, const sr_strand_instance& in, uint32_t seg) {
  // get_segment_vertices_from_index_unchecked, strand_rasterizer.wgsl:289-303
  const sr_asset_slabs* is = slab(F, in.index_id);
  const sr_asset_slabs* vs = slab(F, in.vertex_id);
  uint32_t i0 = is->indices[seg], i1 = is->indices[seg + 1u];
  return SegWorld{mat_vec(in.world_from_local, vertex_at(vs, i0)), mat_vec(in.world_from_local, vertex_at(vs, i1))};
}
inline const sr_strand_material& material_of(const Frame& F, const sr_strand_instance& in, uint32_t idx) {
  return slab(F, in.material_id)->materials[idx];
}

// ----------------------------------------------- K25 shadow DOM raster
// strand_rasterizer.wgsl:391-711, without VSMS paging (a dense map per
// light) and without opaque shadow geometry (opaque depth 0).  Every pixel
// of a tile is rasterised (the reference serves 64 pixel slots per tile).
void shadow_raster(Frame& F) {
  uint32_t L = F.st.dom_layers;
  F.dom_depth_base.assign(F.F, 0); F.dom_opacity_base.assign(F.F, 0);
  uint64_t dsz = 0, osz = 0;
  for (uint32_t fi = 0; fi < F.F; ++fi)
    if (F.fr[fi].kind == 1u) {
      F.dom_depth_base[fi] = dsz; F.dom_opacity_base[fi] = osz;
      uint64_t px = (uint64_t)F.fr[fi].screen_width * F.fr[fi].screen_height;
      dsz += px; osz += px * L;
    }
  F.dom_depth.assign(dsz, 0.0f);
  F.dom_opacity.assign(osz, 0.0f);
  parallel_for(F.threads, F.shadow_runs.size(), [&](size_t rb, size_t re) {
    std::vector<V3> bp0, bp1; std::vector<V2> brad; std::vector<float> babs; std::vector<char> bvalid;
    for (size_t ri = rb; ri < re; ++ri) {
      const sr_raster_tile_run& run = F.shadow_runs[ri];
      uint32_t fi = run.frustum_id;
      const sr_frustum_desc& fd = F.fr[fi];
      uint32_t layer;
      if (!light_layer_ok(F, fi, &layer)) continue;
      const float* lcfw = F.views[fi].clip_from_world;
      float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
      uint32_t ftx = ceil_div(fd.screen_width, fd.froxel_size_x);
      uint32_t tx = run.screen_tile_id % ftx, ty = run.screen_tile_id / ftx;
      uint32_t npx = fd.froxel_size_x * fd.froxel_size_y;
      // stage every work item's refs once (projection + material)
      std::vector<uint32_t> wstart(run.work_count + 1, 0);
      bp0.clear(); bp1.clear(); brad.clear(); babs.clear(); bvalid.clear();
      for (uint32_t w = 0; w < run.work_count; ++w) {
        const sr_raster_work_item& wi = F.work[run.work_base + w];
        bool valid_item = wi.frustum_id == fi && wi.screen_tile_id == run.screen_tile_id;
        wstart[w] = (uint32_t)bp0.size();
        for (uint32_t r = 0; r < wi.seg_ref_count; ++r) {
          const sr_fine_seg_ref& sr = F.refs[wi.seg_ref_base + r];
          const sr_strand_instance& in = F.inst[sr.inst_id];
          SegWorld sw = seg_world(F, in, sr.seg_id);
          V4 a = world_to_screen_raw(sw.v0, lcfw, vw, vh), b = world_to_screen_raw(sw.v1, lcfw, vw, vh);
          const sr_strand_material& m = material_of(F, in, sr.material_idx);
          float minr = std::fmax(m.min_radius_pixels, 1e-4f);
          bp0.push_back(V3{a.x, a.y, normalize_depth01(a.z)});
          bp1.push_back(V3{b.x, b.y, normalize_depth01(b.z)});
          brad.push_back(V2{minr, std::fmax(m.max_radius_pixels, minr)});
          babs.push_back(m.absorption_color[3]);
          bvalid.push_back(valid_item ? 1 : 0);
        }
      }
      wstart[run.work_count] = (uint32_t)bp0.size();
      for (uint32_t pidx = 0; pidx < npx; ++pidx) {
        uint32_t pxu = tx * fd.froxel_size_x + pidx % fd.froxel_size_x;
        uint32_t pyu = ty * fd.froxel_size_y + pidx / fd.froxel_size_x;
        if (pxu >= fd.screen_width || pyu >= fd.screen_height) continue;
        float pfx = (float)pxu + 0.5f, pfy = (float)pyu + 0.5f;
        // coverage test shared by both passes
        auto cover = [&](uint32_t k, float* pz, float* cov) -> bool {
          if (!bvalid[k]) return false;
          V3 p0 = bp0[k], p1 = bp1[k];
          // fragment_position_line_relative, strand_rasterizer.wgsl:138-147
          float abx = p1.x - p0.x, aby = p1.y - p0.y;
          float l2 = abx * abx + aby * aby;
          float t = 0.0f;
          if (!(l2 == 0.0f)) t = ((pfx - p0.x) * abx + (pfy - p0.y) * aby) / l2;
          if (t < 0.0f || t > 1.0f) return false;
          V3 p{wmix(p0.x, p1.x, t), wmix(p0.y, p1.y, t), wmix(p0.z, p1.z, t)};
          // in_front_of_opaque_shadow_depth with opaque depth 0 -> always true
          float r = wmix(brad[k].x, brad[k].y, wclamp(p.z, 0.0f, 1.0f));
          float dx = pfx - p.x, dy = pfy - p.y;
          float c = wclamp(1.0f - std::sqrt(dx * dx + dy * dy) / r, 0.0f, 1.0f);
          *pz = p.z; *cov = c;
          return c > 0.0f;
        };
        // pass 1: z0 = max depth among covering refs of the first work item
        // (near to far) with any hit
        float z0 = -1.0f;
        bool found = false;
        for (uint32_t w = 0; w < run.work_count && !found; ++w) {
          for (uint32_t k = wstart[w]; k < wstart[w + 1]; ++k) {
            float pz, cov;
            if (cover(k, &pz, &cov)) { z0 = std::fmax(z0, pz); found = true; }
          }
        }
        size_t dpx = F.dom_depth_base[fi] + (size_t)pyu * fd.screen_width + pxu;
        size_t plane = (size_t)fd.screen_width * fd.screen_height;
        size_t opx = F.dom_opacity_base[fi] + (size_t)pyu * fd.screen_width + pxu;
        if (!found) {  // no opaque shadow depth: clear depth and layers
          F.dom_depth[dpx] = 0.0f;
          for (uint32_t i = 0; i < L; ++i) F.dom_opacity[opx + i * plane] = 0.0f;
          continue;
        }
        float part[SR_MAX_DOM_LAYERS];
        for (uint32_t i = 0; i < L; ++i) part[i] = 0.0f;
        float inv_span = 1.0f / std::fmax(1e-6f, z0);
        for (uint32_t k = 0; k < wstart[run.work_count]; ++k) {
          float pz, cov;
          if (!cover(k, &pz, &cov)) continue;
          float dzp = std::fmax(0.0f, z0 - pz) * inv_span;
          float u = std::pow(wclamp(dzp, 0.0f, 1.0f), 1.0f);  // DOM_GAMMA = 1
          float tL = u * (float)L;
          uint32_t si = std::min(f2u(std::floor(tL)), L - 1u);
          float w = wfract(tL);
          float a = cov * babs[k] * 0.5f;
          float a0 = part[si];
          part[si] = wclamp(a0 + (1.0f - a0) * (1.0f - w) * a, 0.0f, 1.0f);
          if (si + 1u < L) {
            float a1 = part[si + 1u];
            part[si + 1u] = wclamp(a1 + (1.0f - a1) * w * a, 0.0f, 1.0f);
          }
        }
        F.dom_depth[dpx] = z0;
        float acc = 0.0f;  // opaque_terminal_slice == L without opaque depth
        for (uint32_t i = 0; i < L; ++i) {
          acc = acc + (1.0f - acc) * part[i];
          F.dom_opacity[opx + i * plane] = acc;
        }
      }
    }
  });
}

// ------------------------------------------------ virtual DOM paging
// The reference's VSMS surfaces (src/plugin.rs:1776-1854) with page
// requests as request_shadow_dom_pages makes them
// (strand_prepass.wgsl:2537-2579: every shadow tile run requests the pages
// its pixels fall in and their neighbours within DOM_PREFETCH_PAGE_BORDER
// pages).  The residency policy lives in the absent bevy_vsms crate; this
// restatement makes requested pages resident light by light, in page order,
// up to dom_resident_pages per light (0 = all).  A texel of a page that is
// not resident is never stored (strand_rasterizer.wgsl:484-517) and reads as
// depth 0 / opacity 0 (shadow_dom_stampback.wgsl:58-72), where shading lets
// all light through (strand_shading.wgsl:238-257) as it does for depth 0; so
// the dense DOM with those texels cleared is the same frame.
void dom_pages(Frame& F) {
  const uint32_t P = SR_DOM_PAGE_XY, B = SR_DOM_PREFETCH_PAGE_BORDER, L = F.st.dom_layers;
  std::vector<uint32_t> base(F.F, 0), pages_x(F.F, 0), pages_y(F.F, 0);
  uint32_t total = 0;
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    if (F.fr[fi].kind != 1u) continue;
    base[fi] = total;
    pages_x[fi] = ceil_div(F.fr[fi].screen_width, P);
    pages_y[fi] = ceil_div(F.fr[fi].screen_height, P);
    total += pages_x[fi] * pages_y[fi];
  }
  F.dom_requests.assign(total, 0u);
  F.dom_table.assign(total, kInvalid);
  for (const sr_raster_tile_run& run : F.shadow_runs) {
    const uint32_t fi = run.frustum_id;
    if (run.work_count == 0u || !light_layer_ok(F, fi, nullptr)) continue;
    const sr_frustum_desc& fd = F.fr[fi];
    const uint32_t ftx = ceil_div(fd.screen_width, fd.froxel_size_x);
    const uint32_t x0 = (run.screen_tile_id % ftx) * fd.froxel_size_x, y0 = (run.screen_tile_id / ftx) * fd.froxel_size_y;
    const uint32_t x1 = std::min(x0 + fd.froxel_size_x - 1u, fd.screen_width - 1u);
    const uint32_t y1 = std::min(y0 + fd.froxel_size_y - 1u, fd.screen_height - 1u);
    for (uint32_t py = y0 / P; py <= y1 / P; ++py)      // request_shadow_dom_pages
      for (uint32_t px = x0 / P; px <= x1 / P; ++px) {  // mark_shadow_dom_page_request_with_border
        const uint32_t mx = px - std::min(px, B), my = py - std::min(py, B);
        const uint32_t Mx = std::min(px + B, pages_x[fi] - 1u), My = std::min(py + B, pages_y[fi] - 1u);
        for (uint32_t y = my; y <= My; ++y)
          for (uint32_t x = mx; x <= Mx; ++x) F.dom_requests[base[fi] + y * pages_x[fi] + x] = 1u;
      }
  }
  uint32_t phys = 0;
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    if (F.fr[fi].kind != 1u) continue;
    uint32_t resident = 0;
    for (uint32_t v = 0; v < pages_x[fi] * pages_y[fi]; ++v) {
      if (!F.dom_requests[base[fi] + v]) continue;
      if (F.st.dom_resident_pages != 0u && resident >= F.st.dom_resident_pages) continue;
      F.dom_table[base[fi] + v] = phys++;
      ++resident;
    }
    const sr_frustum_desc& fd = F.fr[fi];
    const size_t plane = (size_t)fd.screen_width * fd.screen_height;
    for (uint32_t y = 0; y < fd.screen_height; ++y)
      for (uint32_t x = 0; x < fd.screen_width; ++x) {
        if (F.dom_table[base[fi] + (y / P) * pages_x[fi] + x / P] != kInvalid) continue;
        const size_t pix = (size_t)y * fd.screen_width + x;
        F.dom_depth[F.dom_depth_base[fi] + pix] = 0.0f;
        for (uint32_t i = 0; i < L; ++i) F.dom_opacity[F.dom_opacity_base[fi] + i * plane + pix] = 0.0f;
      }
  }
  F.ctr.dom_pages = phys;
}

// ------------------------------------------------- K26 DOM stampback
// shadow_dom_stampback.wgsl:74-123: max DOM depth over the target texel's
// footprint, 0 (discard) when empty.
void stampback(Frame& F) {
  F.stamp.clear();
  for (uint32_t fi = 0; fi < F.F; ++fi) {
    const sr_frustum_desc& fd = F.fr[fi];
    if (fd.kind != 1u) continue;
    uint32_t tw = F.st.stamp_target_width ? F.st.stamp_target_width : fd.screen_width;
    uint32_t th = F.st.stamp_target_height ? F.st.stamp_target_height : fd.screen_height;
    size_t base = F.stamp.size();
    F.stamp.resize(base + (size_t)tw * th, 0.0f);
    for (uint32_t y = 0; y < th; ++y)
      for (uint32_t x = 0; x < tw; ++x) {
        uint32_t x0 = std::min((x * fd.screen_width) / tw, fd.screen_width - 1u);
        uint32_t y0 = std::min((y * fd.screen_height) / th, fd.screen_height - 1u);
        uint32_t x1 = std::min(((x + 1u) * fd.screen_width + tw - 1u) / tw, fd.screen_width);
        uint32_t y1 = std::min(((y + 1u) * fd.screen_height + th - 1u) / th, fd.screen_height);
        float z = 0.0f;
        for (uint32_t yy = y0; yy < y1; ++yy)
          for (uint32_t xx = x0; xx < x1; ++xx)
            z = std::fmax(z, F.dom_depth[F.dom_depth_base[fi] + (size_t)yy * fd.screen_width + xx]);
        F.stamp[base + (size_t)y * tw + x] = z;  // discarded texels keep the cleared 0
      }
  }
}

// --------------------------------------------------------- K27 shading
constexpr double kPI = 3.14159265359;
const float PI_F = (float)kPI;
const float PI_HALF_LO = (float)(-kPI / 2.0 + 1e-6);
const float PI_HALF_HI = (float)(kPI / 2.0 - 1e-6);
const float SQRT_2PI_F = (float)2.5066282746310002;
const float TWO_PI_F = (float)(2.0 * kPI);
const float LEG_ROOTS[10] = {-0.9739065285171717f, -0.8650633666889844f, -0.6794095682990244f, -0.4333953941292472f,
                             -0.14887433898163116f, 0.14887433898163116f, 0.4333953941292472f, 0.6794095682990244f,
                             0.8650633666889844f, 0.9739065285171717f};
const float LEG_WEIGHTS[10] = {0.06667134430868714f, 0.14945134915058053f, 0.21908636251598224f, 0.26926671930999674f,
                               0.2955242247147533f, 0.2955242247147533f, 0.26926671930999674f, 0.21908636251598224f,
                               0.14945134915058053f, 0.06667134430868714f};
const float FACT[10] = {1.0f, 1.0f, 2.0f, 6.0f, 24.0f, 120.0f, 720.0f, 5040.0f, 40320.0f, 362880.0f};

inline V3 v3sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 v3scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 v3cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 v3norm(V3 a) { float l = std::sqrt(v3dot(a, a)); return V3{a.x / l, a.y / l, a.z / l}; }
inline float wsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
// acos/asin arguments are kept inside [-1, 1]; WGSL leaves them undefined
// outside, where both the GPU and libm would return NaN.
inline float sacos(float x) { return std::acos(wclamp(x, -1.0f, 1.0f)); }
inline float sasin(float x) { return std::asin(wclamp(x, -1.0f, 1.0f)); }

// strand_shading.wgsl:335-343
float bessel_first_approx(float x, uint32_t steps) {
  float y = 0.0f;
  for (uint32_t i = 0; i < steps; ++i) {
    float num = std::pow(x, 2.0f * (float)i);
    float den = std::pow(4.0f, (float)i) * FACT[i] * FACT[i];
    y = y + num / den;
  }
  return y;
}
// strand_shading.wgsl:360-376 (clamp(0, 1, c) = min(1, c))
float fresnel_reflectance(float eta_i, float eta_t, float cos_theta_i) {
  float c = std::fabs(cos_theta_i);
  float cc = wclamp(0.0f, 1.0f, c);
  float sin2 = 1.0f - cc * cc;
  float ratio = eta_i / eta_t;
  float sin_t2 = ratio * ratio * sin2;
  float cos_t = std::sqrt(1.0f - sin_t2);
  float rpar = ((eta_t * c) - (eta_i * cos_t)) / ((eta_t * c) + (eta_i * cos_t));
  float rper = ((eta_i * c) - (eta_t * cos_t)) / ((eta_i * c) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) * 0.5f;
}
// strand_shading.wgsl:383-413
V3 A_term(uint32_t p, float h_, float eta, float eta_prime, V3 mu, float theta_d) {
  float h = wclamp(-1.0f + 1e-6f, 1.0f - 1e-6f, h_);
  float gamma_t = std::asin(h / eta_prime);
  float gamma_i = std::asin(h);
  float f = fresnel_reflectance(1.0f, eta, sacos(std::cos(theta_d) * std::cos(gamma_i)));
  float e = 1.0f + std::cos(2.0f * gamma_t);
  V3 T{std::exp(-2.0f * mu.x * e), std::exp(-2.0f * mu.y * e), std::exp(-2.0f * mu.z * e)};
  if (p == 0u) return V3{f * T.x * T.x, f * T.y * T.y, f * T.z * T.z};
  float f_ = 1.0f - f;
  float fp = std::pow(f, (float)(p - 1u));
  float s = (f_ * f_) * fp;
  return V3{s * std::pow(T.x, (float)p), s * std::pow(T.y, (float)p), s * std::pow(T.z, (float)p)};
}
// strand_shading.wgsl:440-454
float gaussian_detector(float beta, float phi_delta, uint32_t k_limit) {
  float var = beta * beta;
  if (var < 1e-6f) return std::fabs(std::fmod(phi_delta, TWO_PI_F)) < 1e-3f ? 0.0f : 1e6f;
  float norm = 1.0f / (SQRT_2PI_F * beta);
  float total = 0.0f;
  int ik = (int)k_limit;
  for (int k = -ik; k <= ik; ++k) {
    float ang = phi_delta - TWO_PI_F * (float)k;
    total = total + std::exp(-ang * ang / (2.0f * var));
  }
  return norm * total;
}
// strand_shading.wgsl:464-468
float Phi(uint32_t p, float h_, float eta_prime) {
  float h = wclamp(-1.0f + 1e-6f, 1.0f - 1e-6f, h_);
  float gamma_t = std::asin(wclamp(h / eta_prime, -1.0f, 1.0f));
  return 2.0f * (float)p * gamma_t - 2.0f * std::asin(h) + (float)p * PI_F;
}
// strand_shading.wgsl:470-523
float Mp(float v, float theta_i_, float theta_r, float shift) {
  float trs = theta_r + shift;
  float theta_i = wclamp(PI_HALF_LO, PI_HALF_HI, theta_i_);
  trs = wclamp(PI_HALF_LO, PI_HALF_HI, trs);
  float cti = std::cos(-theta_i), ctr = std::cos(trs);
  float sti = std::sin(-theta_i), str = std::sin(trs);
  if (v < 1e-6f) {
    const float tol = (float)(0.1 * kPI / 180.0);
    return std::fabs(trs + theta_i) < tol ? 0.0f : 1e6f / tol;
  }
  float barg = (cti * ctr) / v;
  float earg = (sti * str) / v;
  return bessel_first_approx(barg, 5u) * std::exp(earg);
}
// strand_shading.wgsl:525-562
V3 Np(uint32_t p, float phi, float theta_i, float theta_r, float eta, V3 mu, float beta_azim) {
  float theta_d = (theta_r - theta_i) * 0.5f;
  float sti = std::sin(theta_i);
  float ctd = std::cos(theta_d);
  float ep_arg = std::sqrt(std::fmax(0.0f, eta * eta - sti * sti));
  float eta_prime = ep_arg / ctd;
  V3 mup{mu.x / ctd, mu.y / ctd, mu.z / ctd};
  V3 acc{0.0f, 0.0f, 0.0f};
  for (uint32_t i = 0; i < 10u; ++i) {
    V3 att = A_term(p, LEG_ROOTS[i], eta, eta_prime, mup, theta_d);
    float sh = Phi(p, LEG_ROOTS[i], eta_prime);
    float det = gaussian_detector(beta_azim, phi - sh, 10u);
    acc.x = acc.x + att.x * det * LEG_WEIGHTS[i];
    acc.y = acc.y + att.y * det * LEG_WEIGHTS[i];
    acc.z = acc.z + att.z * det * LEG_WEIGHTS[i];
  }
  return v3scale(acc, 0.5f);
}
// strand_shading.wgsl:565-605 (marschner is called with direction = L,
// view_normal = V, light_normal = U, as at :686)
V3 marschner(V3 dir, V3 view_n, V3 light_n, const sr_strand_material& m, float occlusion) {
  float theta_i = sacos(v3dot(dir, light_n));
  float theta_r = sacos(v3dot(dir, view_n));
  V3 lp = v3norm(v3sub(light_n, v3scale(dir, v3dot(light_n, dir))));
  V3 vp = v3norm(v3sub(view_n, v3scale(dir, v3dot(view_n, dir))));
  float phi = sacos(v3dot(v3norm(lp), v3norm(vp))) * wsign(v3dot(v3cross(lp, vp), dir));
  V3 sigma{1.0f - m.absorption_color[0], 1.0f - m.absorption_color[1], 1.0f - m.absorption_color[2]};
  float v_long = m.alpha * m.alpha, v_azim = m.beta * m.beta;
  // weta_strand_bsdf
  float beta_azim = std::sqrt(v_azim);
  float scat = wclamp(1.0f - occlusion, 0.0f, 1.0f);
  float direct = wmix(0.18f, 1.0f, scat);
  float mp = Mp(v_long, theta_i, theta_r, m.shift);
  V3 total{mp * m.specular_color[0] * m.specular_color[3] * direct, mp * m.specular_color[1] * m.specular_color[3] * direct,
           mp * m.specular_color[2] * m.specular_color[3] * direct};
  for (uint32_t p = 0; p < 3u; ++p) {
    V3 n = Np(p, phi, theta_i, theta_r, m.eta, sigma, beta_azim);
    float s = ((float)(p + 1u) * scat) / 3.0f;
    total.x = total.x + n.x * s; total.y = total.y + n.y * s; total.z = total.z + n.z * s;
  }
  return total;
}

// strand_shading.wgsl:198-277 against the dense DOM (clamped >= 0.08)
float dom_visibility(const Frame& F, V3 pw, uint32_t layer) {
  uint32_t fi = kInvalid, l = 0;
  for (uint32_t i = 0; i < F.F; ++i) {
    if (F.fr[i].kind != 1u) continue;
    if (l == layer) { fi = i; break; }
    ++l;
  }
  if (fi == kInvalid) return 1.0f;
  const sr_frustum_desc& fd = F.fr[fi];
  float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
  V4 raw = world_to_screen_raw(V4{pw.x, pw.y, pw.z, 1.0f}, F.views[fi].clip_from_world, vw, vh);
  if (raw.x < 0.0f || raw.y < 0.0f || raw.x >= vw || raw.y >= vh || raw.z < 0.0f || raw.z > 1.0f) return 1.0f;
  uint32_t px = f2u(raw.x), py = f2u(raw.y);
  float z0 = F.dom_depth[F.dom_depth_base[fi] + (size_t)py * fd.screen_width + px];
  if (z0 <= 0.0f) return 1.0f;
  float z = normalize_depth01(raw.z);
  if (z > z0 - 1e-3f) return 1.0f;
  uint32_t L = F.st.dom_layers;
  float u = std::pow(wclamp(std::fmax(0.0f, z0 - z) / std::fmax(z0, 1e-6f), 0.0f, 1.0f), 1.0f);
  uint32_t s = std::min(f2u(std::floor(u * (float)L)), L - 1u);
  float op = F.dom_opacity[F.dom_opacity_base[fi] + (size_t)s * fd.screen_width * fd.screen_height +
                           (size_t)py * fd.screen_width + px];
  return wclamp(1.0f - op, 0.08f, 1.0f);
}

struct ShadeOut { float rgba[4]; float hist[4]; };

// strand_shading.wgsl:608-695 for one camera task
ShadeOut shade_task(const Frame& F, uint32_t t, uint32_t fi, const std::vector<float>& prev_hist, size_t hist_off) {
  const sr_binning_task& tk = F.tasks[t];
  const sr_strand_instance& in = F.inst[tk.id_info];
  const sr_asset_slabs* ms = slab(F, in.meta_id);
  const sr_strand_meta& sm = ms->meta[tk.chunk_id];
  SegWorld sw = seg_world(F, in, tk.seg_idx);
  V3 v0{sw.v0.x, sw.v0.y, sw.v0.z}, v1{sw.v1.x, sw.v1.y, sw.v1.z};
  V3 U = v3norm(v3sub(v1, v0));
  const float* wp = F.views[fi].world_position;
  V3 cam_dir = v3norm(v3sub(V3{wp[0], wp[1], wp[2]}, v0));
  V3 bin = v3norm(v3cross(U, cam_dir));
  V3 Vv = v3norm(v3cross(bin, U));
  const sr_strand_material& m = material_of(F, in, sm.material_idx);
  ShadeOut o;
  o.rgba[0] = o.rgba[1] = o.rgba[2] = 0.0f;
  o.hist[2] = o.hist[3] = 0.0f;
  o.rgba[3] = m.absorption_color[3];
  V3 mid{wmix(v0.x, v1.x, 0.5f), wmix(v0.y, v1.y, 0.5f), wmix(v0.z, v1.z, 0.5f)};
  uint32_t nl = std::min<uint32_t>((uint32_t)F.lights.size(), 4u);
  float cur = 1.0f;
  if (nl > 0) {
    float vis = 0.0f;
    for (uint32_t i = 0; i < nl; ++i) vis = vis + dom_visibility(F, mid, i);
    cur = vis / (float)nl;
  }
  float prev = F.st.shadow_history_valid ? prev_hist[hist_off] : 0.0f;
  float scat_vis = prev > 0.0f ? wmix(cur, prev, 0.88f) : cur;
  o.hist[0] = scat_vis; o.hist[1] = cur; o.hist[2] = 0.0f; o.hist[3] = 1.0f;
  float occ = 1.0f - scat_vis;
  for (uint32_t j = 0; j < F.lights.size(); ++j) {
    const sr_directional_light& lt = F.lights[j];
    V3 L = v3norm(V3{lt.direction_to_light[0], lt.direction_to_light[1], lt.direction_to_light[2]});
    V3 bc = marschner(L, Vv, U, m, occ);
    float amb = m.ambient_factor;
    float cx = bc.x * (lt.color[0] / 255.0f), cy = bc.y * (lt.color[1] / 255.0f), cz = bc.z * (lt.color[2] / 255.0f);
    float ax = m.absorption_color[0] * amb + (F.st.ambient_color[0] / 255.0f) * amb;
    float ay = m.absorption_color[1] * amb + (F.st.ambient_color[1] / 255.0f) * amb;
    float az = m.absorption_color[2] * amb + (F.st.ambient_color[2] / 255.0f) * amb;
    o.rgba[0] += wmix(cx, ax, amb);
    o.rgba[1] += wmix(cy, ay, amb);
    o.rgba[2] += wmix(cz, az, amb);
  }
  return o;
}

void shading(Frame& F) {
  // atlas per camera frustum x instance, texels = max index count over instances
  uint32_t max_idx = 1;
  for (auto& in : F.inst) {
    const sr_asset_slabs* is = slab(F, in.index_id);
    if (is) max_idx = std::max(max_idx, is->index_count);
  }
  uint32_t w = std::min(std::max(max_idx, 1u), SR_MAX_TEXTURE_EXTENT);
  uint32_t h = std::min(std::max(ceil_div(max_idx, w), 1u), SR_MAX_TEXTURE_EXTENT);
  F.atlas_texels = w * h;
  F.atlas_layer_base.assign(F.F, 0);
  uint32_t cams = 0;
  for (uint32_t fi = 0; fi < F.F; ++fi)
    if (F.fr[fi].kind == 0u) F.atlas_layer_base[fi] = (cams++) * F.NI;
  size_t layers = (size_t)cams * F.NI;
  F.atlas.assign(layers * F.atlas_texels * 4, 0.0f);
  F.history.assign(layers * F.atlas_texels * 2, 0.0f);
  // shadow_history_prev: the caller's previous history when its layout
  // matches, else cleared (frame 0)
  std::vector<float> prev_hist(layers * F.atlas_texels, 0.0f);
  if (F.prev_history && F.prev_history_texels == prev_hist.size())
    for (size_t i = 0; i < prev_hist.size(); ++i) prev_hist[i] = F.prev_history[2 * i];
  // segment starts that are shaded this frame, per layer
  std::vector<uint8_t> shaded(layers * F.atlas_texels, 0);
  for (uint32_t t : F.shading_queue) {
    const sr_binning_task& tk = F.tasks[t];
    uint32_t fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
    size_t layer = F.atlas_layer_base[fi] + tk.id_info;
    if (tk.seg_idx < F.atlas_texels) shaded[layer * F.atlas_texels + tk.seg_idx] = 1;
  }
  parallel_for(F.threads, F.shading_queue.size(), [&](size_t b, size_t e) {
    for (size_t q = b; q < e; ++q) {
      uint32_t t = F.shading_queue[q];
      const sr_binning_task& tk = F.tasks[t];
      uint32_t fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
      size_t layer = F.atlas_layer_base[fi] + tk.id_info;
      uint32_t idx = tk.seg_idx;
      if (idx >= F.atlas_texels) continue;
      ShadeOut o = shade_task(F, t, fi, prev_hist, layer * F.atlas_texels + idx);
      size_t a = (layer * F.atlas_texels + idx) * 4;
      F.history[(layer * F.atlas_texels + idx) * 2 + 0] = o.hist[0];
      F.history[(layer * F.atlas_texels + idx) * 2 + 1] = o.hist[1];
      // unorm store: values saturate to [0, 1]; the 8-bit rounding is not applied
      float col[4];
      for (int c = 0; c < 4; ++c) col[c] = wclamp(o.rgba[c], 0.0f, 1.0f);
      for (int c = 0; c < 4; ++c) F.atlas[a + c] = col[c];
      uint32_t nxt = idx + 1u;
      if (nxt < F.atlas_texels && !shaded[layer * F.atlas_texels + nxt]) {
        size_t an = (layer * F.atlas_texels + nxt) * 4;
        for (int c = 0; c < 4; ++c) F.atlas[an + c] = col[c];
      }
    }
  });
}

// ------------------------------------------------------ K28 camera raster
// strand_rasterizer.wgsl:855-1044: per pixel, work items near to far, refs in
// cell order, premultiplied front-to-back over, early out at alpha > 0.98
// after a work item.  Every pixel of a tile is rasterised.
void camera_raster(Frame& F) {
  const bool sorted = (F.st.flags & SR_FRAME_DEPTH_SORTED_BLEND) != 0u;
  F.image_base.assign(F.F, 0);
  uint64_t sz = 0;
  for (uint32_t fi = 0; fi < F.F; ++fi)
    if (F.fr[fi].kind == 0u) { F.image_base[fi] = sz; sz += (uint64_t)F.fr[fi].screen_width * F.fr[fi].screen_height; }
  F.image.assign(sz * 4, 0.0f);
  F.depth.assign(sz, 0.0f);
  parallel_for(F.threads, F.cam_runs.size(), [&](size_t rb, size_t re) {
    struct Rec { bool valid; V2 p0, ab; float inv_l2; V2 dep, cw, rad, mn, mx; float s0[4], s1[4]; float vis; };
    std::vector<Rec> recs;
    std::vector<uint32_t> wstart;
    for (size_t ri = rb; ri < re; ++ri) {
      const sr_raster_tile_run& run = F.cam_runs[ri];
      uint32_t fi = run.frustum_id;
      const sr_frustum_desc& fd = F.fr[fi];
      if (fd.kind != 0u || run.work_count == 0u) continue;
      const float* cfw = F.views[fi].clip_from_world;
      float vw = (float)fd.screen_width, vh = (float)fd.screen_height;
      uint32_t ftx = ceil_div(fd.screen_width, fd.froxel_size_x);
      uint32_t tx = run.screen_tile_id % ftx, ty = run.screen_tile_id / ftx;
      recs.clear(); wstart.assign(run.work_count + 1, 0);
      for (uint32_t w = 0; w < run.work_count; ++w) {
        const sr_raster_work_item& wi = F.work[run.work_base + w];
        wstart[w] = (uint32_t)recs.size();
        for (uint32_t r = 0; r < wi.seg_ref_count; ++r) {
          const sr_fine_seg_ref& sr = F.refs[wi.seg_ref_base + r];
          const sr_strand_instance& in = F.inst[sr.inst_id];
          SegWorld sw = seg_world(F, in, sr.seg_id);
          V4 a = world_to_screen_raw(sw.v0, cfw, vw, vh), b = world_to_screen_raw(sw.v1, cfw, vw, vh);
          Rec rc{};
          rc.valid = !(a.x < 0.0f && b.x < 0.0f);
          if (rc.valid) {
            const sr_strand_material& m = material_of(F, in, sr.material_idx);
            float minr = std::fmax(m.min_radius_pixels, 1e-4f);
            float maxr = std::fmax(m.max_radius_pixels, minr);
            size_t layer = F.atlas_layer_base[fi] + sr.inst_id;
            uint32_t i0 = sr.seg_id, i1 = std::min(i0 + 1u, F.atlas_texels - 1u);
            rc.p0 = V2{a.x, a.y};
            rc.ab = V2{b.x - a.x, b.y - a.y};
            rc.inv_l2 = 1.0f / (rc.ab.x * rc.ab.x + rc.ab.y * rc.ab.y);
            rc.dep = V2{a.z, b.z};
            rc.cw = V2{a.w, b.w};
            rc.rad = V2{minr, maxr};
            rc.mn = V2{std::fmin(a.x, b.x) - maxr, std::fmin(a.y, b.y) - maxr};
            rc.mx = V2{std::fmax(a.x, b.x) + maxr, std::fmax(a.y, b.y) + maxr};
            for (int c = 0; c < 4; ++c) {
              rc.s0[c] = F.atlas[((layer * F.atlas_texels) + i0) * 4 + c];
              rc.s1[c] = F.atlas[((layer * F.atlas_texels) + i1) * 4 + c];
            }
            rc.vis = F.history[((layer * F.atlas_texels) + i0) * 2 + 0];
          }
          recs.push_back(rc);
        }
      }
      wstart[run.work_count] = (uint32_t)recs.size();
      uint32_t npx = fd.froxel_size_x * fd.froxel_size_y;
      for (uint32_t pidx = 0; pidx < npx; ++pidx) {
        uint32_t pxu = tx * fd.froxel_size_x + pidx % fd.froxel_size_x;
        uint32_t pyu = ty * fd.froxel_size_y + pidx / fd.froxel_size_x;
        if (pxu >= fd.screen_width || pyu >= fd.screen_height) continue;
        float cx = (float)pxu + 0.5f, cy = (float)pyu + 0.5f;
        float fc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float gdepth = 0.0f;
        bool done = false;
        // final = blend_over_premult(final, src), strand_rasterizer.wgsl:156-162
        auto blend = [&fc](const float* src) {
          float om = 1.0f - fc[3];
          for (int c = 0; c < 3; ++c) fc[c] = fc[c] + src[c] * om;
          fc[3] = fc[3] + src[3] * om;
        };
        // SR_FRAME_DEPTH_SORTED_BLEND: the work item's fragments nearest
        // first (larger depth; equal depths in arrival order), at most
        // SR_BLEND_SORT_FRAGMENTS held; one more blends the nearest held
        struct Frag { float z, c[4]; };
        std::vector<Frag> held;
        for (uint32_t w = 0; w < run.work_count && !done; ++w) {
          for (uint32_t k = wstart[w]; k < wstart[w + 1]; ++k) {
            const Rec& rc = recs[k];
            if (!rc.valid) continue;
            if (cx < rc.mn.x || cy < rc.mn.y || cx > rc.mx.x || cy > rc.mx.y) continue;
            float t = ((cx - rc.p0.x) * rc.ab.x + (cy - rc.p0.y) * rc.ab.y) * rc.inv_l2;
            if (!(t >= 0.0f && t <= 1.0f)) continue;
            float fx = rc.p0.x + rc.ab.x * t, fy = rc.p0.y + rc.ab.y * t;
            float fz = wmix(rc.dep.x, rc.dep.y, t);
            float minr = std::fmax(rc.rad.x, 1e-4f);
            float r = wmix(minr, std::fmax(rc.rad.y, minr), wclamp(normalize_depth01(fz), 0.0f, 1.0f));
            float dx = cx - fx, dy = cy - fy;
            float d2 = dx * dx + dy * dy;
            if (!(d2 < r * r)) continue;
            float cov = wclamp(1.0f - std::sqrt(d2) / r, 0.0f, 1.0f);
            // perspective_correct_line_t, strand_rasterizer.wgsl:149-154
            float aw0 = std::fmax(std::fabs(rc.cw.x), 1e-6f), aw1 = std::fmax(std::fabs(rc.cw.y), 1e-6f);
            float tw = (t * aw0) / ((1.0f - t) * aw1 + t * aw0);
            float sh[4];
            for (int c = 0; c < 4; ++c) sh[c] = wmix(rc.s0[c], rc.s1[c], tw);
            float alpha = sh[3] * cov;
            float src[4] = {sh[0] * rc.vis * alpha, sh[1] * rc.vis * alpha, sh[2] * rc.vis * alpha, alpha};
            if (sorted) {
              Frag f{fz, {src[0], src[1], src[2], src[3]}};
              size_t at = held.size();  // ahead of every farther fragment
              while (at > 0 && fz > held[at - 1].z) --at;
              held.insert(held.begin() + (long)at, f);
              if (held.size() > SR_BLEND_SORT_FRAGMENTS) {
                blend(held.front().c);
                held.erase(held.begin());
              }
            } else {
              blend(src);
            }
            gdepth = std::fmax(gdepth, fz);
          }
          for (const Frag& f : held) blend(f.c);
          held.clear();
          if (fc[3] > 0.98f) {
            for (int c = 0; c < 3; ++c) fc[c] = fc[c] / fc[3];
            fc[3] = 1.0f;
            done = true;
          }
        }
        float out[4] = {fc[0], fc[1], fc[2], fc[3]};
        if (out[3] > 1e-6f && out[3] < 1.0f)
          for (int c = 0; c < 3; ++c) out[c] = out[c] / out[3];
        size_t o = F.image_base[fi] + (size_t)pyu * fd.screen_width + pxu;
        for (int c = 0; c < 4; ++c) F.image[o * 4 + c] = wclamp(out[c], 0.0f, 1.0f);  // unorm store
        F.depth[o] = gdepth;
      }
    }
  });
}

// --------------------------------------------------------- K29 composite
// strand_composite.wgsl:12-33 over a constant headless scene
void composite(Frame& F) {
  F.composite.assign(F.image.size(), 0.0f);
  const float* sc = F.st.scene_clear_color;
  float sd = F.st.scene_clear_depth;
  for (size_t i = 0; i < F.depth.size(); ++i) {
    float* o = &F.composite[i * 4];
    if (sd > F.depth[i]) { for (int c = 0; c < 4; ++c) o[c] = sc[c]; continue; }
    float a = F.image[i * 4 + 3];
    for (int c = 0; c < 3; ++c) o[c] = wmix(sc[c], F.image[i * 4 + c], a);
    o[3] = sc[3];
  }
}

template <class Fn>
void timed(Frame& F, const char* name, Fn fn) {
  auto t0 = std::chrono::steady_clock::now();
  fn();
  auto t1 = std::chrono::steady_clock::now();
  F.stage_names.push_back(name);
  F.stage_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
}

}  // namespace

extern "C" {

struct or_frame { Frame f; };

int oracle_render_history(const sr_asset_slabs* assets, uint32_t asset_count, const sr_frame_desc* fd, int threads,
                          uint32_t ref_capacity, int stop_after_binning, const float* prev_history,
                          uint64_t prev_history_texels, or_frame** out);

int oracle_render(const sr_asset_slabs* assets, uint32_t asset_count, const sr_frame_desc* fd, int threads,
                  uint32_t ref_capacity, int stop_after_binning, or_frame** out) {
  return oracle_render_history(assets, asset_count, fd, threads, ref_capacity, stop_after_binning, nullptr, 0, out);
}

// One frame whose shading reads `prev_history` (float2 per atlas texel, the
// previous frame's history dump) as shadow_history_prev
// (strand_shading.wgsl:660-679).
int oracle_render_history(const sr_asset_slabs* assets, uint32_t asset_count, const sr_frame_desc* fd, int threads,
                          uint32_t ref_capacity, int stop_after_binning, const float* prev_history,
                          uint64_t prev_history_texels, or_frame** out) {
  if (!assets || !fd || !out) return -1;
  or_frame* of = new or_frame();
  Frame& F = of->f;
  F.prev_history = prev_history;
  F.prev_history_texels = prev_history_texels;
  if (fd->opaque_depth) {
    if (fd->opaque_samples == 0u || fd->opaque_samples > 4u) { delete of; return -1; }
    F.opaque_depth = fd->opaque_depth;
    F.opaque_samples = fd->opaque_samples;
  }
  F.assets = assets; F.asset_count = asset_count;
  F.inst.assign(fd->instances, fd->instances + fd->instance_count);
  F.fr.assign(fd->frusta, fd->frusta + fd->frustum_count);
  F.views.assign(fd->views, fd->views + fd->frustum_count);
  if (fd->light_count) F.lights.assign(fd->lights, fd->lights + fd->light_count);
  F.st = fd->settings;
  if (F.st.dom_layers == 0u || F.st.dom_layers > SR_MAX_DOM_LAYERS) { delete of; return -1; }
  {
    uint32_t S = F.st.coarse_depth_slices;
    if (S != 0u && S != 8u && S != 16u && S != 32u) { delete of; return -1; }
  }
  F.threads = std::max(1, threads);
  F.F = fd->frustum_count; F.NI = fd->instance_count;
  F.coarse_total = 0;
  for (auto& f : F.fr) F.coarse_total = std::max(F.coarse_total, f.coarse_depth_tile_base + f.coarse_depth_tile_count);
  F.ref_capacity = ref_capacity;
  timed(F, "opaque_reduce", [&] { reduce_opaque_depth(F); });
  timed(F, "broad", [&] { broad_prepass(F); });
  timed(F, "project", [&] { fine_prepass(F); });
  timed(F, "depth_lut", [&] { build_luts(F); });
  timed(F, "mark_count", [&] { bin_count(F); });
  timed(F, "prefix", [&] { prefix_pages(F); });
  timed(F, "fill", [&] { bin_fill(F); });
  timed(F, "compact", [&] { compact_active(F); });
  timed(F, "emit", [&] { emit_work(F); });
  F.ctr.binning_tasks = (uint32_t)F.tasks.size();
  F.ctr.camera_tasks = (uint32_t)F.shading_queue.size();
  F.ctr.coarse_tiles = F.coarse_total;
  F.ctr.ref_capacity = ref_capacity;
  if (!stop_after_binning) {
    timed(F, "shadow", [&] { shadow_raster(F); });
    if (F.st.flags & SR_FRAME_VIRTUAL_DOMS) timed(F, "dom_pages", [&] { dom_pages(F); });
    timed(F, "stampback", [&] { stampback(F); });
    timed(F, "shading", [&] { shading(F); });
    timed(F, "raster", [&] { camera_raster(F); });
    timed(F, "composite", [&] { composite(F); });
  }
  *out = of;
  return 0;
}

int oracle_get(or_frame* of, int kind, const void** data, uint64_t* bytes) {
  if (!of || !data || !bytes) return -1;
  Frame& F = of->f;
#define RET(vec) do { *data = (vec).data(); *bytes = (uint64_t)(vec).size() * sizeof((vec)[0]); return 0; } while (0)
  switch (kind) {
    case SR_DUMP_COUNTERS: *data = &F.ctr; *bytes = sizeof(F.ctr); return 0;
    case SR_DUMP_BROAD_META: RET(F.broad);
    case SR_DUMP_KEEP: RET(F.keep);
    case SR_DUMP_RANGES: RET(F.ranges);
    case SR_DUMP_TASKS: RET(F.tasks);
    case SR_DUMP_PROJECTED: RET(F.proj);
    case SR_DUMP_SHADING_QUEUE: RET(F.shading_queue);
    case SR_DUMP_LUT: RET(F.lut);
    case SR_DUMP_PAGE_TABLE: RET(F.page_table);
    case SR_DUMP_PAGE_META: RET(F.page_meta);
    case SR_DUMP_CELL_COUNTS: RET(F.cell_counts);
    case SR_DUMP_CELL_OFFSETS: RET(F.cell_offsets);
    case SR_DUMP_REFS: RET(F.refs);
    case SR_DUMP_ACTIVE_TILES: RET(F.active_queue);
    case SR_DUMP_WORK_ITEMS: RET(F.work);
    case SR_DUMP_CAMERA_RUNS: RET(F.cam_runs);
    case SR_DUMP_SHADOW_RUNS: RET(F.shadow_runs);
    case SR_DUMP_DOM_DEPTH: RET(F.dom_depth);
    case SR_DUMP_DOM_OPACITY: RET(F.dom_opacity);
    case SR_DUMP_ATLAS: RET(F.atlas);
    case SR_DUMP_HISTORY: RET(F.history);
    case SR_DUMP_OPAQUE_TILES: RET(F.opaque_tiles);
    case SR_DUMP_IMAGE: RET(F.image);
    case SR_DUMP_DEPTH: RET(F.depth);
    case SR_DUMP_COMPOSITE: RET(F.composite);
    case SR_DUMP_STAMP: RET(F.stamp);
    case SR_DUMP_DOM_REQUESTS: RET(F.dom_requests);
    case SR_DUMP_DOM_PAGE_TABLE: RET(F.dom_table);
    default: return -1;
  }
#undef RET
}

int oracle_stage_times(or_frame* of, double* ms, int cap, char* names, int names_cap) {
  if (!of) return -1;
  Frame& F = of->f;
  int n = (int)F.stage_ms.size();
  std::string joined;
  for (int i = 0; i < n; ++i) {
    if (i < cap && ms) ms[i] = F.stage_ms[i];
    joined += F.stage_names[i];
    joined += '\n';
  }
  if (names && names_cap > 0) {
    std::strncpy(names, joined.c_str(), (size_t)names_cap - 1);
    names[names_cap - 1] = 0;
  }
  return n;
}

void oracle_free(or_frame* of) { delete of; }

// Primitive entry points for the pin tests.
uint32_t oracle_wang_hash(uint32_t x) { return wang_hash(x); }
float oracle_hash_to_unit_float(uint32_t x) { return hash_to_unit_float(x); }
uint32_t oracle_pack2x16unorm(float a, float b) { return pack2x16unorm(a, b); }
uint32_t oracle_quantize_depth01(float z) { return quantize_depth01(z); }
float oracle_normalize_depth01(float z) { return normalize_depth01(z); }
void oracle_build_lut(const uint32_t* zmin_zmax_newest_first, uint32_t n, sr_coarse_depth_lut_entry* out,
                      uint32_t slices) {
  std::vector<std::pair<uint32_t, uint32_t>> iv;
  for (uint32_t i = 0; i < n; ++i) iv.push_back({zmin_zmax_newest_first[2 * i], zmin_zmax_newest_first[2 * i + 1]});
  build_lut_for_tile(iv, out, slices ? slices : kSlices);
}
int oracle_clip_segment(const float* p0, const float* p1, const float* bmin, const float* bmax, float* out6) {
  Clipped c = clip_segment_to_box(V3{p0[0], p0[1], p0[2]}, V3{p1[0], p1[1], p1[2]}, V3{bmin[0], bmin[1], bmin[2]},
                                  V3{bmax[0], bmax[1], bmax[2]});
  out6[0] = c.p0.x; out6[1] = c.p0.y; out6[2] = c.p0.z; out6[3] = c.p1.x; out6[4] = c.p1.y; out6[5] = c.p1.z;
  return c.ok ? 1 : 0;
}
float oracle_gaussian_detector(float beta, float phi, uint32_t k) { return gaussian_detector(beta, phi, k); }
float oracle_fresnel(float eta_i, float eta_t, float c) { return fresnel_reflectance(eta_i, eta_t, c); }
float oracle_bessel(float x, uint32_t steps) { return bessel_first_approx(x, steps); }
float oracle_mp(float v, float ti, float tr, float s) { return Mp(v, ti, tr, s); }
void oracle_np(uint32_t p, float phi, float ti, float tr, float eta, const float* mu, float beta, float* out3) {
  V3 r = Np(p, phi, ti, tr, eta, V3{mu[0], mu[1], mu[2]}, beta);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float oracle_phi(uint32_t p, float h, float eta_prime) { return Phi(p, h, eta_prime); }
void oracle_a(uint32_t p, float h, float eta, float eta_prime, const float* mu, float theta_d, float* out3) {
  V3 r = A_term(p, h, eta, eta_prime, V3{mu[0], mu[1], mu[2]}, theta_d);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
// marschner(direction, view_normal, light_normal, material, occlusion)
void oracle_marschner(const float* dir3, const float* view3, const float* light3, const sr_strand_material* m,
                      float occlusion, float* out3) {
  V3 r = marschner(V3{dir3[0], dir3[1], dir3[2]}, V3{view3[0], view3[1], view3[2]},
                   V3{light3[0], light3[1], light3[2]}, *m, occlusion);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

}  // extern "C"
