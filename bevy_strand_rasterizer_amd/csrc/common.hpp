// common.hpp -- device-side definitions shared by the strand kernels.
//
// The arithmetic helpers reproduce WGSL semantics exactly: f32 only, no FMA
// contraction (every translation unit is built with -ffp-contract=off),
// correctly rounded division, ties-to-even round, saturating f32 -> int
// conversions, clamp(e, lo, hi) = min(max(e, lo), hi).  Binning results
// (counts, offsets, per-cell reference lists) are therefore bit-identical
// to any implementation that follows the same rules on the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/strand_rt.h"

#define SR_HD __host__ __device__ __forceinline__

namespace srk {

constexpr uint32_t kExt = SR_COARSE_FINE_TILE_EXTENT;   // 4 fine tiles per coarse tile edge
constexpr uint32_t kSlices = SR_COARSE_DEPTH_SLICES;    // 16 fine depth slices per page (and default LUT slices)
constexpr uint32_t kMaxSlices = SR_MAX_COARSE_DEPTH_SLICES;  // 32: most depth-warp LUT slices per coarse tile
constexpr uint32_t kPage = SR_COARSE_COUNT_PAGE_SIZE;   // 256 cells per count page
constexpr uint32_t kQuant = SR_DEPTH_QUANT_MAX;
constexpr uint32_t kInvalid = SR_INVALID_PTR;
constexpr uint32_t kMarked = 0xFFFFFFFEu;
constexpr uint32_t kOrderBlocks = 64u;  // blocks per run kind of the raster run-order sort
constexpr uint32_t kDomPage = SR_DOM_PAGE_XY;                  // virtual DOM page edge, texels
constexpr uint32_t kDomPageTexels = kDomPage * kDomPage;
constexpr uint32_t kDomBorder = SR_DOM_PREFETCH_PAGE_BORDER;    // neighbouring pages each request adds

// ------------------------------------------------------------- records
// Per-frustum record built on the host for every frame.
struct FrustumRec {
  sr_frustum_desc d;
  float clip[16];
  float eye[4];
  uint32_t light_layer;    // kInvalid unless a light with a bound light slot
  uint32_t cam_index;      // kInvalid unless a camera
  uint32_t enabled;        // produces binning tasks
  uint32_t fine_tiles_x;
  uint32_t fine_tiles_y;
  uint32_t stamp_w;
  uint32_t stamp_h;
  uint32_t dom_pages_x;      // virtual DOMs: pages per row of this light's DOM
  uint32_t dom_pages_y;
  uint32_t dom_vpage_base;   // first entry of this light's pages in the request and page tables
  uint64_t dom_depth_off;    // floats (dense DOMs)
  uint64_t dom_opacity_off;  // floats (dense DOMs)
  uint64_t image_off;        // pixels
  uint64_t stamp_off;        // floats
};

// Per-instance record: the instance's slab pointers resolved from its
// DevicePtr ids, plus its transform.
struct InstRec {
  const float4* vertices;
  const uint32_t* indices;
  const sr_strand_meta* meta;
  const sr_strand_material* materials;
  const uint32_t* pos_strand;  // index position -> strand, kInvalid if not a segment start
  uint32_t vertex_count;
  uint32_t index_count;
  uint32_t meta_count;
  uint32_t material_count;
  uint32_t geo_strand_count;
  uint32_t pos_base;  // first slot of this instance inside a frustum's slot range
  uint32_t valid;
  uint32_t pad;
  sr_aabb aabb;
  float world_from_local[16];
};

// Telemetry words (strand_prepass.wgsl:72-75, prepass.rs:1345-1396).
constexpr uint32_t kTelHistBins = 32, kTelHistOff = 8, kTelPageOff = kTelHistOff + kTelHistBins, kTelFrustumStride = 11;
enum TelFrustumField : uint32_t {
  TF_VISIBLE_INSTANCES, TF_VISIBLE_STRANDS, TF_RETAINED_STRANDS, TF_EMITTED_SEGMENTS, TF_ALLOCATED_PAGES,
  TF_PAGE_CANDIDATES, TF_FINE_REFS, TF_RASTER_RUNS, TF_RASTER_LOAD, TF_MAX_PAGE_CANDIDATES, TF_MAX_RASTER_LOAD
};

// Device counters of one frame (uint32 words).
enum Counter : uint32_t {
  C_TASKS = 0,
  C_CAM_TASKS = 1,
  C_PAGES = 2,
  C_REFS = 3,
  C_ACTIVE = 4,
  C_WORK = 5,
  C_CAM_RUNS = 6,
  C_SHADOW_RUNS = 7,
  C_OVERFLOW = 8,
  C_MARKED = 9,       // marked pages before the page capacity is applied
  C_PAGE_DROPS = 10,  // marked pages dropped by the page capacity
  C_CAMERA_HEAD = 11, // camera raster run queue head
  C_SHADOW_HEAD = 12, // shadow raster run queue head
  C_CANDS = 13,       // fine-cell hits of the page CSR backend
  C_DOM_PAGES = 14,   // resident virtual DOM pages
  C_DOM_WANT = 15,    // requested pages within every light's budget (resident with an unbounded pool)
  C_DOM_DROPS = 16,   // of those, pages past the pool capacity (not resident)
  C_SORT_PAIR = 17,   // page CSR: the (key, slot) buffer pair holding the page-sorted hits
  C_COUNT = 20
};

struct FrameParams {
  const FrustumRec* fr;
  const InstRec* inst;
  const sr_directional_light* lights;
  uint32_t F, NI, NL, cams;
  uint32_t slots;         // slot count per frustum (multiple of 64)
  uint32_t coarse_total;  // coarse tiles over all frusta
  uint32_t task_cap, ref_cap, work_cap, atlas_texels;
  uint32_t dom_layers, page_cap;
  uint32_t mark_waves;     // waves per workgroup of the projection / marking pass (1, 2 or 4)
  uint32_t raster_waves;   // waves per CU of the persistent rasterisers (0 = 16)
  uint32_t slices, slice_shift;  // depth-warp LUT slices per coarse tile (8, 16 or 32) and their log2
  sr_frame_settings st;

  uint32_t* counters;  // Counter words
  // broad
  sr_broad_instance_meta* broad;
  float* keep;                       // [NI * F]
  sr_coarse_asset_range* ranges;     // [NI * F]
  // projection
  unsigned long long* slot_masks;    // [F * slots / 64]
  uint32_t* wave_offsets;            // [F * slots / 64]
  uint32_t* frustum_task_base;       // [F + 1]
  uint32_t* frustum_cam_base;        // [F]
  uint8_t* shaded;                   // [cams * slots] camera segment starts shaded this frame
  sr_binning_task* tasks;            // [task_cap]
  sr_projected_segment* proj;        // [task_cap]
  uint32_t* shading_queue;           // [task_cap]
  // binning
  sr_coarse_depth_lut_entry* lut;    // [coarse_total * slices]
  uint2* lut_z;                      // [coarse_total * slices] (z_min_q, z_max_q) per slice; unused: (~0u, 0)
  uint2* lut_iv;                     // [coarse_total] a tile's one depth interval spanning >= slices quantised
                                     // depths (its slices split it evenly), else (1, 0)
  uint32_t* page_table;              // [coarse_total * slices]
  sr_fine_page_meta* page_meta;      // [page_cap]
  uint32_t* page_totals;             // [page_cap]
  uint32_t* cell_counts;             // [pages * 256]
  uint32_t* cell_offsets;            // [pages * 256]
  uint32_t* cell_cursors;            // [pages * 256]
  sr_fine_seg_ref* refs;             // [ref_cap]
  uint32_t* active_flags;            // [coarse_total * 16]
  uint32_t* active_queue;            // [coarse_total * 16]
  uint4* tile_work;                  // [coarse_total * 16]: (work count, cam run, shadow run, load)
  uint4* tile_bases;                 // [coarse_total * 16]: exclusive scans of tile_work
  sr_raster_work_item* work;         // [work_cap]
  sr_raster_tile_run* cam_runs;      // [coarse_total * 16]
  sr_raster_tile_run* shadow_runs;   // [coarse_total * 16]
  float4* slot_screen;               // [F * slots][2] raw screen x, y, z, w of each valid slot's endpoints
  uint32_t* cam_order;               // [coarse_total * 16] camera runs, heaviest first
  uint32_t* shadow_order;            // [coarse_total * 16] shadow runs, heaviest first
  uint32_t* run_hist;                // [2 * kOrderBlocks * 32] run-order bucket counts
  uint32_t* page_bounds;             // [page_cap + 1] first sorted hit of each page (CSR)
  uint32_t* page_order;              // [page_cap] pages, heaviest first (CSR)
  uint32_t* page_hist;               // [kOrderBlocks * 32] page-order bucket counts
  // page CSR backend
  // page CSR backend: every (page, cell, slot) hit is written in slot order
  // at scanned offsets, then one stable radix sort by (page, cell) puts the
  // references in canonical order; cell runs of the sorted keys give the
  // cell counts and offsets
  uint32_t* slot_hit_total;          // [F * slots] visible fine cells of each slot (0 when invalid)
  uint32_t* hit_base;                // [F * slots] exclusive scan of slot_hit_total
  uint32_t* slot_material;           // [slots] material index of the segment at each index position
  uint32_t* hkeys_w;                 // [ref_cap] page * 256 + cell of each hit as written
  uint32_t* hvals_w;                 // [ref_cap] index position of each hit's segment, as written
  const uint32_t* hkeys;             // [ref_cap] keys after the page sort
  uint32_t* page_xy_mask;            // [page_cap] fine tiles with work per page
  // fine cells each slot's segment hits, recorded while marking (CSR backend):
  // slot_hits[j * F * slots + s] = page-table entry << 8 | cell, in traversal
  // order (slot-minor, so a wave's slots touch consecutive words);
  // slot_hit_count[s] = count, or kHitOverflow past hit_k (those slots are
  // traced again).  hit_k = 0 disables recording.
  // telemetry (finalize_telemetry layout): 8 global words, 32 histogram
  // bins, one candidate count per page, 11 words per frustum; null when off
  uint32_t* tel;
  // opaque scene depth (per camera, opaque_samples per pixel) and its
  // fine-tile reduction (covered, quantised farthest); null when absent
  const float* opaque_depth;
  uint2* opaque_tiles;
  uint32_t opaque_samples, fine_total;
  const sr_projected_segment* slot_proj;  // [F * slots]
  uint32_t* slot_hits;
  uint8_t* slot_hit_count;
  uint32_t hit_k;
  uint32_t mark_queue;  // fine walks a wave queues while marking (<= kMarkQ)
  uint32_t tile_queue;  // coarse tiles a wave queues per marking round (1 .. kTileQ)
  // shadow raster: span columns of each workgroup's first batches, kept
  // from its first pass for the second (kShadowColBytes per frame)
  unsigned long long* shadow_cols;
  // shadow / shading / raster.  Dense DOMs: per light frustum W * H depths
  // and layer-major L * W * H opacities at its offsets.  Virtual DOMs
  // (dom_table set): a pool of pages, depth texel t of physical page e at
  // dom_depth[e * kDomPageTexels + t], opacity layer i at
  // dom_opacity[(e * L + i) * kDomPageTexels + t].
  float* dom_depth;
  float* dom_opacity;
  uint32_t* dom_requests;  // [dom_vpages] 1 = requested this frame, or null (dense DOMs)
  uint32_t* dom_table;     // [dom_vpages] physical page or kInvalid, or null (dense DOMs)
  uint32_t dom_vpages;     // pages over every light frustum
  uint32_t dom_pool_cap;   // physical pages the pool holds
  float* stamp;
  float4* atlas;      // [cams * NI * atlas_texels]
  float2* history;    // [cams * NI * atlas_texels] (visibility, current visibility)
  const float2* history_prev;
  float4* image;
  float* depth;
  float4* composite;
  uint32_t* image8;      // rgba8unorm image, or null
  uint32_t* composite8;  // rgba8unorm composite, or null
};

// ------------------------------------------------------------- vectors
struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

SR_HD float wclamp(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }
SR_HD float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
SR_HD int32_t iclamp(int32_t e, int32_t lo, int32_t hi) { return min(max(e, lo), hi); }
SR_HD uint32_t ceil_div(uint32_t x, uint32_t y) { return (x + y - 1u) / y; }

SR_HD uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
SR_HD int32_t f2i(float x) {
  if (x != x) return 0;
  if (x <= -2147483648.0f) return INT32_MIN;
  if (x >= 2147483648.0f) return INT32_MAX;
  return (int32_t)x;
}

// mat4 (column-major) x vec4 summed as ((c0*x + c1*y) + c2*z) + c3*w
SR_HD V4 mat_vec(const float* m, V4 v) {
  V4 r;
  r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
  r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
  r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
  r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
  return r;
}

SR_HD uint32_t pack2x16unorm(float a, float b) {
  uint32_t lo = (uint32_t)floorf(0.5f + 65535.0f * wclamp(a, 0.0f, 1.0f));
  uint32_t hi = (uint32_t)floorf(0.5f + 65535.0f * wclamp(b, 0.0f, 1.0f));
  return lo | (hi << 16);
}
// i / 65535 for a 16-bit i: the product with the rounded reciprocal,
// corrected once by an FMA, equals the correctly rounded quotient for every
// i (tests/unorm_check.cpp checks all 65,536) at a quarter of the cost of
// the division sequence.
SR_HD float unorm16_to_float(uint32_t i) {
  const float x = (float)i, r = 1.0f / 65535.0f;
  const float q = x * r;
  return fmaf(fmaf(-q, 65535.0f, x), r, q);
}
SR_HD V2 unpack2x16unorm(uint32_t v) { return V2{unorm16_to_float(v & 0xFFFFu), unorm16_to_float(v >> 16)}; }

SR_HD uint32_t wang_hash(uint32_t seed) {
  uint32_t x = seed;
  x = (x ^ 61u) ^ (x >> 16u);
  x = x + (x << 3u);
  x = x ^ (x >> 4u);
  x = x * 0x27d4eb2du;
  x = x ^ (x >> 15u);
  return x;
}
SR_HD float hash_to_unit_float(uint32_t x) { return (float)x / 4294967296.0f; }

SR_HD float normalize_depth01(float z) {
  if (z >= 0.0f && z <= 1.0f) return z;
  return wclamp(z * 0.5f + 0.5f, 0.0f, 1.0f);
}
SR_HD uint32_t quantize_depth01(float z) { return f2u(rintf(wclamp(z, 0.0f, 1.0f) * (float)kQuant)); }
// q / 16777215 as the product with the rounded reciprocal corrected once by
// an FMA: the correctly rounded quotient for every q < 2^24
// (tests/unorm_check.cpp checks them all), and so for every float (float)q,
// whose quotients are those of a 24-bit significand scaled by a power of two.
SR_HD float dequantize_depth01(uint32_t q) {
  const float x = (float)q, r = 1.0f / (float)kQuant;
  const float p = x * r;
  return fmaf(fmaf(-p, (float)kQuant, x), r, p);
}

// The coarse depth slice of quantised depth q in a tile whose one depth
// interval [zmin, zmax] (span = zmax - zmin + 1 >= S, S = 2^shift slices)
// build_depth_warp_lut splits into S slices
// [zmin + span * i / S, zmin + span * (i + 1) / S - 1]: the last i whose
// start is <= q (zmin <= q <= zmax), i.e. floor((S d + S - 1) / span) for
// d = q - zmin.  rcp_span is 1 / span within a few ulp; the float estimate
// of that quotient is then within one of it (relative error near 2^-22 on a
// quotient below S <= 32) and is settled on the exact integer starts
// (tests/even_slice_check.cpp compares it with the slice scan).
SR_HD uint32_t even_slice(uint32_t zmin, uint32_t zmax, uint32_t q, float rcp_span, uint32_t shift) {
  const uint32_t S = 1u << shift, span = zmax - zmin + 1u, d = q - zmin;
  uint32_t c = (uint32_t)((float)((d << shift) + S - 1u) * rcp_span);
  c = c < S - 1u ? c : S - 1u;
  if (c > 0u && ((span * c) >> shift) > d) --c;
  else if (c + 1u < S && ((span * (c + 1u)) >> shift) <= d) ++c;
  return c;
}

SR_HD V4 world_to_screen_raw(V4 pos, const float* clip_from_world, float vw, float vh) {
  V4 clip = mat_vec(clip_from_world, pos);
  if (clip.w <= 0.0f) return V4{-1.0f, -1.0f, -1.0f, -1.0f};
  float nx = clip.x / clip.w, ny = clip.y / clip.w, nz = clip.z / clip.w;
  float sx = 0.0f + (nx * 0.5f + 0.5f) * vw;
  float sy = 0.0f + (ny * -0.5f + 0.5f) * vh;
  return V4{sx, sy, nz, clip.w};
}

SR_HD int32_t sgn_i32(float f) { return f > 1e-6f ? 1 : (f < -1e-6f ? -1 : 0); }

struct Clipped { bool ok; V3 p0, p1; };

// A segment's direction and its componentwise reciprocal, computed once for
// every box the segment is clipped against (the reciprocal of a component
// below 1e-6 in magnitude is never used).
struct SegDir { V3 d, inv; };

SR_HD SegDir seg_dir(V3 p0, V3 p1) {
  SegDir s;
  s.d = V3{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
  s.inv = V3{1.0f / s.d.x, 1.0f / s.d.y, 1.0f / s.d.z};
  return s;
}

SR_HD bool clip_axis(float p, float d, float inv, float mn, float mx, float& tmin, float& tmax) {
  if (fabsf(d) < 1e-6f) return !(p < mn || p > mx);
  float a = (mn - p) * inv, b = (mx - p) * inv;
  if (a > b) { float t = a; a = b; b = t; }
  float omin = fmaxf(tmin, a), omax = fminf(tmax, b);
  tmin = omin; tmax = omax;
  return !(omin > omax);
}

SR_HD Clipped clip_segment_dir(V3 p0, V3 p1, const SegDir& s, V3 bmin, V3 bmax) {
  const V3 d = s.d;
  float tmin = 0.0f, tmax = 1.0f;
  Clipped c;
  c.ok = false; c.p0 = p0; c.p1 = p1;
  if (!clip_axis(p0.x, d.x, s.inv.x, bmin.x, bmax.x, tmin, tmax)) return c;
  if (!clip_axis(p0.y, d.y, s.inv.y, bmin.y, bmax.y, tmin, tmax)) return c;
  if (!clip_axis(p0.z, d.z, s.inv.z, bmin.z, bmax.z, tmin, tmax)) return c;
  c.ok = true;
  c.p0 = V3{p0.x + d.x * tmin, p0.y + d.y * tmin, p0.z + d.z * tmin};
  c.p1 = V3{p0.x + d.x * tmax, p0.y + d.y * tmax, p0.z + d.z * tmax};
  return c;
}

SR_HD Clipped clip_segment_to_box(V3 p0, V3 p1, V3 bmin, V3 bmax) {
  return clip_segment_dir(p0, p1, seg_dir(p0, p1), bmin, bmax);
}

// Where a DOM texel of light frustum `fr` lives: its depth at
// dom_depth[depth], opacity layer i at dom_opacity[opacity + i * stride].
// False when the DOMs are virtual and the texel's page is not resident.
struct DomTexel { size_t depth, opacity, stride; };

__device__ __forceinline__ bool dom_texel(const FrameParams& fp, const FrustumRec& fr, uint32_t x, uint32_t y,
                                          DomTexel* t) {
  if (!fp.dom_table) {
    const size_t pix = (size_t)y * fr.d.screen_width + x;
    t->depth = fr.dom_depth_off + pix;
    t->opacity = fr.dom_opacity_off + pix;
    t->stride = (size_t)fr.d.screen_width * fr.d.screen_height;
    return true;
  }
  const uint32_t e = fp.dom_table[fr.dom_vpage_base + (y / kDomPage) * fr.dom_pages_x + x / kDomPage];
  if (e == kInvalid) return false;
  const size_t local = (size_t)(y % kDomPage) * kDomPage + x % kDomPage;
  t->depth = (size_t)e * kDomPageTexels + local;
  t->opacity = (size_t)e * fp.dom_layers * kDomPageTexels + local;
  t->stride = kDomPageTexels;
  return true;
}

SR_HD uint32_t pack_binning_field(uint32_t is_shadow, uint32_t frustum) {
  return (is_shadow & 1u) << SR_BINNING_LEVEL_BITS | (frustum << SR_BINNING_FRUSTUM_SHIFT);
}

}  // namespace srk
