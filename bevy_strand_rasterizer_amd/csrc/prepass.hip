// prepass.hip -- culling, projection, froxel binning, scans, compaction and
// raster-work emission on MI355X (gfx950, wave64).
//
// Reference kernels restated here (paths relative to the reference root):
//   broad_prepass            src/shaders/strand_prepass.wgsl:684-760
//   broad_strand_prepass     :762-782   (fused into the projection)
//   fine_prepass             :1016-1140
//   coarse_interval_pass     :835-882   (fused into the LUT build)
//   build_depth_warp_lut     :884-1014
//   mark / allocate pages    :1928-1969
//   binning_queue_pass       :1986-1990 (count)
//   prefix_fine_pages        :2051-2143
//   fill_fine_seg_refs       :2190-2194
//   prefix_active_tile_blocks / compact_active_tiles  :2581-2676
//   emit_raster_work         :2678-2834
//
// Differences of design (results are the canonical ones the oracle defines):
//   * every queue position comes from a device-wide scan instead of an
//     atomic tail, so task order, page ids, page bases, the active-tile queue,
//     work bases and run indices are deterministic;
//   * projection runs over a dense (frustum, index position) slot space:
//     pass 1 writes a wave64 ballot mask and the packed endpoints per slot,
//     a scan of mask popcounts gives each wave its queue base, pass 2
//     compacts with ballot ranks;
//   * the fill pass appends with cell cursors (as the reference) and a
//     per-cell sort puts every cell in (inst, seg) order afterwards.
#include "common.hpp"
#include "kernels.hpp"
#include "scan.hpp"

namespace srk {

// ----------------------------------------------------------- K2 broad
struct CornerBounds { V2 smin, smax; float zmin, zmax; bool any, near_clip; };

__device__ CornerBounds project_aabb(const FrustumRec& fr, const sr_aabb& bb, const float* w) {
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  CornerBounds cb{{1e30f, 1e30f}, {-1e30f, -1e30f}, 1.0f, 0.0f, false, false};
  for (int i = 0; i < 8; ++i) {
    V4 c{(i & 4) ? bb.max[0] : bb.min[0], (i & 2) ? bb.max[1] : bb.min[1], (i & 1) ? bb.max[2] : bb.min[2], 1.0f};
    V4 wc = mat_vec(w, c);
    V4 clip = mat_vec(fr.clip, wc);
    if (clip.w <= 1e-6f) { cb.near_clip = true; continue; }
    float nx = clip.x / clip.w, ny = clip.y / clip.w, nz = clip.z / clip.w;
    if (nz < 0.0f) cb.near_clip = true;
    float rx = 0.0f + (nx * 0.5f + 0.5f) * vw;
    float ry = 0.0f + (ny * -0.5f + 0.5f) * vh;
    cb.any = true;
    cb.smin.x = fminf(cb.smin.x, rx); cb.smin.y = fminf(cb.smin.y, ry);
    cb.smax.x = fmaxf(cb.smax.x, rx); cb.smax.y = fmaxf(cb.smax.y, ry);
    float z = normalize_depth01(nz);
    cb.zmin = fminf(cb.zmin, z); cb.zmax = fmaxf(cb.zmax, z);
  }
  return cb;
}

__device__ float screen_area(const FrustumRec& fr, const CornerBounds& cb) {
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  if (cb.near_clip) return vw * vh;
  if (!cb.any) return -1.0f;
  if (cb.smax.x < 0.0f || cb.smax.y < 0.0f || cb.smin.x >= vw || cb.smin.y >= vh) return -1.0f;
  if (cb.zmax < 0.0f || cb.zmin > 1.0f) return -1.0f;
  float ex = fmaxf(wclamp(cb.smax.x, 0.0f, vw - 1.0f) - wclamp(cb.smin.x, 0.0f, vw - 1.0f), 1.0f);
  float ey = fmaxf(wclamp(cb.smax.y, 0.0f, vh - 1.0f) - wclamp(cb.smin.y, 0.0f, vh - 1.0f), 1.0f);
  return ex * ey;
}

// The instance's coarse asset range and keep probability in frustum fi
// (emit_coarse_asset_range_for_aabb + the stochastic keep); returns whether
// the instance is visible there.
__device__ bool broad_frustum(const FrameParams& fp, uint32_t ii, const InstRec& in, uint32_t strand_count,
                              uint32_t fi, float* keep) {
  const FrustumRec& fr = fp.fr[fi];
  const sr_frame_settings& st = fp.st;
  *keep = 0.0f;
  if (fr.d.kind > 1u) return false;
  bool vis = false;
  CornerBounds cb = project_aabb(fr, in.aabb, in.world_from_local);
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  if (fr.d.coarse_depth_tile_count != 0u && !(fr.d.kind == 1u && fr.light_layer == kInvalid)) {
    CornerBounds r = cb;
    if (r.near_clip) {
      r.smin = V2{0.0f, 0.0f};
      r.smax = V2{vw - 1.0f, vh - 1.0f};
      r.zmax = 1.0f;
      if (!r.any) r.zmin = 0.0f;
      r.any = true;
    }
    if (r.any && !(r.smax.x < 0.0f || r.smax.y < 0.0f || r.smin.x >= vw || r.smin.y >= vh) &&
        !(r.zmax < 0.0f || r.zmin > 1.0f)) {
      uint32_t cpx = max(1u, fr.d.froxel_size_x * kExt), cpy = max(1u, fr.d.froxel_size_y * kExt);
      sr_coarse_asset_range rg;
      rg.geo_id = ii;
      rg.frustum_id = fi;
      rg.min_x = min(f2u(wclamp(r.smin.x, 0.0f, vw - 1.0f)) / cpx, fr.d.coarse_tiles_x - 1u);
      rg.max_x = min(f2u(wclamp(r.smax.x, 0.0f, vw - 1.0f)) / cpx, fr.d.coarse_tiles_x - 1u);
      rg.min_y = min(f2u(wclamp(r.smin.y, 0.0f, vh - 1.0f)) / cpy, fr.d.coarse_tiles_y - 1u);
      rg.max_y = min(f2u(wclamp(r.smax.y, 0.0f, vh - 1.0f)) / cpy, fr.d.coarse_tiles_y - 1u);
      rg.z_min_q = quantize_depth01(r.zmin);
      rg.z_max_q = quantize_depth01(r.zmax);
      fp.ranges[(size_t)ii * fp.F + fi] = rg;
      vis = true;
    }
  }
  if (vis) {
    float area = screen_area(fr, cb);
    float target = fmaxf(st.target_strands_per_pixel, 1e-5f);
    float min_keep = wclamp(st.min_keep_probability, 0.0f, 1.0f);
    float k;
    if (st.stochastic_cull_enabled == 0u) {
      k = 1.0f;
    } else if (area < 0.0f) {
      k = 0.0f;
    } else {
      float area_keep = wclamp(area * target / fmaxf((float)strand_count, 1.0f), min_keep, 1.0f);
      if (fr.d.kind == 0u) {
        k = area_keep;
      } else {
        float cap = wclamp(st.shadow_keep_probability, 0.0f, 1.0f);
        float scale = 1.0f / exp2f((float)fr.d.cascade_index);
        k = wclamp(area_keep * area_keep * scale * cap, min_keep, 1.0f);
      }
    }
    *keep = k;
  }
  return vis;
}

// One wave per instance, its lanes over the frusta: every (instance,
// frustum) range and keep probability, then the instance's broad record
// (visible anywhere; the largest camera and shadow keep, a max that does not
// depend on the order).
__global__ __launch_bounds__(64) void k_broad(FrameParams fp) {
  __shared__ uint32_t kmax[2];  // largest camera / shadow keep, as the bits of a float >= 0
  const uint32_t lane = threadIdx.x;
  for (uint32_t ii = blockIdx.x; ii < fp.NI; ii += gridDim.x) {
    const InstRec& in = fp.inst[ii];
    for (uint32_t fi = lane; fi < fp.F; fi += 64u) {
      sr_coarse_asset_range none{};
      none.geo_id = kInvalid;
      fp.ranges[(size_t)ii * fp.F + fi] = none;
      fp.keep[(size_t)ii * fp.F + fi] = 0.0f;
    }
    if (!in.valid || in.meta_count == 0u || in.geo_strand_count == 0u) {
      if (lane == 0u) fp.broad[ii] = sr_broad_instance_meta{0u, 0u, 0.0f, 0.0f};
      continue;
    }
    if (lane < 2u) kmax[lane] = 0u;
    __syncthreads();
    const uint32_t strand_count = min(in.meta_count, in.geo_strand_count);
    bool visible = false;
    for (uint32_t fi = lane; fi < fp.F; fi += 64u) {
      float k;
      const bool vis = broad_frustum(fp, ii, in, strand_count, fi, &k);
      visible = visible || vis;
      fp.keep[(size_t)ii * fp.F + fi] = k;
      if (vis && k > 0.0f) atomicMax(&kmax[fp.fr[fi].d.kind == 0u ? 0 : 1], __float_as_uint(k));
    }
    const bool any = __ballot(visible) != 0ull;
    __syncthreads();
    if (lane == 0u)
      fp.broad[ii] = sr_broad_instance_meta{strand_count, any ? 1u : 0u, __uint_as_float(kmax[0]),
                                            __uint_as_float(kmax[1])};
    __syncthreads();
  }
}

// -------------------------------------------------- projection (K3 + K5)
__device__ __forceinline__ uint32_t find_instance(const FrameParams& fp, uint32_t pos) {
  // instances own consecutive position ranges starting at pos_base
  uint32_t lo = 0, hi = fp.NI;
  while (hi - lo > 1u) {
    uint32_t mid = (lo + hi) >> 1;
    if (fp.inst[mid].pos_base <= pos) lo = mid; else hi = mid;
  }
  return lo;
}

// Evaluates one (frustum, index position) slot.  Returns true and the
// packed endpoints when the segment survives culling and clipping.
// scr receives the endpoints' raw screen coordinates, which the rasterizers
// use as they are: the world point (w included) through world_to_screen_raw
// (get_segment_vertices_from_index_unchecked, strand_rasterizer.wgsl:289-303).
__device__ bool project_slot(const FrameParams& fp, uint32_t fi, uint32_t pos, uint32_t* ii_out, uint32_t* strand_out,
                             uint32_t* seg_out, sr_projected_segment* out, float4* scr) {
  if (fp.NI == 0u) return false;
  const FrustumRec& fr = fp.fr[fi];
  if (!fr.enabled) return false;
  uint32_t ii = find_instance(fp, pos);
  const InstRec& in = fp.inst[ii];
  if (!in.valid) return false;
  uint32_t p = pos - in.pos_base;
  if (p >= in.index_count) return false;
  const sr_broad_instance_meta b = fp.broad[ii];
  if (b.visible == 0u) return false;
  uint32_t strand = in.pos_strand[p];
  if (strand == kInvalid || strand >= b.strand_count || strand >= in.meta_count) return false;
  float rnd = hash_to_unit_float(wang_hash((ii * 16777619u) ^ strand));
  if (rnd > fmaxf(b.camera_keep_probability, b.shadow_keep_probability)) return false;
  if (rnd > fp.keep[(size_t)ii * fp.F + fi]) return false;
  if (p + 1u >= in.index_count) return false;
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  float4 a = in.vertices[in.indices[p]];
  float4 bq = in.vertices[in.indices[p + 1u]];
  V4 p0w = mat_vec(in.world_from_local, V4{a.x, a.y, a.z, 1.0f});
  V4 p1w = mat_vec(in.world_from_local, V4{bq.x, bq.y, bq.z, 1.0f});
  V4 r0 = world_to_screen_raw(p0w, fr.clip, vw, vh);
  V4 r1 = world_to_screen_raw(p1w, fr.clip, vw, vh);
  V3 q0{r0.x, r0.y, normalize_depth01(r0.z)};
  V3 q1{r1.x, r1.y, normalize_depth01(r1.z)};
  Clipped c = clip_segment_to_box(q0, q1, V3{0.0f, 0.0f, 0.0f},
                                  V3{fmaxf(0.0f, vw - 1e-3f), fmaxf(0.0f, vh - 1e-3f), 1.0f});
  if (!c.ok) return false;
  float ssx = fmaxf(1.0f, vw), ssy = fmaxf(1.0f, vh);
  out->p0_xy = pack2x16unorm(wclamp(c.p0.x / ssx, 0.0f, 1.0f), wclamp(c.p0.y / ssy, 0.0f, 1.0f));
  out->p1_xy = pack2x16unorm(wclamp(c.p1.x / ssx, 0.0f, 1.0f), wclamp(c.p1.y / ssy, 0.0f, 1.0f));
  out->depths = pack2x16unorm(wclamp(c.p0.z, 0.0f, 1.0f), wclamp(c.p1.z, 0.0f, 1.0f));
  scr[0] = make_float4(r0.x, r0.y, r0.z, r0.w);
  scr[1] = make_float4(r1.x, r1.y, r1.z, r1.w);
  *ii_out = ii;
  *strand_out = strand;
  *seg_out = p;
  return true;
}

// ------------------------------------------------------- K7 + K8 LUT
// Interval merge of build_depth_warp_lut, fed one interval at a time in
// linked-list walk order, for S slices per coarse tile.
template <uint32_t S>
struct LutBuilder {
  uint32_t mins[S], maxs[S];
  uint32_t count;
  __device__ void add(uint32_t a, uint32_t b) {
    uint32_t zmin = min(a, b), zmax = max(a, b);
    bool merged = false;
    for (uint32_t i = 0; i < count; ++i) {
      if (zmin <= maxs[i] + 1u && zmax + 1u >= mins[i]) {
        mins[i] = min(mins[i], zmin);
        maxs[i] = max(maxs[i], zmax);
        merged = true;
      }
    }
    if (merged) return;
    if (count < S) {
      mins[count] = zmin; maxs[count] = zmax; ++count;
      return;
    }
    uint32_t best = 0, best_gap = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < S; ++i) {
      uint32_t gap = maxs[i] < zmin ? zmin - maxs[i] : (zmax < mins[i] ? mins[i] - zmax : 0u);
      if (gap < best_gap) { best_gap = gap; best = i; }
    }
    mins[best] = min(mins[best], zmin);
    maxs[best] = max(maxs[best], zmax);
  }
  __device__ void finish(sr_coarse_depth_lut_entry* out) {
    if (count == 0u) {
      for (uint32_t i = 0; i < S; ++i) out[i] = sr_coarse_depth_lut_entry{0, 0, 0, 0};
      return;
    }
    for (uint32_t i = 1; i < count; ++i) {
      uint32_t j = i, kmin = mins[i], kmax = maxs[i];
      while (j != 0u && mins[j - 1u] > kmin) { mins[j] = mins[j - 1u]; maxs[j] = maxs[j - 1u]; --j; }
      mins[j] = kmin; maxs[j] = kmax;
    }
    uint32_t out_slice = 0;
    for (uint32_t iv = 0; iv < count; ++iv) {
      uint32_t slice_count = max(1u, S / count) + (iv < S % count ? 1u : 0u);
      uint32_t zmin = mins[iv], zmax = maxs[iv];
      uint32_t span = max(1u, zmax - zmin + 1u);
      for (uint32_t ls = 0; ls < slice_count && out_slice < S; ++ls) {
        uint32_t start = zmin + (span * ls) / slice_count;
        uint32_t ex_end = zmin + (span * (ls + 1u)) / slice_count;
        uint32_t end = min(ex_end > start ? ex_end - 1u : start, zmax);
        out[out_slice++] = sr_coarse_depth_lut_entry{start, end, start, max(1u, end - start + 1u)};
      }
    }
    while (out_slice < S) out[out_slice++] = sr_coarse_depth_lut_entry{0, 0, 0, 0};
  }
};

__device__ __forceinline__ uint32_t frustum_for_coarse_tile(const FrameParams& fp, uint32_t tile) {
  for (uint32_t fi = 0; fi < fp.F; ++fi) {
    const sr_frustum_desc& d = fp.fr[fi].d;
    if (tile >= d.coarse_depth_tile_base && tile < d.coarse_depth_tile_base + d.coarse_depth_tile_count) return fi;
  }
  return kInvalid;
}

// One thread per coarse tile: walk the intervals of every asset range that
// covers it (newest-first over (instance, frustum) insertion order), build
// the S-slice depth-warp LUT and clear the tile's page-table and active
// words for this frame.
template <uint32_t S>
__global__ void k_lut(FrameParams fp) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fp.coarse_total) return;
  for (uint32_t z = 0; z < S; ++z) fp.page_table[t * S + z] = 0u;
  for (uint32_t z = 0; z < kExt * kExt; ++z) fp.active_flags[t * kExt * kExt + z] = 0u;
  uint32_t fi = frustum_for_coarse_tile(fp, t);
  LutBuilder<S> b;
  b.count = 0;
  if (fi != kInvalid) {
    const sr_frustum_desc& d = fp.fr[fi].d;
    uint32_t local = t - d.coarse_depth_tile_base;
    uint32_t cx = local % d.coarse_tiles_x, cy = local / d.coarse_tiles_x;
    for (int32_t ii = (int32_t)fp.NI - 1; ii >= 0; --ii) {
      const sr_coarse_asset_range r = fp.ranges[(size_t)ii * fp.F + fi];
      if (r.geo_id == kInvalid || r.max_x < r.min_x || r.max_y < r.min_y) continue;
      if (cx < r.min_x || cx > r.max_x || cy < r.min_y || cy > r.max_y) continue;
      b.add(r.z_min_q, r.z_max_q);
    }
  }
  sr_coarse_depth_lut_entry out[S];
  b.finish(out);
  for (uint32_t z = 0; z < S; ++z) {
    fp.lut[(size_t)t * S + z] = out[z];
    fp.lut_z[(size_t)t * S + z] =
        out[z].virtual_count_q ? make_uint2(out[z].z_min_q, out[z].z_max_q) : make_uint2(~0u, 0u);
  }
  // one interval of at least S depths: its slices are S contiguous ranges
  // [zmin + span * i / S, zmin + span * (i + 1) / S - 1] (finish above),
  // which the marking pass finds arithmetically
  const bool single = b.count == 1u && b.maxs[0] - b.mins[0] + 1u >= S;
  fp.lut_iv[t] = single ? make_uint2(b.mins[0], b.maxs[0]) : make_uint2(1u, 0u);
}

// ------------------------------------------------ binning traversal
// coarse_depth_slice_for_depth: the first used slice whose range holds the
// quantised depth, else the first nearest one.  Both endpoint depths are
// looked up in one pass over the tile's S (z_min_q, z_max_q) pairs, two
// pairs per load (a used slice has z_min_q <= z_max_q; unused ones are
// (~0, 0)).  The LUT slice of a quantised depth: the first used slice whose
// range holds it, else the used slice nearest to it (first on ties), else
// invalid.  The holding slices come from one mask per depth; the
// nearest-slice scan runs only for a depth no slice holds.
__device__ uint32_t nearest_slice(const uint4* l4, uint32_t zq, uint32_t S) {
  uint32_t best = kInvalid, best_d = 0xFFFFFFFFu;
  for (uint32_t i = 0; i < S; i += 2) {
    const uint4 q = l4[i / 2];
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t ex = h ? q.z : q.x, ey = h ? q.w : q.y;
      if (ex > ey) continue;
      const uint32_t d = zq > ey ? zq - ey : ex - zq;
      if (d < best_d) {
        best_d = d;
        best = i + h;
      }
    }
  }
  return best;
}

__device__ __forceinline__ void lut_slices_for_depths(const uint2* lz, uint2 iv, float za, float zb, uint32_t* sa,
                                                      uint32_t* sb, uint32_t S, uint32_t shift) {
  const uint32_t qa = quantize_depth01(za), qb = quantize_depth01(zb);
  if (qa >= iv.x && qa <= iv.y && qb >= iv.x && qb <= iv.y) {  // both inside one evenly split interval
    const float r = __builtin_amdgcn_rcpf((float)(iv.y - iv.x + 1u));  // within 1 ulp
    *sa = even_slice(iv.x, iv.y, qa, r, shift);
    *sb = even_slice(iv.x, iv.y, qb, r, shift);
    return;
  }
  const uint4* l4 = reinterpret_cast<const uint4*>(lz);
  uint32_t ma = 0u, mb = 0u;  // slices holding qa / qb (a used slice has x <= y; S <= 32 bits)
#pragma unroll 2
  for (uint32_t i = 0; i < S; i += 2) {
    const uint4 q = l4[i / 2];
    ma |= ((qa >= q.x && qa <= q.y) ? (1u << i) : 0u) | ((qa >= q.z && qa <= q.w) ? (2u << i) : 0u);
    mb |= ((qb >= q.x && qb <= q.y) ? (1u << i) : 0u) | ((qb >= q.z && qb <= q.w) ? (2u << i) : 0u);
  }
  *sa = ma ? (uint32_t)(__ffs(ma) - 1) : nearest_slice(l4, qa, S);
  *sb = mb ? (uint32_t)(__ffs(mb) - 1) : nearest_slice(l4, qb, S);
}

enum { MODE_MARK = 0, MODE_COUNT = 1, MODE_FILL = 2, MODE_SCATTER = 3 };

// x / n for a tile edge n in pixels: a power of two divides as the product
// with its exact reciprocal 2^-k (both are x / n correctly rounded), any
// other n by the division.
__device__ __forceinline__ float div_by(float x, uint32_t n) {
  if (n != 0u && (n & (n - 1u)) == 0u) return x * ldexpf(1.0f, -(int)(31u - (uint32_t)__clz(n)));
  return x / (float)n;
}

// The fine tiles of one coarse tile as fine_tile_occluded_by_opaque
// (strand_prepass.wgsl:551-573) sees them: a camera fine tile fully covered
// by opaque depth hides a slice whose nearest depth (z_max, reverse-Z) lies
// behind the tile's farthest opaque sample.
struct Occluder {
  const uint2* tiles;  // null: nothing is occluded
  uint32_t base, ftx, fty, fx0, fy0;
  __device__ __forceinline__ bool hides(int32_t fx, int32_t fy, float nearest) const {
    if (!tiles) return false;
    const uint32_t x = fx0 + (uint32_t)fx, y = fy0 + (uint32_t)fy;
    if (x >= ftx || y >= fty) return false;
    const uint2 t = tiles[base + y * ftx + x];
    return t.x != 0u && nearest <= dequantize_depth01(t.y) - 1e-4f;
  }
};

__device__ __forceinline__ Occluder occluder_for(const FrameParams& fp, const sr_frustum_desc& fd, uint32_t local) {
  Occluder o{nullptr, 0u, 0u, 0u, 0u, 0u};
  if (!fp.opaque_tiles || fd.kind != 0u) return o;
  o.tiles = fp.opaque_tiles;
  o.base = fd.fine_depth_tile_base;
  o.ftx = (fd.screen_width + fd.froxel_size_x - 1u) / fd.froxel_size_x;
  o.fty = (fd.screen_height + fd.froxel_size_y - 1u) / fd.froxel_size_y;
  o.fx0 = (local % fd.coarse_tiles_x) * kExt;
  o.fy0 = (local / fd.coarse_tiles_x) * kExt;
  return o;
}

// Fine 3D Amanatides-Woo walk through the 4 x 4 x 16 cells of one count page
// (strand_prepass.wgsl:1517-1573) for a segment already clipped to the coarse
// tile and the LUT slice [zmin, zmax]; calls emit(cell) per visited cell that
// opaque depth does not hide (a hidden cell is stepped over, :1541-1555).
template <class Emit>
__device__ __forceinline__ void fine_dda(V3 q0, V3 q1, V2 origin, float zmin, float zmax, const sr_frustum_desc& fd,
                                         const Occluder& occ, Emit&& emit) {
  float zspan = fmaxf(zmax - zmin, 1.0f / (float)kQuant);
  float pa[3] = {div_by(q0.x - origin.x, fd.froxel_size_x), div_by(q0.y - origin.y, fd.froxel_size_y),
                 wclamp((q0.z - zmin) / zspan, 0.0f, 0.999999f) * (float)kSlices};
  float pb[3] = {div_by(q1.x - origin.x, fd.froxel_size_x), div_by(q1.y - origin.y, fd.froxel_size_y),
                 wclamp((q1.z - zmin) / zspan, 0.0f, 0.999999f) * (float)kSlices};
  const int32_t hi[3] = {(int32_t)kExt - 1, (int32_t)kExt - 1, (int32_t)kSlices - 1};
  int32_t f[3], f1[3], step[3];
  float delta[3], tmax[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[k] = iclamp(f2i(floorf(pa[k])), 0, hi[k]);
    f1[k] = iclamp(f2i(floorf(pb[k])), 0, hi[k]);
    float d = pb[k] - pa[k];
    step[k] = sgn_i32(d);
    bool tiny = fabsf(d) < 1e-6f;
    float safe = tiny ? 1e-6f : d;
    delta[k] = fabsf(1.0f / safe);
    float fl = floorf(pa[k]);
    tmax[k] = step[k] > 0 ? ((fl + 1.0f) - pa[k]) / safe : (fl - pa[k]) / safe;
    if (tiny) tmax[k] = 1e38f;
  }
  for (uint32_t safety = 0; safety <= kExt + kExt + kSlices + 3u; ++safety) {
    if (!occ.hides(f[0], f[1], zmax)) emit(((uint32_t)f[2] * kExt + (uint32_t)f[1]) * kExt + (uint32_t)f[0]);
    if (f[0] == f1[0] && f[1] == f1[1] && f[2] == f1[2]) break;
    float mn = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
    int k = tmax[0] == mn ? 0 : (tmax[1] == mn ? 1 : 2);  // canonical_min_mask: x, then y, then z
    f[k] += step[k];
    tmax[k] += delta[k];
    if (f[0] < 0 || f[1] < 0 || f[2] < 0 || f[0] > hi[0] || f[1] > hi[1] || f[2] > hi[2]) break;
  }
}

// Coarse-tile box of a frustum's coarse tile (strand_prepass.wgsl:1422-1433).
__device__ __forceinline__ void coarse_tile_box(const sr_frustum_desc& fd, uint32_t cx, uint32_t cy, V2* origin,
                                                V3* bmax) {
  uint32_t cpx = max(1u, fd.froxel_size_x * kExt), cpy = max(1u, fd.froxel_size_y * kExt);
  *origin = V2{(float)(cx * cpx), (float)(cy * cpy)};
  float tmx = fminf(origin->x + (float)cpx, (float)fd.screen_width) - 1e-3f;
  float tmy = fminf(origin->y + (float)cpy, (float)fd.screen_height) - 1e-3f;
  *bmax = V3{fmaxf(origin->x, tmx), fmaxf(origin->y, tmy), 1.0f};
}

constexpr uint32_t kHitOverflow = 0xFFu;

// Marking defers each (slot, page) fine walk to a queue of its wave: the
// lanes walk their segments' coarse tiles and slices, then the wave runs the
// queued fine walks one per lane, so a slot with many pages does not hold the
// other 63 lanes idle.  A slot's cells are counted and recorded through its
// lane's LDS counter; their order in the record does not matter (a slot's
// cells are distinct (page, cell) pairs, and the sort orders hits by page,
// cell and slot).  A full queue walks in place.
struct MarkJob {
  float q0x, q0y, q0z, q1x, q1y, q1z, zmin, zmax;  // the segment clipped to the page
  uint32_t pti, lo;  // page-table index; frustum-local coarse tile << 6 | owner lane
};

struct MarkQueue {
  MarkJob* q;
  uint32_t* n;     // walks queued (may pass the capacity: the rest walked in place)
  uint32_t* nrec;  // per lane: cells of its slot so far
  uint32_t owner;  // the lane whose slot the walk belongs to
};

// Per-slot state of a walk (CSR backend).  MARK: the cells recorded (dst;
// every visible cell is counted, also past cap: in n, or with a queue in the
// owner lane's LDS counter).  SCATTER: the slot's next hit goes to base + n.
struct HitRec {
  uint32_t* dst;  // null: not recording; hit j at dst[j * stride]
  uint32_t n;
  uint32_t cap;
  uint32_t base, slot;
  uint32_t stride;
  MarkQueue* mq = nullptr;  // MARK with the CSR backend: the wave's queue
};

__device__ __forceinline__ void mark_cell(HitRec& rec, uint32_t pti, uint32_t cell) {
  const uint32_t n = rec.mq ? atomicAdd(&rec.mq->nrec[rec.mq->owner], 1u) : rec.n++;
  if (rec.dst && n < rec.cap) rec.dst[(size_t)n * rec.stride] = (pti << 8) | cell;
}

// The sort key of a hit: page * 256 + cell; a dropped page's hits carry the
// page id page_cap and sort after every kept one.
__device__ __forceinline__ uint32_t hit_key(const FrameParams& fp, uint32_t page_entry, uint32_t cell) {
  const uint32_t page = (page_entry == 0u || page_entry == kMarked) ? fp.page_cap : page_entry - 1u;
  return (page << 8) | cell;
}

__device__ __forceinline__ void write_hit(const FrameParams& fp, uint64_t idx, uint32_t key, uint32_t slot) {
  if (idx >= fp.ref_cap) return;
  fp.hkeys_w[idx] = key;
  fp.hvals_w[idx] = slot;
}

// trace_segment_into_coarse_count_page (strand_prepass.wgsl:1412-1575).
// MARK also counts page candidates for the CSR backend and records the fine
// cells of every touched page; SCATTER appends the slot to each touched
// page's candidate list.
template <int MODE>
__device__ void trace_coarse_tile(const FrameParams& fp, V3 p0, V3 p1, const SegDir& sd, const sr_frustum_desc& fd,
                                  uint32_t tile, uint32_t local, uint32_t cx, uint32_t cy, const sr_fine_seg_ref& ref,
                                  uint32_t task, HitRec& rec) {
  V2 origin;
  V3 bmax;
  coarse_tile_box(fd, cx, cy, &origin, &bmax);
  Clipped c = clip_segment_dir(p0, p1, sd, V3{origin.x, origin.y, 0.0f}, bmax);
  if (!c.ok) return;
  const uint2* lz = fp.lut_z + (size_t)tile * fp.slices;
  uint32_t z0s, z1s;
  lut_slices_for_depths(lz, fp.lut_iv[tile], c.p0.z, c.p1.z, &z0s, &z1s, fp.slices, fp.slice_shift);
  if (z0s == kInvalid || z1s == kInvalid) return;
  const SegDir cd = seg_dir(c.p0, c.p1);  // the tile's piece, clipped against each slice
  uint32_t zlo = min(z0s, z1s), zhi = max(z0s, z1s);
  for (uint32_t cz = zlo; cz <= zhi; ++cz) {
    const uint2 e = lz[cz];
    if (e.x > e.y) continue;  // unused slice
    float zmin = dequantize_depth01(e.x), zmax = dequantize_depth01(e.y);
    Clipped dc = clip_segment_dir(c.p0, c.p1, cd, V3{origin.x, origin.y, zmin}, V3{bmax.x, bmax.y, zmax});
    if (!dc.ok) continue;
    uint32_t pti = tile * fp.slices + cz;
    if (MODE == MODE_MARK) {
      fp.page_table[pti] = kMarked;
      if (fp.st.fine_binning_backend != 0u) {  // the visible cells: recorded (up to cap) and counted
        if (rec.mq) {
          const uint32_t j = atomicAdd(rec.mq->n, 1u);
          if (j < fp.mark_queue) {
            rec.mq->q[j] = MarkJob{dc.p0.x, dc.p0.y, dc.p0.z, dc.p1.x, dc.p1.y, dc.p1.z, zmin, zmax, pti,
                                   (local << 6) | rec.mq->owner};
            continue;
          }
        }
        fine_dda(dc.p0, dc.p1, origin, zmin, zmax, fd, occluder_for(fp, fd, local),
                 [&](uint32_t cell) { mark_cell(rec, pti, cell); });
      }
      continue;
    }
    uint32_t h = fp.page_table[pti];
    if (MODE == MODE_SCATTER) {  // the slot's hits in walk order
      const uint32_t before = rec.n;
      fine_dda(dc.p0, dc.p1, origin, zmin, zmax, fd, occluder_for(fp, fd, local), [&](uint32_t cell) {
        write_hit(fp, (uint64_t)rec.base + rec.n++, hit_key(fp, h, cell), rec.slot);
      });
      if (fp.tel && rec.n > before && h != 0u && h != kMarked)  // (segment, page) candidates
        atomicAdd(&fp.tel[kTelPageOff + h - 1u], 1u);
      continue;
    }
    if (h == 0u || h == kMarked) continue;
    uint32_t page = h - 1u;
    uint32_t ref_base = 0;
    if (MODE == MODE_FILL) {
      ref_base = fp.page_meta[page].seg_ref_base;
      if (fp.page_meta[page].seg_ref_count == 0u) continue;
    }
    const uint32_t page_cells = page * kPage;
    uint32_t emitted = 0;
    fine_dda(dc.p0, dc.p1, origin, zmin, zmax, fd, occluder_for(fp, fd, local), [&](uint32_t cell) {
      ++emitted;
      if (MODE == MODE_COUNT) {
        atomicAdd(&fp.cell_counts[page_cells + cell], 1u);
      } else {
        uint32_t w = atomicAdd(&fp.cell_cursors[page_cells + cell], 1u);
        uint64_t dst = (uint64_t)ref_base + fp.cell_offsets[page_cells + cell] + w;
        if (dst < fp.ref_cap) fp.refs[dst] = ref;
      }
    });
    // (segment, page) candidates for telemetry: pairs with a visible cell
    if (MODE == MODE_COUNT && fp.tel && emitted) atomicAdd(&fp.tel[kTelPageOff + page], 1u);
  }
}

// The coarse 2D DDA of visit_segment_asset_coarse_tiles
// (strand_prepass.wgsl:1591-1696) over the asset range's coarse tiles,
// resumable: init clips the segment to the range's tile rectangle and finds
// the first tile, next() the one after it.  `pending` is the frustum-local
// tile to visit; tiles outside the frustum's coarse-depth tiles are skipped
// as the reference skips them.
struct CoarseWalk {
  int32_t f[2], f1[2], step[2], lo[2], hi[2];
  float delta[2], tmax[2];
  uint32_t left;     // DDA iterations left (the reference's safety bound)
  uint32_t pending;  // frustum-local coarse tile to visit next
  bool live;         // pending is valid
};

__device__ __forceinline__ bool coarse_walk_next(const sr_frustum_desc& fd, CoarseWalk& cw) {
  while (cw.left != 0u) {
    --cw.left;
    if (cw.f[0] < cw.lo[0] || cw.f[0] > cw.hi[0] || cw.f[1] < cw.lo[1] || cw.f[1] > cw.hi[1]) break;
    const uint32_t local = (uint32_t)cw.f[1] * fd.coarse_tiles_x + (uint32_t)cw.f[0];
    if (cw.f[0] == cw.f1[0] && cw.f[1] == cw.f1[1]) {
      cw.left = 0u;  // the last tile
    } else if (cw.tmax[0] <= cw.tmax[1]) {
      cw.f[0] += cw.step[0];
      cw.tmax[0] += cw.delta[0];
    } else {
      cw.f[1] += cw.step[1];
      cw.tmax[1] += cw.delta[1];
    }
    if (local < fd.coarse_depth_tile_count) {
      cw.pending = local;
      cw.live = true;
      return true;
    }
  }
  cw.live = false;
  return false;
}

__device__ bool coarse_walk_init(const FrameParams& fp, uint32_t fi, uint32_t ii, V3 p0, V3 p1, const SegDir& sd,
                                 CoarseWalk& cw) {
  cw.live = false;
  cw.left = 0u;
  const sr_frustum_desc& fd = fp.fr[fi].d;
  const sr_coarse_asset_range r = fp.ranges[(size_t)ii * fp.F + fi];
  if (r.geo_id != ii || r.frustum_id != fi) return false;
  float W = (float)fd.screen_width, H = (float)fd.screen_height;
  uint32_t cpx = max(1u, fd.froxel_size_x * kExt), cpy = max(1u, fd.froxel_size_y * kExt);
  uint32_t min_cx = max(r.min_x, min(f2u(fmaxf(fminf(p0.x, p1.x), 0.0f)) / cpx, fd.coarse_tiles_x - 1u));
  uint32_t max_cx = min(r.max_x, min(f2u(fmaxf(fmaxf(p0.x, p1.x), 0.0f)) / cpx, fd.coarse_tiles_x - 1u));
  uint32_t min_cy = max(r.min_y, min(f2u(fmaxf(fminf(p0.y, p1.y), 0.0f)) / cpy, fd.coarse_tiles_y - 1u));
  uint32_t max_cy = min(r.max_y, min(f2u(fmaxf(fmaxf(p0.y, p1.y), 0.0f)) / cpy, fd.coarse_tiles_y - 1u));
  if (max_cx < min_cx || max_cy < min_cy) return false;
  V3 bmin{(float)(min_cx * cpx), (float)(min_cy * cpy), 0.0f};
  V3 bmax{fminf((float)((max_cx + 1u) * cpx), W) - 1e-3f, fminf((float)((max_cy + 1u) * cpy), H) - 1e-3f, 1.0f};
  Clipped c = clip_segment_dir(p0, p1, sd, bmin, bmax);
  if (!c.ok) return false;
  float pa[2] = {div_by(c.p0.x, cpx), div_by(c.p0.y, cpy)};
  float pb[2] = {div_by(c.p1.x, cpx), div_by(c.p1.y, cpy)};
  cw.lo[0] = (int32_t)min_cx; cw.lo[1] = (int32_t)min_cy;
  cw.hi[0] = (int32_t)max_cx; cw.hi[1] = (int32_t)max_cy;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    cw.f[k] = iclamp(f2i(floorf(pa[k])), cw.lo[k], cw.hi[k]);
    cw.f1[k] = iclamp(f2i(floorf(pb[k])), cw.lo[k], cw.hi[k]);
    float d = pb[k] - pa[k];
    cw.step[k] = sgn_i32(d);
    bool tiny = fabsf(d) < 1e-6f;
    float safe = tiny ? 1e-6f : d;
    cw.delta[k] = fabsf(1.0f / safe);
    float fl = floorf(pa[k]);
    cw.tmax[k] = cw.step[k] > 0 ? ((fl + 1.0f) - pa[k]) / safe : (fl - pa[k]) / safe;
    if (tiny) cw.tmax[k] = 1e38f;
  }
  cw.left = (max_cx - min_cx + 1u) + (max_cy - min_cy + 1u) + 3u;
  return coarse_walk_next(fd, cw);
}

// Every coarse tile of the segment in walk order, each traced in place.
template <int MODE>
__device__ void visit_segment(const FrameParams& fp, uint32_t fi, uint32_t ii, V3 p0, V3 p1,
                              const sr_fine_seg_ref& ref, uint32_t task, HitRec& rec) {
  const sr_frustum_desc& fd = fp.fr[fi].d;
  const SegDir sd = seg_dir(p0, p1);  // shared with every coarse tile's clip
  CoarseWalk cw;
  if (!coarse_walk_init(fp, fi, ii, p0, p1, sd, cw)) return;
  do {
    const uint32_t local = cw.pending;
    trace_coarse_tile<MODE>(fp, p0, p1, sd, fd, fd.coarse_depth_tile_base + local, local, local % fd.coarse_tiles_x,
                            local / fd.coarse_tiles_x, ref, task, rec);
  } while (coarse_walk_next(fd, cw));
}

// Endpoints as every binning pass sees them: the packed ProjectedSegment
// unpacked and scaled by the frustum's screen size (process_binning_task,
// strand_prepass.wgsl:1894-1926).
__device__ __forceinline__ void task_endpoints(const sr_frustum_desc& fd, const sr_projected_segment& ps, V3* p0,
                                               V3* p1) {
  float W = (float)fd.screen_width, H = (float)fd.screen_height;
  V2 a = unpack2x16unorm(ps.p0_xy), b = unpack2x16unorm(ps.p1_xy), dz = unpack2x16unorm(ps.depths);
  *p0 = V3{a.x * W, a.y * H, dz.x};
  *p1 = V3{b.x * W, b.y * H, dz.y};
}

template <int MODE>
__device__ void visit_task(const FrameParams& fp, uint32_t t) {
  const sr_binning_task tk = fp.tasks[t];
  const sr_projected_segment ps = fp.proj[t];
  uint32_t fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
  if (fi >= fp.F) return;
  uint32_t ii = tk.id_info;
  V3 p0, p1;
  task_endpoints(fp.fr[fi].d, ps, &p0, &p1);
  sr_fine_seg_ref ref{ii, tk.seg_idx, 0u};
  if (MODE == MODE_FILL) {
    const InstRec& in = fp.inst[ii];
    if (tk.chunk_id < in.meta_count) ref.material_idx = in.meta[tk.chunk_id].material_idx;
  }
  HitRec none{nullptr, 0u, 0u, 0u, 0u, 0u};
  visit_segment<MODE>(fp, fi, ii, p0, p1, ref, t, none);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_visit(FrameParams fp) {
  const uint32_t n = min(fp.counters[C_TASKS], fp.task_cap);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) visit_task<MODE>(fp, t);
}

// Pass 1: per slot validity ballot + packed endpoints in slot layout; valid
// slots also mark their pages (the LUT is built before projection) and, for
// the CSR backend, record their fine cells and count page candidates.
// Each wave balances its 64 segments' work in rounds: the lanes queue their
// segments' coarse tiles (a few DDA steps each), then every lane traces
// queued tiles (clip, LUT slices, page marks) and queues their fine walks,
// then every lane runs queued fine walks.  A segment crossing many tiles no
// longer holds the other 63 lanes; a lane whose tiles do not fit the queue
// resumes its walk in the next round.  The wave is converged at every flush.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_project_mark(FrameParams fp, sr_projected_segment* slot_proj) {
  __shared__ MarkJob jobs[WAVES][kMarkQ];
  __shared__ uint32_t tjobs[WAVES][kTileQ];  // frustum-local coarse tile << 6 | owner lane
  __shared__ float lseg[WAVES][9][64];       // per lane: endpoints and 1 / direction
  __shared__ uint32_t jn[WAVES], tn[WAVES], nrec[WAVES][64];
  const uint64_t total = (uint64_t)fp.F * fp.slots;
  const bool csr = fp.st.fine_binning_backend != 0u;
  const bool record = csr && fp.hit_k != 0u;
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += (uint64_t)gridDim.x * blockDim.x) {
    if (lane == 0u) {
      jn[w] = 0u;
      tn[w] = 0u;
    }
    nrec[w][lane] = 0u;
    const uint32_t fi = (uint32_t)(s / fp.slots), pos = (uint32_t)(s % fp.slots);  // one frustum per wave
    const sr_frustum_desc& fd = fp.fr[fi].d;
    uint32_t ii, strand, seg;
    sr_projected_segment ps;
    float4 scr[2];
    const bool ok = project_slot(fp, fi, pos, &ii, &strand, &seg, &ps, scr);
    CoarseWalk cw;
    cw.live = false;
    if (ok) {
      slot_proj[s] = ps;
      fp.slot_screen[2 * s] = scr[0];
      fp.slot_screen[2 * s + 1] = scr[1];
      // K10: mark the (coarse tile, LUT slice) pages the segment touches,
      // from the same packed endpoints the count and fill passes unpack
      V3 p0, p1;
      task_endpoints(fd, ps, &p0, &p1);
      const SegDir sd = seg_dir(p0, p1);
      float* ls = &lseg[w][0][lane];
      ls[0 * 64] = p0.x; ls[1 * 64] = p0.y; ls[2 * 64] = p0.z;
      ls[3 * 64] = p1.x; ls[4 * 64] = p1.y; ls[5 * 64] = p1.z;
      ls[6 * 64] = sd.inv.x; ls[7 * 64] = sd.inv.y; ls[8 * 64] = sd.inv.z;
      coarse_walk_init(fp, fi, ii, p0, p1, sd, cw);
    }
    const uint32_t cpx = max(1u, fd.froxel_size_x * kExt), cpy = max(1u, fd.froxel_size_y * kExt);
    while (__ballot(cw.live) != 0ull) {
      while (cw.live) {  // queue this lane's tiles while they fit
        const uint32_t j = atomicAdd(&tn[w], 1u);
        if (j >= fp.tile_queue) break;  // the pending tile waits for the next round
        tjobs[w][j] = (cw.pending << 6) | lane;
        coarse_walk_next(fd, cw);
      }
      const uint32_t nt = min(tn[w], fp.tile_queue);
      for (uint32_t j = lane; j < nt; j += 64u) {  // the queued tiles, one per lane
        const uint32_t job = tjobs[w][j], owner = job & 63u, local = job >> 6;
        const float* os = &lseg[w][0][owner];
        const V3 p0{os[0 * 64], os[1 * 64], os[2 * 64]}, p1{os[3 * 64], os[4 * 64], os[5 * 64]};
        const SegDir sd{V3{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, V3{os[6 * 64], os[7 * 64], os[8 * 64]}};
        MarkQueue omq{jobs[w], &jn[w], nrec[w], owner};
        HitRec orec{record ? fp.slot_hits + (s - lane + owner) : nullptr, 0u, fp.hit_k, 0u, 0u, (uint32_t)total};
        orec.mq = &omq;
        trace_coarse_tile<MODE_MARK>(fp, p0, p1, sd, fd, fd.coarse_depth_tile_base + local, local,
                                     local % fd.coarse_tiles_x, local / fd.coarse_tiles_x, sr_fine_seg_ref{}, 0u,
                                     orec);
      }
      const uint32_t nj = min(jn[w], fp.mark_queue);
      for (uint32_t j = lane; j < nj; j += 64u) {  // their queued fine walks, one per lane
        const MarkJob jb = jobs[w][j];
        const uint32_t owner = jb.lo & 63u, local = jb.lo >> 6;
        MarkQueue omq{nullptr, nullptr, nrec[w], owner};
        HitRec orec{record ? fp.slot_hits + (s - lane + owner) : nullptr, 0u, fp.hit_k, 0u, 0u, (uint32_t)total};
        orec.mq = &omq;
        const V2 origin{(float)((local % fd.coarse_tiles_x) * cpx), (float)((local / fd.coarse_tiles_x) * cpy)};
        fine_dda(V3{jb.q0x, jb.q0y, jb.q0z}, V3{jb.q1x, jb.q1y, jb.q1z}, origin, jb.zmin, jb.zmax, fd,
                 occluder_for(fp, fd, local), [&](uint32_t cell) { mark_cell(orec, jb.pti, cell); });
      }
      if (lane == 0u) {  // every lane has read both counts
        tn[w] = 0u;
        jn[w] = 0u;
      }
    }
    if (csr) {
      const uint32_t hits = ok ? nrec[w][lane] : 0u;
      if (record && ok) fp.slot_hit_count[s] = (uint8_t)(hits > fp.hit_k ? kHitOverflow : hits);
      fp.slot_hit_total[s] = hits;  // offsets by a scan: no device atomic
    }
    unsigned long long m = __ballot(ok);
    if ((threadIdx.x & 63) == 0) fp.slot_masks[s >> 6] = m;
    const FrustumRec& fr = fp.fr[fi];
    if (fr.cam_index != kInvalid) fp.shaded[(size_t)fr.cam_index * fp.slots + pos] = ok ? 1 : 0;
  }
}

struct MaskSrc {
  const unsigned long long* masks;
  __device__ uint32_t load(uint32_t i) const { return (uint32_t)__popcll(masks[i]); }
};
struct WaveOffsetDst {
  uint32_t* offsets;
  uint32_t* counters;
  __device__ void store(uint32_t i, uint32_t ex, uint32_t) const { offsets[i] = ex; }
  __device__ void total(uint32_t t) const { counters[C_TASKS] = t; }
};

__global__ void k_frustum_bases(FrameParams fp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t waves_per_frustum = fp.slots / 64u;
  uint32_t cam = 0;
  for (uint32_t fi = 0; fi < fp.F; ++fi) {
    uint32_t base = fp.wave_offsets[(size_t)fi * waves_per_frustum];
    fp.frustum_task_base[fi] = base;
  }
  fp.frustum_task_base[fp.F] = fp.counters[C_TASKS];
  for (uint32_t fi = 0; fi < fp.F; ++fi) {
    fp.frustum_cam_base[fi] = cam;
    if (fp.fr[fi].d.kind == 0u) cam += fp.frustum_task_base[fi + 1] - fp.frustum_task_base[fi];
  }
  fp.counters[C_CAM_TASKS] = cam;
}

// Pass 2: compaction with ballot ranks; the queue order is slot order.
// CSR backend: each slot joins the candidate lists of its pages (K17).
__global__ __launch_bounds__(256) void k_project_write(FrameParams fp, const sr_projected_segment* slot_proj) {
  const uint64_t total = (uint64_t)fp.F * fp.slots;
  const uint32_t lane = threadIdx.x & 63;
  const unsigned long long lower = (1ull << lane) - 1ull;
  const bool csr = fp.st.fine_binning_backend != 0u;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long m = fp.slot_masks[s >> 6];
    const uint32_t t = fp.wave_offsets[s >> 6] + (uint32_t)__popcll(m & lower);
    const bool valid = ((m >> lane) & 1ull) && t < fp.task_cap;
    if (valid) {
      // every load first (recorded cells up to kPre and their page-table
      // entries, index clamped; values past the slot's count are unused),
      // then the stores, so the latencies overlap
      constexpr uint32_t kPre = 6u;
      uint32_t fi = (uint32_t)(s / fp.slots), pos = (uint32_t)(s % fp.slots);
      uint32_t ii = find_instance(fp, pos);
      const InstRec& in = fp.inst[ii];
      uint32_t p = pos - in.pos_base;
      const FrustumRec& fr = fp.fr[fi];
      const uint32_t strand = in.pos_strand[p];
      const sr_projected_segment ps = slot_proj[s];
      uint32_t nh = kHitOverflow, base = 0u, mat = 0u;
      uint32_t hv[kPre], he[kPre];
      const uint32_t* h = fp.slot_hits + s;
      const uint32_t pt_last = fp.coarse_total * fp.slices - 1u;
      if (csr) {
        nh = fp.hit_k ? fp.slot_hit_count[s] : kHitOverflow;
        base = fp.hit_base[s];
#pragma unroll
        for (uint32_t k = 0; k < kPre; ++k) hv[k] = k < fp.hit_k ? h[(size_t)k * total] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kPre; ++k) he[k] = fp.page_table[min(hv[k] >> 8, pt_last)];
        mat = strand < in.meta_count ? in.meta[strand].material_idx : 0u;
      }
      fp.tasks[t] = sr_binning_task{ii, strand, p, pack_binning_field(fr.d.kind == 1u ? 1u : 0u, fi)};
      fp.proj[t] = ps;
      if (fr.d.kind == 0u) fp.shading_queue[fp.frustum_cam_base[fi] + (t - fp.frustum_task_base[fi])] = t;
      if (csr) {  // K17: the slot's hits at its scanned offset
        fp.slot_material[pos] = mat;  // the same value from every frustum
        if (nh != kHitOverflow) {  // the recorded cells, in any order
#pragma unroll
          for (uint32_t k = 0; k < kPre; ++k) {
            if (k >= nh) break;
            write_hit(fp, (uint64_t)base + k, hit_key(fp, he[k], hv[k] & 255u), pos);
          }
          for (uint32_t k0 = kPre; k0 < nh; k0 += kPre) {  // the rest, kPre at a time, loads first
#pragma unroll
            for (uint32_t k = 0; k < kPre; ++k) hv[k] = k0 + k < nh ? h[(size_t)(k0 + k) * total] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kPre; ++k) he[k] = fp.page_table[min(hv[k] >> 8, pt_last)];
#pragma unroll
            for (uint32_t k = 0; k < kPre; ++k)
              if (k0 + k < nh) write_hit(fp, (uint64_t)base + k0 + k, hit_key(fp, he[k], hv[k] & 255u), pos);
          }
          if (fp.tel) {  // (segment, page) candidates: each recorded page once
            for (uint32_t k = 0; k < nh; ++k) {
              const uint32_t pti = h[(size_t)k * total] >> 8;
              bool seen = false;
              for (uint32_t k2 = 0; k2 < k && !seen; ++k2) seen = (h[(size_t)k2 * total] >> 8) == pti;
              const uint32_t e = seen ? 0u : fp.page_table[pti];
              if (e != 0u && e != kMarked) atomicAdd(&fp.tel[kTelPageOff + e - 1u], 1u);
            }
          }
        } else {  // too many cells to record: trace again
          V3 p0, p1;
          task_endpoints(fr.d, ps, &p0, &p1);
          HitRec rec{nullptr, 0u, 0u, base, pos, 0u};
          visit_segment<MODE_SCATTER>(fp, fi, ii, p0, p1, sr_fine_seg_ref{ii, p, 0u}, (uint32_t)s, rec);
        }
      }
    }
  }
}

// ------------------------------------------------------ K11 allocate
struct PageMarkSrc {
  const uint32_t* table;
  __device__ uint32_t load(uint32_t i) const { return table[i] == kMarked ? 1u : 0u; }
};
struct PageAllocDst {
  uint32_t* table;
  sr_fine_page_meta* meta;
  uint32_t* counters;
  uint32_t page_cap;
  uint32_t slice_shift;  // page-table entry e is slice e % 2^shift of coarse tile e >> shift
  __device__ void store(uint32_t e, uint32_t ex, uint32_t v) const {
    if (!v) return;
    if (ex >= page_cap) {  // the reference drops pages past its capacity (entry 0)
      table[e] = 0u;
      atomicAdd(&counters[C_PAGE_DROPS], 1u);
      return;
    }
    table[e] = ex + 1u;
    meta[ex] = sr_fine_page_meta{0u, 0u, 0u, 0u, e >> slice_shift, e & ((1u << slice_shift) - 1u), 0u, 0u};
  }
  __device__ void total(uint32_t t) const {
    counters[C_MARKED] = t;
    counters[C_PAGES] = min(t, page_cap);
  }
};

// Clears one per-cell array over the allocated pages.
__global__ __launch_bounds__(256) void k_zero_cells(FrameParams fp, uint32_t* cells) {
  const uint32_t n = fp.counters[C_PAGES] * (kPage / 4u);
  uint4* c = reinterpret_cast<uint4*>(cells);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ------------------------------------------------- K14 page prefix
// One 256-lane workgroup per page: page-local exclusive scan of the cell
// counts (wave64 DPP scans + one LDS hop across the four waves), cursor reset,
// page total and the 16-bit mask of fine tiles with work.
__global__ __launch_bounds__(256) void k_page_prefix(FrameParams fp, uint32_t* page_xy_mask) {
  __shared__ uint32_t lds4[4];
  __shared__ uint32_t colmask;
  const uint32_t P = fp.counters[C_PAGES];
  for (uint32_t page = blockIdx.x; page < P; page += gridDim.x) {
    uint32_t c = threadIdx.x;
    uint32_t cnt = fp.cell_counts[page * kPage + c];
    if (c == 0) colmask = 0u;
    uint32_t total;
    uint32_t ex = block_exclusive<uint32_t>(cnt, &total, lds4);
    fp.cell_offsets[page * kPage + c] = ex;
    fp.cell_cursors[page * kPage + c] = 0u;
    if (cnt) atomicOr(&colmask, 1u << (c & 15u));
    __syncthreads();
    if (c == 0) {
      fp.page_totals[page] = total;
      page_xy_mask[page] = colmask;
    }
    __syncthreads();
  }
}

struct PageTotalSrc {
  const uint32_t* totals;
  __device__ uint32_t load(uint32_t i) const { return totals[i]; }
};
struct PageBaseDst {
  FrameParams fp;
  __device__ void store(uint32_t p, uint32_t ex, uint32_t v) const {
    sr_fine_page_meta& pm = fp.page_meta[p];
    uint32_t flags = (v > 0u && (uint64_t)ex + v > fp.ref_cap) ? 1u : 0u;
    pm.seg_ref_base = v > 0u ? ex : 0u;
    pm.seg_ref_count = flags ? 0u : v;
    pm.work_base = 0u; pm.work_count = 0u; pm.first_depth_key = 0u;
    pm.flags = flags;
    if (flags) atomicAdd(&fp.counters[C_OVERFLOW], 1u);
  }
  __device__ void total(uint32_t t) const { fp.counters[C_REFS] = t; }
};

// The fine tiles with work of every page that kept its references become
// active (one thread per page, after the page bases).
__global__ __launch_bounds__(256) void k_page_active(FrameParams fp, const uint32_t* xy_mask) {
  const uint32_t P = fp.counters[C_PAGES];
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
    const sr_fine_page_meta pm = fp.page_meta[p];
    uint32_t m = xy_mask[p];
    if (pm.flags || m == 0u) continue;
    const uint32_t fi = frustum_for_coarse_tile(fp, pm.coarse_tile_id);
    if (fi == kInvalid) continue;
    const FrustumRec& fr = fp.fr[fi];
    const uint32_t lc = pm.coarse_tile_id - fr.d.coarse_depth_tile_base;
    const uint32_t cx = lc % fr.d.coarse_tiles_x, cy = lc / fr.d.coarse_tiles_x;
    while (m) {
      const uint32_t xy = __ffs(m) - 1u;
      m &= m - 1u;
      const uint32_t fx = cx * kExt + xy % kExt, fy = cy * kExt + xy / kExt;
      if (fx >= fr.fine_tiles_x || fy >= fr.fine_tiles_y) continue;
      fp.active_flags[pm.coarse_tile_id * kExt * kExt + xy] = 1u;
    }
  }
}

// ------------------------------------------------------ page CSR backend
// K16-K18 (prefix_page_candidates, scatter_page_candidates,
// build_fine_pages_csr; strand_prepass.wgsl:1698-2049, 2240-2507) build each
// page's candidate list with atomics and count / fill per page.  Here the
// same references come out of a stable sort, without a device-scope atomic
// (on this part those go to memory and serialise on hot pages): marking
// counts each slot's visible fine cells, a scan gives every slot its offset,
// the write pass stores one (page * 256 + cell, index position) hit per
// visible cell in slot order, a stable LSD radix sort on the page bits groups
// the hits by page (slot order kept; a page belongs to one frustum, so the
// index position names the segment), and one workgroup per page sorts its
// hits by cell (stable) and writes the references -- the canonical order
// (page, cell, slot).  The page-base scan is the segment-scatter backend's.
struct HitTotalSrc {
  const uint32_t* totals;
  __device__ uint32_t load(uint32_t i) const { return totals[i]; }
};
struct HitBaseDst {
  uint32_t* base;
  uint32_t* counters;
  __device__ void store(uint32_t s, uint32_t ex, uint32_t) const { base[s] = ex; }
  __device__ void total(uint32_t t) const { counters[C_CANDS] = t; }
};

// ---- stable LSD radix sort of (key, value) pairs, 8-bit digits
// A block sorts kRadixBlock items; wave w of its four owns the w-th quarter,
// read 64 at a time.  Lanes holding the same digit find each other with eight
// ballots (one per digit bit), so a wave ranks its items without LDS atomics
// or barriers: rank = the wave's running count of the digit + the peers in
// lower lanes.
constexpr uint32_t kRadixRounds = kRadixBlock / 256u;  // rounds of 64 per wave
constexpr uint32_t kRadixChunk = kRadixRounds * 64u;   // items per wave

__device__ __forceinline__ uint32_t radix_count(const uint32_t* counters, uint32_t cap) {
  return min(counters[C_CANDS], cap);
}

// Whether the LSD pass over key bits [shift, shift + 8) reorders anything.
// Keys are page * 256 + cell; with no dropped page every page id is below
// C_PAGES, so a pass whose digit is zero for every page would leave the
// order as it is, and it is skipped (the pass count follows the page
// capacity, which frames fill to a fraction).  The first pass always runs.
__device__ __forceinline__ bool radix_pass_needed(const uint32_t* counters, uint32_t shift) {
  if (shift <= 8u) return true;
  return counters[C_PAGE_DROPS] != 0u || counters[C_PAGES] > (1u << (shift - 8u));
}

// Lanes of the wave whose (valid) digit equals this lane's.
__device__ __forceinline__ unsigned long long digit_peers(uint32_t d, bool has) {
  unsigned long long m = __ballot(has);
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const bool set = (d >> bit) & 1u;
    const unsigned long long bb = __ballot(set);
    m &= set ? bb : ~bb;
  }
  return m;
}

// Per block holding hits: digit histogram (all keys loaded first; per round
// the wave's first two digit groups add their sizes with one LDS atomic each,
// the other lanes one by one), stored block-major (hist[b * 256 + d], one
// coalesced row per block).
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, const uint32_t* counters, uint32_t cap,
                                                    uint32_t shift, uint32_t* hist) {
  __shared__ uint32_t cnt[256];
  const uint32_t n = radix_count(counters, cap), b = blockIdx.x;
  if (b * kRadixBlock >= n || !radix_pass_needed(counters, shift)) return;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t chunk = b * kRadixBlock + w * kRadixChunk;
  if (chunk < n) {
    uint32_t key[kRadixRounds];
#pragma unroll
    for (uint32_t k = 0; k < kRadixRounds; ++k) {
      const uint32_t i = chunk + k * 64u + lane;
      key[k] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kRadixRounds; ++k) {
      bool pending = chunk + k * 64u + lane < n;
      const uint32_t d = (key[k] >> shift) & 255u;
      // the first two digit groups of the wave add their sizes at once
      // (sorted keys repeat digits); the rest count one by one
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const unsigned long long act = __ballot(pending);
        if (act == 0ull) break;
        const int first = __ffsll((long long)act) - 1;
        const uint32_t dl = (uint32_t)__builtin_amdgcn_readlane((int)d, first);  // first is wave-uniform
        const bool same = pending && d == dl;
        const unsigned long long grp = __ballot(same);
        if ((int)lane == first) atomicAdd(&cnt[dl], (uint32_t)__popcll(grp));
        if (same) pending = false;
      }
      if (pending) atomicAdd(&cnt[d], 1u);
    }
  }
  __syncthreads();
  hist[b * 256u + threadIdx.x] = cnt[threadIdx.x];
}

// Output bases of every (block, digit): base[b][d] = the items of smaller
// digits + the items of digit d in earlier blocks.  The block-major counts
// are scanned down their columns in kDigitChunks chunks of rows: chunk sums,
// one workgroup that scans the chunks per digit and adds the digit offsets,
// then each chunk writes its rows' bases.  Every access is a coalesced row.

__device__ __forceinline__ uint32_t radix_nb(const uint32_t* counters, uint32_t cap) {
  return (radix_count(counters, cap) + kRadixBlock - 1u) / kRadixBlock;
}

__global__ __launch_bounds__(256) void k_digit_chunk_sums(const uint32_t* hist, const uint32_t* counters, uint32_t cap,
                                                         uint32_t shift, uint32_t* chunk) {
  if (!radix_pass_needed(counters, shift)) return;
  const uint32_t nb = radix_nb(counters, cap), rows = (nb + kDigitChunks - 1u) / kDigitChunks;
  const uint32_t c = blockIdx.x, d = threadIdx.x, b0 = c * rows, b1 = min(nb, b0 + rows);
  uint32_t sum = 0u;
  for (uint32_t bl = b0; bl < b1; bl += 8u) {  // eight rows' loads in flight
    uint32_t v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8u; ++j) v[j] = bl + j < b1 ? hist[(bl + j) * 256u + d] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < 8u; ++j) sum += v[j];
  }
  chunk[c * 256u + d] = sum;
}

__global__ __launch_bounds__(256) void k_digit_chunk_scan(const uint32_t* counters, uint32_t shift, uint32_t* chunk) {
  __shared__ uint32_t lds4[4];
  if (!radix_pass_needed(counters, shift)) return;
  const uint32_t d = threadIdx.x;
  uint32_t run = 0u;
  uint32_t ex[kDigitChunks / 16u][16];
  for (uint32_t g = 0; g < kDigitChunks / 16u; ++g) {  // sixteen loads in flight
    uint32_t v[16];
#pragma unroll
    for (uint32_t j = 0; j < 16u; ++j) v[j] = chunk[(g * 16u + j) * 256u + d];
#pragma unroll
    for (uint32_t j = 0; j < 16u; ++j) {
      ex[g][j] = run;
      run += v[j];
    }
  }
  uint32_t total;
  const uint32_t off = block_exclusive<uint32_t>(run, &total, lds4);
  for (uint32_t g = 0; g < kDigitChunks / 16u; ++g)
#pragma unroll
    for (uint32_t j = 0; j < 16u; ++j) chunk[(g * 16u + j) * 256u + d] = ex[g][j] + off;
}

__global__ __launch_bounds__(256) void k_digit_bases(const uint32_t* hist, const uint32_t* counters, uint32_t cap,
                                                    uint32_t shift, const uint32_t* chunk, uint32_t* base) {
  if (!radix_pass_needed(counters, shift)) return;
  const uint32_t nb = radix_nb(counters, cap), rows = (nb + kDigitChunks - 1u) / kDigitChunks;
  const uint32_t c = blockIdx.x, d = threadIdx.x, b0 = c * rows, b1 = min(nb, b0 + rows);
  uint32_t run = chunk[c * 256u + d];
  for (uint32_t bl = b0; bl < b1; bl += 8u) {  // eight rows' loads, then their stores
    uint32_t v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8u; ++j) v[j] = bl + j < b1 ? hist[(bl + j) * 256u + d] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < 8u; ++j) {
      if (bl + j < b1) base[(bl + j) * 256u + d] = run;
      run += v[j];
    }
  }
}

// Scatter: ranks by ballots, then the block's items are reordered by digit in
// LDS and written out in that order, so consecutive lanes write consecutive
// addresses of one digit's output range.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* keys, const uint32_t* vals,
                                                       uint32_t* okeys, uint32_t* ovals, uint32_t* counters,
                                                       uint32_t cap, uint32_t shift, const uint32_t* bases,
                                                       uint32_t dst_pair) {
  __shared__ uint32_t lkeys[kRadixBlock];
  __shared__ uint32_t lvals[kRadixBlock];
  __shared__ uint32_t wcnt[4][256];  // per wave: running digit count, then the wave's start of each digit
  __shared__ uint32_t dstart[256];   // block-local start of each digit
  __shared__ uint32_t gbase[256];    // output start of each digit for this block
  __shared__ uint32_t lds4[4];
  const uint32_t n = radix_count(counters, cap), b = blockIdx.x, base = b * kRadixBlock;
  if (!radix_pass_needed(counters, shift)) return;
  if (b == 0u && threadIdx.x == 0u) counters[C_SORT_PAIR] = dst_pair;  // the pair now holding the sort
  if (base >= n) return;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t nvalid = min(kRadixBlock, n - base);
  const unsigned long long lower = (1ull << lane) - 1ull;
#pragma unroll
  for (int q = 0; q < 4; ++q) wcnt[q][t] = 0u;
  gbase[t] = bases[b * 256u + t];
  __syncthreads();
  const uint32_t chunk = base + w * kRadixChunk;
  uint32_t key[kRadixRounds], val[kRadixRounds], rank[kRadixRounds];
#pragma unroll
  for (uint32_t k = 0; k < kRadixRounds; ++k) {
    const uint32_t i = chunk + k * 64u + lane;
    key[k] = i < n ? keys[i] : 0u;
    val[k] = i < n ? vals[i] : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < kRadixRounds; ++k) {
    const bool has = chunk + k * 64u + lane < n;
    const uint32_t d = (key[k] >> shift) & 255u;
    const unsigned long long m = digit_peers(d, has);
    uint32_t r = 0u;
    if (has) {
      // every lane reads the count before the group's lowest lane updates it
      const uint32_t before = wcnt[w][d];
      r = before + (uint32_t)__popcll(m & lower);
      if ((m & lower) == 0ull) wcnt[w][d] = before + (uint32_t)__popcll(m);
    }
    rank[k] = r;
  }
  __syncthreads();
  const uint32_t c0 = wcnt[0][t], c1 = wcnt[1][t], c2 = wcnt[2][t], c3 = wcnt[3][t];
  uint32_t total;
  const uint32_t ex = block_exclusive<uint32_t>(c0 + c1 + c2 + c3, &total, lds4);
  wcnt[0][t] = ex;
  wcnt[1][t] = ex + c0;
  wcnt[2][t] = ex + c0 + c1;
  wcnt[3][t] = ex + c0 + c1 + c2;
  dstart[t] = ex;
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kRadixRounds; ++k) {
    if (chunk + k * 64u + lane < n) {
      const uint32_t pos = wcnt[w][(key[k] >> shift) & 255u] + rank[k];
      lkeys[pos] = key[k];
      lvals[pos] = val[k];
    }
  }
  __syncthreads();
  for (uint32_t pos = t; pos < nvalid; pos += 256u) {
    const uint32_t k = lkeys[pos], d = (k >> shift) & 255u;
    const uint32_t g = gbase[d] + (pos - dstart[d]);
    okeys[g] = k;
    ovals[g] = lvals[pos];
  }
}

// The two (key, slot) buffer pairs of the hit sort; counters[C_SORT_PAIR]
// names the one its last executed pass wrote.
struct HitPairs {
  const uint32_t* keys[2];
  const uint32_t* vals[2];
};

// Each page's hits [bounds[p], bounds[p + 1]) in the page-sorted keys: the
// lower bound of p << 8, one thread per page boundary (dropped pages' hits,
// page >= page_cap, sort after every kept page).
__global__ __launch_bounds__(256) void k_page_bounds(FrameParams fp, HitPairs hp) {
  const uint32_t* keys = hp.keys[fp.counters[C_SORT_PAIR] & 1u];
  const uint32_t P = fp.counters[C_PAGES];
  const uint32_t n = min(fp.counters[C_CANDS], fp.ref_cap);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P; p += gridDim.x * blockDim.x) {
    const uint64_t target = (uint64_t)p << 8;
    uint32_t lo = 0u, hi = n;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if ((uint64_t)keys[mid] < target) lo = mid + 1u; else hi = mid;
    }
    fp.page_bounds[p] = lo;
  }
}

// Page order for the per-page cell sort: pages by descending log2 of their
// hit count (a counting sort over 32 buckets, kOrderBlocks blocks counting
// their chunk of pages, then placing them), so the heaviest pages start
// first and none is left for the end.  The order inside a bucket does not
// matter: every page writes only its own references and cells.
__device__ __forceinline__ uint32_t page_bucket(const FrameParams& fp, uint32_t p) {
  return 31u - (uint32_t)__clz((fp.page_bounds[p + 1u] - fp.page_bounds[p]) | 1u);
}

__device__ __forceinline__ uint2 page_chunk(const FrameParams& fp) {
  const uint32_t P = fp.counters[C_PAGES], chunk = (P + kOrderBlocks - 1u) / kOrderBlocks;
  const uint32_t lo = min(P, blockIdx.x * chunk);
  return make_uint2(lo, min(P, lo + chunk));
}

__global__ __launch_bounds__(256) void k_page_order_hist(FrameParams fp) {
  __shared__ uint32_t cnt[32];
  const uint2 c = page_chunk(fp);
  if (threadIdx.x < 32u) cnt[threadIdx.x] = 0u;
  __syncthreads();
  for (uint32_t p = c.x + threadIdx.x; p < c.y; p += blockDim.x) atomicAdd(&cnt[page_bucket(fp, p)], 1u);
  __syncthreads();
  if (threadIdx.x < 32u) fp.page_hist[blockIdx.x * 32u + threadIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_page_order_scatter(FrameParams fp) {
  __shared__ uint32_t cur[32], tot[32];
  const uint2 c = page_chunk(fp);
  if (threadIdx.x < 32u) {  // bucket totals and this block's start inside its bucket
    uint32_t t = 0u, before = 0u;
#pragma unroll 8
    for (uint32_t k = 0; k < kOrderBlocks; ++k) {
      const uint32_t v = fp.page_hist[k * 32u + threadIdx.x];
      t += v;
      before += k < blockIdx.x ? v : 0u;
    }
    tot[threadIdx.x] = t;
    cur[threadIdx.x] = before;
  }
  __syncthreads();
  if (threadIdx.x < 32u) {  // heavier buckets come first
    uint32_t heavier = 0u;
    for (uint32_t k = threadIdx.x + 1u; k < 32u; ++k) heavier += tot[k];
    cur[threadIdx.x] += heavier;
  }
  __syncthreads();
  for (uint32_t p = c.x + threadIdx.x; p < c.y; p += blockDim.x) fp.page_order[atomicAdd(&cur[page_bucket(fp, p)], 1u)] = p;
}

// K14 + the final order for the page-sorted hits, one workgroup per page:
// the page's cell histogram, the page-local exclusive scan (cell counts,
// offsets, page total, mask of fine tiles with work), then a stable counting
// sort by cell that writes each hit's reference (make_fine_seg_ref,
// strand_prepass.wgsl:1313-1327) at its final position.  Hits arrive in slot
// order; kPageStep per step, wave w ranks the w-th quarter by ballots in
// rounds of 64, so each cell keeps slot order = (inst, seg) order; the step
// is reordered by cell in LDS before the references are written.  Pages are
// taken heaviest first (page_order), their hits from page_bounds.
constexpr uint32_t kPageRounds = 8u;                    // rounds of 64 per wave and step
constexpr uint32_t kPageStep = 4u * 64u * kPageRounds;  // hits per step

__global__ __launch_bounds__(256) void k_page_cells(FrameParams fp, HitPairs hp) {
  const uint32_t pair = fp.counters[C_SORT_PAIR] & 1u;
  const uint32_t* keys = hp.keys[pair];
  const uint32_t* vals = hp.vals[pair];
  __shared__ uint32_t cur[256];       // output position of each cell's next reference
  __shared__ uint32_t wcnt[4][256];   // per wave: digit counts, then its start per cell in the step
  __shared__ uint32_t lstart[256], gbase[256];
  __shared__ uint32_t lval[kPageStep];  // the step's hits in cell order
  __shared__ uint8_t lcell[kPageStep];   // and their cells
  __shared__ uint32_t lds4[4];
  __shared__ uint32_t colmask;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const unsigned long long lower = (1ull << lane) - 1ull;
  const uint32_t P = fp.counters[C_PAGES];
  for (uint32_t pi = blockIdx.x; pi < P; pi += gridDim.x) {
    const uint32_t page = fp.page_order[pi];
    const uint32_t start = fp.page_bounds[page], end = fp.page_bounds[page + 1u];
    cur[t] = 0u;
    if (t == 0) colmask = 0u;
    __syncthreads();
    for (uint32_t i = start + t; i < end; i += 256u) atomicAdd(&cur[keys[i] & 255u], 1u);
    __syncthreads();
    const uint32_t cnt = cur[t];
    uint32_t total;
    const uint32_t ex = block_exclusive<uint32_t>(cnt, &total, lds4);
    const uint32_t g = page * kPage + t;
    fp.cell_counts[g] = cnt;
    fp.cell_offsets[g] = ex;
    if (cnt) atomicOr(&colmask, 1u << (t & 15u));
    cur[t] = start + ex;  // absolute position of the cell's next reference
    __syncthreads();
    if (t == 0) {
      fp.page_totals[page] = total;
      fp.page_xy_mask[page] = colmask;
    }
    for (uint32_t b = start; b < end; b += kPageStep) {
      const uint32_t nstep = min(kPageStep, end - b);
#pragma unroll
      for (int q = 0; q < 4; ++q) wcnt[q][t] = 0u;
      __syncthreads();
      const uint32_t chunk = b + w * 64u * kPageRounds;
      uint32_t key[kPageRounds], val[kPageRounds], rank[kPageRounds];
#pragma unroll
      for (uint32_t r = 0; r < kPageRounds; ++r) {
        const uint32_t i = chunk + r * 64u + lane;
        key[r] = i < end ? keys[i] : 0u;
        val[r] = i < end ? vals[i] : 0u;
      }
#pragma unroll
      for (uint32_t r = 0; r < kPageRounds; ++r) {
        const bool has = chunk + r * 64u + lane < end;
        const uint32_t d = key[r] & 255u;
        const unsigned long long m = digit_peers(d, has);
        uint32_t rk = 0u;
        if (has) {  // every lane reads the count before the group's lowest lane updates it
          const uint32_t before = wcnt[w][d];
          rk = before + (uint32_t)__popcll(m & lower);
          if ((m & lower) == 0ull) wcnt[w][d] = before + (uint32_t)__popcll(m);
        }
        rank[r] = rk;
      }
      __syncthreads();
      // per cell: its start within the step (cell-major), each wave's start
      // after it, and the output position of the step's first hit of the cell
      const uint32_t c0 = wcnt[0][t], c1 = wcnt[1][t], c2 = wcnt[2][t], c3 = wcnt[3][t];
      uint32_t step_total;
      const uint32_t ls = block_exclusive<uint32_t>(c0 + c1 + c2 + c3, &step_total, lds4);
      wcnt[0][t] = ls;
      wcnt[1][t] = ls + c0;
      wcnt[2][t] = ls + c0 + c1;
      wcnt[3][t] = ls + c0 + c1 + c2;
      lstart[t] = ls;
      gbase[t] = cur[t];
      cur[t] += c0 + c1 + c2 + c3;
      __syncthreads();
#pragma unroll
      for (uint32_t r = 0; r < kPageRounds; ++r) {
        if (chunk + r * 64u + lane >= end) continue;
        const uint32_t d = key[r] & 255u, l = wcnt[w][d] + rank[r];
        lcell[l] = (uint8_t)d;
        lval[l] = val[r];
      }
      __syncthreads();
      // consecutive lanes write consecutive references of a cell
      for (uint32_t l = t; l < nstep; l += 256u) {
        const uint32_t d = lcell[l], pos = gbase[d] + (l - lstart[d]);
        if (pos >= fp.ref_cap) continue;
        const uint32_t v = lval[l], ii = find_instance(fp, v);  // v: the segment's index position
        fp.refs[pos] = sr_fine_seg_ref{ii, v - fp.inst[ii].pos_base, fp.slot_material[v]};
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------ canonical cell order
// Sorts every cell's references by (inst, seg).  The mirror-bitonic network
// compares i with i ^ (k - 1) and then i with i ^ j, always ascending, so a
// non-power-of-two length behaves as if padded with +inf at the end.
__device__ __forceinline__ unsigned long long ref_key(const sr_fine_seg_ref& r) {
  return ((unsigned long long)r.inst_id << 32) | r.seg_id;
}

// Sorts refs[0, n) by (inst, seg) in global memory with the whole workgroup
// (cells larger than an LDS chunk; rare).
__device__ void sort_refs_global(sr_fine_seg_ref* refs, uint32_t n) {
  uint32_t m = 1;
  while (m < n) m <<= 1;
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t partner = (j == (k >> 1)) ? (i ^ (k - 1u)) : (i ^ j);
        if (partner <= i || partner >= n) continue;
        sr_fine_seg_ref a = refs[i], b = refs[partner];
        if (ref_key(b) < ref_key(a)) { refs[i] = b; refs[partner] = a; }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
}

constexpr uint32_t kSortChunk = 2048;

// One workgroup per page.  The page's cells are cut into chunks of at most
// kSortChunk references at cell boundaries; each chunk is loaded into LDS
// with key (cell, inst, seg), sorted by one mirror-bitonic network and
// written back, which orders every cell and keeps the cell layout.
__global__ __launch_bounds__(256) void k_sort_cells(FrameParams fp) {
  __shared__ unsigned long long skey[kSortChunk];
  __shared__ uint32_t smat[kSortChunk];
  __shared__ uint32_t offs[kPage], cnts[kPage];
  __shared__ uint32_t cstart[kPage + 1];
  __shared__ uint32_t cneed[kPage];
  __shared__ uint32_t nchunks;
  const uint32_t P = fp.counters[C_PAGES];
  for (uint32_t page = blockIdx.x; page < P; page += gridDim.x) {
    const sr_fine_page_meta pm = fp.page_meta[page];
    if (pm.seg_ref_count == 0u) continue;
    __syncthreads();
    offs[threadIdx.x] = fp.cell_offsets[page * kPage + threadIdx.x];
    cnts[threadIdx.x] = fp.cell_counts[page * kPage + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t nc = 0, acc = 0, need = 0;
      cstart[0] = 0;
      for (uint32_t c = 0; c < kPage; ++c) {
        uint32_t n = cnts[c];
        if (acc > 0 && acc + n > kSortChunk) {
          cneed[nc] = need;
          cstart[++nc] = c;
          acc = 0;
          need = 0;
        }
        acc += n;
        need |= n > 1u ? 1u : 0u;
      }
      cneed[nc] = need;
      cstart[++nc] = kPage;
      nchunks = nc;
    }
    __syncthreads();
    sr_fine_seg_ref* page_refs = fp.refs + pm.seg_ref_base;
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
      if (!cneed[ch]) continue;
      const uint32_t c0 = cstart[ch], c1 = cstart[ch + 1];
      const uint32_t start = offs[c0];
      const uint32_t n = offs[c1 - 1] + cnts[c1 - 1] - start;
      if (n > kSortChunk) {  // a single cell above the chunk size
        sort_refs_global(page_refs + start, n);
        continue;
      }
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t lo = c0, hi = c1;  // offs[lo] - start <= i
        while (hi - lo > 1u) {
          uint32_t mid = (lo + hi) >> 1;
          if (offs[mid] - start <= i) lo = mid; else hi = mid;
        }
        const sr_fine_seg_ref r = page_refs[start + i];
        skey[i] = ((unsigned long long)lo << 56) | ((unsigned long long)(r.inst_id & 0xFFFFFFu) << 32) | r.seg_id;
        smat[i] = r.material_idx;
      }
      __syncthreads();
      uint32_t m = 1;
      while (m < n) m <<= 1;
      for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
          for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            uint32_t partner = (j == (k >> 1)) ? (i ^ (k - 1u)) : (i ^ j);
            if (partner <= i || partner >= n) continue;
            unsigned long long a = skey[i], bk = skey[partner];
            if (bk < a) {
              skey[i] = bk; skey[partner] = a;
              uint32_t tm = smat[i]; smat[i] = smat[partner]; smat[partner] = tm;
            }
          }
          __syncthreads();
        }
      }
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        unsigned long long k = skey[i];
        page_refs[start + i] = sr_fine_seg_ref{(uint32_t)(k >> 32) & 0xFFFFFFu, (uint32_t)k, smat[i]};
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------- K20 + K21 active tiles
struct FlagSrc {
  const uint32_t* flags;
  __device__ uint32_t load(uint32_t i) const { return flags[i]; }
};
struct ActiveDst {
  uint32_t* queue;
  uint32_t* counters;
  __device__ void store(uint32_t i, uint32_t ex, uint32_t v) const { if (v) queue[ex] = i; }
  __device__ void total(uint32_t t) const { counters[C_ACTIVE] = t; }
};

// ------------------------------------------------------ K23 emission
__device__ __forceinline__ uint32_t raster_depth_key(const sr_coarse_depth_lut_entry& e, uint32_t fz) {
  if (e.virtual_count_q == 0u) return fz;
  return e.virtual_start_q + (e.virtual_count_q * fz) / kSlices;
}

struct TileLane {
  uint32_t count, base, page, cell, pti, fz;
};

// A fine tile's 16 S depth lanes (S LUT slices of 16 fine slices), PER =
// S / 4 per thread of a wave: lane l = PER * thread + i serves depth lane
// 16 S - 1 - l, so lanes run near (large keys) to far.  A thread's PER
// depth lanes share one page (a page holds the 16 fine slices of a LUT
// slice), so its page-table and page-meta words are read once and the
// cells' counts and offsets are read together.
template <uint32_t PER>
__device__ __forceinline__ void tile_lanes(const FrameParams& fp, uint32_t tile, uint32_t lx, uint32_t ly,
                                           TileLane (&tl)[PER]) {
  constexpr uint32_t S = 4u * PER;
  const uint32_t dl0 = (S * kSlices - 1u) - threadIdx.x * PER;
  const uint32_t pti = tile * S + dl0 / kSlices;
  const uint32_t h = fp.page_table[pti];
  const bool paged = h != 0u && h != kMarked;
  const uint32_t page = paged ? h - 1u : 0u;
  uint32_t ref_base = 0u;
  bool ok = false;
  if (paged) {
    const sr_fine_page_meta& pm = fp.page_meta[page];
    ref_base = pm.seg_ref_base;
    ok = pm.seg_ref_count > 0u && pm.flags == 0u;
  }
  uint32_t n[PER], off[PER];
#pragma unroll
  for (uint32_t i = 0; i < PER; ++i) {
    const uint32_t fz = (dl0 - i) % kSlices, cell = (fz * kExt + ly) * kExt + lx;
    tl[i] = TileLane{0u, 0u, paged ? page : kInvalid, paged ? cell : 0u, pti, fz};
    n[i] = ok ? fp.cell_counts[page * kPage + cell] : 0u;
    off[i] = ok ? fp.cell_offsets[page * kPage + cell] : 0u;
  }
#pragma unroll
  for (uint32_t i = 0; i < PER; ++i) {
    tl[i].count = n[i];
    tl[i].base = ref_base + off[i];
  }
}

template <uint32_t PER>
__global__ __launch_bounds__(64) void k_emit_count(FrameParams fp) {
  const uint32_t A = fp.counters[C_ACTIVE];
  for (uint32_t a = blockIdx.x; a < A; a += gridDim.x) {
    uint32_t at = fp.active_queue[a];
    uint32_t tile = at / (kExt * kExt), lf = at % (kExt * kExt);
    TileLane tl[PER];
    tile_lanes<PER>(fp, tile, lf % kExt, lf / kExt, tl);
    uint32_t works = 0u, load = 0u;
#pragma unroll
    for (uint32_t i = 0; i < PER; ++i) {
      works += tl[i].count ? 1u : 0u;
      load += tl[i].count;
    }
    works = (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive<uint32_t>(works), 63);
    load = (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive<uint32_t>(load), 63);
    if (threadIdx.x == 0) {
      uint32_t fi = frustum_for_coarse_tile(fp, tile);
      uint32_t kind = fi == kInvalid ? 2u : fp.fr[fi].d.kind;
      bool has = works > 0u && fi != kInvalid;
      fp.tile_work[a] = make_uint4(has ? works : 0u, has && kind == 0u ? 1u : 0u, has && kind == 1u ? 1u : 0u, load);
    }
  }
}

struct TileWorkSrc {
  const uint4* w;
  __device__ U4 load(uint32_t i) const { uint4 v = w[i]; return U4{v.x, v.y, v.z, 0u}; }
};
struct TileBaseDst {
  uint4* bases;
  uint32_t* counters;
  __device__ void store(uint32_t i, U4 ex, U4) const { bases[i] = make_uint4(ex.x, ex.y, ex.z, 0u); }
  __device__ void total(U4 t) const {
    counters[C_WORK] = t.x;
    counters[C_CAM_RUNS] = t.y;
    counters[C_SHADOW_RUNS] = t.z;
  }
};

template <uint32_t PER>
__global__ __launch_bounds__(64) void k_emit_write(FrameParams fp) {
  const uint32_t A = fp.counters[C_ACTIVE];
  for (uint32_t a = blockIdx.x; a < A; a += gridDim.x) {
    uint4 tw = fp.tile_work[a];
    if (tw.x == 0u) continue;  // uniform per block
    uint4 tb = fp.tile_bases[a];
    uint32_t at = fp.active_queue[a];
    uint32_t tile = at / (kExt * kExt), lf = at % (kExt * kExt);
    uint32_t lx = lf % kExt, ly = lf / kExt;
    uint32_t fi = frustum_for_coarse_tile(fp, tile);
    const FrustumRec& fr = fp.fr[fi];
    uint32_t lc = tile - fr.d.coarse_depth_tile_base;
    uint32_t cx = lc % fr.d.coarse_tiles_x, cy = lc / fr.d.coarse_tiles_x;
    uint32_t screen_tile = (cy * kExt + ly) * fr.fine_tiles_x + (cx * kExt + lx);
    TileLane tl[PER];
    tile_lanes<PER>(fp, tile, lx, ly, tl);
    uint32_t mine = 0u;
#pragma unroll
    for (uint32_t i = 0; i < PER; ++i) mine += tl[i].count ? 1u : 0u;
    uint32_t w = tb.x + wave_inclusive<uint32_t>(mine) - mine;  // lane order
    const sr_coarse_depth_lut_entry e = fp.lut[tl[0].pti];      // the thread's one page slice
#pragma unroll
    for (uint32_t i = 0; i < PER; ++i) {
      if (!tl[i].count) continue;
      if (w < fp.work_cap)
        fp.work[w] = sr_raster_work_item{tl[i].base, tl[i].count, fi, screen_tile,
                                         raster_depth_key(e, tl[i].fz), tl[i].count, tl[i].page,
                                         tl[i].cell};
      ++w;
    }
    if (threadIdx.x == 0) {
      sr_raster_tile_run run{tb.x, tw.x, fi, screen_tile, tw.w, 0u, 0u, 0u};
      if (fr.d.kind == 1u) fp.shadow_runs[tb.z] = run; else fp.cam_runs[tb.y] = run;
    }
  }
}

// Raster run order: each kind's runs by descending log2 of their load
// (references), so a persistent raster starts the longest runs first.  A
// counting sort over the 32 buckets in two launches of kOrderBlocks blocks
// per kind: k_run_hist counts each block's chunk of runs per bucket,
// k_run_scatter places them.  The order inside a bucket does not matter:
// every run writes only its own tile.
__device__ __forceinline__ uint32_t run_bucket(uint32_t load) { return 31u - (uint32_t)__clz(load | 1u); }

struct RunChunk {
  const sr_raster_tile_run* runs;
  uint32_t* order;
  uint32_t* hist;  // [kOrderBlocks][32] bucket counts of this kind
  uint32_t blk, lo, hi;
};

__device__ __forceinline__ RunChunk run_chunk(const FrameParams& fp) {
  const bool cam = blockIdx.x < kOrderBlocks;
  RunChunk c;
  c.runs = cam ? fp.cam_runs : fp.shadow_runs;
  c.order = cam ? fp.cam_order : fp.shadow_order;
  c.hist = fp.run_hist + (cam ? 0u : kOrderBlocks * 32u);
  c.blk = blockIdx.x % kOrderBlocks;
  const uint32_t n = fp.counters[cam ? C_CAM_RUNS : C_SHADOW_RUNS];
  const uint32_t chunk = (n + kOrderBlocks - 1u) / kOrderBlocks;
  c.lo = min(n, c.blk * chunk);
  c.hi = min(n, c.lo + chunk);
  return c;
}

// Adds this wave's runs [i0, hi) to the bucket counters cnt, one LDS atomic
// per distinct bucket; with order set, also writes each run at its place
// (the lanes sharing a bucket take consecutive places).
__device__ __forceinline__ void run_bucket_add(const RunChunk& c, uint32_t i0, uint32_t* cnt, bool place) {
  const uint32_t lane = threadIdx.x & 63u, i = i0 + threadIdx.x;
  const bool valid = i < c.hi;
  const uint32_t b = valid ? run_bucket(c.runs[i].load_score) : 0u;
  uint64_t todo = __ballot(valid);
  while (todo) {
    const int first = __ffsll((long long)todo) - 1;
    const uint32_t fb = (uint32_t)__builtin_amdgcn_readlane((int)b, first);  // first is wave-uniform
    const uint64_t m = __ballot(valid && b == fb) & todo;
    uint32_t at = 0u;
    if ((int)lane == first) at = atomicAdd(&cnt[fb], (uint32_t)__popcll(m));
    if (place) {
      at = (uint32_t)__builtin_amdgcn_readlane((int)at, first);
      if ((m >> lane) & 1ull) c.order[at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    }
    todo &= ~m;
  }
}

__global__ __launch_bounds__(256) void k_run_hist(FrameParams fp) {
  __shared__ uint32_t cnt[32];
  const RunChunk c = run_chunk(fp);
  if (threadIdx.x < 32u) cnt[threadIdx.x] = 0u;
  __syncthreads();
  for (uint32_t i0 = c.lo; i0 < c.hi; i0 += blockDim.x) run_bucket_add(c, i0, cnt, false);
  __syncthreads();
  if (threadIdx.x < 32u) c.hist[c.blk * 32u + threadIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_run_scatter(FrameParams fp) {
  __shared__ uint32_t cur[32], tot[32];
  const RunChunk c = run_chunk(fp);
  if (c.lo >= c.hi) return;
  if (threadIdx.x < 32u) {  // bucket totals and this block's start inside its bucket
    uint32_t t = 0u, before = 0u;
#pragma unroll 8
    for (uint32_t k = 0; k < kOrderBlocks; ++k) {
      const uint32_t v = c.hist[k * 32u + threadIdx.x];
      t += v;
      before += k < c.blk ? v : 0u;
    }
    tot[threadIdx.x] = t;
    cur[threadIdx.x] = before;
  }
  __syncthreads();
  if (threadIdx.x < 32u) {  // heavier buckets come first
    uint32_t heavier = 0u;
    for (uint32_t k = threadIdx.x + 1u; k < 32u; ++k) heavier += tot[k];
    cur[threadIdx.x] += heavier;
  }
  __syncthreads();
  for (uint32_t i0 = c.lo; i0 < c.hi; i0 += blockDim.x) run_bucket_add(c, i0, cur, true);
}

// ------------------------------------------------------------ sequencer
static inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t max_blocks) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (uint32_t)g;
}

// ------------------------------------------------------ K1 opaque depth
// reduce_opaque_depth (opaque_depth_reduce.wgsl:35-86): one thread per camera
// fine tile, over its pixels: covered when every pixel's nearest sample
// (min over samples, reverse-Z) is > 0, and the quantised farthest of them.
__global__ __launch_bounds__(256) void k_opaque_reduce(FrameParams fp) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < fp.fine_total; g += gridDim.x * blockDim.x) {
    uint32_t fi = kInvalid;
    for (uint32_t i = 0; i < fp.F; ++i) {
      const sr_frustum_desc& d = fp.fr[i].d;
      const uint32_t n = fp.fr[i].fine_tiles_x * fp.fr[i].fine_tiles_y;
      if (g >= d.fine_depth_tile_base && g < d.fine_depth_tile_base + n) { fi = i; break; }
    }
    if (fi == kInvalid || fp.fr[fi].d.kind != 0u) {
      fp.opaque_tiles[g] = make_uint2(0u, 0u);
      continue;
    }
    const FrustumRec& fr = fp.fr[fi];
    const uint32_t t = g - fr.d.fine_depth_tile_base;
    const uint32_t x0 = (t % fr.fine_tiles_x) * fr.d.froxel_size_x, y0 = (t / fr.fine_tiles_x) * fr.d.froxel_size_y;
    const uint32_t x1 = min(x0 + fr.d.froxel_size_x, fr.d.screen_width);
    const uint32_t y1 = min(y0 + fr.d.froxel_size_y, fr.d.screen_height);
    const uint32_t S = fp.opaque_samples;
    const float* img = fp.opaque_depth + fr.image_off * S;
    uint32_t covered = 0u, total = 0u;
    float farthest = 1.0f;
    for (uint32_t y = y0; y < y1; ++y)
      for (uint32_t x = x0; x < x1; ++x) {
        ++total;
        const float* d = img + ((size_t)y * fr.d.screen_width + x) * S;
        float depth = d[0];
        for (uint32_t k = 1; k < S; ++k) depth = fminf(depth, d[k]);
        if (depth > 0.0f) {
          ++covered;
          farthest = fminf(farthest, depth);
        }
      }
    fp.opaque_tiles[g] = make_uint2((covered == total && total > 0u) ? 1u : 0u, quantize_depth01(farthest));
  }
}

// ------------------------------------------------------ K19 telemetry
// finalize_telemetry (strand_prepass.wgsl:2145-2188) plus the per-frustum
// words the reference adds along the way (:736-737, 1085, 1136, 2830-2832):
// pages, tasks, candidates, references, the log2 histogram of candidates
// per page, and per frustum visible instances / strands, retained strands,
// emitted segments, pages, candidates, references, raster runs and loads.
__device__ __forceinline__ void tel_add(const FrameParams& fp, uint32_t fi, uint32_t field, uint32_t v) {
  atomicAdd(&fp.tel[kTelPageOff + fp.page_cap + fi * kTelFrustumStride + field], v);
}
__device__ __forceinline__ void tel_max(const FrameParams& fp, uint32_t fi, uint32_t field, uint32_t v) {
  atomicMax(&fp.tel[kTelPageOff + fp.page_cap + fi * kTelFrustumStride + field], v);
}

__global__ __launch_bounds__(256) void k_telemetry(FrameParams fp) {
  const uint32_t P = min(fp.counters[C_PAGES], fp.page_cap);
  const uint32_t stride = gridDim.x * blockDim.x, gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) {
    fp.tel[0] = P;
    fp.tel[1] = fp.counters[C_TASKS];
    fp.tel[3] = fp.counters[C_REFS];
    for (uint32_t fi = 0; fi < fp.F; ++fi)
      tel_add(fp, fi, TF_EMITTED_SEGMENTS, fp.frustum_task_base[fi + 1] - fp.frustum_task_base[fi]);
  }
  for (uint32_t p = gid; p < P; p += stride) {
    const sr_fine_page_meta pm = fp.page_meta[p];
    const uint32_t fi = frustum_for_coarse_tile(fp, pm.coarse_tile_id);
    const uint32_t cand = fp.tel[kTelPageOff + p];
    if (fi != kInvalid) tel_add(fp, fi, TF_ALLOCATED_PAGES, 1u);
    if (cand == 0u) continue;
    atomicAdd(&fp.tel[2], cand);
    atomicAdd(&fp.tel[4], 1u);
    atomicMax(&fp.tel[5], cand);
    if (fi != kInvalid) {
      tel_add(fp, fi, TF_PAGE_CANDIDATES, cand);
      tel_add(fp, fi, TF_FINE_REFS, pm.seg_ref_count);
      tel_max(fp, fi, TF_MAX_PAGE_CANDIDATES, cand);
    }
    uint32_t bin = 0, rem = cand;
    while (rem > 1u && bin + 1u < kTelHistBins) { rem >>= 1; ++bin; }
    atomicAdd(&fp.tel[kTelHistOff + bin], 1u);
  }
  const uint32_t Rc = fp.counters[C_CAM_RUNS], Rs = fp.counters[C_SHADOW_RUNS];
  for (uint32_t r = gid; r < Rc + Rs; r += stride) {
    const sr_raster_tile_run run = r < Rc ? fp.cam_runs[r] : fp.shadow_runs[r - Rc];
    if (run.frustum_id >= fp.F) continue;
    tel_add(fp, run.frustum_id, TF_RASTER_RUNS, 1u);
    tel_add(fp, run.frustum_id, TF_RASTER_LOAD, run.load_score);
    tel_max(fp, run.frustum_id, TF_MAX_RASTER_LOAD, run.load_score);
  }
  // visible instances / strands and retained strands per (instance, frustum)
  for (uint32_t ii = 0; ii < fp.NI; ++ii) {
    const sr_broad_instance_meta b = fp.broad[ii];
    if (b.visible == 0u) continue;
    const uint64_t n = (uint64_t)fp.F * b.strand_count;
    for (uint64_t k = gid; k < n; k += stride) {
      const uint32_t fi = (uint32_t)(k / b.strand_count), strand = (uint32_t)(k % b.strand_count);
      const FrustumRec& fr = fp.fr[fi];
      if (strand == 0u && fp.ranges[(size_t)ii * fp.F + fi].geo_id == ii) {
        tel_add(fp, fi, TF_VISIBLE_INSTANCES, 1u);
        tel_add(fp, fi, TF_VISIBLE_STRANDS, b.strand_count);
      }
      if (fr.d.kind > 1u || (fr.d.kind == 1u && fr.light_layer == kInvalid)) continue;
      const float rnd = hash_to_unit_float(wang_hash((ii * 16777619u) ^ strand));
      if (rnd > fmaxf(b.camera_keep_probability, b.shadow_keep_probability)) continue;
      if (rnd > fp.keep[(size_t)ii * fp.F + fi]) continue;
      tel_add(fp, fi, TF_RETAINED_STRANDS, 1u);
    }
  }
}

// Raster work emission: one wave per active fine tile, PER = S / 4 depth
// lanes per thread, then the run bases, then the work items.
template <uint32_t PER>
static void launch_emit(const FrameParams& fp, const PrepassScratch& ps, hipStream_t s, uint32_t max_grid) {
  const dim3 g(grid_for(fp.coarse_total * kExt * kExt, 1, max_grid));
  hipLaunchKernelGGL(k_emit_count<PER>, g, dim3(64), 0, s, fp);
  launch_scan<U4>(TileWorkSrc{fp.tile_work}, TileBaseDst{fp.tile_bases, fp.counters}, fp.coarse_total * kExt * kExt,
                  fp.counters + C_ACTIVE, reinterpret_cast<U4*>(ps.partials4), s);
  hipLaunchKernelGGL(k_emit_write<PER>, g, dim3(64), 0, s, fp);
}

void launch_prepass(const FrameParams& fp, const PrepassScratch& ps, hipStream_t s, StageTimer* timer) {
  const uint32_t kMaxGrid = 256u * 16u;
  const uint64_t slots_total = (uint64_t)fp.F * fp.slots;
  const uint32_t ptab = fp.coarse_total * fp.slices;
  const bool csr = fp.st.fine_binning_backend != 0u;

  if (fp.opaque_tiles) {
    timer->mark(s, "opaque_reduce");
    hipLaunchKernelGGL(k_opaque_reduce, dim3(grid_for(fp.fine_total, 256, kMaxGrid)), dim3(256), 0, s, fp);
  }
  timer->mark(s, "broad");
  hipLaunchKernelGGL(k_broad, dim3(grid_for(fp.NI, 1, 4096)), dim3(64), 0, s, fp);
  timer->mark(s, "depth_lut");
  {
    const dim3 g(grid_for(fp.coarse_total, 64, 1u << 20));
    if (fp.slices == 8u) hipLaunchKernelGGL(k_lut<8>, g, dim3(64), 0, s, fp);
    else if (fp.slices == 32u) hipLaunchKernelGGL(k_lut<32>, g, dim3(64), 0, s, fp);
    else hipLaunchKernelGGL(k_lut<16>, g, dim3(64), 0, s, fp);
  }
  timer->mark(s, "project_mark");
  // one to four waves per workgroup, the same number of threads in flight
  switch (fp.mark_waves) {
    case 1u:
      hipLaunchKernelGGL(k_project_mark<1>, dim3(grid_for(slots_total, 64, 4u * kMaxGrid)), dim3(64), 0, s, fp,
                         ps.slot_proj);
      break;
    case 2u:
      hipLaunchKernelGGL(k_project_mark<2>, dim3(grid_for(slots_total, 128, 2u * kMaxGrid)), dim3(128), 0, s, fp,
                         ps.slot_proj);
      break;
    default:
      hipLaunchKernelGGL(k_project_mark<4>, dim3(grid_for(slots_total, 256, kMaxGrid)), dim3(256), 0, s, fp,
                         ps.slot_proj);
  }
  timer->mark(s, "project_compact");
  launch_scan<uint32_t>(MaskSrc{fp.slot_masks}, WaveOffsetDst{fp.wave_offsets, fp.counters},
                        (uint32_t)(slots_total / 64u), nullptr, ps.partials, s);
  hipLaunchKernelGGL(k_frustum_bases, dim3(1), dim3(64), 0, s, fp);
  timer->mark(s, "allocate");
  launch_scan<uint32_t>(PageMarkSrc{fp.page_table}, PageAllocDst{fp.page_table, fp.page_meta, fp.counters, fp.page_cap, fp.slice_shift},
                        ptab, nullptr, ps.partials, s);
  if (csr) {
    timer->mark(s, "hit_offsets");
    launch_scan<uint32_t>(HitTotalSrc{fp.slot_hit_total}, HitBaseDst{fp.hit_base, fp.counters}, (uint32_t)slots_total,
                          nullptr, ps.partials, s);
  }
  timer->mark(s, "project_write");
  // one slot per thread: the pass is latency-bound, and a thread that owns
  // one slot has no later slot waiting behind its loads
  hipLaunchKernelGGL(k_project_write, dim3(grid_for(slots_total, 256, 1u << 20)), dim3(256), 0, s, fp, ps.slot_proj);
  if (csr) {
    timer->mark(s, "sort_hits");
    // stable LSD passes over the page bits of (key, position) pairs, as many
    // as the page capacity needs; a pass above the frame's highest page digit
    // returns at once, and the page kernels read the pair the last executed
    // pass wrote (all passes executed: pair 0)
    int src = (ps.radix_passes & 1u) ? 1 : 0;  // the pair the hits were written to
    for (uint32_t k = 0; k < ps.radix_passes; ++k, src ^= 1) {
      const uint32_t shift = 8u + 8u * k;
      const int dst = k + 1u == ps.radix_passes ? 0 : (src ^ 1);
      hipLaunchKernelGGL(k_radix_hist, dim3(ps.radix_blocks), dim3(256), 0, s, ps.hkeys[src], fp.counters,
                         fp.ref_cap, shift, ps.radix_hist);
      hipLaunchKernelGGL(k_digit_chunk_sums, dim3(kDigitChunks), dim3(256), 0, s, ps.radix_hist, fp.counters,
                         fp.ref_cap, shift, ps.partials);
      hipLaunchKernelGGL(k_digit_chunk_scan, dim3(1), dim3(256), 0, s, fp.counters, shift, ps.partials);
      hipLaunchKernelGGL(k_digit_bases, dim3(kDigitChunks), dim3(256), 0, s, ps.radix_hist, fp.counters,
                         fp.ref_cap, shift, ps.partials, ps.radix_base);
      hipLaunchKernelGGL(k_radix_scatter, dim3(ps.radix_blocks), dim3(256), 0, s, ps.hkeys[src], ps.hvals[src],
                         ps.hkeys[dst], ps.hvals[dst], fp.counters, fp.ref_cap, shift, ps.radix_base, (uint32_t)dst);
    }
    timer->mark(s, "page_cells");
    const HitPairs hp{{ps.hkeys[0], ps.hkeys[1]}, {ps.hvals[0], ps.hvals[1]}};
    hipLaunchKernelGGL(k_page_bounds, dim3(grid_for((uint64_t)fp.page_cap + 1u, 256, 1u << 16)), dim3(256), 0, s, fp,
                       hp);
    hipLaunchKernelGGL(k_page_order_hist, dim3(kOrderBlocks), dim3(256), 0, s, fp);
    hipLaunchKernelGGL(k_page_order_scatter, dim3(kOrderBlocks), dim3(256), 0, s, fp);
    hipLaunchKernelGGL(k_page_cells, dim3(grid_for(fp.page_cap, 1, 1u << 20)), dim3(256), 0, s, fp, hp);
    timer->mark(s, "page_bases");
    launch_scan<uint32_t>(PageTotalSrc{fp.page_totals}, PageBaseDst{fp}, fp.page_cap,
                          fp.counters + C_PAGES, ps.partials, s);
  } else {
    hipLaunchKernelGGL(k_zero_cells, dim3(grid_for((uint64_t)fp.page_cap * 64u, 256, kMaxGrid)), dim3(256), 0, s, fp,
                       fp.cell_counts);
    timer->mark(s, "count");
    hipLaunchKernelGGL((k_visit<MODE_COUNT>), dim3(grid_for(fp.task_cap, 256, kMaxGrid)), dim3(256), 0, s, fp);
    timer->mark(s, "prefix");
    hipLaunchKernelGGL(k_page_prefix, dim3(grid_for(fp.page_cap, 1, kMaxGrid)), dim3(256), 0, s, fp, fp.page_xy_mask);
    launch_scan<uint32_t>(PageTotalSrc{fp.page_totals}, PageBaseDst{fp}, fp.page_cap,
                          fp.counters + C_PAGES, ps.partials, s);
    timer->mark(s, "fill");
    hipLaunchKernelGGL((k_visit<MODE_FILL>), dim3(grid_for(fp.task_cap, 256, kMaxGrid)), dim3(256), 0, s, fp);
    timer->mark(s, "sort_cells");
    hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(fp.page_cap, 1, kMaxGrid)), dim3(256), 0, s, fp);
  }
  timer->mark(s, "compact");
  hipLaunchKernelGGL(k_page_active, dim3(grid_for(fp.page_cap, 256, kMaxGrid)), dim3(256), 0, s, fp, fp.page_xy_mask);
  launch_scan<uint32_t>(FlagSrc{fp.active_flags}, ActiveDst{fp.active_queue, fp.counters},
                        fp.coarse_total * kExt * kExt, nullptr, ps.partials, s);
  timer->mark(s, "emit");
  if (fp.slices == 8u) launch_emit<2>(fp, ps, s, 4u * kMaxGrid);
  else if (fp.slices == 32u) launch_emit<8>(fp, ps, s, 4u * kMaxGrid);
  else launch_emit<4>(fp, ps, s, 4u * kMaxGrid);
  hipLaunchKernelGGL(k_run_hist, dim3(2 * kOrderBlocks), dim3(256), 0, s, fp);
  hipLaunchKernelGGL(k_run_scatter, dim3(2 * kOrderBlocks), dim3(256), 0, s, fp);
  if (fp.tel) {
    timer->mark(s, "telemetry");
    hipLaunchKernelGGL(k_telemetry, dim3(1024), dim3(256), 0, s, fp);
  }
}

}  // namespace srk
