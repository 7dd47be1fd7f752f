// render.hip -- deep-shadow-map build, stampback, hair shading, strand
// raster and composite on MI355X (gfx950, wave64).
//
// Reference kernels restated here (paths relative to the reference root):
//   rasterize_strands [SHADOWS]   src/shaders/strand_rasterizer.wgsl:391-711
//   stampback fragment            src/shaders/shadow_dom_stampback.wgsl:74-123
//   shade_strands                 src/shaders/strand_shading.wgsl:608-695
//   rasterize_strands [LINEAR]    src/shaders/strand_rasterizer.wgsl:855-1044
//   composite fragment            src/shaders/strand_composite.wgsl:12-33
//
// Design: one workgroup per tile run, sized to the tile (64 lanes for 8x8
// froxels, 256 for 16x16); every lane owns one pixel (all pixels of a tile
// are rasterised) and all lanes stage a batch of projected references in
// LDS, so each reference is fetched and projected once per tile.  Runs come
// from a device queue head for load balance; batches span work items, whose
// boundaries each pixel tracks for the early-out rules.  Work items arrive
// near to far; a workgroup vote ends the walk once every pixel is opaque
// (camera) or has its front depth (shadow pass 1).  Deep-shadow partial opacities live in
// LDS (pixel x layer), so no dynamic register indexing spills.  Shadow maps
// are dense per light or virtual (pages of 256^2 texels made resident on
// request, through a page table), outputs stay FP32.
#include "common.hpp"
#include "coverage.hpp"
#include "kernels.hpp"
#include "scan.hpp"

#include <algorithm>

namespace srk {

// Fills every span of a FillSpans in one launch: a flat grid-stride loop
// over the spans' 16-byte chunks (span bases are allocation bases, so
// 16-byte aligned; a span's last chunk may be partial).
__global__ void k_fill_spans(FillSpans f) {
  const uint64_t total = f.start[f.count];
  const uint4 v4 = make_uint4(f.value, f.value, f.value, f.value);
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t k = 0;
    while (k + 1u < f.count && q >= f.start[k + 1u]) ++k;
    const uint64_t w = (q - f.start[k]) * 4u;
    uint32_t* d = f.dst[k];
    if (w + 4u <= f.n[k]) {
      *reinterpret_cast<uint4*>(d + w) = v4;
    } else {
      for (uint64_t j = w; j < f.n[k]; ++j) d[j] = f.value;
    }
  }
}

void FillSpans::add(uint32_t* p, uint64_t words) {
  if (!p || words == 0 || count == kFillSpans) return;
  dst[count] = p;
  n[count] = words;
  start[count + 1] = start[count] + (words + 3u) / 4u;
  ++count;
}

void launch_fill_spans(const FillSpans& f, hipStream_t s) {
  if (f.count == 0) return;
  uint64_t g = (f.start[f.count] + 255u) / 256u;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL(k_fill_spans, dim3((uint32_t)g), dim3(256), 0, s, f);
}

__device__ __forceinline__ V3 world_v3(const InstRec& in, uint32_t vi) {
  float4 v = in.vertices[vi];
  V4 w = mat_vec(in.world_from_local, V4{v.x, v.y, v.z, 1.0f});
  return V3{w.x, w.y, w.z};
}

// ------------------------------------------------ tile-run work tables
// A workgroup owns one tile run at a time, taken from a device queue head
// (dynamic balance: run loads vary by orders of magnitude).  The run's work
// items (one per non-empty cell of the fine tile's 16 x S depth lanes: at
// most 512, near to far) are staged as an exclusive prefix of their
// reference counts, so reference batches of TPX entries span work-item
// boundaries: one barrier pair per TPX references instead of per cell.
constexpr uint32_t kRunItems = 16u * kMaxSlices;
template <int TPX>
struct RunTable {
  uint32_t pref[kRunItems + 1];  // exclusive prefix of counts; pref[n] = total
  uint32_t base[kRunItems];      // seg_ref_base per work item
  uint32_t n;
  uint32_t run;
};

// The workgroup's next tile run: tickets from `head` index `order` (runs
// sorted heaviest first, so the longest runs start early and none is left
// for the end).
template <int TPX>
__device__ bool next_run(uint32_t* head, uint32_t count, const uint32_t* order, RunTable<TPX>& t) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t k = atomicAdd(head, 1u);
    t.run = k < count ? order[k] : count;
  }
  __syncthreads();
  return t.run < count;
}

// Stages the work table of `run`; items failing `keep` get count 0.
template <int TPX, bool CHECK>
__device__ void stage_run(const FrameParams& fp, const sr_raster_tile_run& run, uint32_t fi, RunTable<TPX>& t) {
  __shared__ uint32_t lds4[4];
  const uint32_t n = min(run.work_count, kRunItems);
  constexpr uint32_t kPer = kRunItems / TPX;
  uint32_t cnt[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    uint32_t w = threadIdx.x * kPer + k;
    cnt[k] = 0u;
    if (w < n) {
      const sr_raster_work_item wi = fp.work[run.work_base + w];
      bool ok = !CHECK || (wi.frustum_id == fi && wi.screen_tile_id == run.screen_tile_id);
      cnt[k] = ok ? wi.seg_ref_count : 0u;
      t.base[w] = wi.seg_ref_base;
    }
  }
  uint32_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) local += cnt[k];
  // exclusive scan of per-thread sums across the workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = wave_inclusive<uint32_t>(local);
  if (TPX > 64) {
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    uint32_t wb = 0;
    for (int w = 0; w < wave; ++w) wb += lds4[w];
    inc += wb;
  }
  uint32_t ex = inc - local;
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    uint32_t w = threadIdx.x * kPer + k;
    if (w <= n) t.pref[w] = ex;
    ex += cnt[k];
  }
  if (threadIdx.x == 0) t.n = n;
  __syncthreads();
  if (threadIdx.x == TPX - 1) t.pref[n] = ex;  // total
  __syncthreads();
}

// Work item holding concatenated reference i (pref[w] <= i < pref[w + 1]).
template <int TPX>
__device__ __forceinline__ uint32_t item_of(const RunTable<TPX>& t, uint32_t i) {
  uint32_t lo = 0, hi = t.n;  // pref[lo] <= i < pref[hi]
  while (hi - lo > 1u) {
    uint32_t mid = (lo + hi) >> 1;
    if (t.pref[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// --------------------------------------------------------- K25 shadow
// One reference of a shadow batch, as three 16-byte LDS records so a pixel
// reads it with three broadcast loads.
template <int TPX>
struct ShadowBatch {
  float4 a[TPX];  // p0.x, p0.y, p0.z (depth01), |p1 - p0|^2 in xy
  float4 c[TPX];  // p1.x, p1.y, p1.z, absorption alpha
  float4 r[TPX];  // min radius, max radius, p1.x - p0.x, p1.y - p0.y
  float inv[TPX]; // 1 / |p1 - p0|^2, correctly rounded (0 when the length is 0)
  uint32_t item[TPX];
};

// The endpoints' raw screen coordinates of a reference's segment in frustum
// fi (world_from_local times the vertex with w = 1, then world_to_screen_raw
// of that world point, w included), as the projection pass stored them for
// the segment's slot.
__device__ __forceinline__ void ref_screen(const FrameParams& fp, uint32_t fi, const InstRec& in, uint32_t seg,
                                           float4* a, float4* c) {
  const float4* sc = fp.slot_screen + 2ull * ((size_t)fi * fp.slots + in.pos_base + seg);
  *a = sc[0];
  *c = sc[1];
}

// A reference of a shadow batch is fetched in two stages, a batch apart, so
// a batch's loads are in flight while the batch before it is rasterised:
// its record (stage 1), then its segment's screen endpoints and material
// (stage 2), which put_shadow_ref turns into the batch's LDS records.
struct ShadowRefAddr {
  sr_fine_seg_ref sr;
  uint32_t w;
  bool has;
};
struct ShadowRefData {
  float4 a, c;
  float minr, maxr, alpha;
  uint32_t w;
  bool has;
};

template <int TPX>
__device__ __forceinline__ ShadowRefAddr shadow_ref_addr(const FrameParams& fp, const RunTable<TPX>& t, uint32_t i,
                                                         uint32_t total) {
  ShadowRefAddr r{sr_fine_seg_ref{0u, 0u, 0u}, 0u, i < total};
  if (r.has) {
    r.w = item_of(t, i);
    r.sr = fp.refs[t.base[r.w] + (i - t.pref[r.w])];
  }
  return r;
}

__device__ __forceinline__ ShadowRefData shadow_ref_data(const FrameParams& fp, uint32_t fi, const ShadowRefAddr& r) {
  ShadowRefData d{make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, 0.0f, 0.0f, r.w,
                  r.has};
  if (r.has) {
    const InstRec& in = fp.inst[r.sr.inst_id];
    ref_screen(fp, fi, in, r.sr.seg_id, &d.a, &d.c);
    const sr_strand_material& m = in.materials[r.sr.material_idx];
    d.minr = m.min_radius_pixels;
    d.maxr = m.max_radius_pixels;
    d.alpha = m.absorption_color[3];
  }
  return d;
}

template <int TPX>
__device__ __forceinline__ void put_shadow_ref(const ShadowRefData& d, ShadowBatch<TPX>& b) {
  const uint32_t k = threadIdx.x;
  float minr = fmaxf(d.minr, 1e-4f);
  float z0 = normalize_depth01(d.a.z), z1 = normalize_depth01(d.c.z);
  float abx = d.c.x - d.a.x, aby = d.c.y - d.a.y;
  float l2 = abx * abx + aby * aby;  // divided per pixel: t = dot / l2 as the reference
  b.a[k] = make_float4(d.a.x, d.a.y, z0, l2);
  b.c[k] = make_float4(d.c.x, d.c.y, z1, d.alpha);
  b.r[k] = make_float4(minr, fmaxf(d.maxr, minr), abx, aby);
  b.inv[k] = l2 == 0.0f ? 0.0f : 1.0f / l2;
  b.item[k] = d.w;
}

// fragment_position_line_relative + coverage (strand_rasterizer.wgsl:138-147,
// 553-566); p = mix(p0, p1, t) evaluated as p0 * (1 - t) + p1 * t.
template <int TPX>
__device__ __forceinline__ bool shadow_cover(const ShadowBatch<TPX>& b, uint32_t k, float px, float py, float* pz,
                                             float* cov) {
  // t = num / l2 correctly rounded, from the reference's reciprocal: q0 is
  // within an ulp of the quotient, num - q0 l2 is exact (FMA) and q0 + that
  // remainder / l2 rounds to the correctly rounded quotient (Markstein).  A
  // zero-length segment (inv = 0) gets t = 0, as in the reference.  Every
  // candidate runs the whole test and one predicate decides it: the lanes
  // of a wave test different references, so early exits only added branches.
  const float4 A = b.a[k], Cc = b.c[k], Rr = b.r[k];
  const float num = (px - A.x) * Rr.z + (py - A.y) * Rr.w;
  const float inv = b.inv[k];
  const float q0 = num * inv;
  const float t = __builtin_fmaf(__builtin_fmaf(-q0, A.w, num), inv, q0);
  const bool in = t >= 0.0f && t <= 1.0f;
  float qx = wmix(A.x, Cc.x, t), qy = wmix(A.y, Cc.y, t), qz = wmix(A.z, Cc.z, t);
  float r = wmix(Rr.x, Rr.y, wclamp(qz, 0.0f, 1.0f));
  float dx = px - qx, dy = py - qy;
  const float s2 = dx * dx + dy * dy;
  float q = __builtin_amdgcn_sqrtf(s2) * __builtin_amdgcn_rcpf(r);
  if (in && !(q < 0.99999f || q > 1.00001f)) q = sqrtf(s2) / r;
  float c = wclamp(1.0f - q, 0.0f, 1.0f);
  *pz = qz;
  *cov = c;
  return in && c > 0.0f;
}

// The value lane l ^ SH holds, with every lane of the wave active (DPP
// within a row for SH < 16; the gfx950 permlane swaps across rows): no LDS
// round trip.  quad_perm [1,0,3,2] and [2,3,0,1] exchange within a quad;
// row_half_mirror (lane i of an 8-lane half takes 7 - i) then quad_perm
// [3,2,1,0] (i takes i ^ 3) gives i ^ 4; row_ror:8 gives i ^ 8 in a 16-lane
// row.  permlane16_swap(v, v) swaps the odd 16-lane rows of its first
// operand with the even rows of its second, so the first result holds, in
// the odd rows, the rows below them and the second, in the even rows, the
// rows above (permlane32_swap likewise with halves of the wave).
template <uint32_t SH>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, uint32_t lane) {
  if constexpr (SH == 1u) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
  } else if constexpr (SH == 2u) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
  } else if constexpr (SH == 4u) {
    const int m = __builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp(m, 0x1B, 0xF, 0xF, false);
  } else if constexpr (SH == 8u) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
  } else if constexpr (SH == 16u) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16u) ? r[0] : r[1];
  } else {
    static_assert(SH == 32u, "a lane distance of the wave");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32u) ? r[0] : r[1];
  }
}

// One butterfly stage of the transpose: lanes l and l ^ SH swap the
// off-diagonal SH-bit blocks (selected by KEEP in the lower lane) of their
// 64-bit rows, without branches.
template <uint32_t SH, unsigned long long KEEP>
__device__ __forceinline__ unsigned long long transpose_stage(unsigned long long x, uint32_t lane) {
  const uint32_t ylo = lane_xor<SH>((uint32_t)x, lane), yhi = lane_xor<SH>((uint32_t)(x >> 32), lane);
  const unsigned long long y = ((unsigned long long)yhi << 32) | ylo;
  const unsigned long long upper = 0ull - (unsigned long long)((lane >> __builtin_ctz(SH)) & 1u);  // all ones in the upper lane
  const unsigned long long mk = KEEP ^ upper;  // x's bits that stay: KEEP (lower lane) or ~KEEP (upper lane)
  const unsigned long long t = ((y >> SH) & upper) | ((y << SH) & ~upper);
  return (x & mk) | (t & ~mk);
}

// 64 x 64 bit-matrix transpose across the wave: lane l holds row l on entry
// and column l on return (bit j = bit l of lane j's row).  Six butterfly
// stages swap the off-diagonal blocks of halving size.  Every lane of the
// wave must be active.
__device__ __forceinline__ unsigned long long wave_transpose64(unsigned long long x) {
  const uint32_t l = threadIdx.x & 63u;
  x = transpose_stage<32u, 0x00000000FFFFFFFFull>(x, l);
  x = transpose_stage<16u, 0x0000FFFF0000FFFFull>(x, l);
  x = transpose_stage<8u, 0x00FF00FF00FF00FFull>(x, l);
  x = transpose_stage<4u, 0x0F0F0F0F0F0F0F0Full>(x, l);
  x = transpose_stage<2u, 0x3333333333333333ull>(x, l);
  return transpose_stage<1u, 0x5555555555555555ull>(x, l);
}

// The hardware reciprocal and square root (within 1 ulp) for the span setup.
struct HwMath {
  __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  __device__ static float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// Per pixel of the tile, one bit per reference of the batch that may cover
// it (cols[p][w]: references w*64 .. w*64+63): each lane builds its
// reference's conservative pixel mask from row spans, and a wave transpose
// turns the 64 masks into per-pixel columns.
template <int TPX>
__device__ __forceinline__ void batch_columns(bool valid, float x0, float y0, float abx, float aby, float rmax,
                                              uint32_t tx0, uint32_t ty0, uint32_t fx, uint32_t fy,
                                              unsigned long long (*cols)[TPX / 64]) {
  constexpr int W = TPX / 64;
  const uint32_t k = threadIdx.x, w = k >> 6, l = k & 63u;
  uint64_t m[W];
  if (valid) {
    span_mask<W>(span_setup<HwMath>(x0, y0, abx, aby, rmax, tx0, ty0, fy), fx, m);
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) m[j] = 0ull;
  }
#pragma unroll
  for (int j = 0; j < W; ++j) cols[j * 64 + l][w] = wave_transpose64(m[j]);
}

template <int TPX>
__device__ __forceinline__ void shadow_columns(const ShadowBatch<TPX>& b, uint32_t nb, uint32_t tx0, uint32_t ty0,
                                               uint32_t fx, uint32_t fy, unsigned long long (*cols)[TPX / 64]) {
  const uint32_t k = threadIdx.x;
  const bool valid = k < nb;
  float4 A = make_float4(0.0f, 0.0f, 0.0f, 0.0f), Rr = A;
  float rb = 0.0f;
  if (valid) {
    A = b.a[k];
    Rr = b.r[k];
    rb = span_radius(Rr.x, Rr.y, fmaxf(A.z, b.c[k].z));
  }
  batch_columns<TPX>(valid, A.x, A.y, Rr.z, Rr.w, rb, tx0, ty0, fx, fy, cols);
}

// K25 (strand_shadow_raster.wgsl): per pixel of a light tile, pass 1 finds
// the front depth z0 (max depth over the covering references of the nearest
// work item with a hit), pass 2 accumulates each covering reference's alpha
// into the two DOM layers around its relative depth (per-pixel partials in
// LDS, indexed by layer).
template <int TPX>
__global__ __launch_bounds__(TPX) void k_shadow_raster(FrameParams fp) {
  __shared__ RunTable<TPX> t;
  __shared__ ShadowBatch<TPX> b;
  extern __shared__ float part[];  // [dom_layers * TPX], layer-major: layer i of pixel p at part[i * TPX + p]
  __shared__ unsigned long long cols[TPX][TPX / 64];
  const uint32_t R = fp.counters[C_SHADOW_RUNS];
  const uint32_t L = fp.dom_layers;
  // this workgroup's slice of the column cache: kKeep batches of TPX pixels
  // x TPX / 64 words; pixel p's words of batch k at [k][p][.]
  constexpr uint32_t kWords = TPX / 64u;
  const uint32_t kKeep = (uint32_t)(kShadowColBytes / (8ull * TPX * kWords * gridDim.x));
  unsigned long long* keep = fp.shadow_cols + (size_t)blockIdx.x * kKeep * TPX * kWords;
  while (next_run<TPX>(&fp.counters[C_SHADOW_HEAD], R, fp.shadow_order, t)) {
    const sr_raster_tile_run run = fp.shadow_runs[t.run];
    const uint32_t fi = run.frustum_id;
    const FrustumRec& fr = fp.fr[fi];
    if (fr.light_layer == kInvalid || run.work_count == 0u) continue;
    if ((uint64_t)run.work_base + run.work_count > fp.work_cap) continue;  // emission overflowed
    stage_run<TPX, true>(fp, run, fi, t);
    const uint32_t total = t.pref[t.n];
    const uint32_t fx = fr.d.froxel_size_x, fy = fr.d.froxel_size_y;
    const uint32_t tx = run.screen_tile_id % fr.fine_tiles_x, ty = run.screen_tile_id / fr.fine_tiles_x;
    const uint32_t pidx = threadIdx.x;
    const uint32_t pxu = tx * fx + pidx % fx, pyu = ty * fy + pidx / fx;
    // a pixel whose virtual DOM page is not resident is neither rasterised nor stored
    DomTexel dt{0, 0, 0};
    const bool active =
        pidx < fx * fy && pxu < fr.d.screen_width && pyu < fr.d.screen_height && dom_texel(fp, fr, pxu, pyu, &dt);
    const float pfx = (float)pxu + 0.5f, pfy = (float)pyu + 0.5f;
    // pass 1
    float z0 = -1.0f;
    bool found = !active, hit = false;
    uint32_t cur = 0, kept = 0;
    // (references a pixel's column leaves out do not cover it: skipping them
    // changes no result, item boundaries included)
    ShadowRefAddr ra = shadow_ref_addr(fp, t, threadIdx.x, total);
    ShadowRefData rd = shadow_ref_data(fp, fi, ra);
    ra = shadow_ref_addr(fp, t, TPX + threadIdx.x, total);
    for (uint32_t bs = 0; bs < total; bs += TPX) {
      __syncthreads();
      if (rd.has) put_shadow_ref(rd, b);
      __syncthreads();
      const uint32_t nb = min((uint32_t)TPX, total - bs), bi = bs / TPX;
      shadow_columns<TPX>(b, nb, tx * fx, ty * fy, fx, fy, cols);
      __syncthreads();
      if (bi < kKeep) {  // each pixel keeps its own column words for pass 2
#pragma unroll
        for (uint32_t cw = 0; cw < kWords; ++cw) keep[((size_t)bi * TPX + pidx) * kWords + cw] = cols[pidx][cw];
        kept = bi + 1u;
      }
      rd = shadow_ref_data(fp, fi, ra);                             // the next batch
      ra = shadow_ref_addr(fp, t, bs + 2u * TPX + threadIdx.x, total);  // the one after it
      for (uint32_t cw = 0; cw < (uint32_t)TPX / 64u && !found; ++cw) {
        for (unsigned long long bits = cols[pidx][cw]; bits;) {
          const uint32_t k = cw * 64u + (uint32_t)(__ffsll((long long)bits) - 1);
          bits &= bits - 1ull;
          uint32_t w = b.item[k];
          if (w != cur) {
            if (hit) { found = true; break; }
            cur = w;
          }
          float pz, cov;
          if (shadow_cover(b, k, pfx, pfy, &pz, &cov)) { z0 = fmaxf(z0, pz); hit = true; }
        }
      }
      if (__syncthreads_count(!found) == 0) break;
    }
    if (hit) found = true;
    found = found && active;
    float* part_px = part + threadIdx.x;
    for (uint32_t i = 0; i < L; ++i) part_px[i * TPX] = 0.0f;
    const float inv_span = 1.0f / fmaxf(1e-6f, found ? z0 : 1.0f);
    if (__syncthreads_count(found) != 0) {
      ShadowRefAddr ra = shadow_ref_addr(fp, t, threadIdx.x, total);
      ShadowRefData rd = shadow_ref_data(fp, fi, ra);
      ra = shadow_ref_addr(fp, t, TPX + threadIdx.x, total);
      for (uint32_t bs = 0; bs < total; bs += TPX) {
        __syncthreads();
        if (rd.has) put_shadow_ref(rd, b);
        __syncthreads();
        const uint32_t nb = min((uint32_t)TPX, total - bs), bi = bs / TPX;
        if (bi >= kept) shadow_columns<TPX>(b, nb, tx * fx, ty * fy, fx, fy, cols);
        __syncthreads();
        unsigned long long kw[kWords];  // this pixel's column words, read before the next batch's loads
#pragma unroll
        for (uint32_t cw = 0; cw < kWords; ++cw)
          kw[cw] = bi < kept ? keep[((size_t)bi * TPX + pidx) * kWords + cw] : cols[pidx][cw];
        rd = shadow_ref_data(fp, fi, ra);
        ra = shadow_ref_addr(fp, t, bs + 2u * TPX + threadIdx.x, total);
        if (found) {
#pragma unroll
          for (uint32_t cw = 0; cw < kWords; ++cw)
          for (unsigned long long bits = kw[cw]; bits;) {
            const uint32_t k = cw * 64u + (uint32_t)(__ffsll((long long)bits) - 1);
            bits &= bits - 1ull;
            float pz, cov;
            if (!shadow_cover(b, k, pfx, pfy, &pz, &cov)) continue;
            float dzp = fmaxf(0.0f, z0 - pz) * inv_span;
            float u = wclamp(dzp, 0.0f, 1.0f);  // pow(u, DOM_GAMMA = 1)
            float tL = u * (float)L;
            uint32_t si = min((uint32_t)floorf(tL), L - 1u);  // tL in [0, L]: u is clamped, never NaN
            float wgt = tL - floorf(tL);
            const float a = cov * b.c[k].w * 0.5f;
            const float a0 = part_px[si * TPX];
            part_px[si * TPX] = wclamp(a0 + (1.0f - a0) * (1.0f - wgt) * a, 0.0f, 1.0f);
            if (si + 1u < L) {
              const float a1 = part_px[(si + 1u) * TPX];
              part_px[(si + 1u) * TPX] = wclamp(a1 + (1.0f - a1) * wgt * a, 0.0f, 1.0f);
            }
          }
        }
      }
    }
    if (active) {
      fp.dom_depth[dt.depth] = found ? z0 : 0.0f;
      float acc = 0.0f;
      for (uint32_t i = 0; i < L; ++i) {
        acc = acc + (1.0f - acc) * part_px[i * TPX];
        fp.dom_opacity[dt.opacity + i * dt.stride] = found ? acc : 0.0f;
      }
    }
  }
}

// ------------------------------------------------------ K26 stampback
__global__ void k_stampback(FrameParams fp, uint32_t fi) {
  const FrustumRec& fr = fp.fr[fi];
  const uint32_t tw = fr.stamp_w, th = fr.stamp_h;
  const uint32_t W = fr.d.screen_width, H = fr.d.screen_height;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < tw * th; i += gridDim.x * blockDim.x) {
    uint32_t x = i % tw, y = i / tw;
    uint32_t x0 = min((x * W) / tw, W - 1u), y0 = min((y * H) / th, H - 1u);
    uint32_t x1 = min(((x + 1u) * W + tw - 1u) / tw, W), y1 = min(((y + 1u) * H + th - 1u) / th, H);
    float z = 0.0f;
    for (uint32_t yy = y0; yy < y1; ++yy)
      for (uint32_t xx = x0; xx < x1; ++xx) {
        DomTexel dt;  // a texel of a page that is not resident reads 0 (sample_dom_depth, :58-72)
        if (dom_texel(fp, fr, xx, yy, &dt)) z = fmaxf(z, fp.dom_depth[dt.depth]);
      }
    fp.stamp[fr.stamp_off + i] = z;
  }
}

// --------------------------------------------- virtual DOM pages (VSMS)
// request_shadow_dom_pages (strand_prepass.wgsl:2537-2579): each shadow tile
// run requests the pages its pixels fall in and their neighbours within
// kDomBorder pages of the light's page grid.  Every writer stores 1, so
// plain stores serve (no atomics).
__global__ __launch_bounds__(256) void k_dom_requests(FrameParams fp) {
  const uint32_t n = min(fp.counters[C_SHADOW_RUNS], fp.coarse_total * kExt * kExt);
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const sr_raster_tile_run run = fp.shadow_runs[r];
    if (run.frustum_id >= fp.F || run.work_count == 0u) continue;
    const FrustumRec& fr = fp.fr[run.frustum_id];
    if (fr.d.kind != 1u || fr.light_layer == kInvalid) continue;
    const uint32_t tx = run.screen_tile_id % fr.fine_tiles_x, ty = run.screen_tile_id / fr.fine_tiles_x;
    const uint32_t x0 = tx * fr.d.froxel_size_x, y0 = ty * fr.d.froxel_size_y;
    if (x0 >= fr.d.screen_width || y0 >= fr.d.screen_height) continue;
    const uint32_t x1 = min(x0 + fr.d.froxel_size_x - 1u, fr.d.screen_width - 1u);
    const uint32_t y1 = min(y0 + fr.d.froxel_size_y - 1u, fr.d.screen_height - 1u);
    const uint32_t px0 = x0 / kDomPage, py0 = y0 / kDomPage, px1 = x1 / kDomPage, py1 = y1 / kDomPage;
    const uint32_t bx0 = px0 - min(px0, kDomBorder), by0 = py0 - min(py0, kDomBorder);
    const uint32_t bx1 = min(px1 + kDomBorder, fr.dom_pages_x - 1u), by1 = min(py1 + kDomBorder, fr.dom_pages_y - 1u);
    for (uint32_t py = by0; py <= by1; ++py)
      for (uint32_t px = bx0; px <= bx1; ++px) fp.dom_requests[fr.dom_vpage_base + py * fr.dom_pages_x + px] = 1u;
  }
}

// Residency, the demand-driven mode of the reference's VSMS allocator
// (src/plugin.rs:1776-1854): light by light in frustum order, each light's
// requested pages in page order, at most st.dom_resident_pages of them (0:
// all), take the pool's next physical pages; the rest, and pages past the
// pool capacity (counted, so the frame reports its capacity exceeded), are
// not resident.  One workgroup (a few thousand pages at most).
__global__ __launch_bounds__(256) void k_dom_residency(FrameParams fp) {
  __shared__ uint32_t lds4[4];
  const uint32_t budget = fp.st.dom_resident_pages;
  uint32_t want = 0;  // pages within their light's budget so far (block-uniform)
  for (uint32_t fi = 0; fi < fp.F; ++fi) {
    const FrustumRec& fr = fp.fr[fi];
    if (fr.d.kind != 1u) continue;
    const uint32_t n = fr.dom_pages_x * fr.dom_pages_y;
    uint32_t rank = 0;  // this light's requested pages so far
    for (uint32_t b = 0; b < n; b += 256u) {
      const uint32_t i = b + threadIdx.x, v = fr.dom_vpage_base + i;
      const uint32_t req = i < n && fp.dom_requests[v] != 0u ? 1u : 0u;
      uint32_t total;
      const uint32_t ex = block_exclusive<uint32_t>(req, &total, lds4);
      if (i < n) {
        const bool kept = req && (budget == 0u || rank + ex < budget);
        fp.dom_table[v] = kept && want + ex < fp.dom_pool_cap ? want + ex : kInvalid;
      }
      want += budget == 0u ? total : min(total, budget - min(rank, budget));
      rank += total;
    }
  }
  if (threadIdx.x == 0) {
    fp.counters[C_DOM_WANT] = want;
    fp.counters[C_DOM_PAGES] = min(want, fp.dom_pool_cap);
    fp.counters[C_DOM_DROPS] = want > fp.dom_pool_cap ? want - fp.dom_pool_cap : 0u;
  }
}

// --------------------------------------------------------- K27 shading
constexpr float kPiF = 3.14159265359f;
#define PI_HALF_LO ((float)(-3.14159265359 / 2.0 + 1e-6))
#define PI_HALF_HI ((float)(3.14159265359 / 2.0 - 1e-6))
#define TWO_PI_F ((float)(2.0 * 3.14159265359))
#define SQRT_2PI_F ((float)2.5066282746310002)
#define DELTA_TOL_F ((float)(0.1 * 3.14159265359 / 180.0))

__constant__ float kLegRoots[10] = {-0.9739065285171717f, -0.8650633666889844f, -0.6794095682990244f,
                                    -0.4333953941292472f, -0.14887433898163116f, 0.14887433898163116f,
                                    0.4333953941292472f, 0.6794095682990244f, 0.8650633666889844f,
                                    0.9739065285171717f};
__constant__ float kLegWeights[10] = {0.06667134430868714f, 0.14945134915058053f, 0.21908636251598224f,
                                      0.26926671930999674f, 0.2955242247147533f, 0.2955242247147533f,
                                      0.26926671930999674f, 0.21908636251598224f, 0.14945134915058053f,
                                      0.06667134430868714f};

__device__ __forceinline__ V3 v3sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 v3scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 v3cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// The BSDF runs at the shading tolerance, not bit for bit: it uses the
// hardware reciprocal, square root and exp (about 1 ulp) and sin / cos
// (absolute error near 2^-21 for the angles here).  The atlas differs from
// the correctly rounded evaluation by far less than its 1e-3 parity bound;
// visibility (which picks DOM texels and slices) stays exact.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fsin(float x) { return __sinf(x); }
__device__ __forceinline__ float fcos(float x) { return __cosf(x); }
__device__ __forceinline__ float fexp(float x) { return __expf(x); }

// (normalisation stays correctly rounded: acos of a dot product near +-1
// would magnify its error)
__device__ __forceinline__ V3 v3norm(V3 a) {
  float l = sqrtf(v3dot(a, a));
  return V3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float sacos(float x) { return acosf(wclamp(x, -1.0f, 1.0f)); }

__device__ float fresnel_reflectance(float eta_t, float c_in) {  // eta_i = 1
  float c = fabsf(c_in);
  float cc = wclamp(0.0f, 1.0f, c);
  float sin2 = 1.0f - cc * cc;
  float ratio = __builtin_amdgcn_rcpf(eta_t);
  float cos_t = fsqrt(1.0f - ratio * ratio * sin2);
  float rpar = fdiv((eta_t * c) - (1.0f * cos_t), (eta_t * c) + (1.0f * cos_t));
  float rper = fdiv((1.0f * c) - (eta_t * cos_t), (1.0f * c) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) * 0.5f;
}

// sum_{i<5} x^(2i) / (4^i (i!)^2), bessel_first_approx(x, 5)
__device__ __forceinline__ float bessel5(float x) {
  float x2 = x * x;
  float y = 1.0f;
  float term = 1.0f;
  const float inv_den[5] = {1.0f, 1.0f / 4.0f, 1.0f / 64.0f, 1.0f / 2304.0f, 1.0f / 147456.0f};  // 1 / (4^i (i!)^2)
  float xp = 1.0f;
#pragma unroll
  for (int i = 1; i < 5; ++i) {
    xp = xp * x2;
    term = xp * inv_den[i];
    y = y + term;
  }
  return y;
}

__device__ float Mp(float v, float theta_i_, float theta_r, float shift) {
  float trs = theta_r + shift;
  float theta_i = wclamp(PI_HALF_LO, PI_HALF_HI, theta_i_);
  trs = wclamp(PI_HALF_LO, PI_HALF_HI, trs);
  float cti = fcos(-theta_i), ctr = fcos(trs);
  float sti = fsin(-theta_i), str = fsin(trs);
  if (v < 1e-6f) return fabsf(trs + theta_i) < DELTA_TOL_F ? 0.0f : 1e6f / DELTA_TOL_F;
  const float iv = __builtin_amdgcn_rcpf(v);
  return bessel5((cti * ctr) * iv) * fexp((sti * str) * iv);
}

// Wrapped Gaussian sum_{k=-10..10} exp(-(d - 2 pi k)^2 / (2 var)) over the
// k with |d - 2 pi k| <= r = sqrt(32 var): every term left out is below
// e^-16 (1e-7 of the largest term).  For r < pi that is at most one term,
// the k nearest d / (2 pi), and for r < 2 pi at most two; both are
// evaluated without a loop.
__device__ __forceinline__ float wrapped_gaussian(float pd, float neg_inv_2var, float r) {
  const float inv_2pi = (float)(1.0 / (2.0 * 3.14159265359));
  if (r < kPiF) {
    const float k = fminf(10.0f, fmaxf(-10.0f, rintf(pd * inv_2pi)));
    const float ang = pd - TWO_PI_F * k;
    return fabsf(ang) <= r ? fexp(ang * ang * neg_inv_2var) : 0.0f;
  }
  int k0 = max(-10, (int)ceilf((pd - r) * inv_2pi)), k1 = min(10, (int)floorf((pd + r) * inv_2pi));
  if (r < TWO_PI_F) {  // at most two terms, k0 and k0 + 1, summed as the loop below sums them
    const float a0 = pd - TWO_PI_F * (float)k0, a1 = pd - TWO_PI_F * (float)(k0 + 1);
    const float t0 = k0 <= k1 ? fexp(a0 * a0 * neg_inv_2var) : 0.0f;
    const float t1 = k0 + 1 <= k1 ? fexp(a1 * a1 * neg_inv_2var) : 0.0f;
    return t0 + t1;
  }
  float total = 0.0f;
  for (int k = k0; k <= k1; ++k) {
    float ang = pd - TWO_PI_F * (float)k;
    total = total + fexp(ang * ang * neg_inv_2var);
  }
  return total;
}

// Per Gauss-Legendre point: the clamped offset h, gamma_i = asin(h), its
// cosine and 2 gamma_i -- the same for every segment, computed once per
// workgroup with the functions the per-segment code would call.
struct LegPoint {
  float h, gamma_i, cos_gamma_i, ai;
};

__device__ __forceinline__ LegPoint leg_point(int i) {
  const float h = wclamp(-1.0f + 1e-6f, 1.0f - 1e-6f, kLegRoots[i]);
  const float gi = asinf(h);
  return LegPoint{h, gi, cosf(gi), 2.0f * asinf(h)};
}

// The three Np lobes of weta_strand_bsdf (strand_shading.wgsl:525-577) in
// one pass over the 10 quadrature points: the per-point terms of A and Phi
// (gamma_i, gamma_t, Fresnel, absorption) are shared by the three paths.
__device__ void np_lobes(float phi, float theta_i, float theta_r, float eta, V3 mu, float beta, float scat,
                         const LegPoint* leg, V3* out) {
  float theta_d = (theta_r - theta_i) * 0.5f;
  float sti = fsin(theta_i);
  float ctd = fcos(theta_d);
  const float ictd = __builtin_amdgcn_rcpf(ctd);
  float eta_prime = fsqrt(fmaxf(0.0f, eta * eta - sti * sti)) * ictd;
  const float ieta = __builtin_amdgcn_rcpf(eta_prime);
  V3 mup{mu.x * ictd, mu.y * ictd, mu.z * ictd};
  float var = beta * beta;
  bool narrow = var < 1e-6f;
  float norm = __builtin_amdgcn_rcpf(SQRT_2PI_F * beta);
  float neg_inv_2var = -__builtin_amdgcn_rcpf(2.0f * var);
  // a term with |d - 2 pi k| above r = sqrt(16 * 2 var) is below e^-16
  const float r_win = sqrtf(32.0f * var);
  V3 acc[3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
  for (int i = 0; i < 10; ++i) {
    const LegPoint lp = leg[i];
    const float x = lp.h * ieta;
    float gamma_t = asinf(x);
    float f = fresnel_reflectance(eta, sacos(ctd * lp.cos_gamma_i));
    // 1 + cos(2 asin x) = 2 - 2 x^2 while x is in asin's domain
    float e = fabsf(x) <= 1.0f ? 2.0f - 2.0f * (x * x) : 1.0f + fcos(2.0f * gamma_t);
    V3 T{fexp(-2.0f * mup.x * e), fexp(-2.0f * mup.y * e), fexp(-2.0f * mup.z * e)};
    // asin(clamp(x, -1, 1)) is gamma_t whenever x is already in range
    float gt_phi = fabsf(x) <= 1.0f ? gamma_t : asinf(wclamp(x, -1.0f, 1.0f));
    float ai = lp.ai;
    float f_ = 1.0f - f;
    float wgt = kLegWeights[i];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      V3 att;
      if (p == 0) {
        att = V3{f * T.x * T.x, f * T.y * T.y, f * T.z * T.z};
      } else if (p == 1) {
        float s = f_ * f_;
        att = V3{s * T.x, s * T.y, s * T.z};
      } else {
        float s = (f_ * f_) * f;
        att = V3{s * T.x * T.x, s * T.y * T.y, s * T.z * T.z};
      }
      float shift = 2.0f * (float)p * gt_phi - ai + (float)p * kPiF;
      float pd = phi - shift;
      float det;
      if (narrow) det = fabsf(fmodf(pd, TWO_PI_F)) < 1e-3f ? 0.0f : 1e6f;
      else det = norm * wrapped_gaussian(pd, neg_inv_2var, r_win);
      const float dw = det * wgt;
      acc[p].x = acc[p].x + att.x * dw;
      acc[p].y = acc[p].y + att.y * dw;
      acc[p].z = acc[p].z + att.z * dw;
    }
  }
  V3 total{0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    float s = ((float)(p + 1) * scat) * (1.0f / 3.0f);
    total.x = total.x + (acc[p].x * 0.5f) * s;
    total.y = total.y + (acc[p].y * 0.5f) * s;
    total.z = total.z + (acc[p].z * 0.5f) * s;
  }
  *out = total;
}

__device__ V3 marschner(V3 dir, V3 view_n, V3 light_n, const sr_strand_material& m, float occlusion,
                        const LegPoint* leg) {
  float theta_i = sacos(v3dot(dir, light_n));
  float theta_r = sacos(v3dot(dir, view_n));
  V3 lp = v3norm(v3sub(light_n, v3scale(dir, v3dot(light_n, dir))));
  V3 vp = v3norm(v3sub(view_n, v3scale(dir, v3dot(view_n, dir))));
  float sg = v3dot(v3cross(lp, vp), dir);
  float phi = sacos(v3dot(v3norm(lp), v3norm(vp))) * (sg > 0.0f ? 1.0f : (sg < 0.0f ? -1.0f : 0.0f));
  V3 sigma{1.0f - m.absorption_color[0], 1.0f - m.absorption_color[1], 1.0f - m.absorption_color[2]};
  float v_long = m.alpha * m.alpha, v_azim = m.beta * m.beta;
  float beta_azim = fsqrt(v_azim);
  float scat = wclamp(1.0f - occlusion, 0.0f, 1.0f);
  float direct = wmix(0.18f, 1.0f, scat);
  float mp = Mp(v_long, theta_i, theta_r, m.shift);
  V3 total{mp * m.specular_color[0] * m.specular_color[3] * direct, mp * m.specular_color[1] * m.specular_color[3] * direct,
           mp * m.specular_color[2] * m.specular_color[3] * direct};
  V3 lobes;
  np_lobes(phi, theta_i, theta_r, m.eta, sigma, beta_azim, scat, leg, &lobes);
  total.x = total.x + lobes.x;
  total.y = total.y + lobes.y;
  total.z = total.z + lobes.z;
  return total;
}

__device__ float dom_visibility(const FrameParams& fp, V3 pw, uint32_t layer) {
  uint32_t fi = kInvalid, l = 0;
  for (uint32_t i = 0; i < fp.F; ++i) {
    if (fp.fr[i].d.kind != 1u) continue;
    if (l == layer) { fi = i; break; }
    ++l;
  }
  if (fi == kInvalid) return 1.0f;
  const FrustumRec& fr = fp.fr[fi];
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  V4 raw = world_to_screen_raw(V4{pw.x, pw.y, pw.z, 1.0f}, fr.clip, vw, vh);
  if (raw.x < 0.0f || raw.y < 0.0f || raw.x >= vw || raw.y >= vh || raw.z < 0.0f || raw.z > 1.0f) return 1.0f;
  uint32_t px = f2u(raw.x), py = f2u(raw.y);
  DomTexel dt;  // a page that is not resident lets all light through (strand_shading.wgsl:238-257)
  if (!dom_texel(fp, fr, px, py, &dt)) return 1.0f;
  float z0 = fp.dom_depth[dt.depth];
  if (z0 <= 0.0f) return 1.0f;
  float z = normalize_depth01(raw.z);
  if (z > z0 - 1e-3f) return 1.0f;
  uint32_t L = fp.dom_layers;
  float u = wclamp(fmaxf(0.0f, z0 - z) / fmaxf(z0, 1e-6f), 0.0f, 1.0f);
  uint32_t s = min(f2u(floorf(u * (float)L)), L - 1u);
  float op = fp.dom_opacity[dt.opacity + (size_t)s * dt.stride];
  return wclamp(1.0f - op, 0.08f, 1.0f);
}

__global__ __launch_bounds__(128) void k_shading(FrameParams fp) {
  __shared__ LegPoint leg[10];
  if (threadIdx.x < 10u) leg[threadIdx.x] = leg_point((int)threadIdx.x);
  __syncthreads();
  const uint32_t n = min(fp.counters[C_CAM_TASKS], fp.task_cap);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const uint32_t t = fp.shading_queue[q];
    const sr_binning_task tk = fp.tasks[t];
    const uint32_t fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
    const FrustumRec& fr = fp.fr[fi];
    const InstRec& in = fp.inst[tk.id_info];
    const uint32_t idx = tk.seg_idx;
    if (idx >= fp.atlas_texels) continue;
    const sr_strand_meta sm = in.meta[tk.chunk_id];
    V3 v0 = world_v3(in, in.indices[idx]);
    V3 v1 = world_v3(in, in.indices[idx + 1u]);
    V3 U = v3norm(v3sub(v1, v0));
    V3 cam_dir = v3norm(v3sub(V3{fr.eye[0], fr.eye[1], fr.eye[2]}, v0));
    V3 bin = v3norm(v3cross(U, cam_dir));
    V3 Vv = v3norm(v3cross(bin, U));
    const sr_strand_material& m = in.materials[sm.material_idx];
    V3 mid{wmix(v0.x, v1.x, 0.5f), wmix(v0.y, v1.y, 0.5f), wmix(v0.z, v1.z, 0.5f)};
    const uint32_t nl = min(fp.NL, 4u);
    float cur = 1.0f;
    if (nl > 0u) {
      float vis = 0.0f;
      for (uint32_t i = 0; i < nl; ++i) vis = vis + dom_visibility(fp, mid, i);
      cur = vis / (float)nl;
    }
    const size_t layer = (size_t)fr.cam_index * fp.NI + tk.id_info;
    const size_t texel = layer * fp.atlas_texels + idx;
    float prev = (fp.st.shadow_history_valid && fp.history_prev) ? fp.history_prev[texel].x : 0.0f;
    float scat_vis = prev > 0.0f ? wmix(cur, prev, 0.88f) : cur;
    fp.history[texel] = make_float2(scat_vis, cur);
    float occ = 1.0f - scat_vis;
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (uint32_t j = 0; j < fp.NL; ++j) {
      const sr_directional_light lt = fp.lights[j];
      V3 L = v3norm(V3{lt.direction_to_light[0], lt.direction_to_light[1], lt.direction_to_light[2]});
      V3 bc = marschner(L, Vv, U, m, occ, leg);
      float amb = m.ambient_factor;
      float cx = bc.x * (lt.color[0] / 255.0f), cy = bc.y * (lt.color[1] / 255.0f), cz = bc.z * (lt.color[2] / 255.0f);
      float ax = m.absorption_color[0] * amb + (fp.st.ambient_color[0] / 255.0f) * amb;
      float ay = m.absorption_color[1] * amb + (fp.st.ambient_color[1] / 255.0f) * amb;
      float az = m.absorption_color[2] * amb + (fp.st.ambient_color[2] / 255.0f) * amb;
      r += wmix(cx, ax, amb);
      g += wmix(cy, ay, amb);
      bl += wmix(cz, az, amb);
    }
    float4 col = make_float4(wclamp(r, 0.0f, 1.0f), wclamp(g, 0.0f, 1.0f), wclamp(bl, 0.0f, 1.0f),
                             wclamp(m.absorption_color[3], 0.0f, 1.0f));
    fp.atlas[texel] = col;
    // texel idx + 1 is also written by the segment starting there; the
    // segment that starts at a texel owns it
    uint32_t nxt = idx + 1u;
    if (nxt < fp.atlas_texels) {
      bool nxt_shaded = nxt < in.index_count && fp.shaded[(size_t)fr.cam_index * fp.slots + in.pos_base + nxt];
      if (!nxt_shaded) fp.atlas[texel + 1] = col;
    }
  }
}

// --------------------------------------------------- K28 camera raster
// One reference of a camera batch as 16-byte LDS records, in the order a
// pixel's coverage test reads them (one load per record).
template <int TPX>
struct ColorBatch {
  float4 box[TPX];  // screen box grown by the max radius: min x, min y, max x, max y
  float4 seg[TPX];  // p0.x, p0.y, p1.x - p0.x, p1.y - p0.y
  float4 geo[TPX];  // 1 / |p1 - p0|^2, raw depths at p0 and p1, min radius
  float4 wgt[TPX];  // max radius, |w| at p0 and p1, shadow visibility
  float4 s0[TPX], s1[TPX];
  uint32_t item[TPX];  // work item, with bit 31 set when the reference is invalid
};

template <int TPX>
__device__ void load_color_ref(const FrameParams& fp, const FrustumRec& fr, uint32_t fi, const RunTable<TPX>& t,
                               uint32_t i, ColorBatch<TPX>& b) {
  const uint32_t k = threadIdx.x;
  const uint32_t w = item_of(t, i);
  const sr_fine_seg_ref sr = fp.refs[t.base[w] + (i - t.pref[w])];
  const InstRec& in = fp.inst[sr.inst_id];
  float4 a, c;
  ref_screen(fp, fi, in, sr.seg_id, &a, &c);
  b.item[k] = w | ((a.x < 0.0f && c.x < 0.0f) ? 0x80000000u : 0u);
  if (a.x < 0.0f && c.x < 0.0f) return;
  const sr_strand_material& m = in.materials[sr.material_idx];
  float minr = fmaxf(m.min_radius_pixels, 1e-4f);
  float maxr = fmaxf(m.max_radius_pixels, minr);
  const size_t lt = ((size_t)fr.cam_index * fp.NI + sr.inst_id) * fp.atlas_texels;
  uint32_t i0 = sr.seg_id, i1 = min(i0 + 1u, fp.atlas_texels - 1u);
  float abx = c.x - a.x, aby = c.y - a.y;
  b.box[k] = make_float4(fminf(a.x, c.x) - maxr, fminf(a.y, c.y) - maxr, fmaxf(a.x, c.x) + maxr,
                         fmaxf(a.y, c.y) + maxr);
  b.seg[k] = make_float4(a.x, a.y, abx, aby);
  b.geo[k] = make_float4(1.0f / (abx * abx + aby * aby), a.z, c.z, minr);
  b.wgt[k] = make_float4(maxr, fmaxf(fabsf(a.w), 1e-6f), fmaxf(fabsf(c.w), 1e-6f), fp.history[lt + i0].x);
  b.s0[k] = fp.atlas[lt + i0];
  b.s1[k] = fp.atlas[lt + i1];
}

// Depth-sorted blending (SR_FRAME_DEPTH_SORTED_BLEND): a pixel's fragments
// of the current work item, nearest first (larger depth), premultiplied
// colour and alpha, in registers (each lane owns its pixel; indices are
// compile-time, so nothing spills).  One slot more than the list holds
// takes the arriving fragment before the nearest is blended.
constexpr int kSortK = (int)SR_BLEND_SORT_FRAGMENTS;

struct FragList {
  float z[kSortK + 1], r[kSortK + 1], g[kSortK + 1], b[kSortK + 1], a[kSortK + 1];
  int n;
};

__device__ __forceinline__ void blend_under(float& f0, float& f1, float& f2, float& f3, float r, float g, float b,
                                            float a) {
  const float om = 1.0f - f3;
  f0 = f0 + r * om;
  f1 = f1 + g * om;
  f2 = f2 + b * om;
  f3 = f3 + a * om;
}

__device__ __forceinline__ void frag_push(FragList& L, float z, float r, float g, float b, float a, float& f0,
                                          float& f1, float& f2, float& f3) {
#pragma unroll
  for (int i = 0; i <= kSortK; ++i)
    if (i == L.n) { L.z[i] = z; L.r[i] = r; L.g[i] = g; L.b[i] = b; L.a[i] = a; }
  // the new fragment moves ahead of every farther one (equal depths keep arrival order)
#pragma unroll
  for (int i = kSortK; i >= 1; --i) {
    if (i <= L.n && L.z[i] > L.z[i - 1]) {
      float t;
      t = L.z[i]; L.z[i] = L.z[i - 1]; L.z[i - 1] = t;
      t = L.r[i]; L.r[i] = L.r[i - 1]; L.r[i - 1] = t;
      t = L.g[i]; L.g[i] = L.g[i - 1]; L.g[i - 1] = t;
      t = L.b[i]; L.b[i] = L.b[i - 1]; L.b[i - 1] = t;
      t = L.a[i]; L.a[i] = L.a[i - 1]; L.a[i - 1] = t;
    }
  }
  if (++L.n > kSortK) {  // full: the nearest goes first
    blend_under(f0, f1, f2, f3, L.r[0], L.g[0], L.b[0], L.a[0]);
#pragma unroll
    for (int i = 0; i < kSortK; ++i) {
      L.z[i] = L.z[i + 1]; L.r[i] = L.r[i + 1]; L.g[i] = L.g[i + 1]; L.b[i] = L.b[i + 1]; L.a[i] = L.a[i + 1];
    }
    L.n = kSortK;
  }
}

__device__ __forceinline__ void frag_flush(FragList& L, float& f0, float& f1, float& f2, float& f3) {
#pragma unroll
  for (int i = 0; i < kSortK; ++i)
    if (i < L.n) blend_under(f0, f1, f2, f3, L.r[i], L.g[i], L.b[i], L.a[i]);
  L.n = 0;
}

template <int TPX, bool SORTED>
__global__ __launch_bounds__(TPX) void k_camera_raster(FrameParams fp) {
  __shared__ RunTable<TPX> t;
  __shared__ ColorBatch<TPX> b;
  __shared__ unsigned long long cols[TPX][TPX / 64];
  const uint32_t R = fp.counters[C_CAM_RUNS];
  while (next_run<TPX>(&fp.counters[C_CAMERA_HEAD], R, fp.cam_order, t)) {
    const sr_raster_tile_run run = fp.cam_runs[t.run];
    const uint32_t fi = run.frustum_id;
    const FrustumRec& fr = fp.fr[fi];
    if (fr.d.kind != 0u || run.work_count == 0u) continue;
    if ((uint64_t)run.work_base + run.work_count > fp.work_cap) continue;  // emission overflowed
    stage_run<TPX, false>(fp, run, fi, t);
    const uint32_t total = t.pref[t.n];
    const uint32_t fx = fr.d.froxel_size_x, fy = fr.d.froxel_size_y;
    const uint32_t tx = run.screen_tile_id % fr.fine_tiles_x, ty = run.screen_tile_id / fr.fine_tiles_x;
    const uint32_t pidx = threadIdx.x;
    const uint32_t pxu = tx * fx + pidx % fx, pyu = ty * fy + pidx / fx;
    const bool active = pidx < fx * fy && pxu < fr.d.screen_width && pyu < fr.d.screen_height;
    const float cx = (float)pxu + 0.5f, cy = (float)pyu + 0.5f;
    float fc0 = 0.0f, fc1 = 0.0f, fc2 = 0.0f, fc3 = 0.0f, gdepth = 0.0f;
    bool done = !active;
    uint32_t cur = 0;
    FragList fl;
    fl.n = 0;
    for (uint32_t bs = 0; bs < total; bs += TPX) {
      __syncthreads();
      if (bs + threadIdx.x < total) load_color_ref(fp, fr, fi, t, bs + threadIdx.x, b);
      __syncthreads();
      {  // references a pixel's column leaves out do not cover it
        const uint32_t nb = min((uint32_t)TPX, total - bs), k = threadIdx.x;
        const bool valid = k < nb && !(b.item[k] & 0x80000000u);
        float x0 = 0.0f, y0 = 0.0f, ax = 0.0f, ay = 0.0f, rmx = 0.0f;
        if (valid) {
          const float4 sg = b.seg[k], ge = b.geo[k];
          x0 = sg.x; y0 = sg.y; ax = sg.z; ay = sg.w;
          rmx = span_radius(ge.w, b.wgt[k].x, span_depth01_max(ge.y, ge.z));  // radii clamped by the load
        }
        batch_columns<TPX>(valid, x0, y0, ax, ay, rmx, tx * fx, ty * fy, fx, fy, cols);
      }
      __syncthreads();
      if (!done) {
        for (uint32_t cw = 0; cw < (uint32_t)TPX / 64u && !done; ++cw)
        for (unsigned long long bits = cols[pidx][cw]; bits;) {
          const uint32_t k = cw * 64u + (uint32_t)(__ffsll((long long)bits) - 1);
          bits &= bits - 1ull;
          const uint32_t item = b.item[k];
          const uint32_t w = item & 0x7FFFFFFFu;
          if (w != cur) {  // end of a work item: early out above 0.98 alpha
            if (SORTED) frag_flush(fl, fc0, fc1, fc2, fc3);
            cur = w;
            if (fc3 > 0.98f) {
              fc0 = fc0 / fc3; fc1 = fc1 / fc3; fc2 = fc2 / fc3; fc3 = 1.0f;
              done = true;
              break;
            }
          }
          const float4 bx = b.box[k];
          const float4 sg = b.seg[k], ge = b.geo[k];
          float tt = ((cx - sg.x) * sg.z + (cy - sg.y) * sg.w) * ge.x;
          float qx = sg.x + sg.z * tt, qy = sg.y + sg.w * tt;
          float qz = wmix(ge.y, ge.z, tt);
          const float4 wg = b.wgt[k];
          // minr >= 1e-4 and maxr >= minr already (load_color_ref)
          float r = wmix(ge.w, wg.x, wclamp(normalize_depth01(qz), 0.0f, 1.0f));
          float dx = cx - qx, dy = cy - qy;
          float d2 = dx * dx + dy * dy;
          // one predicate for the box, projection and radius tests (an
          // invalid reference has an empty column: it is never a candidate)
          const bool hit = !(cx < bx.x || cy < bx.y || cx > bx.z || cy > bx.w) && (tt >= 0.0f && tt <= 1.0f) &&
                           d2 < r * r;
          if (!hit) continue;
          // coverage and the perspective weight from the hardware square root
          // and reciprocal (within 1 ulp): they scale colour and alpha (1e-3
          // tolerance); whether the fragment covers the pixel is d2 < r^2, exact
          float cov = wclamp(1.0f - __builtin_amdgcn_sqrtf(d2) * __builtin_amdgcn_rcpf(r), 0.0f, 1.0f);
          float aw0 = wg.y, aw1 = wg.z;
          float tw = (tt * aw0) * __builtin_amdgcn_rcpf((1.0f - tt) * aw1 + tt * aw0);
          float4 s0 = b.s0[k], s1 = b.s1[k];
          float alpha = wmix(s0.w, s1.w, tw) * cov;
          float vis = wg.w;
          if (SORTED) {
            frag_push(fl, qz, wmix(s0.x, s1.x, tw) * vis * alpha, wmix(s0.y, s1.y, tw) * vis * alpha,
                      wmix(s0.z, s1.z, tw) * vis * alpha, alpha, fc0, fc1, fc2, fc3);
          } else {
            float om = 1.0f - fc3;
            fc0 = fc0 + (wmix(s0.x, s1.x, tw) * vis * alpha) * om;
            fc1 = fc1 + (wmix(s0.y, s1.y, tw) * vis * alpha) * om;
            fc2 = fc2 + (wmix(s0.z, s1.z, tw) * vis * alpha) * om;
            fc3 = fc3 + alpha * om;
          }
          gdepth = fmaxf(gdepth, qz);
        }
      }
      if (__syncthreads_count(!done) == 0) break;
    }
    if (SORTED && !done) frag_flush(fl, fc0, fc1, fc2, fc3);
    if (!done && fc3 > 0.98f) {  // end of the last work item
      fc0 = fc0 / fc3; fc1 = fc1 / fc3; fc2 = fc2 / fc3; fc3 = 1.0f;
    }
    if (active) {
      float o0 = fc0, o1 = fc1, o2 = fc2, o3 = fc3;
      if (o3 > 1e-6f && o3 < 1.0f) { o0 = o0 / o3; o1 = o1 / o3; o2 = o2 / o3; }
      size_t o = fr.image_off + (size_t)pyu * fr.d.screen_width + pxu;
      fp.image[o] = make_float4(wclamp(o0, 0.0f, 1.0f), wclamp(o1, 0.0f, 1.0f), wclamp(o2, 0.0f, 1.0f),
                                wclamp(o3, 0.0f, 1.0f));
      fp.depth[o] = gdepth;
    }
  }
}

// ------------------------------------------------------- K29 composite
__global__ void k_composite(FrameParams fp, uint64_t pixels) {
  const float4 sc = make_float4(fp.st.scene_clear_color[0], fp.st.scene_clear_color[1], fp.st.scene_clear_color[2],
                                fp.st.scene_clear_color[3]);
  const float sd = fp.st.scene_clear_depth;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += (uint64_t)gridDim.x * blockDim.x) {
    float d = fp.depth[i];
    if (sd > d) { fp.composite[i] = sc; continue; }
    float4 im = fp.image[i];
    fp.composite[i] = make_float4(wmix(sc.x, im.x, im.w), wmix(sc.y, im.y, im.w), wmix(sc.z, im.z, im.w), sc.w);
  }
}

// ------------------------------------------- reference framebuffer format
// rgba8unorm stores of the strand image and the composite (the reference's
// render target, src/pipelines/raster.rs:500-513): each channel clamped to
// [0, 1], scaled by 255 and rounded to nearest, ties to even (as WGSL
// pack4x8unorm), red in the lowest byte.
__device__ __forceinline__ uint32_t unorm8(float x) { return (uint32_t)rintf(wclamp(x, 0.0f, 1.0f) * 255.0f); }
__device__ __forceinline__ uint32_t pack_rgba8(float4 c) {
  return unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (unorm8(c.w) << 24);
}

__global__ void k_pack_rgba8(FrameParams fp, uint64_t pixels) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += (uint64_t)gridDim.x * blockDim.x) {
    fp.image8[i] = pack_rgba8(fp.image[i]);
    fp.composite8[i] = pack_rgba8(fp.composite[i]);
  }
}

// ------------------------------------------------------------ launchers
// Persistent raster workgroups: raster_waves waves per CU (256 CUs, 16 by
// default) in workgroups of `waves` waves.
static inline uint32_t raster_grid(const FrameParams& fp, uint32_t waves) {
  const uint32_t per_cu = fp.raster_waves ? fp.raster_waves : 16u;
  return 256u * (per_cu > waves ? per_cu / waves : 1u);
}

void launch_shadow(const FrameParams& fp, hipStream_t s, uint32_t max_tile_px) {
  if (max_tile_px <= 64)
    hipLaunchKernelGGL(k_shadow_raster<64>, dim3(raster_grid(fp, 1u)), dim3(64), 64u * 4u * fp.dom_layers, s, fp);
  else
    hipLaunchKernelGGL(k_shadow_raster<256>, dim3(raster_grid(fp, 4u)), dim3(256), 256u * 4u * fp.dom_layers, s, fp);
}

void launch_dom_pages(const FrameParams& fp, hipStream_t s) {
  if (!fp.dom_table) return;
  const uint32_t runs = fp.coarse_total * kExt * kExt;  // the shadow run queue's capacity
  const uint32_t g = std::min<uint32_t>(1024u, std::max<uint32_t>(1u, (runs + 255u) / 256u));
  hipLaunchKernelGGL(k_dom_requests, dim3(g), dim3(256), 0, s, fp);
  hipLaunchKernelGGL(k_dom_residency, dim3(1), dim3(256), 0, s, fp);
}

// Atlas and shadow-history texels that no camera segment start shades this
// frame read as zero (the reference leaves them stale; dumps must be
// deterministic): a texel of layer (camera, instance) whose slot is not a
// shaded segment start is cleared.  Shading writes every texel of a shaded
// start, and the texel after a strand's last one, afterwards on the same
// stream, so clearing that last one too changes nothing; only the texels it
// leaves are written here, instead of the whole atlas.
__global__ __launch_bounds__(256) void k_clear_unshaded(FrameParams fp) {
  const uint32_t T = fp.atlas_texels, layers = fp.cams * fp.NI;
  for (uint32_t layer = 0; layer < layers; ++layer) {
    const uint32_t cam = layer / fp.NI, ii = layer - cam * fp.NI;
    const InstRec& in = fp.inst[ii];
    const uint32_t n = min(in.index_count, T);
    const uint8_t* sh = fp.shaded + (size_t)cam * fp.slots + in.pos_base;
    float4* atlas = fp.atlas + (size_t)layer * T;
    float2* hist = fp.history + (size_t)layer * T;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < T; t += gridDim.x * 256u) {
      if (t < n && sh[t] != 0u) continue;
      atlas[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      hist[t] = make_float2(0.0f, 0.0f);
    }
  }
}

void launch_shading(const FrameParams& fp, hipStream_t s) {
  const uint32_t T = fp.atlas_texels;
  if (T && fp.cams && fp.NI)
    hipLaunchKernelGGL(k_clear_unshaded, dim3(std::min<uint32_t>(4096u, (T + 255u) / 256u)), dim3(256), 0, s, fp);
  // 16k workgroups: each thread shades a few segments, so the uneven per-segment
  // cost (lights, lobe windows) leaves a short tail
  hipLaunchKernelGGL(k_shading, dim3(256u * 64u), dim3(128), 0, s, fp);
}

void launch_camera_raster(const FrameParams& fp, hipStream_t s, uint32_t max_tile_px) {
  const bool sorted = (fp.st.flags & SR_FRAME_DEPTH_SORTED_BLEND) != 0u;
  if (max_tile_px <= 64) {
    if (sorted)
      hipLaunchKernelGGL((k_camera_raster<64, true>), dim3(raster_grid(fp, 1u)), dim3(64), 0, s, fp);
    else
      hipLaunchKernelGGL((k_camera_raster<64, false>), dim3(raster_grid(fp, 1u)), dim3(64), 0, s, fp);
  } else {
    if (sorted)
      hipLaunchKernelGGL((k_camera_raster<256, true>), dim3(raster_grid(fp, 4u)), dim3(256), 0, s, fp);
    else
      hipLaunchKernelGGL((k_camera_raster<256, false>), dim3(raster_grid(fp, 4u)), dim3(256), 0, s, fp);
  }
}

void launch_stampback(const FrameParams& fp, uint32_t fi, uint32_t texels, hipStream_t s) {
  if (texels == 0) return;
  uint32_t g = (texels + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_stampback, dim3(g), dim3(256), 0, s, fp, fi);
}

void launch_pack_rgba8(const FrameParams& fp, uint64_t pixels, hipStream_t s) {
  if (pixels == 0 || !fp.image8 || !fp.composite8) return;
  uint64_t g = (pixels + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_pack_rgba8, dim3((uint32_t)g), dim3(256), 0, s, fp, pixels);
}

void launch_composite(const FrameParams& fp, uint64_t pixels, hipStream_t s) {
  if (pixels == 0) return;
  uint64_t g = (pixels + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_composite, dim3((uint32_t)g), dim3(256), 0, s, fp, pixels);
}

}  // namespace srk
