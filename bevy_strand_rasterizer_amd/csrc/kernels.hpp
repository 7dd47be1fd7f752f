// kernels.hpp -- host-side launch interface of the strand kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace srk {

// Named HIP events recorded between pipeline stages when profiling is on.
struct StageTimer {
  bool enabled = false;
  std::vector<hipEvent_t> events;
  std::vector<std::string> names;
  size_t used = 0;
  void begin() { used = 0; names.clear(); }
  void mark(hipStream_t s, const char* name) {
    if (!enabled) return;
    if (used == events.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      events.push_back(e);
    }
    if (hipEventRecord(events[used], s) != hipSuccess) return;
    names.emplace_back(name);
    ++used;
  }
};

// Items per block of the page-CSR hit radix sort (256 threads x 16).
constexpr uint32_t kRadixBlock = 256u * 16u;
// Row chunks of the radix digit-base scan; its chunk sums (kDigitChunks x
// 256 words) live in PrepassScratch::partials.
constexpr uint32_t kDigitChunks = 128u;

// Scratch owned by the runtime for one frame's prepass.
struct PrepassScratch {
  sr_projected_segment* slot_proj;  // [F * slots]
  uint32_t* partials;               // scan tile partials
  uint4* partials4;                 // scan tile partials (4-wide)
  uint32_t* page_xy_mask;           // [page_cap]
  // page-CSR hit sort: the two (key, slot) buffer pairs (pair 0 holds the
  // hits as written and, after the last pass, the sorted keys), digit
  // histograms, block and pass counts
  uint32_t* hkeys[2];
  uint32_t* hvals[2];
  uint32_t* radix_hist;             // [radix_blocks * 256] digit counts, block-major
  uint32_t* radix_base;             // [radix_blocks * 256] output bases, block-major
  uint32_t radix_blocks, radix_passes;
};

// Fine walks (40 B each) and coarse-tile visits (4 B each) each wave of the
// marking pass can queue, beside 36 B of segment state per lane: four waves
// per workgroup hold 38 KB of LDS, so four workgroups fit a CU.
// FrameParams::mark_queue and tile_queue may lower them.
constexpr uint32_t kMarkQ = 160;
constexpr uint32_t kTileQ = 128;

// Bytes of the shadow raster's column cache (FrameParams::shadow_cols).
constexpr uint64_t kShadowColBytes = 64ull << 20;

void launch_prepass(const FrameParams& fp, const PrepassScratch& ps, hipStream_t s, StageTimer* timer);
void launch_shadow(const FrameParams& fp, hipStream_t s, uint32_t max_tile_px);
void launch_stampback(const FrameParams& fp, uint32_t frustum, uint32_t texels, hipStream_t s);
// Virtual DOMs: the page requests of this frame's shadow tile runs, then the
// page table (nothing with dense DOMs).
void launch_dom_pages(const FrameParams& fp, hipStream_t s);
void launch_shading(const FrameParams& fp, hipStream_t s);
void launch_camera_raster(const FrameParams& fp, hipStream_t s, uint32_t max_tile_px);
void launch_composite(const FrameParams& fp, uint64_t pixels, hipStream_t s);
// rgba8unorm copies of the image and the composite (FrameParams::image8 /
// composite8; nothing when they are null).
void launch_pack_rgba8(const FrameParams& fp, uint64_t pixels, hipStream_t s);
// Up to kFillSpans buffers filled with one 32-bit value by one launch.
constexpr uint32_t kFillSpans = 12;
struct FillSpans {
  uint32_t* dst[kFillSpans] = {};
  uint64_t n[kFillSpans] = {};          // words
  uint64_t start[kFillSpans + 1] = {};  // first 16-byte chunk of each span
  uint32_t count = 0, value = 0;
  void add(uint32_t* p, uint64_t words);  // ignored when full, empty or null
};
void launch_fill_spans(const FillSpans& f, hipStream_t s);

}  // namespace srk
