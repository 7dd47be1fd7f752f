// runtime.hip -- host side of the strand runtime behind the C ABI.
//
// Replaces the reference's plugin orchestration (src/plugin.rs:264-1209
// use_prepass_buffers: frustum and instance tables, capacity planning, queue
// resets) and its four run_* sequencers (src/pipelines/prepass.rs:818-1397,
// shadows.rs:621-747, shading.rs:470-567, raster.rs:540-641).
//
// Assets live in device paging slabs addressed by DevicePtr{slab, offset,
// size} (src/shaders/types.wgsl:23-27).  Every frame resolves each instance's
// five slab ids to raw device pointers once on the host, so kernels never
// chase page tables.  Per-frame buffers grow and never shrink.  A frame is
// one stream-ordered sequence of launches with every data-dependent size kept
// on the device; nothing is read back until sr_frame_wait / sr_frame_dump.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "scan.hpp"

using namespace srk;

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string& e) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = e;
}

inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------ no-allocation scratch
// Every intermediate buffer of a frame drawn by sr_frame_submit_noalloc lives
// in this array.  Each GPU the library runs on holds its own copy until the
// process ends; host code never passes the array's name to a kernel (on the
// host it is a stand-in, not a device address) but the address that
// hipGetSymbolAddress returned on that device, kept in g_scratch_ptr.
constexpr uint64_t kScratchBytes = SR_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's device address on `dev`, looked up once per device
// with that device current.
hipError_t scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// ------------------------------------------------------------ slabs
struct SlabAllocator {
  struct Slab {
    char* base = nullptr;
    uint64_t bytes = 0;
    std::vector<std::pair<uint64_t, uint64_t>> free;  // (offset, size), sorted by offset
  };
  uint64_t slab_bytes = 256ull << 20;
  std::vector<Slab> slabs;

  ~SlabAllocator() {
    for (auto& s : slabs)
      if (s.base) (void)hipFree(s.base);
  }
  // returns false on device OOM
  bool alloc(uint64_t size, sr_device_ptr* out) {
    size = align_up(std::max<uint64_t>(size, 1), 256);
    for (uint32_t si = 0; si < slabs.size(); ++si) {
      auto& fl = slabs[si].free;
      for (size_t k = 0; k < fl.size(); ++k) {
        if (fl[k].second >= size) {
          uint64_t off = fl[k].first;
          fl[k].first += size;
          fl[k].second -= size;
          if (fl[k].second == 0) fl.erase(fl.begin() + k);
          if (off > 0xFFFFFFFFull || size > 0xFFFFFFFFull) return false;
          *out = sr_device_ptr{si, (uint32_t)off, (uint32_t)size};
          return true;
        }
      }
    }
    uint64_t bytes = std::max(slab_bytes, size);
    Slab s;
    if (hipMalloc((void**)&s.base, bytes) != hipSuccess) return false;
    s.bytes = bytes;
    if (bytes > size) s.free.push_back({size, bytes - size});
    slabs.push_back(std::move(s));
    *out = sr_device_ptr{(uint32_t)(slabs.size() - 1), 0u, (uint32_t)size};
    return true;
  }
  void release(const sr_device_ptr& p) {
    if (p.slab >= slabs.size() || p.size == 0) return;
    auto& fl = slabs[p.slab].free;
    fl.push_back({p.offset, p.size});
    std::sort(fl.begin(), fl.end());
    std::vector<std::pair<uint64_t, uint64_t>> merged;
    for (auto& f : fl) {
      if (!merged.empty() && merged.back().first + merged.back().second == f.first) merged.back().second += f.second;
      else merged.push_back(f);
    }
    fl.swap(merged);
  }
  char* ptr(const sr_device_ptr& p) const { return slabs[p.slab].base + p.offset; }
  uint64_t capacity() const {
    uint64_t c = 0;
    for (auto& s : slabs) c += s.bytes;
    return c;
  }
};

struct AssetRec {
  bool live = false;
  sr_asset_info info{};
  sr_device_ptr pos_strand{};
  sr_strand_geo geo{};
  uint32_t vertex_count = 0, material_count = 0;
};

// grow-only device buffer, or (borrowed) a span of memory the buffer does
// not own: the no-allocation scratch or a caller's output
struct DBuf {
  void* p = nullptr;
  uint64_t bytes = 0;
  bool borrowed = false;
  bool ensure(uint64_t need) {
    need = std::max<uint64_t>(need, 256);
    if (need <= bytes && !borrowed) return true;
    free_();
    uint64_t b = align_up(need + need / 4, 4096);
    if (hipMalloc(&p, b) != hipSuccess) { p = nullptr; return false; }
    bytes = b;
    return true;
  }
  void borrow(void* q, uint64_t n) {
    free_();
    p = q;
    bytes = n;
    borrowed = true;
  }
  void free_() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

enum BufId {
  B_COUNTERS, B_BROAD, B_KEEP, B_RANGES, B_SLOT_MASKS, B_WAVE_OFF, B_FR_TASK_BASE, B_FR_CAM_BASE, B_SHADED,
  B_TASKS, B_PROJ, B_SHADING_Q, B_LUT, B_LUT_Z, B_PAGE_TABLE, B_PAGE_META, B_PAGE_TOTALS, B_CELL_COUNTS, B_CELL_OFFSETS,
  B_CELL_CURSORS, B_REFS, B_ACTIVE_FLAGS, B_ACTIVE_Q, B_TILE_WORK, B_TILE_BASES, B_WORK, B_CAM_RUNS, B_SHADOW_RUNS, B_CAM_ORDER, B_SHADOW_ORDER, B_RUN_HIST, B_PAGE_BOUNDS, B_PAGE_ORDER, B_PAGE_HIST,
  B_DOM_DEPTH, B_DOM_OPACITY, B_STAMP, B_ATLAS, B_HIST0, B_HIST1, B_IMAGE, B_DEPTH, B_COMPOSITE, B_SHADOW_COLS, B_SLOT_PROJ, B_SLOT_SCREEN,
  B_PARTIALS, B_PARTIALS4, B_XY_MASK, B_SLOT_HIT_TOTAL, B_HIT_BASE, B_SLOT_MATERIAL, B_HKEYS0,
  B_HKEYS1, B_HVALS0, B_HVALS1, B_RADIX_HIST, B_RADIX_BASE, B_SLOT_HITS, B_SLOT_HIT_COUNT, B_TELEMETRY, B_OPAQUE_DEPTH, B_OPAQUE_TILES, B_TABLES,
  B_IMAGE8, B_COMPOSITE8, B_DOM_REQ, B_DOM_TABLE,
  B_COUNT
};

// Shape of the last submitted frame, for dumps.
struct FrameShape {
  bool valid = false;
  uint32_t F = 0, NI = 0, NL = 0, cams = 0, slots = 0, coarse_total = 0, slices = 16;
  uint32_t task_cap = 0, ref_cap = 0, work_cap = 0, page_cap = 0, hit_cap = 0, atlas_texels = 0, dom_layers = 0;
  uint64_t dom_px = 0, img_px = 0, stamp_px = 0, tel_words = 0, opaque_tiles = 0;
  int hist_cur = 0;
  bool rgba8 = false;  // the frame stored rgba8unorm outputs
  // virtual DOMs: page grid of every light frustum (width, height, pages x,
  // pages y, first page), pages over all of them, pool capacity
  bool dom_virtual = false;
  uint32_t dom_vpages = 0, dom_pool_cap = 0;
  std::vector<uint32_t> dom_lights;
};

// Everything one in-flight frame owns: its stream, grow-only buffers, the
// staging ring for the per-frame tables, the shape of its last frame and the
// grow-only capacity hints.  The context's slot 0 serves sr_frame_submit; a
// server owns further slots so that frames overlap on the device.
struct FrameSlot {
  hipStream_t stream = nullptr;
  DBuf buf[B_COUNT];
  // pinned host staging for per-frame tables: a ring, so the host can run
  // several frames ahead of the device without overwriting pending copies
  static constexpr int kRing = 8;
  void* staging[kRing] = {};
  uint64_t staging_bytes[kRing] = {};
  hipEvent_t staging_ev[kRing] = {};
  bool staging_used[kRing] = {};
  int ring_pos = 0;
  FrameShape shape;
  StageTimer timer;
  hipEvent_t done = nullptr;
  bool submitted = false;
  uint32_t ref_cap_hint = 0, page_cap_hint = 0, work_cap_hint = 0, dom_pool_hint = 0;
  int hist_cur = 0;
  // captured launch sequences, keyed by everything the launches bake in
  struct Graph {
    std::vector<uint8_t> key;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
  };
  std::vector<Graph> graphs;
  uint64_t graph_clock = 0, graph_captures = 0, graph_replays = 0;
  // deep shadow maps left by the last frame that built its lights: what they
  // were built from (light frusta and views, lights, instances, culling
  // settings), and whether that frame completed without overflow
  std::vector<uint8_t> dom_key;
  bool dom_pending = false, dom_ready = false;

  void drop_graphs() {
    for (Graph& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  uint64_t held_bytes() const {
    uint64_t b = 0;
    for (const DBuf& d : buf) b += d.bytes;
    return b;
  }
  void release_buffers() {
    drop_graphs();  // they point into the buffers
    for (DBuf& d : buf) d.free_();
    shape.valid = false;
    dom_pending = dom_ready = false;
  }
};

FrameSlot* new_slot(std::string* err) {
  FrameSlot* sl = new FrameSlot();
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&sl->done, hipEventDisableTiming)) != hipSuccess) {
    if (err) *err = std::string("stream creation: ") + hipGetErrorString(e);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    delete sl;
    return nullptr;
  }
  return sl;
}

void free_slot(FrameSlot* sl) {
  if (!sl) return;
  if (sl->stream) (void)hipStreamSynchronize(sl->stream);
  sl->release_buffers();  // and the graphs
  for (int k = 0; k < FrameSlot::kRing; ++k) {
    if (sl->staging[k]) (void)hipHostFree(sl->staging[k]);
    if (sl->staging_ev[k]) (void)hipEventDestroy(sl->staging_ev[k]);
  }
  for (auto e : sl->timer.events) (void)hipEventDestroy(e);
  if (sl->done) (void)hipEventDestroy(sl->done);
  if (sl->stream) (void)hipStreamDestroy(sl->stream);
  delete sl;
}

}  // namespace

struct sr_context {
  int device = 0;
  std::string err;
  std::mutex mu;  // assets and slabs (a server thread reads them while the caller uploads)
  SlabAllocator slabs;
  std::vector<AssetRec> assets;
  std::vector<FrameSlot*> slots;  // slots[0]: sr_frame_submit
  uint32_t flags = 0;             // SR_CONFIG_* bits
  uint64_t hbm_budget = 0;
  FrameSlot& main() { return *slots[0]; }
  // sr_frame_submit_noalloc: a slot whose buffers are carved out of this
  // device's scratch array (set up at sr_create)
  FrameSlot* na = nullptr;
  void* scratch = nullptr;
  std::string scratch_err;
};

namespace {

int fail(sr_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  set_global_error(msg);
  return code;
}

int hip_fail(sr_context* ctx, hipError_t e, const char* what) {
  return fail(ctx, SR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_OKS(expr, what)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      err = std::string(what) + ": " + hipGetErrorString(_e);                 \
      return SR_ERR_DEVICE;                                                   \
    }                                                                         \
  } while (0)

inline int efail(std::string& err, int code, const std::string& msg) {
  err = msg;
  return code;
}

#define HIP_OK(expr, what)                        \
  do {                                            \
    hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
  } while (0)

// Returns the next ring entry with at least `bytes`, after the device has
// consumed the copy that last used it.
char* next_staging(FrameSlot& sl, uint64_t bytes, int* slot_out) {
  int k = sl.ring_pos;
  sl.ring_pos = (k + 1) % FrameSlot::kRing;
  if (sl.staging_used[k]) (void)hipEventSynchronize(sl.staging_ev[k]);
  if (!sl.staging_ev[k] && hipEventCreateWithFlags(&sl.staging_ev[k], hipEventDisableTiming) != hipSuccess)
    return nullptr;
  if (bytes > sl.staging_bytes[k]) {
    if (sl.staging[k]) (void)hipHostFree(sl.staging[k]);
    sl.staging[k] = nullptr;
    sl.staging_bytes[k] = 0;
    uint64_t b = align_up(bytes * 2, 4096);
    if (hipHostMalloc(&sl.staging[k], b, hipHostMallocDefault) != hipSuccess) return nullptr;
    sl.staging_bytes[k] = b;
  }
  *slot_out = k;
  return (char*)sl.staging[k];
}

inline uint32_t next_pow2(uint32_t x) {
  uint32_t p = 1;
  while (p < x && p < 0x80000000u) p <<= 1;
  return p;
}

}  // namespace

extern "C" {

const char* sr_version(void) { return "bevy_strand_rasterizer_amd 0.1.0 (gfx950)"; }

const char* sr_last_error(sr_context* ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

int sr_create(const sr_device_config* cfg, sr_context** out) {
  if (!out) return fail(nullptr, SR_ERR_INVALID_ARGUMENT, "sr_create: out is NULL");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(nullptr, SR_ERR_DEVICE, "sr_create: no HIP device available");
  int dev = cfg ? cfg->device_index : 0;
  if (dev < 0 || dev >= count) return fail(nullptr, SR_ERR_INVALID_ARGUMENT, "sr_create: device index out of range");
  if ((e = hipSetDevice(dev)) != hipSuccess) return fail(nullptr, SR_ERR_DEVICE, hipGetErrorString(e));
  sr_context* ctx = new sr_context();
  ctx->device = dev;
  if (cfg && cfg->slab_bytes) ctx->slabs.slab_bytes = cfg->slab_bytes;
  if (cfg) ctx->hbm_budget = cfg->hbm_budget_bytes;
  if (cfg) ctx->flags = cfg->flags;
  std::string err;
  FrameSlot* sl = new_slot(&err);
  if (!sl) {
    delete ctx;
    return fail(nullptr, SR_ERR_DEVICE, "sr_create: " + err);
  }
  ctx->slots.push_back(sl);
  // the no-allocation path: this device's scratch address and its own slot
  // (stream and completion event), so that sr_frame_submit_noalloc creates
  // nothing itself
  if ((e = scratch_address(dev, &ctx->scratch)) != hipSuccess) {
    ctx->scratch = nullptr;
    ctx->scratch_err = std::string("scratch address: ") + hipGetErrorString(e);
  } else if (!(ctx->na = new_slot(&err))) {
    ctx->scratch_err = "no-allocation slot: " + err;
  }
  *out = ctx;
  return SR_OK;
}

int sr_destroy(sr_context* ctx) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  for (FrameSlot* sl : ctx->slots) free_slot(sl);
  ctx->slots.clear();
  free_slot(ctx->na);  // its buffers are borrowed: nothing is freed but the stream and events
  ctx->na = nullptr;
  delete ctx;
  return SR_OK;
}

// validate_cache-style checks (src/strand_cache.rs:174-209) plus the slot
// layout requirement: strands occupy disjoint index ranges.
int sr_asset_upload(sr_context* ctx, const sr_asset_slabs* a, sr_asset_info* out) {
  if (!ctx || !a || !out) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "sr_asset_upload: NULL argument");
  if (!a->vertices || !a->indices || !a->meta || !a->geo || !a->materials || a->vertex_count == 0 ||
      a->index_count == 0 || a->strand_count == 0 || a->material_count == 0)
    return fail(ctx, SR_ERR_INVALID_ASSET, "sr_asset_upload: empty slab");
  for (uint32_t i = 0; i < a->index_count; ++i)
    if (a->indices[i] >= a->vertex_count) return fail(ctx, SR_ERR_INVALID_ASSET, "index points past the vertex slab");
  std::vector<std::pair<uint32_t, uint32_t>> spans;
  spans.reserve(a->strand_count);
  uint32_t max_verts = 0;
  for (uint32_t s = 0; s < a->strand_count; ++s) {
    const sr_strand_meta& m = a->meta[s];
    if ((uint64_t)m.offset + m.count > a->index_count)
      return fail(ctx, SR_ERR_INVALID_ASSET, "strand range outside the index slab");
    if (m.material_idx >= a->material_count) return fail(ctx, SR_ERR_INVALID_ASSET, "material index out of range");
    if (m.count > 0) spans.push_back({m.offset, m.count});
    max_verts = std::max(max_verts, m.count);
  }
  std::sort(spans.begin(), spans.end());
  for (size_t i = 1; i < spans.size(); ++i)
    if (spans[i - 1].first + spans[i - 1].second > spans[i].first)
      return fail(ctx, SR_ERR_INVALID_ASSET, "strands overlap in the index slab");
  std::vector<uint32_t> pos_strand(a->index_count, kInvalid);
  for (uint32_t s = 0; s < a->strand_count; ++s) {
    const sr_strand_meta& m = a->meta[s];
    if (m.count < 2) continue;
    for (uint32_t p = m.offset; p + 1 < m.offset + m.count; ++p) pos_strand[p] = s;
  }
  (void)hipSetDevice(ctx->device);
  std::lock_guard<std::mutex> lk(ctx->mu);
  AssetRec rec;
  rec.live = true;
  rec.geo = *a->geo;
  rec.vertex_count = a->vertex_count;
  rec.material_count = a->material_count;
  struct Up { const void* src; uint64_t bytes; sr_device_ptr* dst; };
  sr_asset_info& info = rec.info;
  Up ups[6] = {{a->vertices, 16ull * a->vertex_count, &info.vertices},
               {a->indices, 4ull * a->index_count, &info.indices},
               {a->meta, 16ull * a->strand_count, &info.meta},
               {a->geo, 64ull, &info.geo},
               {a->materials, 64ull * a->material_count, &info.materials},
               {pos_strand.data(), 4ull * a->index_count, &rec.pos_strand}};
  for (auto& u : ups) {
    if (!ctx->slabs.alloc(u.bytes, u.dst)) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "sr_asset_upload: slab allocation failed");
    u.dst->size = (uint32_t)u.bytes;  // DevicePtr.size carries the payload size
    HIP_OK(hipMemcpy(ctx->slabs.ptr(*u.dst), u.src, u.bytes, hipMemcpyHostToDevice), "sr_asset_upload: copy");
  }
  uint32_t id = 0;
  while (id < ctx->assets.size() && ctx->assets[id].live) ++id;
  info.vertex_id = info.index_id = info.meta_id = info.geo_id = info.material_id = id;
  info.strand_count = a->strand_count;
  info.index_count = a->index_count;
  info.max_segments_in_strand = max_verts;
  if (id == ctx->assets.size()) ctx->assets.push_back(rec); else ctx->assets[id] = rec;
  *out = info;
  return SR_OK;
}

int sr_asset_release(sr_context* ctx, uint32_t id) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (id >= ctx->assets.size() || !ctx->assets[id].live)
    return fail(ctx, SR_ERR_INVALID_ARGUMENT, "sr_asset_release: unknown asset");
  (void)hipSetDevice(ctx->device);
  for (FrameSlot* sl : ctx->slots) (void)hipStreamSynchronize(sl->stream);  // frames in flight may read it
  if (ctx->na) (void)hipStreamSynchronize(ctx->na->stream);
  AssetRec& r = ctx->assets[id];
  const sr_device_ptr* owned[6] = {&r.info.vertices, &r.info.indices, &r.info.meta,
                                   &r.info.geo,      &r.info.materials, &r.pos_strand};
  for (const sr_device_ptr* p : owned) {
    sr_device_ptr q = *p;
    q.size = (uint32_t)align_up(q.size, 256);
    ctx->slabs.release(q);
  }
  r.live = false;
  return SR_OK;
}

int sr_set_profiling(sr_context* ctx, int32_t enabled) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  ctx->main().timer.enabled = enabled != 0;
  return SR_OK;
}

}  // extern "C"

namespace {

// Where a no-allocation frame puts its buffers and outputs: every buffer is a
// 256-byte aligned span of `base` (the scratch array, `bytes` long), except
// the composite and depth, which go straight to the caller's device memory;
// counters (optional) receives the frame's sr_frame_counters on the device.
struct NoAllocTarget {
  char* base;
  uint64_t bytes;
  float4* composite;
  float* depth;
  sr_frame_counters* counters;
};

// Table bytes handed to k_put_words per launch (kernel arguments, so that no
// staging memory is needed).
constexpr uint32_t kPutWords = 480;
struct PutWords {
  uint32_t w[kPutWords];
  uint32_t n;
};

__global__ void k_put_words(uint32_t* dst, PutWords c) {
  for (uint32_t i = threadIdx.x; i < c.n; i += blockDim.x) dst[i] = c.w[i];
}

// sr_frame_counters of a frame from its device counter words and the
// capacities it ran with (the same rule as frame_wait_slot).
struct CounterCaps {
  uint32_t coarse_total, task_cap, ref_cap, work_cap, hit_cap;
};

__host__ __device__ inline sr_frame_counters counters_from_words(const uint32_t* c, CounterCaps k) {
  sr_frame_counters fc;
  fc.binning_tasks = c[C_TASKS];
  fc.camera_tasks = c[C_CAM_TASKS];
  fc.allocated_pages = c[C_PAGES];
  fc.fine_refs = c[C_REFS];
  fc.active_tiles = c[C_ACTIVE];
  fc.work_items = c[C_WORK];
  fc.camera_runs = c[C_CAM_RUNS];
  fc.shadow_runs = c[C_SHADOW_RUNS];
  fc.overflow_pages = c[C_OVERFLOW];
  fc.coarse_tiles = k.coarse_total;
  fc.task_capacity = k.task_cap;
  fc.ref_capacity = k.ref_cap;
  fc.page_drops = c[C_PAGE_DROPS];
  fc.page_candidates = k.hit_cap ? c[C_CANDS] : 0u;
  fc.capacity_exceeded = (c[C_REFS] > k.ref_cap || c[C_PAGE_DROPS] > 0u || c[C_WORK] > k.work_cap ||
                          c[C_OVERFLOW] > 0u || (k.hit_cap && c[C_CANDS] > k.hit_cap) || c[C_DOM_DROPS] > 0u)
                             ? 1u
                             : 0u;
  fc.dom_pages = c[C_DOM_PAGES];
  return fc;
}

__global__ void k_counters_out(const uint32_t* c, CounterCaps k, sr_frame_counters* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = counters_from_words(c, k);
}

// Plans and launches one frame on `sl`.  need_bytes (optional) receives the
// device bytes the slot holds once its buffers fit this frame; dry_run stops
// after planning.  With `na`, the slot's buffers are carved out of na->base
// (nothing is allocated; a frame that does not fit fails with
// SR_ERR_CAPACITY), the per-frame tables travel as kernel arguments, no graph
// is captured and shadow history is off.  Asset records are read under
// ctx->mu.
int frame_submit_slot(sr_context* ctx, FrameSlot& sl, const sr_frame_desc* fd, hipStream_t stream_in,
                      std::string& err, uint64_t* need_bytes, bool dry_run, const NoAllocTarget* na = nullptr,
                      uint64_t* own_bytes = nullptr) {
  if (!fd) return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: NULL argument");
  if (na && fd->opaque_depth)
    return efail(err, SR_ERR_UNSUPPORTED, "sr_frame_submit_noalloc: opaque depth needs a staging upload; use sr_frame_submit");
  if (fd->frustum_count == 0 || !fd->frusta || !fd->views)
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: no frusta");
  if (fd->instance_count && !fd->instances) return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: no instances");
  if (fd->instance_count >= (1u << 24)) return efail(err, SR_ERR_UNSUPPORTED, "sr_frame_submit: at most 2^24 instances");
  if (fd->light_count && !fd->lights) return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: no lights");
  if (!(fd->mark_waves == 0u || fd->mark_waves == 1u || fd->mark_waves == 2u || fd->mark_waves == 4u) ||
      fd->raster_waves > 32u)
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: mark_waves must be 1, 2 or 4 and raster_waves at most 32");
  const sr_frame_settings& st = fd->settings;
  if (st.dom_layers == 0 || st.dom_layers > SR_MAX_DOM_LAYERS)
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: dom_layers must be in 1..16");
  if (st.fine_binning_backend > 1) return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: unknown binning backend");
  if (!(st.coarse_depth_slices == 0u || st.coarse_depth_slices == 8u || st.coarse_depth_slices == 16u ||
        st.coarse_depth_slices == 32u))
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: coarse_depth_slices must be 8, 16 or 32 (0 = 16)");
  const uint32_t slices = st.coarse_depth_slices ? st.coarse_depth_slices : kSlices;
  if (st.flags & ~(SR_FRAME_REUSE_LIGHT_DOMS | SR_FRAME_RGBA8_OUTPUTS | SR_FRAME_VIRTUAL_DOMS |
                   SR_FRAME_DEPTH_SORTED_BLEND | SR_FRAME_TUNE_KEEP_BACKEND))
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: unknown frame flags");
  const bool reuse = (st.flags & SR_FRAME_REUSE_LIGHT_DOMS) != 0u;
  const bool rgba8 = (st.flags & SR_FRAME_RGBA8_OUTPUTS) != 0u;
  const bool vdom = (st.flags & SR_FRAME_VIRTUAL_DOMS) != 0u;
  if (reuse && na)
    return efail(err, SR_ERR_UNSUPPORTED, "sr_frame_submit_noalloc: the scratch array keeps no shadow maps between calls");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = stream_in ? stream_in : sl.stream;

  const uint32_t F = fd->frustum_count, NI = fd->instance_count, NL = fd->light_count;
  // ---- frustum records
  std::vector<FrustumRec> fr(F);
  uint32_t lights_seen = 0, cams = 0, coarse_total = 0, max_cam_tile_px = 1, max_light_tile_px = 1;
  uint64_t fine_total = 0;
  uint64_t dom_px = 0, img_px = 0, stamp_px = 0;
  // virtual DOMs: every light frustum's page grid, and the pages the pool
  // would need for every light's budget
  uint32_t dom_vpages = 0;
  uint64_t dom_cap_all = 0;
  std::vector<uint32_t> dom_lights;
  for (uint32_t i = 0; i < F; ++i) {
    const sr_frustum_desc& d = fd->frusta[i];
    if (d.screen_width == 0 || d.screen_height == 0 || d.froxel_size_x == 0 || d.froxel_size_y == 0)
      return efail(err, SR_ERR_INVALID_ARGUMENT, "frustum with zero extent");
    if (d.froxel_size_x * d.froxel_size_y > 256)
      return efail(err, SR_ERR_UNSUPPORTED, "froxel tiles above 256 pixels are not supported");
    uint32_t ftx = (d.screen_width + d.froxel_size_x - 1) / d.froxel_size_x;
    uint32_t fty = (d.screen_height + d.froxel_size_y - 1) / d.froxel_size_y;
    if (d.coarse_tiles_x * kExt < ftx || d.coarse_tiles_y * kExt < fty ||
        d.coarse_depth_tile_count < d.coarse_tiles_x * d.coarse_tiles_y)
      return efail(err, SR_ERR_INVALID_ARGUMENT, "frustum coarse tiling does not cover its fine tiles");
    FrustumRec& r = fr[i];
    std::memset(&r, 0, sizeof(r));
    r.d = d;
    std::memcpy(r.clip, fd->views[i].clip_from_world, sizeof(r.clip));
    std::memcpy(r.eye, fd->views[i].world_position, sizeof(r.eye));
    r.fine_tiles_x = ftx;
    r.fine_tiles_y = fty;
    r.light_layer = kInvalid;
    r.cam_index = kInvalid;
    if (d.kind == 1u) {
      if (lights_seen < NL) r.light_layer = lights_seen;
      ++lights_seen;
      if (vdom) {
        r.dom_pages_x = (d.screen_width + kDomPage - 1u) / kDomPage;
        r.dom_pages_y = (d.screen_height + kDomPage - 1u) / kDomPage;
        r.dom_vpage_base = dom_vpages;
        const uint32_t n = r.dom_pages_x * r.dom_pages_y;
        dom_vpages += n;
        dom_cap_all += st.dom_resident_pages ? std::min(n, st.dom_resident_pages) : n;
        dom_lights.insert(dom_lights.end(), {d.screen_width, d.screen_height, r.dom_pages_x, r.dom_pages_y, r.dom_vpage_base});
      } else {
        r.dom_depth_off = dom_px;
        r.dom_opacity_off = dom_px * st.dom_layers;
        dom_px += (uint64_t)d.screen_width * d.screen_height;
      }
      r.stamp_w = st.stamp_target_width ? st.stamp_target_width : d.screen_width;
      r.stamp_h = st.stamp_target_height ? st.stamp_target_height : d.screen_height;
      r.stamp_off = stamp_px;
      stamp_px += (uint64_t)r.stamp_w * r.stamp_h;
    } else if (d.kind == 0u) {
      r.cam_index = cams++;
      r.image_off = img_px;
      img_px += (uint64_t)d.screen_width * d.screen_height;
    }
    r.enabled = (d.kind == 0u) || (d.kind == 1u && r.light_layer != kInvalid && !reuse);
    if (d.kind == 0u) max_cam_tile_px = std::max(max_cam_tile_px, d.froxel_size_x * d.froxel_size_y);
    if (d.kind == 1u) max_light_tile_px = std::max(max_light_tile_px, d.froxel_size_x * d.froxel_size_y);
    coarse_total = std::max(coarse_total, d.coarse_depth_tile_base + d.coarse_depth_tile_count);
    fine_total = std::max<uint64_t>(fine_total, (uint64_t)d.fine_depth_tile_base + (uint64_t)ftx * fty);
  }
  if (fd->opaque_depth && (fd->opaque_samples == 0u || fd->opaque_samples > 4u))
    return efail(err, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit: opaque_samples must be in 1..4");
  if (fine_total > 0xFFFFFFFFull) return efail(err, SR_ERR_CAPACITY, "sr_frame_submit: too many fine tiles");
  const uint64_t opaque_floats = fd->opaque_depth ? img_px * fd->opaque_samples : 0ull;
  // ---- instance records
  std::unique_lock<std::mutex> asset_lock(ctx->mu);
  std::vector<InstRec> ir(NI);
  uint64_t pos = 0;
  uint32_t max_idx = 1;
  for (uint32_t i = 0; i < NI; ++i) {
    const sr_strand_instance& in = fd->instances[i];
    InstRec& r = ir[i];
    std::memset(&r, 0, sizeof(r));
    std::memcpy(r.world_from_local, in.world_from_local, sizeof(r.world_from_local));
    r.pos_base = (uint32_t)pos;
    auto live = [&](uint32_t id) { return id < ctx->assets.size() && ctx->assets[id].live; };
    if (!live(in.vertex_id) || !live(in.index_id) || !live(in.meta_id) || !live(in.geo_id) || !live(in.material_id))
      return efail(err, SR_ERR_INVALID_ARGUMENT, "instance references an unknown asset id");
    const AssetRec& va = ctx->assets[in.vertex_id];
    const AssetRec& ia = ctx->assets[in.index_id];
    const AssetRec& ma = ctx->assets[in.meta_id];
    const AssetRec& ga = ctx->assets[in.geo_id];
    const AssetRec& ta = ctx->assets[in.material_id];
    r.vertices = reinterpret_cast<const float4*>(ctx->slabs.ptr(va.info.vertices));
    r.indices = reinterpret_cast<const uint32_t*>(ctx->slabs.ptr(ia.info.indices));
    r.pos_strand = reinterpret_cast<const uint32_t*>(ctx->slabs.ptr(ia.pos_strand));
    r.meta = reinterpret_cast<const sr_strand_meta*>(ctx->slabs.ptr(ma.info.meta));
    r.materials = reinterpret_cast<const sr_strand_material*>(ctx->slabs.ptr(ta.info.materials));
    r.vertex_count = va.vertex_count;
    r.index_count = ia.info.index_count;
    r.meta_count = ma.info.strand_count;
    r.material_count = ta.material_count;
    r.geo_strand_count = ga.geo.strand_count;
    r.aabb = ga.geo.aabb;
    // kernels index vertices/materials through the meta/index slabs of the
    // same instance; mixing assets requires matching slab shapes
    if (in.index_id != in.meta_id || in.vertex_id != in.index_id || in.material_id != in.meta_id)
      return efail(err, SR_ERR_UNSUPPORTED, "an instance must take all slabs from one asset");
    r.valid = 1u;
    pos += r.index_count;
    max_idx = std::max(max_idx, r.index_count);
  }
  // what this frame's shadow maps are built from: each light frustum's
  // tiling, view and DOM placement (not its table bases, which follow the
  // cameras before it), the lights, the instances and the culling settings
  std::vector<uint8_t> dom_key;
  auto put_key = [&dom_key](const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    dom_key.insert(dom_key.end(), b, b + n);
  };
  for (uint32_t i = 0; i < F; ++i) {
    const FrustumRec& r = fr[i];
    if (r.d.kind != 1u) continue;
    const uint32_t k[11] = {r.d.screen_width, r.d.screen_height, r.d.froxel_size_x, r.d.froxel_size_y,
                            r.d.depth_slices, r.d.coarse_tiles_x, r.d.coarse_tiles_y, r.d.cascade_index,
                            r.light_layer, r.stamp_w, r.stamp_h};
    put_key(k, sizeof(k));
    put_key(r.clip, sizeof(r.clip));
    const uint64_t o[3] = {r.dom_depth_off, r.dom_opacity_off, r.stamp_off};
    put_key(o, sizeof(o));
    const uint32_t vp[3] = {r.dom_pages_x, r.dom_pages_y, r.dom_vpage_base};
    put_key(vp, sizeof(vp));
  }
  if (NL) put_key(fd->lights, sizeof(sr_directional_light) * NL);
  if (NI) put_key(ir.data(), sizeof(InstRec) * NI);
  // virtual DOM pool: the pages every light's budget allows, or half of them
  // to start with (a frame that wants more reports its capacity exceeded and
  // the next one gets what it wanted)
  uint32_t dom_pool_cap = 0;
  if (vdom) {
    const uint32_t all = (uint32_t)std::min<uint64_t>(dom_cap_all, 1ull << 24);
    const uint32_t start = (ctx->flags & SR_CONFIG_TIGHT_CAPACITY) ? 1u : std::max(1u, all / 2u);
    dom_pool_cap = std::min(all, std::max(start, sl.dom_pool_hint));
  }
  const uint32_t ks[8] = {st.stochastic_cull_enabled, st.dom_layers, NL, NI, lights_seen, vdom ? 1u : 0u,
                          vdom ? st.dom_resident_pages : 0u, dom_pool_cap};
  put_key(ks, sizeof(ks));
  const float kf[3] = {st.target_strands_per_pixel, st.min_keep_probability, st.shadow_keep_probability};
  put_key(kf, sizeof(kf));
  if (reuse && !(sl.dom_ready && sl.dom_key == dom_key))
    return efail(err, SR_ERR_INVALID_ARGUMENT,
                 sl.dom_ready ? "sr_frame_submit: SR_FRAME_REUSE_LIGHT_DOMS with lights, instances or culling "
                                "settings other than those the shadow maps were built from"
                              : "sr_frame_submit: SR_FRAME_REUSE_LIGHT_DOMS before a light-building frame "
                                "completed without overflow (wait for it first)");
  const uint32_t slots = (uint32_t)align_up(std::max<uint64_t>(pos, 64), 64);
  if ((uint64_t)F * slots > 0x7FFFFFFFull) return efail(err, SR_ERR_CAPACITY, "frame exceeds 2^31 segment slots");
  const uint32_t task_cap = F * slots;
  const uint32_t ptab = coarse_total * slices;
  const bool tight = (ctx->flags & SR_CONFIG_TIGHT_CAPACITY) != 0u;
  uint32_t page_cap = std::max<uint32_t>(std::min<uint32_t>(ptab, tight ? 16u : std::max(4096u, ptab / 2)),
                                         sl.page_cap_hint);
  page_cap = std::min(page_cap, ptab);
  // page-CSR hit keys are page * 256 + cell in 32 bits, the dropped-page key
  // page_cap included: at most 2^24 - 1 pages (further pages are dropped and
  // counted, as past any capacity)
  if (st.fine_binning_backend != 0u) page_cap = std::min(page_cap, (1u << 24) - 1u);
  // reference slab: 8 references per task slot (the reference sizes it at 4x
  // its binning queue, src/plugin.rs:540-546); work items: half of that.
  // Both grow when a frame reports an overflow.
  // The first frame of a shape starts at 8 per task slot, at most 2^28
  // references (3 GiB); a frame that overflows reports capacity_exceeded and
  // the next one gets the size it needed.
  uint64_t ref_cap64 = std::max<uint64_t>({tight ? 4096ull : std::min<uint64_t>((uint64_t)task_cap * 8ull, 1ull << 28),
                                           tight ? 4096ull : 1ull << 16, (uint64_t)sl.ref_cap_hint});
  const uint32_t ref_cap = (uint32_t)std::min<uint64_t>(ref_cap64, 0x7FFFFFFFull);
  // work items are non-empty cells: never more than the page capacity's cells
  const uint32_t work_cap = (uint32_t)std::max<uint64_t>(
      {std::min<uint64_t>(ref_cap / 2u, (uint64_t)page_cap * kPage), tight ? 1024u : 1u << 15,
       (uint64_t)sl.work_cap_hint});
  // page CSR backend: one (page * 256 + cell, slot) hit per visible fine
  // cell, at most ref_cap of them (a hit is a reference of a kept page)
  const bool csr = st.fine_binning_backend != 0u;
  // cells recorded per slot while marking: 24 words per slot (a segment of
  // the benchmark scalps hits 5 cells on average and more than 24 about once
  // in 10^5), fewer when the table would pass 48 GiB (the stress frames'
  // 400M+ slots record all 24), none when page-table entries do not fit 24
  // bits; a slot with more cells is traced again
  uint32_t hit_k = 0;
  if (csr && ptab < (1u << 24))
    for (uint32_t k : {24u, 16u, 12u, 8u, 6u, 4u})
      if (4ull * k * task_cap <= (48ull << 30)) { hit_k = k; break; }
  if (hit_k && (ctx->flags & SR_CONFIG_CSR_FEW_HITS)) hit_k = 2u;
  const uint64_t hit_bytes = 4ull * hit_k * task_cap;
  const uint32_t hit_cap = csr ? ref_cap : 0u;
  const uint32_t atlas_w = std::min(std::max(max_idx, 1u), SR_MAX_TEXTURE_EXTENT);
  const uint32_t atlas_h = std::min(std::max((max_idx + atlas_w - 1) / atlas_w, 1u), SR_MAX_TEXTURE_EXTENT);
  const uint32_t atlas_texels = atlas_w * atlas_h;
  const uint64_t atlas_n = (uint64_t)cams * NI * atlas_texels;
  const uint32_t tiles16 = coarse_total * kExt * kExt;
  // page-CSR hit sort: 8-bit digits over the page of page * 256 + cell
  // (page_cap = dropped); cells are sorted per page afterwards
  uint32_t page_bits = 1;
  while (page_bits < 24 && (1ull << page_bits) <= page_cap) ++page_bits;
  const uint32_t radix_passes = csr ? (page_bits + 7u) / 8u : 0u;
  const uint32_t radix_blocks = csr ? (hit_cap + kRadixBlock - 1u) / kRadixBlock : 0u;
  const uint64_t scan_len = std::max<uint64_t>({(uint64_t)task_cap, (uint64_t)ptab, (uint64_t)tiles16,
                                                256ull * radix_blocks, 1ull});
  // scan tile partials; the radix digit-base scan keeps its chunk sums there too
  const uint32_t scan_tiles = std::max<uint32_t>((uint32_t)((scan_len + kScanTile - 1) / kScanTile + 1),
                                                 csr ? kDigitChunks * 256u : 0u);

  // history: keep the previous buffer when the atlas layout is unchanged
  const bool hist_valid = !na && st.shadow_history_valid && sl.shape.valid && sl.shape.cams == cams &&
                          sl.shape.NI == NI && sl.shape.atlas_texels == atlas_texels;
  const int hist_next = hist_valid ? (sl.hist_cur ^ 1) : 0;

  struct Need { BufId id; uint64_t bytes; };
  // DOM depth texels: every light's W * H (dense) or the pool's pages
  const uint64_t dom_tex = vdom ? (uint64_t)dom_pool_cap * kDomPageTexels : dom_px;
  // the per-frame tables share one buffer (one upload): frusta, instances,
  // lights, each section 256-byte aligned
  const uint64_t tab_inst = align_up(sizeof(FrustumRec) * F, 256);
  const uint64_t tab_light = tab_inst + align_up(sizeof(InstRec) * std::max(NI, 1u), 256);
  const uint64_t tab_bytes = tab_light + sizeof(sr_directional_light) * std::max(NL, 1u);
  const Need needs[] = {
      {B_COUNTERS, 4ull * C_COUNT}, {B_BROAD, 16ull * NI + 16}, {B_KEEP, 4ull * NI * F + 4},
      {B_RANGES, 32ull * NI * F + 32}, {B_SLOT_MASKS, 8ull * task_cap / 64}, {B_WAVE_OFF, 4ull * task_cap / 64},
      {B_FR_TASK_BASE, 4ull * (F + 1)}, {B_FR_CAM_BASE, 4ull * F}, {B_SHADED, (uint64_t)std::max(cams, 1u) * slots},
      {B_TASKS, 16ull * task_cap}, {B_PROJ, 12ull * task_cap}, {B_SHADING_Q, 4ull * task_cap},
      {B_LUT, 16ull * ptab}, {B_LUT_Z, 8ull * (ptab + coarse_total)}, {B_PAGE_TABLE, 4ull * ptab}, {B_PAGE_META, 32ull * page_cap},
      {B_PAGE_TOTALS, 4ull * page_cap}, {B_CELL_COUNTS, 4ull * page_cap * kPage},
      {B_CELL_OFFSETS, 4ull * page_cap * kPage}, {B_CELL_CURSORS, 4ull * page_cap * kPage},
      {B_REFS, 12ull * ref_cap}, {B_ACTIVE_FLAGS, 4ull * tiles16}, {B_ACTIVE_Q, 4ull * tiles16},
      {B_TILE_WORK, 16ull * tiles16}, {B_TILE_BASES, 16ull * tiles16}, {B_WORK, 32ull * work_cap},
      {B_CAM_RUNS, 32ull * tiles16}, {B_SHADOW_RUNS, 32ull * tiles16}, {B_CAM_ORDER, 4ull * tiles16}, {B_SHADOW_ORDER, 4ull * tiles16}, {B_RUN_HIST, 4ull * 2 * kOrderBlocks * 32}, {B_PAGE_BOUNDS, csr ? 4ull * (page_cap + 1) : 0ull},
      {B_PAGE_ORDER, csr ? 4ull * page_cap : 0ull}, {B_PAGE_HIST, csr ? 4ull * kOrderBlocks * 32 : 0ull}, {B_DOM_DEPTH, 4ull * dom_tex},
      {B_DOM_OPACITY, 4ull * dom_tex * st.dom_layers}, {B_STAMP, 4ull * stamp_px}, {B_ATLAS, 16ull * atlas_n},
      {B_HIST0, 8ull * atlas_n}, {B_HIST1, 8ull * atlas_n}, {B_IMAGE, 16ull * img_px}, {B_DEPTH, 4ull * img_px},
      {B_COMPOSITE, 16ull * img_px}, {B_SHADOW_COLS, kShadowColBytes}, {B_SLOT_PROJ, 12ull * task_cap}, {B_SLOT_SCREEN, 32ull * task_cap}, {B_PARTIALS, 4ull * scan_tiles},
      {B_PARTIALS4, 16ull * scan_tiles}, {B_XY_MASK, 4ull * page_cap},
      {B_SLOT_HIT_TOTAL, csr ? 4ull * task_cap : 0ull}, {B_HIT_BASE, csr ? 4ull * task_cap : 0ull},
      {B_SLOT_MATERIAL, csr ? 4ull * task_cap : 0ull}, {B_HKEYS0, 4ull * hit_cap}, {B_HKEYS1, 4ull * hit_cap},
      {B_HVALS0, 4ull * hit_cap}, {B_HVALS1, 4ull * hit_cap}, {B_RADIX_HIST, 4ull * 256u * radix_blocks}, {B_RADIX_BASE, 4ull * 256u * radix_blocks},
      {B_SLOT_HITS, hit_k ? hit_bytes : 0ull},
      {B_SLOT_HIT_COUNT, hit_k ? (uint64_t)task_cap : 0ull},
      {B_TELEMETRY, st.telemetry_enabled ? 4ull * (kTelPageOff + page_cap + kTelFrustumStride * F) : 0ull},
      {B_OPAQUE_DEPTH, 4ull * opaque_floats}, {B_OPAQUE_TILES, opaque_floats ? 8ull * fine_total : 0ull}, {B_TABLES, tab_bytes},
      {B_IMAGE8, rgba8 ? 4ull * img_px : 0ull}, {B_COMPOSITE8, rgba8 ? 4ull * img_px : 0ull},
      {B_DOM_REQ, 4ull * dom_vpages}, {B_DOM_TABLE, 4ull * dom_vpages}};
  if (need_bytes) {  // what the slot holds once every buffer fits (DBuf::ensure growth rule)
    uint64_t held = sl.held_bytes();
    for (const Need& n : needs) {
      const uint64_t want = std::max<uint64_t>(n.bytes, 256);
      if (want > sl.buf[n.id].bytes) held += align_up(want + want / 4, 4096) - sl.buf[n.id].bytes;
    }
    *need_bytes = held;
  }
  if (own_bytes) {  // what the frame alone would make an empty slot allocate
    uint64_t own = 0;
    for (const Need& n : needs) {
      const uint64_t want = std::max<uint64_t>(n.bytes, 256);
      own += align_up(want + want / 4, 4096);
    }
    *own_bytes = own;
  }
  if (dry_run) return SR_OK;
  const uint64_t tbytes = tab_light + sizeof(sr_directional_light) * NL;
  if (na) {
    // carve every buffer out of the scratch array; the composite and depth
    // are the caller's
    uint64_t off = 0;
    for (const Need& n : needs) {
      if (n.id == B_COMPOSITE || n.id == B_DEPTH) continue;
      off = align_up(off, 256);
      const uint64_t b = std::max<uint64_t>(n.bytes, 256);
      if (off + b <= na->bytes) sl.buf[n.id].borrow(na->base + off, b);
      off += b;
    }
    if (off > na->bytes)
      return efail(err, SR_ERR_CAPACITY,
                   "sr_frame_submit_noalloc: the frame needs " + std::to_string(off) +
                       " bytes of scratch; the no-allocation path holds at most " + std::to_string(na->bytes) +
                       " (SR_NOALLOC_SCRATCH_BYTES)");
    sl.buf[B_COMPOSITE].borrow(na->composite, 16ull * img_px);
    sl.buf[B_DEPTH].borrow(na->depth, 4ull * img_px);
    // per-frame tables as kernel arguments, a few hundred words per launch
    std::vector<uint32_t> words((tbytes + 3) / 4, 0u);
    char* wb = reinterpret_cast<char*>(words.data());
    std::memcpy(wb, fr.data(), sizeof(FrustumRec) * F);
    if (NI) std::memcpy(wb + tab_inst, ir.data(), sizeof(InstRec) * NI);
    if (NL) std::memcpy(wb + tab_light, fd->lights, sizeof(sr_directional_light) * NL);
    uint32_t* dst = sl.buf[B_TABLES].as<uint32_t>();
    for (size_t i = 0; i < words.size(); i += kPutWords) {
      PutWords c;
      c.n = (uint32_t)std::min<size_t>(kPutWords, words.size() - i);
      std::memcpy(c.w, words.data() + i, 4ull * c.n);
      k_put_words<<<1, 256, 0, s>>>(dst + i, c);
    }
  } else {
    bool grow = false;
    for (const Need& n : needs) grow = grow || std::max<uint64_t>(n.bytes, 256) > sl.buf[n.id].bytes || sl.buf[n.id].borrowed;
    if (grow && sl.submitted) {  // in-flight frames may still use the buffers
      (void)hipStreamSynchronize(s);
      (void)hipStreamSynchronize(sl.stream);
    }
    if (grow) sl.drop_graphs();  // captured launches hold the old buffer addresses
    for (const Need& n : needs) {
      if (!sl.buf[n.id].ensure(n.bytes)) return efail(err, SR_ERR_OUT_OF_MEMORY, "sr_frame_submit: device allocation failed");
    }
  }
  // ---- upload per-frame tables through pinned staging
  int slot = 0;
  char* stg = na ? nullptr : next_staging(sl, tab_bytes + 256, &slot);
  if (!na && !stg) return efail(err, SR_ERR_OUT_OF_MEMORY, "staging allocation failed");
  if (stg) {
    std::memset(stg, 0, tbytes);
    std::memcpy(stg, fr.data(), sizeof(FrustumRec) * F);
    if (NI) std::memcpy(stg + tab_inst, ir.data(), sizeof(InstRec) * NI);
    if (NL) std::memcpy(stg + tab_light, fd->lights, sizeof(sr_directional_light) * NL);
    HIP_OKS(hipMemcpyAsync(sl.buf[B_TABLES].p, stg, tbytes, hipMemcpyHostToDevice, s), "upload frame tables");
  }
  if (opaque_floats) {  // the opaque depth travels through its own staging entry
    int oslot = 0;
    char* ostg = next_staging(sl, 4ull * opaque_floats, &oslot);
    if (!ostg) return efail(err, SR_ERR_OUT_OF_MEMORY, "staging allocation failed");
    std::memcpy(ostg, fd->opaque_depth, 4ull * opaque_floats);
    HIP_OKS(hipMemcpyAsync(sl.buf[B_OPAQUE_DEPTH].p, ostg, 4ull * opaque_floats, hipMemcpyHostToDevice, s),
            "upload opaque depth");
    HIP_OKS(hipEventRecord(sl.staging_ev[oslot], s), "staging event");
    sl.staging_used[oslot] = true;
  }
  if (stg) {
    HIP_OKS(hipEventRecord(sl.staging_ev[slot], s), "staging event");
    sl.staging_used[slot] = true;
  }

  FrameParams fp;
  std::memset(&fp, 0, sizeof(fp));
  char* tab = sl.buf[B_TABLES].as<char>();
  fp.fr = reinterpret_cast<FrustumRec*>(tab);
  fp.inst = reinterpret_cast<InstRec*>(tab + tab_inst);
  fp.lights = reinterpret_cast<sr_directional_light*>(tab + tab_light);
  fp.F = F; fp.NI = NI; fp.NL = NL; fp.cams = cams;
  fp.slots = slots; fp.coarse_total = coarse_total;
  fp.task_cap = task_cap; fp.ref_cap = ref_cap; fp.work_cap = work_cap; fp.atlas_texels = atlas_texels;
  fp.dom_layers = st.dom_layers;
  fp.mark_waves = fd->mark_waves ? fd->mark_waves : 4u;
  fp.raster_waves = fd->raster_waves;  // 0: the kernel's default
  fp.slices = slices;
  fp.slice_shift = slices == 8u ? 3u : (slices == 32u ? 5u : 4u);
  fp.st = st;
  fp.st.shadow_history_valid = hist_valid ? 1u : 0u;
  fp.counters = sl.buf[B_COUNTERS].as<uint32_t>();
  fp.broad = sl.buf[B_BROAD].as<sr_broad_instance_meta>();
  fp.keep = sl.buf[B_KEEP].as<float>();
  fp.ranges = sl.buf[B_RANGES].as<sr_coarse_asset_range>();
  fp.slot_masks = sl.buf[B_SLOT_MASKS].as<unsigned long long>();
  fp.wave_offsets = sl.buf[B_WAVE_OFF].as<uint32_t>();
  fp.frustum_task_base = sl.buf[B_FR_TASK_BASE].as<uint32_t>();
  fp.frustum_cam_base = sl.buf[B_FR_CAM_BASE].as<uint32_t>();
  fp.shaded = sl.buf[B_SHADED].as<uint8_t>();
  fp.tasks = sl.buf[B_TASKS].as<sr_binning_task>();
  fp.proj = sl.buf[B_PROJ].as<sr_projected_segment>();
  fp.shading_queue = sl.buf[B_SHADING_Q].as<uint32_t>();
  fp.lut = sl.buf[B_LUT].as<sr_coarse_depth_lut_entry>();
  fp.lut_z = sl.buf[B_LUT_Z].as<uint2>();
  fp.lut_iv = fp.lut_z + ptab;  // after the per-slice pairs
  fp.page_table = sl.buf[B_PAGE_TABLE].as<uint32_t>();
  fp.page_meta = sl.buf[B_PAGE_META].as<sr_fine_page_meta>();
  fp.page_totals = sl.buf[B_PAGE_TOTALS].as<uint32_t>();
  fp.cell_counts = sl.buf[B_CELL_COUNTS].as<uint32_t>();
  fp.cell_offsets = sl.buf[B_CELL_OFFSETS].as<uint32_t>();
  fp.cell_cursors = sl.buf[B_CELL_CURSORS].as<uint32_t>();
  fp.refs = sl.buf[B_REFS].as<sr_fine_seg_ref>();
  fp.active_flags = sl.buf[B_ACTIVE_FLAGS].as<uint32_t>();
  fp.active_queue = sl.buf[B_ACTIVE_Q].as<uint32_t>();
  fp.tile_work = sl.buf[B_TILE_WORK].as<uint4>();
  fp.tile_bases = sl.buf[B_TILE_BASES].as<uint4>();
  fp.work = sl.buf[B_WORK].as<sr_raster_work_item>();
  fp.cam_runs = sl.buf[B_CAM_RUNS].as<sr_raster_tile_run>();
  fp.shadow_runs = sl.buf[B_SHADOW_RUNS].as<sr_raster_tile_run>();
  fp.cam_order = sl.buf[B_CAM_ORDER].as<uint32_t>();
  fp.shadow_order = sl.buf[B_SHADOW_ORDER].as<uint32_t>();
  fp.run_hist = sl.buf[B_RUN_HIST].as<uint32_t>();
  fp.page_bounds = sl.buf[B_PAGE_BOUNDS].as<uint32_t>();
  fp.page_order = sl.buf[B_PAGE_ORDER].as<uint32_t>();
  fp.page_hist = sl.buf[B_PAGE_HIST].as<uint32_t>();
  fp.dom_depth = sl.buf[B_DOM_DEPTH].as<float>();
  fp.dom_opacity = sl.buf[B_DOM_OPACITY].as<float>();
  fp.dom_requests = vdom ? sl.buf[B_DOM_REQ].as<uint32_t>() : nullptr;
  fp.dom_table = vdom ? sl.buf[B_DOM_TABLE].as<uint32_t>() : nullptr;
  fp.dom_vpages = dom_vpages;
  fp.dom_pool_cap = dom_pool_cap;
  fp.stamp = sl.buf[B_STAMP].as<float>();
  fp.atlas = sl.buf[B_ATLAS].as<float4>();
  fp.history = sl.buf[hist_next ? B_HIST1 : B_HIST0].as<float2>();
  fp.history_prev = hist_valid ? sl.buf[hist_next ? B_HIST0 : B_HIST1].as<float2>() : nullptr;
  fp.image = sl.buf[B_IMAGE].as<float4>();
  fp.depth = sl.buf[B_DEPTH].as<float>();
  fp.composite = sl.buf[B_COMPOSITE].as<float4>();
  fp.image8 = rgba8 ? sl.buf[B_IMAGE8].as<uint32_t>() : nullptr;
  fp.composite8 = rgba8 ? sl.buf[B_COMPOSITE8].as<uint32_t>() : nullptr;
  fp.shadow_cols = sl.buf[B_SHADOW_COLS].as<unsigned long long>();
  // page_cap bounds the page arrays; the allocation scan drops pages past it
  PrepassScratch ps{};
  ps.slot_proj = sl.buf[B_SLOT_PROJ].as<sr_projected_segment>();
  ps.partials = sl.buf[B_PARTIALS].as<uint32_t>();
  ps.partials4 = sl.buf[B_PARTIALS4].as<uint4>();
  ps.page_xy_mask = sl.buf[B_XY_MASK].as<uint32_t>();
  ps.hkeys[0] = sl.buf[B_HKEYS0].as<uint32_t>();
  ps.hkeys[1] = sl.buf[B_HKEYS1].as<uint32_t>();
  ps.hvals[0] = sl.buf[B_HVALS0].as<uint32_t>();
  ps.hvals[1] = sl.buf[B_HVALS1].as<uint32_t>();
  ps.radix_hist = sl.buf[B_RADIX_HIST].as<uint32_t>();
  ps.radix_base = sl.buf[B_RADIX_BASE].as<uint32_t>();
  ps.radix_blocks = radix_blocks;
  ps.radix_passes = radix_passes;
  fp.page_cap = page_cap;
  fp.page_xy_mask = ps.page_xy_mask;
  fp.slot_hit_total = sl.buf[B_SLOT_HIT_TOTAL].as<uint32_t>();
  fp.hit_base = sl.buf[B_HIT_BASE].as<uint32_t>();
  fp.slot_material = sl.buf[B_SLOT_MATERIAL].as<uint32_t>();
  // hits are written to the pair the last sort pass does not write; the
  // sorted keys land in pair 0
  fp.hkeys_w = ps.hkeys[(radix_passes & 1u) ? 1 : 0];
  fp.hvals_w = ps.hvals[(radix_passes & 1u) ? 1 : 0];
  fp.hkeys = ps.hkeys[0];
  fp.slot_proj = ps.slot_proj;
  fp.slot_screen = sl.buf[B_SLOT_SCREEN].as<float4>();
  fp.opaque_depth = opaque_floats ? sl.buf[B_OPAQUE_DEPTH].as<float>() : nullptr;
  fp.opaque_tiles = opaque_floats ? sl.buf[B_OPAQUE_TILES].as<uint2>() : nullptr;
  fp.opaque_samples = opaque_floats ? fd->opaque_samples : 0u;
  fp.fine_total = (uint32_t)fine_total;
  const uint64_t tel_words = st.telemetry_enabled ? kTelPageOff + (uint64_t)page_cap + kTelFrustumStride * F : 0ull;
  fp.tel = tel_words ? sl.buf[B_TELEMETRY].as<uint32_t>() : nullptr;
  fp.hit_k = hit_k;
  fp.mark_queue = (ctx->flags & SR_CONFIG_CSR_FEW_HITS) ? 2u : kMarkQ;
  fp.tile_queue = (ctx->flags & SR_CONFIG_CSR_FEW_HITS) ? 2u : kTileQ;
  fp.slot_hits = sl.buf[B_SLOT_HITS].as<uint32_t>();
  fp.slot_hit_count = sl.buf[B_SLOT_HIT_COUNT].as<uint8_t>();

  auto launches = [&]() {
    sl.timer.mark(s, "clear");
    FillSpans clear;  // every frame buffer that starts at zero, in one launch
    clear.add(fp.counters, C_COUNT);
    if (fp.tel) clear.add(fp.tel, tel_words);
    if (!reuse) {  // reused shadow maps stay as their frame left them
      clear.add(reinterpret_cast<uint32_t*>(fp.dom_depth), dom_tex);
      clear.add(reinterpret_cast<uint32_t*>(fp.dom_opacity), dom_tex * st.dom_layers);
      clear.add(reinterpret_cast<uint32_t*>(fp.stamp), stamp_px);
      clear.add(fp.dom_requests, dom_vpages);  // the page table is written whole
    }
    clear.add(reinterpret_cast<uint32_t*>(fp.image), img_px * 4);
    clear.add(reinterpret_cast<uint32_t*>(fp.depth), img_px);
    // (atlas and history texels no segment shades are cleared before the
    // shading pass, from the marking pass's shaded flags)
    launch_fill_spans(clear, s);
    launch_prepass(fp, ps, s, &sl.timer);
    if (vdom && !reuse) {
      sl.timer.mark(s, "dom_pages");
      launch_dom_pages(fp, s);
    }
    sl.timer.mark(s, "shadow");
    if (!reuse) launch_shadow(fp, s, max_light_tile_px);
    sl.timer.mark(s, "stampback");
    for (uint32_t i = 0; i < F && !reuse; ++i)
      if (fr[i].d.kind == 1u) launch_stampback(fp, i, fr[i].stamp_w * fr[i].stamp_h, s);
    sl.timer.mark(s, "shading");
    launch_shading(fp, s);
    sl.timer.mark(s, "raster");
    launch_camera_raster(fp, s, max_cam_tile_px);
    sl.timer.mark(s, "composite");
    launch_composite(fp, img_px, s);
    if (rgba8) {
      sl.timer.mark(s, "rgba8");
      launch_pack_rgba8(fp, img_px, s);
    }
    sl.timer.mark(s, "end");
  };
  sl.timer.begin();
  if ((ctx->flags & SR_CONFIG_HIP_GRAPHS) && !sl.timer.enabled && !na) {
    // The sequence reads every per-frame table from device memory, so one
    // capture serves every frame whose launches bake in the same values:
    // FrameParams (buffers, capacities, settings), the launch extents and
    // the stampback targets.
    std::vector<uint8_t> key(sizeof(fp) + 8 * sizeof(uint64_t) + 2 * sizeof(uint32_t) * F);
    uint8_t* k = key.data();
    std::memcpy(k, &fp, sizeof(fp));
    k += sizeof(fp);
    const uint64_t ext[8] = {dom_tex, stamp_px, img_px, atlas_n, max_light_tile_px, max_cam_tile_px, F, 0};
    std::memcpy(k, ext, sizeof(ext));
    k += sizeof(ext);
    for (uint32_t i = 0; i < F; ++i) {
      const uint32_t st2[2] = {fr[i].d.kind == 1u ? fr[i].stamp_w : 0u, fr[i].d.kind == 1u ? fr[i].stamp_h : 0u};
      std::memcpy(k, st2, sizeof(st2));
      k += sizeof(st2);
    }
    FrameSlot::Graph* g = nullptr;
    for (auto& c : sl.graphs)
      if (c.key == key) g = &c;
    if (!g) {
      if (sl.graphs.size() >= 4) {  // least recently used goes
        auto lru = std::min_element(sl.graphs.begin(), sl.graphs.end(),
                                    [](const FrameSlot::Graph& a, const FrameSlot::Graph& b) { return a.last_use < b.last_use; });
        (void)hipStreamSynchronize(s);  // it may still be running
        if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
        sl.graphs.erase(lru);
      }
      hipGraph_t graph = nullptr;
      HIP_OKS(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "sr_frame_submit: begin capture");
      launches();
      hipError_t ce = hipStreamEndCapture(s, &graph);
      if (ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return efail(err, SR_ERR_DEVICE, std::string("sr_frame_submit: capture: ") + hipGetErrorString(ce));
      }
      FrameSlot::Graph ng;
      ng.key = std::move(key);
      hipError_t ie = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return efail(err, SR_ERR_DEVICE, std::string("sr_frame_submit: instantiate: ") + hipGetErrorString(ie));
      sl.graphs.push_back(std::move(ng));
      g = &sl.graphs.back();
      ++sl.graph_captures;
    } else {
      ++sl.graph_replays;
    }
    g->last_use = ++sl.graph_clock;
    HIP_OKS(hipGraphLaunch(g->exec, s), "sr_frame_submit: graph launch");
  } else {
    launches();
  }
  if (na && na->counters) {
    const CounterCaps caps{coarse_total, task_cap, ref_cap, work_cap, hit_cap};
    k_counters_out<<<1, 64, 0, s>>>(fp.counters, caps, na->counters);
  }
  HIP_OKS(hipGetLastError(), "sr_frame_submit: launch");
  HIP_OKS(hipEventRecord(sl.done, s), "sr_frame_submit: event");

  FrameShape& sh = sl.shape;
  sh.valid = true;
  sh.F = F; sh.NI = NI; sh.NL = NL; sh.cams = cams; sh.slots = slots; sh.coarse_total = coarse_total;
  sh.slices = slices;
  sh.task_cap = task_cap; sh.ref_cap = ref_cap; sh.work_cap = work_cap; sh.page_cap = page_cap;
  sh.hit_cap = hit_cap;
  sh.tel_words = tel_words;
  sh.opaque_tiles = opaque_floats ? fine_total : 0ull;
  sh.atlas_texels = atlas_texels; sh.dom_layers = st.dom_layers;
  sh.dom_px = dom_px; sh.img_px = img_px; sh.stamp_px = stamp_px;
  sh.rgba8 = rgba8;
  sh.dom_virtual = vdom;
  sh.dom_vpages = dom_vpages;
  sh.dom_pool_cap = dom_pool_cap;
  sh.dom_lights = std::move(dom_lights);
  sh.hist_cur = hist_next;
  sl.hist_cur = hist_next;
  sl.submitted = true;
  if (!reuse) {  // this frame built the slot's shadow maps; usable once it completes without overflow
    sl.dom_key = std::move(dom_key);
    sl.dom_pending = true;
    sl.dom_ready = false;
  }
  return SR_OK;
}

// Waits for the slot's last frame; counters and grow-only hints.
int frame_wait_slot(FrameSlot& sl, sr_frame_counters* out, std::string& err) {
  if (!sl.submitted) return efail(err, SR_ERR_NOT_READY, "sr_frame_wait: no frame submitted");
  HIP_OKS(hipEventSynchronize(sl.done), "sr_frame_wait");
  uint32_t c[C_COUNT];
  HIP_OKS(hipMemcpy(c, sl.buf[B_COUNTERS].p, sizeof(c), hipMemcpyDeviceToHost), "sr_frame_wait: counters");
  const FrameShape& sh = sl.shape;
  const sr_frame_counters fc =
      counters_from_words(c, CounterCaps{sh.coarse_total, sh.task_cap, sh.ref_cap, sh.work_cap, sh.hit_cap});
  // grow-only capacity planning for the next frame
  if (c[C_REFS] > sh.ref_cap) sl.ref_cap_hint = next_pow2(c[C_REFS]);
  if (c[C_PAGE_DROPS] > 0) sl.page_cap_hint = next_pow2(c[C_MARKED]);
  if (c[C_WORK] > sh.work_cap) sl.work_cap_hint = next_pow2(c[C_WORK]);
  if (sh.hit_cap && c[C_CANDS] > sh.hit_cap) sl.ref_cap_hint = std::max(sl.ref_cap_hint, next_pow2(c[C_CANDS]));
  if (c[C_DOM_DROPS] > 0) sl.dom_pool_hint = std::max(sl.dom_pool_hint, c[C_DOM_WANT]);
  if (sl.dom_pending) {
    sl.dom_ready = fc.capacity_exceeded == 0u;
    sl.dom_pending = false;
  }
  if (out) *out = fc;
  return SR_OK;
}

}  // namespace

extern "C" {

int sr_frame_submit(sr_context* ctx, const sr_frame_desc* fd, void* stream_ptr) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  std::string err;
  int rc = frame_submit_slot(ctx, ctx->main(), fd, (hipStream_t)stream_ptr, err, nullptr, false);
  return rc == SR_OK ? rc : fail(ctx, rc, err);
}

int sr_frame_submit_noalloc(sr_context* ctx, const sr_frame_desc* fd, float* composite, float* depth,
                            sr_frame_counters* counters, void* stream_ptr) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  if (!composite || !depth) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "sr_frame_submit_noalloc: NULL output");
  if (!ctx->scratch || !ctx->na)
    return fail(ctx, SR_ERR_DEVICE, "sr_frame_submit_noalloc: no scratch on this device (" + ctx->scratch_err + ")");
  (void)hipSetDevice(ctx->device);
  const NoAllocTarget na{static_cast<char*>(ctx->scratch), kScratchBytes, reinterpret_cast<float4*>(composite), depth,
                         counters};
  std::string err;
  int rc = frame_submit_slot(ctx, *ctx->na, fd, (hipStream_t)stream_ptr, err, nullptr, false, &na);
  return rc == SR_OK ? rc : fail(ctx, rc, err);
}

int sr_frame_wait(sr_context* ctx, sr_frame_counters* out) {
  if (!ctx) return SR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  std::string err;
  int rc = frame_wait_slot(ctx->main(), out, err);
  return rc == SR_OK ? rc : fail(ctx, rc, err);
}

int sr_frame_dump(sr_context* ctx, int32_t kind, void* dst, uint64_t capacity, uint64_t* out_bytes) {
  if (!ctx || !out_bytes) return SR_ERR_INVALID_ARGUMENT;
  sr_frame_counters fc;
  int rc = sr_frame_wait(ctx, &fc);
  if (rc != SR_OK) return rc;
  const FrameShape& sh = ctx->main().shape;
  DBuf* buf = ctx->main().buf;
  const void* src = nullptr;
  uint64_t bytes = 0;
  uint32_t P = std::min(fc.allocated_pages, sh.page_cap);
  switch (kind) {
    case SR_DUMP_COUNTERS:
      *out_bytes = sizeof(fc);
      if (dst) {
        if (capacity < sizeof(fc)) return fail(ctx, SR_ERR_CAPACITY, "dump buffer too small");
        std::memcpy(dst, &fc, sizeof(fc));
      }
      return SR_OK;
    case SR_DUMP_BROAD_META: src = buf[B_BROAD].p; bytes = 16ull * sh.NI; break;
    case SR_DUMP_KEEP: src = buf[B_KEEP].p; bytes = 4ull * sh.NI * sh.F; break;
    case SR_DUMP_RANGES: src = buf[B_RANGES].p; bytes = 32ull * sh.NI * sh.F; break;
    case SR_DUMP_TASKS: src = buf[B_TASKS].p; bytes = 16ull * std::min(fc.binning_tasks, sh.task_cap); break;
    case SR_DUMP_PROJECTED: src = buf[B_PROJ].p; bytes = 12ull * std::min(fc.binning_tasks, sh.task_cap); break;
    case SR_DUMP_SHADING_QUEUE: src = buf[B_SHADING_Q].p; bytes = 4ull * fc.camera_tasks; break;
    case SR_DUMP_LUT: src = buf[B_LUT].p; bytes = 16ull * sh.coarse_total * sh.slices; break;
    case SR_DUMP_PAGE_TABLE: src = buf[B_PAGE_TABLE].p; bytes = 4ull * sh.coarse_total * sh.slices; break;
    case SR_DUMP_PAGE_META: src = buf[B_PAGE_META].p; bytes = 32ull * P; break;
    case SR_DUMP_CELL_COUNTS: src = buf[B_CELL_COUNTS].p; bytes = 4ull * P * kPage; break;
    case SR_DUMP_CELL_OFFSETS: src = buf[B_CELL_OFFSETS].p; bytes = 4ull * P * kPage; break;
    case SR_DUMP_REFS: src = buf[B_REFS].p; bytes = 12ull * std::min(fc.fine_refs, sh.ref_cap); break;
    case SR_DUMP_ACTIVE_TILES: src = buf[B_ACTIVE_Q].p; bytes = 4ull * fc.active_tiles; break;
    case SR_DUMP_WORK_ITEMS: src = buf[B_WORK].p; bytes = 32ull * std::min(fc.work_items, sh.work_cap); break;
    case SR_DUMP_CAMERA_RUNS: src = buf[B_CAM_RUNS].p; bytes = 32ull * fc.camera_runs; break;
    case SR_DUMP_SHADOW_RUNS: src = buf[B_SHADOW_RUNS].p; bytes = 32ull * fc.shadow_runs; break;
    case SR_DUMP_DOM_DEPTH:
    case SR_DUMP_DOM_OPACITY: {
      const bool depth = kind == SR_DUMP_DOM_DEPTH;
      if (!sh.dom_virtual) {
        src = buf[depth ? B_DOM_DEPTH : B_DOM_OPACITY].p;
        bytes = 4ull * sh.dom_px * (depth ? 1u : sh.dom_layers);
        break;
      }
      // virtual DOMs in the dense layout: each light's texels from the page
      // table and the pool, 0 where the page is not resident
      const uint32_t L = depth ? 1u : sh.dom_layers;
      uint64_t texels = 0;
      for (size_t i = 0; i < sh.dom_lights.size(); i += 5) texels += (uint64_t)sh.dom_lights[i] * sh.dom_lights[i + 1];
      *out_bytes = 4ull * texels * L;
      if (!dst) return SR_OK;
      if (capacity < *out_bytes) return fail(ctx, SR_ERR_CAPACITY, "dump buffer too small");
      std::vector<uint32_t> table(sh.dom_vpages);
      std::vector<float> pool((size_t)sh.dom_pool_cap * kDomPageTexels * L);
      if (!table.empty())
        HIP_OK(hipMemcpy(table.data(), buf[B_DOM_TABLE].p, 4ull * table.size(), hipMemcpyDeviceToHost), "sr_frame_dump: page table");
      if (!pool.empty())
        HIP_OK(hipMemcpy(pool.data(), buf[depth ? B_DOM_DEPTH : B_DOM_OPACITY].p, 4ull * pool.size(), hipMemcpyDeviceToHost),
               "sr_frame_dump: DOM pool");
      float* out = static_cast<float*>(dst);
      std::memset(out, 0, *out_bytes);
      uint64_t base = 0;
      for (size_t li = 0; li < sh.dom_lights.size(); li += 5) {
        const uint32_t W = sh.dom_lights[li], H = sh.dom_lights[li + 1], pxn = sh.dom_lights[li + 2];
        const uint32_t vb = sh.dom_lights[li + 4];
        for (uint32_t i = 0; i < L; ++i)
          for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
              const uint32_t e = table[vb + (y / kDomPage) * pxn + x / kDomPage];
              if (e == kInvalid) continue;
              out[base + ((uint64_t)i * H + y) * W + x] =
                  pool[((size_t)e * L + i) * kDomPageTexels + (size_t)(y % kDomPage) * kDomPage + x % kDomPage];
            }
        base += (uint64_t)W * H * L;
      }
      return SR_OK;
    }
    case SR_DUMP_DOM_REQUESTS: src = buf[B_DOM_REQ].p; bytes = sh.dom_virtual ? 4ull * sh.dom_vpages : 0ull; break;
    case SR_DUMP_DOM_PAGE_TABLE: src = buf[B_DOM_TABLE].p; bytes = sh.dom_virtual ? 4ull * sh.dom_vpages : 0ull; break;
    case SR_DUMP_ATLAS: src = buf[B_ATLAS].p; bytes = 16ull * sh.cams * sh.NI * sh.atlas_texels; break;
    case SR_DUMP_HISTORY:
      src = buf[sh.hist_cur ? B_HIST1 : B_HIST0].p;
      bytes = 8ull * sh.cams * sh.NI * sh.atlas_texels;
      break;
    case SR_DUMP_IMAGE: src = buf[B_IMAGE].p; bytes = 16ull * sh.img_px; break;
    case SR_DUMP_DEPTH: src = buf[B_DEPTH].p; bytes = 4ull * sh.img_px; break;
    case SR_DUMP_COMPOSITE: src = buf[B_COMPOSITE].p; bytes = 16ull * sh.img_px; break;
    case SR_DUMP_STAMP: src = buf[B_STAMP].p; bytes = 4ull * sh.stamp_px; break;
    case SR_DUMP_TELEMETRY: src = buf[B_TELEMETRY].p; bytes = 4ull * sh.tel_words; break;
    case SR_DUMP_OPAQUE_TILES: src = buf[B_OPAQUE_TILES].p; bytes = 8ull * sh.opaque_tiles; break;
    case SR_DUMP_IMAGE_RGBA8: src = buf[B_IMAGE8].p; bytes = sh.rgba8 ? 4ull * sh.img_px : 0ull; break;
    case SR_DUMP_COMPOSITE_RGBA8: src = buf[B_COMPOSITE8].p; bytes = sh.rgba8 ? 4ull * sh.img_px : 0ull; break;
    default: return fail(ctx, SR_ERR_INVALID_ARGUMENT, "sr_frame_dump: unknown kind");
  }
  *out_bytes = bytes;
  if (!dst) return SR_OK;
  if (capacity < bytes) return fail(ctx, SR_ERR_CAPACITY, "dump buffer too small");
  if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "sr_frame_dump: copy");
  return SR_OK;
}

int sr_stage_times(sr_context* ctx, float* dst_ms, int32_t capacity, int32_t* out_count, char* names,
                   int32_t names_capacity) {
  if (!ctx || !out_count) return SR_ERR_INVALID_ARGUMENT;
  StageTimer& t = ctx->main().timer;
  if (!t.enabled || t.used < 2) { *out_count = 0; return SR_OK; }
  HIP_OK(hipEventSynchronize(t.events[t.used - 1]), "sr_stage_times");
  int n = (int)t.used - 1;
  std::string joined;
  for (int i = 0; i < n; ++i) {
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, t.events[i], t.events[i + 1]), "sr_stage_times: elapsed");
    if (dst_ms && i < capacity) dst_ms[i] = ms;
    joined += t.names[i];
    joined += '\n';
  }
  if (names && names_capacity > 0) {
    std::strncpy(names, joined.c_str(), (size_t)names_capacity - 1);
    names[names_capacity - 1] = 0;
  }
  *out_count = n;
  return SR_OK;
}

}  // extern "C"

// ===================================================================== serving
// A worker thread over frame slots: completes finished frames (copies each
// request's composite and depth out of pinned staging), then admits queued
// requests into free slots.  Admission takes the oldest request and every
// later queued request with the same instance, lights and settings (up to
// max_views_per_frame cameras) as one multi-view frame, when the bytes held
// by assets and every slot after this frame stay within the budget; idle
// slots give their buffers back first.  A request that does not fit the
// budget even alone fails with SR_ERR_CAPACITY.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

// Segment slots (frusta x index positions) a served frame may hold.
constexpr uint64_t kFrameSlots = 1ull << 27;

struct PendingRequest {
  sr_request req;
  uint64_t ticket;
  double submit_ms;
};

struct ServerSlot {
  FrameSlot* fs = nullptr;
  bool busy = false;
  std::vector<PendingRequest> reqs;  // cameras of the frame, in frustum order
  std::vector<uint64_t> out_off;     // byte offset of each request's composite in staging
  void* out = nullptr;               // pinned composite + depth of every camera
  uint64_t out_bytes = 0;
  hipEvent_t out_ev = nullptr;
  double launch_ms = 0.0;
  uint32_t retries = 0;
  sr_frame_counters counters{};
};

bool same_group(const sr_request& a, const sr_request& b) {
  if (std::memcmp(&a.instance, &b.instance, sizeof(a.instance)) != 0) return false;
  if (a.light_count != b.light_count) return false;
  if (std::memcmp(&a.settings, &b.settings, sizeof(a.settings)) != 0) return false;
  for (uint32_t l = 0; l < a.light_count; ++l) {
    if (std::memcmp(&a.light_frusta[l], &b.light_frusta[l], sizeof(a.light_frusta[l])) != 0) return false;
    if (std::memcmp(&a.light_views[l], &b.light_views[l], sizeof(a.light_views[l])) != 0) return false;
    if (std::memcmp(&a.lights[l], &b.lights[l], sizeof(a.lights[l])) != 0) return false;
  }
  return true;
}

// Frame tables of a group: cameras in request order, then the shared lights.
struct GroupFrame {
  std::vector<sr_frustum_desc> frusta;
  std::vector<sr_view> views;
  std::vector<sr_directional_light> lights;
  sr_strand_instance instance;
  sr_frame_desc fd;
};

int build_group_frame(const std::vector<PendingRequest>& g, GroupFrame& out) {
  const sr_request& h = g[0].req;
  const uint32_t nc = (uint32_t)g.size(), nl = h.light_count, F = nc + nl;
  std::vector<sr_froxel_config> cfg(F);
  std::vector<uint32_t> is_light(F, 0u);
  for (uint32_t i = 0; i < nc; ++i) cfg[i] = g[i].req.camera;
  for (uint32_t l = 0; l < nl; ++l) {
    cfg[nc + l] = h.light_frusta[l];
    is_light[nc + l] = 1u;
  }
  out.frusta.resize(F);
  int rc = sr_build_frustum_table(cfg.data(), is_light.data(), nullptr, F, out.frusta.data());
  if (rc != SR_OK) return rc;
  out.views.resize(F);
  for (uint32_t i = 0; i < nc; ++i) out.views[i] = g[i].req.camera_view;
  for (uint32_t l = 0; l < nl; ++l) out.views[nc + l] = h.light_views[l];
  out.lights.assign(h.lights, h.lights + nl);
  out.instance = h.instance;
  std::memset(&out.fd, 0, sizeof(out.fd));
  out.fd.instances = &out.instance;
  out.fd.instance_count = 1;
  out.fd.frustum_count = F;
  out.fd.frusta = out.frusta.data();
  out.fd.views = out.views.data();
  out.fd.lights = nl ? out.lights.data() : nullptr;
  out.fd.light_count = nl;
  out.fd.settings = h.settings;
  for (const PendingRequest& p : g)  // rgba8unorm outputs when any request of the frame asks for them
    if (p.req.flags & SR_REQUEST_RGBA8) out.fd.settings.flags |= SR_FRAME_RGBA8_OUTPUTS;
  return SR_OK;
}

}  // namespace

struct sr_server {
  sr_context* ctx = nullptr;
  sr_server_config cfg{};
  uint64_t budget = 0;
  std::vector<ServerSlot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<PendingRequest> queue;
  std::deque<sr_request_result> results;
  sr_server_stats stats{};
  uint64_t next_ticket = 1;
  bool stopping = false;
  Clock::time_point t0;
  std::thread worker;

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

  uint64_t held_bytes() {
    uint64_t b;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      b = ctx->slabs.capacity();
    }
    for (auto& s : slots) b += s.fs->held_bytes() + s.out_bytes;
    return b;
  }

  void finish(ServerSlot& s, int status) {
    std::vector<sr_request_result> done;
    const double t = now_ms();
    for (size_t i = 0; i < s.reqs.size(); ++i) {
      const PendingRequest& p = s.reqs[i];
      if (status == SR_OK && !(p.req.flags & SR_REQUEST_PINNED_OUTPUTS)) {  // pinned outputs arrived directly
        const uint64_t px = (uint64_t)p.req.camera.screen_width * p.req.camera.screen_height;
        const uint64_t cb = px * ((p.req.flags & SR_REQUEST_RGBA8) ? 4u : 16u);
        const char* base = static_cast<const char*>(s.out) + s.out_off[i];
        if (p.req.out_rgba) std::memcpy(p.req.out_rgba, base, cb);
        if (p.req.out_depth) std::memcpy(p.req.out_depth, base + cb, px * 4u);
      }
      sr_request_result r{};
      r.ticket = p.ticket;
      r.tag = p.req.tag;
      r.status = status;
      r.frame_views = (uint32_t)s.reqs.size();
      r.submit_ms = p.submit_ms;
      r.launch_ms = s.launch_ms;
      r.done_ms = t;
      r.counters = s.counters;
      done.push_back(r);
    }
    s.reqs.clear();
    s.busy = false;
    std::lock_guard<std::mutex> lk(mu);
    for (auto& r : done) {
      results.push_back(r);
      if (r.status == SR_OK) ++stats.completed; else ++stats.failed;
    }
    cv_done.notify_all();
  }

  void fail_requests(std::vector<PendingRequest>& g, int status) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& p : g) {
      sr_request_result r{};
      r.ticket = p.ticket;
      r.tag = p.req.tag;
      r.status = status;
      r.submit_ms = p.submit_ms;
      r.launch_ms = r.done_ms = now_ms();
      results.push_back(r);
      ++stats.failed;
    }
    cv_done.notify_all();
  }

  // Completes finished frames; returns true when any slot is busy.
  bool reap() {
    bool any = false;
    for (auto& s : slots) {
      if (!s.busy) continue;
      hipError_t q = hipEventQuery(s.out_ev);
      if (q == hipErrorNotReady) { any = true; continue; }
      std::string err;
      int rc = q == hipSuccess ? frame_wait_slot(*s.fs, &s.counters, err) : SR_ERR_DEVICE;
      if (rc == SR_OK && s.counters.capacity_exceeded && s.retries < 3) {
        // the slot's hints grew: draw the same frame again
        ++s.retries;
        std::vector<PendingRequest> g = std::move(s.reqs);
        GroupFrame gf;
        rc = build_group_frame(g, gf);
        if (rc == SR_OK) rc = launch(s, g, gf, err);
        if (rc == SR_OK) {
          any = true;
          continue;
        }
        s.reqs = std::move(g);
      }
      if (rc == SR_OK && s.counters.capacity_exceeded) rc = SR_ERR_CAPACITY;
      finish(s, rc);
    }
    return any;
  }

  // Bytes of a request's outputs in the slot's staging: its composite (float
  // RGBA or rgba8unorm) and depth; none when the caller's buffers are pinned.
  static uint64_t staged_bytes(const sr_request& r) {
    if (r.flags & SR_REQUEST_PINNED_OUTPUTS) return 0u;
    const uint64_t px = (uint64_t)r.camera.screen_width * r.camera.screen_height;
    return px * ((r.flags & SR_REQUEST_RGBA8) ? 4u : 16u) + px * 4u;
  }

  // Launches one group on slot s: frame, then the copies of every camera's
  // composite and depth, into the slot's pinned staging or straight into
  // the caller's pinned buffers.
  int launch(ServerSlot& s, std::vector<PendingRequest>& g, GroupFrame& gf, std::string& err) {
    uint64_t bytes = 0;
    s.out_off.clear();
    for (auto& p : g) {
      s.out_off.push_back(bytes);
      bytes += staged_bytes(p.req);
    }
    if (bytes > s.out_bytes) {
      if (s.out) (void)hipHostFree(s.out);
      s.out = nullptr;
      s.out_bytes = 0;
      if (hipHostMalloc(&s.out, bytes, hipHostMallocDefault) != hipSuccess)
        return efail(err, SR_ERR_OUT_OF_MEMORY, "server: pinned output staging");
      s.out_bytes = bytes;
    }
    int rc = frame_submit_slot(ctx, *s.fs, &gf.fd, nullptr, err, nullptr, false);
    if (rc != SR_OK) return rc;
    const FrameSlot& fs = *s.fs;
    const float4* comp = fs.buf[B_COMPOSITE].as<float4>();
    const uint32_t* comp8 = fs.buf[B_COMPOSITE8].as<uint32_t>();
    const float* dep = fs.buf[B_DEPTH].as<float>();
    uint64_t img_off = 0;
    for (size_t i = 0; i < g.size(); ++i) {
      const sr_request& q = g[i].req;
      const uint64_t px = (uint64_t)q.camera.screen_width * q.camera.screen_height;
      const bool r8 = (q.flags & SR_REQUEST_RGBA8) != 0u;
      const void* csrc = r8 ? static_cast<const void*>(comp8 + img_off) : static_cast<const void*>(comp + img_off);
      const uint64_t cb = px * (r8 ? 4u : 16u);
      char* stg = static_cast<char*>(s.out) + s.out_off[i];
      const bool pinned = (q.flags & SR_REQUEST_PINNED_OUTPUTS) != 0u;
      void* cdst = pinned ? static_cast<void*>(q.out_rgba) : static_cast<void*>(stg);
      void* ddst = pinned ? static_cast<void*>(q.out_depth) : static_cast<void*>(stg + cb);
      if (cdst) HIP_OKS(hipMemcpyAsync(cdst, csrc, cb, hipMemcpyDeviceToHost, fs.stream), "server: copy image");
      if (ddst) HIP_OKS(hipMemcpyAsync(ddst, dep + img_off, px * 4u, hipMemcpyDeviceToHost, fs.stream), "server: copy depth");
      img_off += px;
    }
    HIP_OKS(hipEventRecord(s.out_ev, fs.stream), "server: event");
    s.reqs = std::move(g);
    s.busy = true;
    s.launch_ms = now_ms();
    return SR_OK;
  }

  // Admits queued requests into free slots; returns true when it launched.
  bool admit() {
    bool launched = false;
    for (auto& s : slots) {
      if (s.busy) continue;
      std::vector<PendingRequest> g;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (queue.empty()) return launched;
        g.push_back(queue.front());
        queue.pop_front();
        // views per frame: at most max_views_per_frame, and at most what
        // keeps (views + lights) x the asset's index positions within
        // kFrameSlots, so frames of large assets stay a modest share of HBM
        uint32_t vmax = cfg.max_views_per_frame;
        {
          std::lock_guard<std::mutex> alk(ctx->mu);
          const uint32_t id = g[0].req.instance.index_id;
          if (id < ctx->assets.size() && ctx->assets[id].live) {
            const uint64_t idx = std::max<uint64_t>(ctx->assets[id].info.index_count, 1);
            const uint64_t frusta = kFrameSlots / idx;
            vmax = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(vmax, frusta > g[0].req.light_count
                                                                              ? frusta - g[0].req.light_count
                                                                              : 1));
          }
        }
        for (auto it = queue.begin(); it != queue.end() && g.size() < vmax;) {
          if (same_group(g[0].req, it->req)) {
            g.push_back(*it);
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      }
      for (;;) {
        GroupFrame gf;
        int rc = build_group_frame(g, gf);
        std::string err;
        uint64_t need = 0, own = 0;
        if (rc == SR_OK) rc = frame_submit_slot(ctx, *s.fs, &gf.fd, nullptr, err, &need, true, nullptr, &own);
        if (rc != SR_OK) {
          fail_requests(g, rc);
          break;
        }
        // a multi-view frame stays within one slot's share of the budget, so
        // every slot can hold a frame at once and none gives its buffers
        // back: a larger group keeps the views its share holds at its
        // measured bytes per view (the frame's fixed part makes that an
        // underestimate), the tail goes back to the queue
        if (budget && g.size() > 1) {
          uint64_t assets;
          {
            std::lock_guard<std::mutex> alk(ctx->mu);
            assets = ctx->slabs.capacity();
          }
          const uint64_t share = budget > assets ? (budget - assets) / slots.size() : 0u;
          if (own > share) {
            std::lock_guard<std::mutex> lk(mu);
            const size_t fit = (size_t)((double)g.size() * (double)share / (double)own);
            const size_t keep = std::min(std::max<size_t>(fit, 1), g.size() - 1);
            for (size_t k = g.size(); k-- > keep;) queue.push_front(g[k]);
            g.resize(keep);
            ++stats.groups_split;
            continue;
          }
        }
        uint64_t outb = 0;
        for (auto& p : g) outb += staged_bytes(p.req);
        auto total = [&]() { return held_bytes() - s.fs->held_bytes() - s.out_bytes + need + std::max(s.out_bytes, outb); };
        if (budget && total() > budget) {
          for (auto& o : slots)  // idle slots give their buffers back
            if (!o.busy && &o != &s) o.fs->release_buffers();
          if (total() > budget) {
            bool others_busy = false;
            for (auto& o : slots) others_busy = others_busy || (o.busy && &o != &s);
            if (others_busy) {  // wait for memory to free up; requests go back to the front
              std::lock_guard<std::mutex> lk(mu);
              for (auto it = g.rbegin(); it != g.rend(); ++it) queue.push_front(*it);
              ++stats.budget_waits;
              return launched;
            }
            s.fs->release_buffers();
            rc = frame_submit_slot(ctx, *s.fs, &gf.fd, nullptr, err, &need, true);
            if (rc == SR_OK && total() > budget) {
              if (g.size() > 1) {  // split: the tail goes back to the queue
                std::lock_guard<std::mutex> lk(mu);
                for (size_t k = g.size(); k-- > 1;) queue.push_front(g[k]);
                g.resize(1);
                continue;
              }
              std::vector<PendingRequest> one(1, g[0]);
              fail_requests(one, SR_ERR_CAPACITY);
              break;
            }
            if (rc != SR_OK) {
              fail_requests(g, rc);
              break;
            }
          }
        }
        const uint64_t held = total();
        rc = launch(s, g, gf, err);
        if (rc != SR_OK) {
          fail_requests(g, rc);
          break;
        }
        launched = true;
        s.retries = 0;
        std::lock_guard<std::mutex> lk(mu);
        ++stats.frames;
        stats.peak_held_bytes = std::max(stats.peak_held_bytes, held);
        stats.max_frame_views = std::max(stats.max_frame_views, (uint32_t)s.reqs.size());
        uint32_t busy = 0;
        for (auto& o : slots) busy += o.busy ? 1u : 0u;
        stats.max_concurrent_frames = std::max(stats.max_concurrent_frames, busy);
        break;
      }
    }
    return launched;
  }

  void run() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
      bool busy = reap();
      bool launched = admit();
      busy = busy || launched;
      std::unique_lock<std::mutex> lk(mu);
      if (stopping && queue.empty() && !busy) {
        bool any = false;
        for (auto& s : slots) any = any || s.busy;
        if (!any) break;
      }
      if (busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));  // poll the slot events
      } else if (queue.empty() && !stopping) {
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
};

extern "C" {

int sr_server_start(sr_context* ctx, const sr_server_config* config, sr_server** out) {
  if (!ctx || !out) return SR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  sr_server* srv = new sr_server();
  srv->ctx = ctx;
  if (config) srv->cfg = *config;
  if (srv->cfg.slots == 0) srv->cfg.slots = 2;
  if (srv->cfg.max_views_per_frame == 0) srv->cfg.max_views_per_frame = 16;
  srv->budget = srv->cfg.hbm_budget_bytes ? srv->cfg.hbm_budget_bytes : ctx->hbm_budget;
  srv->slots.resize(srv->cfg.slots);
  for (auto& s : srv->slots) {
    std::string err;
    s.fs = new_slot(&err);
    if (!s.fs || hipEventCreateWithFlags(&s.out_ev, hipEventDisableTiming) != hipSuccess) {
      for (auto& o : srv->slots) {
        if (o.fs) free_slot(o.fs);
        if (o.out_ev) (void)hipEventDestroy(o.out_ev);
      }
      delete srv;
      return fail(ctx, SR_ERR_DEVICE, "sr_server_start: " + err);
    }
  }
  srv->t0 = Clock::now();
  srv->worker = std::thread([srv] { srv->run(); });
  *out = srv;
  return SR_OK;
}

int sr_server_submit(sr_server* srv, const sr_request* req, uint64_t* ticket) {
  if (!srv || !req) return SR_ERR_INVALID_ARGUMENT;
  if (req->light_count > SR_SERVER_MAX_LIGHTS || req->camera.screen_width == 0 || req->camera.screen_height == 0)
    return fail(srv->ctx, SR_ERR_INVALID_ARGUMENT, "sr_server_submit: bad request");
  if (req->settings.flags != 0u)  // served frames build their own lights; outputs follow the request's flags
    return fail(srv->ctx, SR_ERR_UNSUPPORTED, "sr_server_submit: frame flags are not served");
  if (req->flags & ~(SR_REQUEST_RGBA8 | SR_REQUEST_PINNED_OUTPUTS))
    return fail(srv->ctx, SR_ERR_INVALID_ARGUMENT, "sr_server_submit: unknown request flags");
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stopping) return fail(srv->ctx, SR_ERR_NOT_READY, "sr_server_submit: server is stopping");
  PendingRequest p{*req, srv->next_ticket++, srv->now_ms()};
  srv->queue.push_back(p);
  ++srv->stats.submitted;
  if (ticket) *ticket = p.ticket;
  srv->cv_work.notify_one();
  return SR_OK;
}

int sr_server_poll(sr_server* srv, sr_request_result* results, uint32_t capacity, uint32_t* count, int32_t wait_ms) {
  if (!srv || !count) return SR_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(srv->mu);
  if (srv->results.empty() && wait_ms > 0)
    srv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !srv->results.empty(); });
  uint32_t n = 0;
  while (n < capacity && !srv->results.empty()) {
    if (results) results[n] = srv->results.front();
    srv->results.pop_front();
    ++n;
  }
  *count = n;
  return SR_OK;
}

int sr_server_stats_get(sr_server* srv, sr_server_stats* out) {
  if (!srv || !out) return SR_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->stats;
  return SR_OK;
}

int sr_server_stop(sr_server* srv) {
  if (!srv) return SR_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stopping = true;
    srv->cv_work.notify_all();
  }
  if (srv->worker.joinable()) srv->worker.join();
  (void)hipSetDevice(srv->ctx->device);
  for (auto& s : srv->slots) {
    free_slot(s.fs);
    if (s.out) (void)hipHostFree(s.out);
    if (s.out_ev) (void)hipEventDestroy(s.out_ev);
  }
  delete srv;
  return SR_OK;
}

}  // extern "C"

// ==================================================================== autotune
// Times candidate configurations of one frame on the context's main slot,
// with the launch sequence in a HIP graph, against the frame as given
// without graphs (the quality reference): fine-binning backend x camera
// froxel size x light froxel size, then for the fastest accepted of those
// the other depth-warp slice counts (8, 16, 32 per coarse tile), then for the
// fastest accepted so far the launch shape (marking waves per workgroup x
// raster waves per CU).  A candidate is accepted when its composite is
// within min_psnr_db of the reference (tile size and slice count change the
// order in which overlapping fragments blend; the launch shape changes no
// result); the fastest accepted one is `best`.
namespace {

struct TuneFrame {
  std::vector<sr_frustum_desc> frusta;
  sr_frame_desc fd;
};

int tune_frame(const sr_frame_desc* src, uint32_t backend, uint32_t cam_fx, uint32_t light_fx, uint32_t slices,
               uint32_t mark_waves, uint32_t raster_waves, TuneFrame& out) {
  const uint32_t F = src->frustum_count;
  std::vector<sr_froxel_config> cfg(F);
  std::vector<uint32_t> is_light(F), cascade(F);
  for (uint32_t i = 0; i < F; ++i) {
    const sr_frustum_desc& d = src->frusta[i];
    const uint32_t fx = d.kind == 1u ? light_fx : cam_fx;
    cfg[i] = sr_froxel_config{d.screen_width, d.screen_height, fx ? fx : d.froxel_size_x, fx ? fx : d.froxel_size_y,
                              d.depth_slices};
    is_light[i] = d.kind == 1u ? 1u : 0u;
    cascade[i] = d.cascade_index;
  }
  out.frusta.resize(F);
  int rc = sr_build_frustum_table(cfg.data(), is_light.data(), cascade.data(), F, out.frusta.data());
  if (rc != SR_OK) return rc;
  out.fd = *src;
  out.fd.frusta = out.frusta.data();
  out.fd.settings.fine_binning_backend = backend;
  out.fd.settings.coarse_depth_slices = slices;
  out.fd.mark_waves = mark_waves;
  out.fd.raster_waves = raster_waves;
  return SR_OK;
}

}  // namespace

extern "C" int sr_autotune(sr_context* ctx, const sr_frame_desc* fd, uint32_t frames_per_trial, float min_psnr_db,
                           sr_tune_result* out) {
  if (!ctx || !fd || !out || fd->frustum_count == 0) return SR_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  FrameSlot& sl = ctx->main();
  const uint32_t saved_flags = ctx->flags;
  const bool saved_timer = sl.timer.enabled;
  sl.timer.enabled = false;
  std::memset(out, 0, sizeof(*out));
  uint32_t cam0 = 0, light0 = 0;
  for (uint32_t i = 0; i < fd->frustum_count; ++i) {
    if (fd->frusta[i].kind == 0u && !cam0) cam0 = fd->frusta[i].froxel_size_x;
    if (fd->frusta[i].kind == 1u && !light0) light0 = fd->frusta[i].froxel_size_x;
  }
  struct Cand { uint32_t backend, cam, light, slices, graphs, mark_waves, raster_waves; };
  std::vector<Cand> cands;
  const uint32_t slices0 = fd->settings.coarse_depth_slices ? fd->settings.coarse_depth_slices : kSlices;
  cands.push_back({fd->settings.fine_binning_backend, cam0, light0, slices0, 0u, fd->mark_waves,
                   fd->raster_waves});  // the reference
  const bool keep_backend = (fd->settings.flags & SR_FRAME_TUNE_KEEP_BACKEND) != 0u;
  for (uint32_t b : {0u, 1u}) {
    if (keep_backend && b != fd->settings.fine_binning_backend) continue;
    for (uint32_t cf : {8u, 16u})
      for (uint32_t lf : {8u, 16u}) cands.push_back({b, cam0 ? cf : 0u, light0 ? lf : 0u, slices0, 1u, 0u, 0u});
  }
  size_t stage_end = cands.size();  // the next stage's candidates follow the best of those before
  int stage = 1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<float4> ref_img, img;
  int rc = SR_OK;
  std::string err;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = SR_ERR_DEVICE;
  const uint32_t frames = frames_per_trial ? frames_per_trial : 3;
  float best_ms = 0.0f;
  for (size_t c = 0; rc == SR_OK && out->trial_count < SR_TUNE_MAX_TRIALS; ++c) {
    if (c == stage_end && stage < 3 && best_ms > 0.0f) {
      const sr_tune_trial b = out->best;
      if (stage == 1) {  // the other depth-warp slice counts of the best configuration so far
        for (uint32_t sc : {8u, 16u, 32u})
          if (sc != b.coarse_depth_slices)
            cands.push_back({b.fine_binning_backend, b.camera_froxel, b.light_froxel, sc, 1u, 0u, 0u});
      } else {  // launch shapes of the best configuration so far
        for (uint32_t mw : {1u, 2u, 4u})
          for (uint32_t rw : {8u, 16u, 24u})
            if (!(mw == 4u && rw == 16u))  // the defaults, timed in the earlier stages
              cands.push_back({b.fine_binning_backend, b.camera_froxel, b.light_froxel, b.coarse_depth_slices, 1u, mw,
                               rw});
      }
      ++stage;
      stage_end = cands.size();
    }
    if (c >= cands.size()) break;
    const Cand k = cands[c];
    TuneFrame tf;
    rc = tune_frame(fd, k.backend, k.cam, k.light, k.slices, k.mark_waves, k.raster_waves, tf);
    if (rc != SR_OK) break;
    ctx->flags = k.graphs ? (saved_flags | SR_CONFIG_HIP_GRAPHS) : (saved_flags & ~SR_CONFIG_HIP_GRAPHS);
    sr_frame_counters fc;
    // warm: grows the buffers and captures the graph
    for (int w = 0; w < 2 && rc == SR_OK; ++w) {
      rc = frame_submit_slot(ctx, sl, &tf.fd, nullptr, err, nullptr, false);
      if (rc == SR_OK) rc = frame_wait_slot(sl, &fc, err);
    }
    if (rc != SR_OK) break;
    (void)hipEventRecord(e0, sl.stream);
    for (uint32_t f = 0; f < frames && rc == SR_OK; ++f) rc = frame_submit_slot(ctx, sl, &tf.fd, nullptr, err, nullptr, false);
    (void)hipEventRecord(e1, sl.stream);
    if (rc == SR_OK) rc = frame_wait_slot(sl, &fc, err);
    if (rc != SR_OK) break;
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    img.resize(sl.shape.img_px);
    if (!img.empty() &&
        hipMemcpy(img.data(), sl.buf[B_COMPOSITE].p, img.size() * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = SR_ERR_DEVICE;
      break;
    }
    double mse = 0.0;
    if (c == 0) {
      ref_img = img;
    } else if (ref_img.size() == img.size()) {
      for (size_t i = 0; i < img.size(); ++i) {
        const float* a = &img[i].x;
        const float* b = &ref_img[i].x;
        for (int q = 0; q < 4; ++q) mse += (double)(a[q] - b[q]) * (a[q] - b[q]);
      }
      mse /= 4.0 * (double)std::max<size_t>(img.size(), 1);
    }
    sr_tune_trial& t = out->trials[out->trial_count++];
    t.fine_binning_backend = k.backend;
    t.camera_froxel = k.cam;
    t.light_froxel = k.light;
    t.hip_graphs = k.graphs;
    t.mark_waves = k.mark_waves;
    t.raster_waves = k.raster_waves;
    t.coarse_depth_slices = k.slices;
    t.ms_per_frame = ms / (float)frames;
    t.psnr_db = mse > 0.0 ? (float)(10.0 * std::log10(1.0 / mse)) : 999.0f;
    t.accepted = t.psnr_db >= min_psnr_db ? 1u : 0u;
    if (t.accepted && (best_ms == 0.0f || t.ms_per_frame < best_ms)) {
      best_ms = t.ms_per_frame;
      out->best = t;
    }
  }
  ctx->flags = saved_flags;
  sl.timer.enabled = saved_timer;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != SR_OK) return fail(ctx, rc, err.empty() ? "sr_autotune: device error" : err);
  return SR_OK;
}
