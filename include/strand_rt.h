/*
 * strand_rt.h -- C ABI of the MI355X strand rendering hot path.
 *
 * Byte layouts mirror the reference's GPU ABI so that a host which already
 * packs strand slabs for the Bevy plugin can hand the same bytes to this
 * library.  Every struct carries a static size check; offsets of the
 * multi-field records are checked in tests/test_abi.py.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   - StrandRasterizerPlugin + render-graph nodes   src/plugin.rs:162-261,
 *     src/nodes/{prepass,shadows,shading,raster,composite}.rs
 *   - run_prepass / run_shadow_pass / run_shading_pass / run_raster_pass
 *                                                   src/pipelines/prepass.rs:818-1397,
 *     src/pipelines/shadows.rs:621-747, src/pipelines/shading.rs:470-567,
 *     src/pipelines/raster.rs:540-641
 *   - GPU ABI structs                               src/shader_types.rs:9-137,
 *     src/pipelines/task_contract.rs:54-160, src/plugin.rs:1545-1593,
 *     src/components.rs:61-123, src/shaders/types.wgsl:23-27
 *
 * Conventions: C linkage, no exceptions cross the boundary, every entry point
 * returns an int status (0 = ok, < 0 = sr_status), handles are out-params.
 * A context is externally synchronised (one host thread at a time).
 */
#ifndef STRAND_RT_H
#define STRAND_RT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define SR_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define SR_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* ------------------------------------------------------------------------ */
/* Constants (src/plugin.rs:57-70, src/shaders/strand_prepass.wgsl:65-69)    */
/* ------------------------------------------------------------------------ */
#define SR_COARSE_FINE_TILE_EXTENT 4u
/* COARSE_DEPTH_SLICES is one shader constant in the reference: the slices of
 * a coarse tile's depth-warp LUT and page table, and the fine depth slices of
 * a count page (4 x 4 x 16 = SR_COARSE_COUNT_PAGE_SIZE cells).  Here a frame
 * may give its coarse tiles 8, 16 or 32 LUT slices
 * (sr_frame_settings.coarse_depth_slices, an autotuner dimension); a page
 * keeps its 16 fine slices, so it spans 1/8, 1/16 or 1/32 of the tile's
 * depth intervals.  At 16 every layout is the reference's. */
#define SR_COARSE_DEPTH_SLICES 16u
#define SR_MAX_COARSE_DEPTH_SLICES 32u
#define SR_COARSE_COUNT_PAGE_SIZE 256u
#define SR_DEPTH_QUANT_MAX 16777215u
#define SR_INVALID_PTR 0xFFFFFFFFu
#define SR_MAX_TEXTURE_EXTENT 8192u
#define SR_MAX_DOM_LAYERS 16u
#define SR_BINNING_LEVEL_BITS 2u
#define SR_BINNING_SHADOW_BIT (1u << SR_BINNING_LEVEL_BITS)
#define SR_BINNING_FRUSTUM_SHIFT (SR_BINNING_LEVEL_BITS + 1u)

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef enum sr_status {
  SR_OK = 0,
  SR_ERR_INVALID_ARGUMENT = -1,
  SR_ERR_OUT_OF_MEMORY = -2,
  SR_ERR_DEVICE = -3,
  SR_ERR_NOT_READY = -4,
  SR_ERR_CAPACITY = -5,
  SR_ERR_INVALID_ASSET = -6,
  SR_ERR_UNSUPPORTED = -7
} sr_status;

/* ------------------------------------------------------------------------ */
/* Asset slabs (src/shader_types.rs, src/components.rs)                      */
/* ------------------------------------------------------------------------ */

/* PushConstants, src/shader_types.rs:9-22 (44 bytes). */
typedef struct sr_push_constants {
  uint32_t workgroup_offset;
  uint32_t num_elements;
  uint32_t frustum_count;
  uint32_t scan_load_base;
  uint32_t scan_save_base;
  uint32_t stochastic_cull_enabled;
  float target_strands_per_pixel;
  float min_keep_probability;
  float shadow_keep_probability;
  uint32_t telemetry_enabled;
  uint32_t fine_binning_backend;
} sr_push_constants;
SR_STATIC_ASSERT(sizeof(sr_push_constants) == 44, "PushConstants is 44 bytes");

/* StrandMeta, src/shader_types.rs:26-31: count = vertices of the strand,
 * offset = first position in the index slab, material_idx into the
 * material slab. */
typedef struct sr_strand_meta {
  uint32_t count;
  uint32_t offset;
  uint32_t material_idx;
  uint32_t pad;
} sr_strand_meta;
SR_STATIC_ASSERT(sizeof(sr_strand_meta) == 16, "StrandMeta is 16 bytes");

/* Aabb, src/shader_types.rs:46-51. */
typedef struct sr_aabb {
  float min[3];
  float pad1;
  float max[3];
  float pad2;
} sr_aabb;
SR_STATIC_ASSERT(sizeof(sr_aabb) == 32, "Aabb is 32 bytes");

/* StrandGeo, src/shader_types.rs:66-72. */
typedef struct sr_strand_geo {
  uint32_t strand_count;
  uint32_t max_segments_in_strand;
  uint32_t offset;
  uint32_t pad;
  sr_aabb aabb;
  uint32_t pad3[4];
} sr_strand_geo;
SR_STATIC_ASSERT(sizeof(sr_strand_geo) == 64, "StrandGeo is 64 bytes");

/* StrandInstance, src/shader_types.rs:77-88.  Matrices are column-major
 * (glam Mat4).  The five ids index the slab tables of the context; an
 * uploaded asset reports its ids in sr_asset_info. */
typedef struct sr_strand_instance {
  uint32_t vertex_id;
  uint32_t index_id;
  uint32_t meta_id;
  uint32_t geo_id;
  uint32_t material_id;
  uint32_t pad0;
  uint32_t pad1;
  uint32_t pad2;
  float world_from_local[16];
  float local_from_world[16];
} sr_strand_instance;
SR_STATIC_ASSERT(sizeof(sr_strand_instance) == 160, "StrandInstance is 160 bytes");

/* StrandMaterial, src/components.rs:61-72. */
typedef struct sr_strand_material {
  float absorption_color[4];
  float specular_color[4];
  float ambient_factor;
  float ao_factor;
  float eta;
  float beta;
  float alpha;
  float shift;
  float min_radius_pixels;
  float max_radius_pixels;
} sr_strand_material;
SR_STATIC_ASSERT(sizeof(sr_strand_material) == 64, "StrandMaterial is 64 bytes");

/* DevicePtr, src/shaders/types.wgsl:23-27: slab index, byte offset, byte
 * size; valid iff slab != 0xFFFFFFFF and size != 0. */
typedef struct sr_device_ptr {
  uint32_t slab;
  uint32_t offset;
  uint32_t size;
} sr_device_ptr;
SR_STATIC_ASSERT(sizeof(sr_device_ptr) == 12, "DevicePtr is 12 bytes");

/* FroxelConfig, src/components.rs:93-123. */
typedef struct sr_froxel_config {
  uint32_t screen_width;
  uint32_t screen_height;
  uint32_t froxel_size_x;
  uint32_t froxel_size_y;
  uint32_t depth_slices;
} sr_froxel_config;
SR_STATIC_ASSERT(sizeof(sr_froxel_config) == 20, "FroxelConfig is 20 bytes");

/* GpuFrustumDesc, src/plugin.rs:1545-1562 (kind 0 = camera, 1 = light). */
typedef struct sr_frustum_desc {
  uint32_t screen_width;
  uint32_t screen_height;
  uint32_t froxel_size_x;
  uint32_t froxel_size_y;
  uint32_t depth_slices;
  uint32_t bucket_base;
  uint32_t bucket_count;
  uint32_t kind;
  uint32_t coarse_depth_tile_base;
  uint32_t coarse_depth_tile_count;
  uint32_t coarse_tiles_x;
  uint32_t coarse_tiles_y;
  uint32_t cascade_index;
  uint32_t fine_depth_tile_base;
  uint32_t pad1;
  uint32_t pad2;
} sr_frustum_desc;
SR_STATIC_ASSERT(sizeof(sr_frustum_desc) == 64, "GpuFrustumDesc is 64 bytes");

/* ------------------------------------------------------------------------ */
/* Queue / task contract (src/pipelines/task_contract.rs:54-135)             */
/* ------------------------------------------------------------------------ */
typedef struct sr_queue_header { uint32_t head; uint32_t tail; } sr_queue_header;
SR_STATIC_ASSERT(sizeof(sr_queue_header) == 8, "QueueHeader is 8 bytes");

typedef struct sr_fine_prepass_task { uint32_t inst_id; uint32_t strand_local; } sr_fine_prepass_task;
SR_STATIC_ASSERT(sizeof(sr_fine_prepass_task) == 8, "FinePrepassTask is 8 bytes");

/* BinningTask: id_info = instance, chunk_id = strand_local, seg_idx = index
 * slab position of the segment's first vertex, packed_field = level bits 0-1,
 * shadow bit 2, frustum bits 3-31. */
typedef struct sr_binning_task {
  uint32_t id_info;
  uint32_t chunk_id;
  uint32_t seg_idx;
  uint32_t packed_field;
} sr_binning_task;
SR_STATIC_ASSERT(sizeof(sr_binning_task) == 16, "BinningTask is 16 bytes");

/* ProjectedSegment: pack2x16unorm(xy / screen) per endpoint, pack2x16unorm(z0, z1). */
typedef struct sr_projected_segment {
  uint32_t p0_xy;
  uint32_t p1_xy;
  uint32_t depths;
} sr_projected_segment;
SR_STATIC_ASSERT(sizeof(sr_projected_segment) == 12, "ProjectedSegment is 12 bytes");

typedef struct sr_raster_work_item {
  uint32_t seg_ref_base;
  uint32_t seg_ref_count;
  uint32_t frustum_id;
  uint32_t screen_tile_id;
  uint32_t depth_key;
  uint32_t load_score;
  uint32_t page_idx;
  uint32_t cell_idx;
} sr_raster_work_item;
SR_STATIC_ASSERT(sizeof(sr_raster_work_item) == 32, "RasterWorkItem is 32 bytes");

typedef struct sr_raster_tile_run {
  uint32_t work_base;
  uint32_t work_count;
  uint32_t frustum_id;
  uint32_t screen_tile_id;
  uint32_t load_score;
  uint32_t flags;
  uint32_t reserved_a;
  uint32_t reserved_b;
} sr_raster_tile_run;
SR_STATIC_ASSERT(sizeof(sr_raster_tile_run) == 32, "RasterTileRun is 32 bytes");

typedef struct sr_fine_seg_ref {
  uint32_t inst_id;
  uint32_t seg_id;
  uint32_t material_idx;
} sr_fine_seg_ref;
SR_STATIC_ASSERT(sizeof(sr_fine_seg_ref) == 12, "FineSegRef is 12 bytes");

typedef struct sr_fine_page_meta {
  uint32_t seg_ref_base;
  uint32_t seg_ref_count;
  uint32_t work_base;
  uint32_t work_count;
  uint32_t coarse_tile_id;
  uint32_t coarse_z;
  uint32_t first_depth_key;
  uint32_t flags; /* 1 = reference range overflowed the capacity */
} sr_fine_page_meta;
SR_STATIC_ASSERT(sizeof(sr_fine_page_meta) == 32, "FinePageMeta is 32 bytes");

/* GpuCoarseAssetRange / GpuCoarseIntervalRef / GpuCoarseDepthLutEntry,
 * src/plugin.rs:1566-1593. */
typedef struct sr_coarse_asset_range {
  uint32_t geo_id;
  uint32_t frustum_id;
  uint32_t min_x;
  uint32_t max_x;
  uint32_t min_y;
  uint32_t max_y;
  uint32_t z_min_q;
  uint32_t z_max_q;
} sr_coarse_asset_range;
SR_STATIC_ASSERT(sizeof(sr_coarse_asset_range) == 32, "CoarseAssetRange is 32 bytes");

typedef struct sr_coarse_interval_ref {
  uint32_t next;
  uint32_t range_id;
  uint32_t z_min_q;
  uint32_t z_max_q;
} sr_coarse_interval_ref;
SR_STATIC_ASSERT(sizeof(sr_coarse_interval_ref) == 16, "CoarseIntervalRef is 16 bytes");

typedef struct sr_coarse_depth_lut_entry {
  uint32_t z_min_q;
  uint32_t z_max_q;
  uint32_t virtual_start_q;
  uint32_t virtual_count_q; /* 0 = empty slice */
} sr_coarse_depth_lut_entry;
SR_STATIC_ASSERT(sizeof(sr_coarse_depth_lut_entry) == 16, "CoarseDepthLutEntry is 16 bytes");

/* BroadInstanceMeta, src/shaders/strand_prepass.wgsl:193-198. */
typedef struct sr_broad_instance_meta {
  uint32_t strand_count;
  uint32_t visible;
  float camera_keep_probability;
  float shadow_keep_probability;
} sr_broad_instance_meta;
SR_STATIC_ASSERT(sizeof(sr_broad_instance_meta) == 16, "BroadInstanceMeta is 16 bytes");

/* ------------------------------------------------------------------------ */
/* Frame description (replaces Bevy's View / Lights uniforms)               */
/* ------------------------------------------------------------------------ */

/* Per-frustum view.  The reference projects every camera frustum with the
 * current view's unjittered_clip_from_world and every light frustum with its
 * cascade's clip_from_world (src/shaders/strand_prepass.wgsl:438-447); here
 * each frustum carries its own matrix so many views batch in one frame.
 * world_position is the camera position used by shading
 * (src/shaders/strand_shading.wgsl:653). */
typedef struct sr_view {
  float clip_from_world[16]; /* column-major */
  float world_position[4];
} sr_view;
SR_STATIC_ASSERT(sizeof(sr_view) == 80, "sr_view is 80 bytes");

/* Directional light, in light-layer order (the i-th kind==1 frustum uses
 * light i).  color follows Bevy's uniform (linear colour x illuminance);
 * shading divides it by 255 (src/shaders/strand_shading.wgsl:687). */
typedef struct sr_directional_light {
  float direction_to_light[4];
  float color[4];
} sr_directional_light;
SR_STATIC_ASSERT(sizeof(sr_directional_light) == 32, "sr_directional_light is 32 bytes");

/* StochasticCullSettings (src/resources.rs:70-104), PrepassTelemetrySettings,
 * FineBinningBackend, plus the shadow and composite parameters that the
 * reference takes from compile-time defines and Bevy resources. */
typedef struct sr_frame_settings {
  uint32_t stochastic_cull_enabled;
  float target_strands_per_pixel;
  float min_keep_probability;
  float shadow_keep_probability;
  uint32_t telemetry_enabled;
  uint32_t fine_binning_backend; /* 0 = SegmentScatter, 1 = PageCsr */
  uint32_t dom_layers;           /* NUM_DOM_SLICES, 1..SR_MAX_DOM_LAYERS */
  uint32_t shadow_history_valid; /* 0 on the first frame (history cleared) */
  float ambient_color[4];        /* lights.ambient_color */
  float scene_clear_color[4];    /* headless scene colour under the composite */
  float scene_clear_depth;       /* headless scene depth (reverse-Z, 0 = far) */
  uint32_t stamp_target_width;   /* light depth-map size for the stampback */
  uint32_t stamp_target_height;
  uint32_t flags;                /* SR_FRAME_* bits */
  uint32_t dom_resident_pages;   /* SR_FRAME_VIRTUAL_DOMS: resident DOM pages per light, 0 = every requested page */
  uint32_t coarse_depth_slices;  /* depth-warp slices per coarse tile: 8, 16 or 32; 0 = 16 (SR_COARSE_DEPTH_SLICES) */
} sr_frame_settings;
SR_STATIC_ASSERT(sizeof(sr_frame_settings) == 88, "sr_frame_settings is 88 bytes");

/* sr_frame_settings.flags: draw the frame's cameras with the deep shadow maps
 * that the context's last light-building frame left (no counterpart in the
 * reference, which renders every light in every frame).  The frame lists the
 * same light frusta, views and lights, the same instances and the same
 * culling settings as that frame (checked: SR_ERR_INVALID_ARGUMENT
 * otherwise), which must have been waited for and must not have reported
 * capacity_exceeded.  Its light frusta are then neither binned nor
 * rasterised, the DOMs and stampback targets stay as they are, and every
 * camera's shading reads them: a camera drawn this way equals the same
 * camera drawn in a frame that builds the lights.  Lets a batch of views too
 * large for one frame share one DOM build. */
#define SR_FRAME_REUSE_LIGHT_DOMS 1u
/* sr_frame_settings.flags: also store the strand image and the composite in
 * the reference's framebuffer format, rgba8unorm (src/pipelines/raster.rs:
 * 500-513): each channel clamped to [0, 1] and rounded to the nearest of the
 * 255 steps, ties to even (WGSL pack4x8unorm), red in the lowest byte
 * (SR_DUMP_IMAGE_RGBA8, SR_DUMP_COMPOSITE_RGBA8). */
#define SR_FRAME_RGBA8_OUTPUTS 2u
/* sr_frame_settings.flags: virtual deep shadow maps (the reference's VSMS
 * surfaces, src/plugin.rs:1776-1854).  Each light's DOM is cut into pages of
 * SR_DOM_PAGE_XY^2 texels (depth plus every opacity layer).  Every shadow
 * tile run requests the pages its pixels fall in and their neighbours within
 * SR_DOM_PREFETCH_PAGE_BORDER pages (request_shadow_dom_pages,
 * src/shaders/strand_prepass.wgsl:2520-2579); requested pages become
 * resident in page order, at most settings.dom_resident_pages per light (0:
 * all of them), in the same frame.  The shadow raster writes, and the
 * stampback and shading read, the DOM through the page table: a texel of a
 * page that is not resident is not written, reads depth 0 and lets all light
 * through (strand_rasterizer.wgsl:484-517, strand_shading.wgsl:198-276,
 * shadow_dom_stampback.wgsl:58-72).  DOM memory then follows the pages in
 * use instead of the light's full resolution; with every requested page
 * resident the frame equals the same frame with dense DOMs.
 * SR_DUMP_DOM_DEPTH / SR_DUMP_DOM_OPACITY still return the dense layout
 * (texels of pages that are not resident read 0); SR_DUMP_DOM_REQUESTS and
 * SR_DUMP_DOM_PAGE_TABLE return the requests and the page table. */
#define SR_FRAME_VIRTUAL_DOMS 4u
/* sr_frame_settings.flags: depth-sorted blending inside a work item.  The
 * reference blends a pixel's fragments of one cell in reference order
 * (strand_rasterizer.wgsl:998-1020, an order its atomics leave open; the
 * default here is (inst, seg)).  With this flag each pixel keeps the
 * fragments of the current work item in a list of SR_BLEND_SORT_FRAGMENTS,
 * nearest first (larger reverse-Z depth; equal depths in arrival order):
 * when one more arrives, the nearest of them is blended, and the rest are
 * blended nearest first when the work item ends.  A cell with at most that
 * many fragments on a pixel is blended in exact depth order. */
#define SR_FRAME_DEPTH_SORTED_BLEND 8u
#define SR_BLEND_SORT_FRAGMENTS 4u
/* sr_frame_settings.flags, read by sr_autotune only: every candidate keeps
 * the frame's fine_binning_backend (froxel sizes and launch shape are still
 * searched).  For frames too large to try the other backend on. */
#define SR_FRAME_TUNE_KEEP_BACKEND 16u
#define SR_DOM_PAGE_XY 256u              /* DOM_PAGE_XY, src/plugin.rs:68 */
#define SR_DOM_PREFETCH_PAGE_BORDER 1u   /* DOM_PREFETCH_PAGE_BORDER, src/plugin.rs:69 */

typedef struct sr_frame_desc {
  const sr_strand_instance* instances;
  uint32_t instance_count;
  uint32_t frustum_count;
  const sr_frustum_desc* frusta; /* [frustum_count] */
  const sr_view* views;          /* [frustum_count] */
  const sr_directional_light* lights;
  uint32_t light_count;
  /* Launch shape (no reference counterpart; results do not depend on it):
   * waves per workgroup of the projection / marking pass (1, 2 or 4; 0 = 4). */
  uint32_t mark_waves;
  sr_frame_settings settings;
  /* Opaque scene depth per camera (the reference's prepass depth texture,
   * src/pipelines/prepass.rs:780-816), host memory: for every camera frustum
   * in frustum order, width * height pixels x opaque_samples floats (the
   * samples of a pixel adjacent), reverse-Z with 0 = nothing drawn.  NULL: no
   * opaque geometry, nothing is occluded. */
  const float* opaque_depth;
  uint32_t opaque_samples; /* 1..4 (MSAA samples) */
  /* Launch shape: waves per CU of the persistent shadow and camera
   * rasterisers (1..32; 0 = 16): one-wave workgroups for tiles of up to 64
   * pixels, four-wave ones (raster_waves / 4 of them) above. */
  uint32_t raster_waves;
} sr_frame_desc;

/* Host-side view of one asset's slabs in the reference layout
 * (src/plugin.rs:1315-1393): vertices are float4 (x, y, z, 1). */
typedef struct sr_asset_slabs {
  const float* vertices;
  uint32_t vertex_count;
  uint32_t index_count;
  const uint32_t* indices;
  const sr_strand_meta* meta;
  uint32_t strand_count;
  uint32_t material_count;
  const sr_strand_geo* geo;
  const sr_strand_material* materials;
} sr_asset_slabs;

/* Slab ids assigned at upload (the five StrandInstance ids) and the
 * DevicePtr of each slab inside the context's paging slabs. */
typedef struct sr_asset_info {
  uint32_t vertex_id;
  uint32_t index_id;
  uint32_t meta_id;
  uint32_t geo_id;
  uint32_t material_id;
  uint32_t strand_count;
  uint32_t index_count;
  uint32_t max_segments_in_strand;
  sr_device_ptr vertices;
  sr_device_ptr indices;
  sr_device_ptr meta;
  sr_device_ptr geo;
  sr_device_ptr materials;
  uint32_t pad;
} sr_asset_info;

/* Frame counters (filled after a frame completes). */
typedef struct sr_frame_counters {
  uint32_t binning_tasks;       /* T */
  uint32_t camera_tasks;        /* shading queue length */
  uint32_t allocated_pages;     /* P */
  uint32_t fine_refs;           /* H = sum of cell counts */
  uint32_t active_tiles;        /* A */
  uint32_t work_items;          /* W */
  uint32_t camera_runs;
  uint32_t shadow_runs;
  uint32_t overflow_pages;      /* pages flagged because refs overflowed */
  uint32_t coarse_tiles;        /* total coarse tiles over all frusta */
  uint32_t task_capacity;
  uint32_t ref_capacity;
  uint32_t page_drops;          /* marked pages past the page capacity (dropped, as the reference) */
  uint32_t page_candidates;     /* fine-cell hits sorted by the page CSR backend (references, dropped pages included); 0 for the cell-atomic backend */
  uint32_t capacity_exceeded;   /* 1: a capacity overflowed; the context grew it, submit the frame again */
  uint32_t dom_pages;           /* SR_FRAME_VIRTUAL_DOMS: resident DOM pages over every light (0 otherwise) */
} sr_frame_counters;
SR_STATIC_ASSERT(sizeof(sr_frame_counters) == 64, "sr_frame_counters is 64 bytes");

/* What sr_frame_dump / the oracle can return, in the layouts below. */
typedef enum sr_dump_kind {
  SR_DUMP_COUNTERS = 0,      /* sr_frame_counters */
  SR_DUMP_BROAD_META = 1,    /* sr_broad_instance_meta[instances] */
  SR_DUMP_KEEP = 2,          /* float[instances * frusta] */
  SR_DUMP_RANGES = 3,        /* sr_coarse_asset_range[instances * frusta], geo_id = INVALID when absent */
  SR_DUMP_TASKS = 4,         /* sr_binning_task[T], order (frustum, instance, seg_idx) */
  SR_DUMP_PROJECTED = 5,     /* sr_projected_segment[T] */
  SR_DUMP_SHADING_QUEUE = 6, /* uint32[camera tasks], ascending task indices */
  SR_DUMP_LUT = 7,           /* sr_coarse_depth_lut_entry[coarse tiles * 16] */
  SR_DUMP_PAGE_TABLE = 8,    /* uint32[coarse tiles * 16]: 0 = none, else page + 1 */
  SR_DUMP_PAGE_META = 9,     /* sr_fine_page_meta[P] */
  SR_DUMP_CELL_COUNTS = 10,  /* uint32[P * 256] */
  SR_DUMP_CELL_OFFSETS = 11, /* uint32[P * 256], page-local exclusive */
  SR_DUMP_REFS = 12,         /* sr_fine_seg_ref[H], each cell sorted by (inst, seg) */
  SR_DUMP_ACTIVE_TILES = 13, /* uint32[A], ascending coarse_tile * 16 + local fine tile */
  SR_DUMP_WORK_ITEMS = 14,   /* sr_raster_work_item[W] */
  SR_DUMP_CAMERA_RUNS = 15,  /* sr_raster_tile_run[] */
  SR_DUMP_SHADOW_RUNS = 16,  /* sr_raster_tile_run[] */
  SR_DUMP_DOM_DEPTH = 17,    /* float per light frustum W*H, concatenated */
  SR_DUMP_DOM_OPACITY = 18,  /* float per light frustum layers*W*H (layer-major) */
  SR_DUMP_ATLAS = 19,        /* float4 per camera, instance, texel */
  SR_DUMP_HISTORY = 20,      /* float4 per camera, instance, texel */
  SR_DUMP_IMAGE = 21,        /* float4 per camera frustum W*H (straight alpha) */
  SR_DUMP_DEPTH = 22,        /* float per camera frustum W*H */
  SR_DUMP_COMPOSITE = 23,    /* float4 per camera frustum W*H */
  SR_DUMP_STAMP = 24,        /* float per light frustum stamp target */
  SR_DUMP_TELEMETRY = 25,    /* uint32 telemetry when settings.telemetry_enabled (finalize_telemetry layout):
                                [0] pages [1] tasks [2] page candidates [3] fine refs [4] pages with candidates
                                [5] max candidates per page, [8..39] log2 histogram of candidates per page,
                                [40 .. 40+page_cap) candidates per page, then 11 words per frustum: visible
                                instances, visible strands, retained strands, emitted segments, pages,
                                candidates, refs, raster runs, raster load, max page candidates, max raster load */
  SR_DUMP_OPAQUE_TILES = 26, /* uint32 (covered, quantised farthest depth) per fine tile when opaque depth was given */
  SR_DUMP_IMAGE_RGBA8 = 27,  /* uint8 x 4 per camera frustum W*H when SR_FRAME_RGBA8_OUTPUTS (the image, rgba8unorm) */
  SR_DUMP_COMPOSITE_RGBA8 = 28, /* uint8 x 4 per camera frustum W*H when SR_FRAME_RGBA8_OUTPUTS (the composite) */
  SR_DUMP_DOM_REQUESTS = 29, /* uint32 per DOM page of every light frustum, row-major ceil(W/256) x ceil(H/256)
                                pages per light, concatenated: 1 = requested (SR_FRAME_VIRTUAL_DOMS) */
  SR_DUMP_DOM_PAGE_TABLE = 30, /* uint32 per DOM page as SR_DUMP_DOM_REQUESTS: its physical page, or
                                  SR_INVALID_PTR when not resident (SR_FRAME_VIRTUAL_DOMS) */
  SR_DUMP_KIND_COUNT = 31     /* one past the largest kind */
} sr_dump_kind;

/* ------------------------------------------------------------------------ */
/* Entry points                                                              */
/* ------------------------------------------------------------------------ */
typedef struct sr_context sr_context;

typedef struct sr_device_config {
  int32_t device_index;     /* HIP device ordinal */
  uint32_t flags;           /* SR_CONFIG_* bits */
  uint64_t slab_bytes;      /* size of each asset paging slab; 0 = default */
  uint64_t hbm_budget_bytes;/* 0 = no budget */
} sr_device_config;

/* sr_device_config.flags: the page CSR backend records only 2 fine cells per
 * segment while marking (instead of up to 24), so most segments take the
 * re-trace path of the later passes, and queues only 2 fine walks per wave,
 * so most walks run in place, and 2 coarse tiles per wave and marking round,
 * so waves walk their segments in many rounds (exercises all three; results
 * are identical). */
#define SR_CONFIG_CSR_FEW_HITS 2u
/* sr_device_config.flags: capture each frame's launch sequence in a HIP graph
 * and replay it for later frames of the same shape (buffers, capacities,
 * settings and launch extents); per-frame tables are uploaded before the
 * replay.  Ignored while profiling is on. */
#define SR_CONFIG_HIP_GRAPHS 4u
/* sr_device_config.flags: a frame slot's first frame of a shape starts with
 * minimal reference, work-item and page capacities (4096, 1024 and 16), so it
 * overflows and reports capacity_exceeded, and the next frame runs with the
 * capacities it needed (exercises the growth path). */
#define SR_CONFIG_TIGHT_CAPACITY 8u

/* Library version string. */
const char* sr_version(void);

/* Create / destroy a context on one device (replaces StrandRasterizerPlugin
 * build/finish, src/plugin.rs:162-261). */
int sr_create(const sr_device_config* config, sr_context** out_ctx);
int sr_destroy(sr_context* ctx);

/* Upload one asset's slabs (replaces insert_packed_strand_geometry,
 * src/plugin.rs:1395-1440).  Validates like validate_cache
 * (src/strand_cache.rs:174-209): strands inside the index slab, indices
 * inside the vertex slab, materials inside the material slab. */
int sr_asset_upload(sr_context* ctx, const sr_asset_slabs* slabs, sr_asset_info* out_info);
int sr_asset_release(sr_context* ctx, uint32_t geo_id);

/* Run one frame: prepass (culling, projection, binning, scans, compaction,
 * raster-work emission), deep-shadow-map build and stampback, shading,
 * camera raster and composite.  Replaces the five render-graph nodes chained
 * at src/plugin.rs:240-260.  stream is a hipStream_t (NULL = the context's
 * stream).  All work is enqueued; nothing is read back. */
int sr_frame_submit(sr_context* ctx, const sr_frame_desc* frame, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the frame of sr_frame_submit, with every intermediate
 * buffer carved out of one file-scope device array of
 * SR_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends.  The call allocates nothing: the per-frame tables travel as
 * kernel arguments, no graph is captured, no staging memory is used.
 *   composite  device memory, float4 per camera pixel in frustum order (the
 *              layout of SR_DUMP_COMPOSITE)
 *   depth      device memory, float per camera pixel (SR_DUMP_DEPTH)
 *   counters   device memory for one sr_frame_counters, or NULL
 *   stream     hipStream_t; NULL = the context's no-allocation stream
 * A frame whose buffers do not fit the array fails with SR_ERR_CAPACITY (at
 * fixed capacities: a frame reporting capacity_exceeded is incomplete and is
 * not regrown); opaque depth is not supported and shadow history is off.
 * Every call on one GPU shares that GPU's array (the last call's contents
 * stay there; what must start at zero is cleared by each call): calls that
 * overlap on one GPU race, so serialise them on one stream or with events.
 * sr_frame_wait and sr_frame_dump do not see these frames. */
#define SR_NOALLOC_SCRATCH_BYTES (256ull << 20)
int sr_frame_submit_noalloc(sr_context* ctx, const sr_frame_desc* frame, float* composite, float* depth,
                            sr_frame_counters* counters, void* stream);

/* Wait for the last submitted frame and fetch its counters. */
int sr_frame_wait(sr_context* ctx, sr_frame_counters* out_counters);

/* Copy one intermediate or output of the last frame to host memory.
 * capacity is in bytes; *out_bytes receives the required size (call with
 * dst = NULL to query). */
int sr_frame_dump(sr_context* ctx, int32_t kind, void* dst, uint64_t capacity, uint64_t* out_bytes);

/* Per-stage GPU time of the last frame (ms), from HIP events; names in
 * stage order, separated by '\n'.  Enabled by sr_set_profiling. */
int sr_set_profiling(sr_context* ctx, int32_t enabled);
int sr_stage_times(sr_context* ctx, float* dst_ms, int32_t capacity, int32_t* out_count, char* names, int32_t names_capacity);

/* Last error message of the context (or the global one when ctx is NULL). */
const char* sr_last_error(sr_context* ctx);

/* ------------------------------------------------------------------------ */
/* Autotuner (no reference counterpart: the reference fixes its tile sizes and
 * binning backend per run, src/resources.rs).  Times one frame under the
 * candidate configurations, replayed from a HIP graph, against the frame as
 * given without graphs (the quality reference), on the context's main slot:
 * first fine-binning backend (segment scatter / page CSR; only the frame's own
 * with SR_FRAME_TUNE_KEEP_BACKEND) x camera froxel size {8, 16} x light
 * froxel size {8, 16}; then, for the fastest accepted of those, the other
 * depth-warp slice counts of {8, 16, 32}; then, for the fastest accepted so
 * far, the launch shape: marking workgroups of 1, 2 or 4 waves x 8, 16 or
 * 24 persistent raster waves per CU.  A candidate whose composite stays
 * within min_psnr_db of the reference is accepted (tile size and slice count
 * change the order in which overlapping fragments blend; the launch shape
 * changes no result); `best` is the fastest accepted one.  The caller
 * rebuilds its frustum table with best's froxel sizes
 * (sr_build_frustum_table), sets the backend, the slice count,
 * SR_CONFIG_HIP_GRAPHS and the frame's mark_waves / raster_waves. */
typedef struct sr_tune_trial {
  uint32_t fine_binning_backend;
  uint32_t camera_froxel;       /* 0 when the frame has no camera */
  uint32_t light_froxel;        /* 0 when the frame has no light */
  uint32_t hip_graphs;
  float ms_per_frame;
  float psnr_db;                /* composite against the reference configuration */
  uint32_t accepted;
  uint32_t mark_waves;          /* sr_frame_desc.mark_waves of the trial (0 = default) */
  uint32_t raster_waves;        /* sr_frame_desc.raster_waves of the trial (0 = default) */
  uint32_t coarse_depth_slices; /* sr_frame_settings.coarse_depth_slices of the trial (8, 16 or 32) */
} sr_tune_trial;

#define SR_TUNE_MAX_TRIALS 24u
typedef struct sr_tune_result {
  sr_tune_trial best;
  uint32_t trial_count;
  uint32_t pad;
  sr_tune_trial trials[SR_TUNE_MAX_TRIALS];
} sr_tune_result;

int sr_autotune(sr_context* ctx, const sr_frame_desc* frame, uint32_t frames_per_trial, float min_psnr_db,
                sr_tune_result* out);

/* ------------------------------------------------------------------------ */
/* Serving: a request loop over device-resident assets (no reference
 * counterpart; the reference renders one frame per Bevy render-graph run,
 * src/plugin.rs:240-260).  A worker thread admits queued requests into frame
 * slots (each slot owns a stream and grow-only buffers, so frames overlap on
 * the device) while the HBM held by assets and slots stays within the budget;
 * requests with the same instance, lights and settings that wait together
 * are merged into one multi-view frame (cameras of different sizes).  Each
 * request's image equals the image of the same request rendered alone. */
#define SR_SERVER_MAX_LIGHTS 4u

typedef struct sr_server sr_server;

typedef struct sr_server_config {
  uint32_t slots;               /* frames in flight; 0 = 2 */
  uint32_t max_views_per_frame; /* requests merged into one frame; 0 = 16 */
  uint64_t hbm_budget_bytes;    /* 0 = the context's budget (0 there = unlimited) */
} sr_server_config;

typedef struct sr_request {
  sr_strand_instance instance;  /* asset ids + transform (one instance) */
  sr_froxel_config camera;      /* resolution and froxel size of the view */
  sr_view camera_view;
  uint32_t light_count;         /* 0..SR_SERVER_MAX_LIGHTS */
  uint32_t flags;               /* SR_REQUEST_* bits */
  sr_froxel_config light_frusta[SR_SERVER_MAX_LIGHTS];
  sr_view light_views[SR_SERVER_MAX_LIGHTS];
  sr_directional_light lights[SR_SERVER_MAX_LIGHTS];
  sr_frame_settings settings;
  float* out_rgba;              /* host, width * height * 4 floats (composite, straight alpha), or with
                                   SR_REQUEST_RGBA8 width * height * 4 bytes (rgba8unorm); or NULL */
  float* out_depth;             /* host, width * height or NULL */
  uint64_t tag;                 /* returned with the result */
} sr_request;

/* sr_request.flags: the composite arrives as rgba8unorm (4 bytes a pixel,
 * SR_FRAME_RGBA8_OUTPUTS) instead of float RGBA. */
#define SR_REQUEST_RGBA8 1u
/* sr_request.flags: out_rgba and out_depth are pinned host memory
 * (hipHostMalloc): the frame copies its outputs straight into them, with no
 * staging copy on the host. */
#define SR_REQUEST_PINNED_OUTPUTS 2u

typedef struct sr_request_result {
  uint64_t ticket;
  uint64_t tag;
  int32_t status;               /* SR_OK, or why the request failed (SR_ERR_CAPACITY: above the budget alone) */
  uint32_t frame_views;         /* cameras in the frame that served it */
  double submit_ms;             /* server clock (ms since start) at sr_server_submit */
  double launch_ms;             /* frame launched */
  double done_ms;               /* outputs written to the caller's buffers */
  sr_frame_counters counters;   /* of the frame that served it */
} sr_request_result;

typedef struct sr_server_stats {
  uint64_t submitted, completed, failed, frames;
  uint64_t peak_held_bytes;     /* assets + slot buffers, the most seen at an admission */
  uint64_t budget_waits;        /* admissions deferred by the budget */
  uint64_t groups_split;        /* multi-view groups halved to fit one slot's share of the budget */
  uint32_t max_frame_views, max_concurrent_frames;
} sr_server_stats;

int sr_server_start(sr_context* ctx, const sr_server_config* config, sr_server** out_server);
/* Copies the request; out buffers must stay valid until its result is polled. */
int sr_server_submit(sr_server* server, const sr_request* request, uint64_t* out_ticket);
/* Up to `capacity` finished results; waits up to wait_ms (0 = no wait) for the first. */
int sr_server_poll(sr_server* server, sr_request_result* results, uint32_t capacity, uint32_t* out_count,
                   int32_t wait_ms);
int sr_server_stats_get(sr_server* server, sr_server_stats* out);
/* Finishes every queued and in-flight request, joins the worker, frees the slots. */
int sr_server_stop(sr_server* server);

/* ------------------------------------------------------------------------ */
/* Synthetic scene generator (configs of the benchmark)                      */
/* ------------------------------------------------------------------------ */
typedef struct sr_scalp_params {
  uint32_t strand_count;
  uint32_t points_per_strand;
  uint64_t seed;
  float scalp_radius;     /* m */
  float length_min;       /* m */
  float length_max;       /* m */
  float curl_radius_max;  /* m */
  float width;            /* m, converted to pixels by the caller */
  float color_mean;
  float color_sigma;
  float radius_pixels;    /* min = max radius in pixels written to materials */
  float pad;
} sr_scalp_params;

/* Fill caller-provided host arrays sized by sr_scalp_sizes. */
int sr_scalp_sizes(const sr_scalp_params* params, uint32_t* vertex_count, uint32_t* index_count,
                   uint32_t* strand_count, uint32_t* material_count);
int sr_scalp_generate(const sr_scalp_params* params, float* vertices, uint32_t* indices,
                      sr_strand_meta* meta, sr_strand_geo* geo, sr_strand_material* materials);

/* Frustum table from FroxelConfigs in frustum order (use_prepass_buffers,
 * src/plugin.rs:386-458): running bucket / coarse / fine tile bases,
 * coarse tiles = ceil(fine tiles / 4). is_light and cascade may be NULL. */
int sr_build_frustum_table(const sr_froxel_config* configs, const uint32_t* is_light,
                           const uint32_t* cascade, uint32_t count, sr_frustum_desc* out);

/* Camera view with Bevy's reverse-Z infinite perspective projection. */
int sr_camera_view(const float* eye3, const float* target3, const float* up3, float fov_y,
                   float aspect, float z_near, sr_view* out);

/* Directional-light view: reverse-Z orthographic projection fitted to a
 * bounding sphere. */
int sr_light_view(const float* direction_to_light3, const float* center3, float radius,
                  sr_view* out);

/* ------------------------------------------------------------------------ */
/* .strands cache (src/strand_cache.rs).                                     */
typedef struct sr_strand_cache_info {
  uint32_t version;                /* 1 */
  float scale;                     /* applied to the vertices written */
  uint64_t vertex_count;
  uint64_t strand_count;
  uint64_t index_count;
  uint32_t max_segments_in_strand; /* max vertex count of a strand (src/plugin.rs:1373) */
  uint32_t format;                 /* 0 = binary "SSRSTRD\0", 1 = JSON */
  sr_aabb aabb;                    /* over every scaled vertex */
} sr_strand_cache_info;
SR_STATIC_ASSERT(sizeof(sr_strand_cache_info) == 72, "sr_strand_cache_info is 72 bytes");

/* Parse, validate and pack a .strands cache held in memory (replaces
 * StrandCacheAssetLoader::load, src/strand_cache.rs:223-232, with
 * parse_cache :165-171 and validate_cache :174-209, plus the packing of
 * set_cached_strand_geometry / pack_indexed_strands, src/plugin.rs:1468-1490,
 * 1357-1393).  Call with NULL outputs to read the counts into `info`, then
 * with host arrays of vertex_count float4 (x, y, z scaled, w = 1),
 * strand_count sr_strand_meta {count, offset, material 0} and index_count
 * uint32.  Errors return SR_ERR_INVALID_ARGUMENT with the reference's message
 * in `error`. */
int sr_strand_cache_parse(const void* bytes, uint64_t size, sr_strand_cache_info* info, float* vertices,
                          sr_strand_meta* meta, uint32_t* indices, char* error, uint64_t error_capacity);

#ifdef __cplusplus
}
#endif

#endif /* STRAND_RT_H */
