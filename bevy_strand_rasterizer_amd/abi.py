"""ctypes / numpy mirrors of ``include/strand_rt.h``.

Record layouts follow the reference's GPU ABI (src/shader_types.rs,
src/pipelines/task_contract.rs, src/plugin.rs:1545-1593,
src/components.rs:61-123).  ``tests/test_abi.py`` checks every size and
field offset against the C header.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libstrand_rt.so"

COARSE_FINE_TILE_EXTENT = 4
COARSE_DEPTH_SLICES = 16
MAX_COARSE_DEPTH_SLICES = 32
COARSE_COUNT_PAGE_SIZE = 256
DEPTH_QUANT_MAX = 16777215
INVALID_PTR = 0xFFFFFFFF
MAX_DOM_LAYERS = 16
BINNING_FRUSTUM_SHIFT = 3
CONFIG_CSR_FEW_HITS = 2  # sr_device_config.flags
CONFIG_HIP_GRAPHS = 4
CONFIG_TIGHT_CAPACITY = 8
NOALLOC_SCRATCH_BYTES = 256 << 20  # SR_NOALLOC_SCRATCH_BYTES
FRAME_REUSE_LIGHT_DOMS = 1  # sr_frame_settings.flags
FRAME_RGBA8_OUTPUTS = 2
FRAME_VIRTUAL_DOMS = 4
FRAME_DEPTH_SORTED_BLEND = 8
FRAME_TUNE_KEEP_BACKEND = 16  # read by sr_autotune only
BLEND_SORT_FRAGMENTS = 4  # SR_BLEND_SORT_FRAGMENTS
DOM_PAGE_XY = 256  # SR_DOM_PAGE_XY
DOM_PREFETCH_PAGE_BORDER = 1
REQUEST_RGBA8 = 1  # sr_request.flags
REQUEST_PINNED_OUTPUTS = 2

# sr_status
SR_OK = 0
STATUS_NAMES = {
    0: "SR_OK", -1: "SR_ERR_INVALID_ARGUMENT", -2: "SR_ERR_OUT_OF_MEMORY", -3: "SR_ERR_DEVICE",
    -4: "SR_ERR_NOT_READY", -5: "SR_ERR_CAPACITY", -6: "SR_ERR_INVALID_ASSET", -7: "SR_ERR_UNSUPPORTED",
}
STATUS_CODES = {v: k for k, v in STATUS_NAMES.items()}

# sr_dump_kind
DUMP = {
    "counters": 0, "broad_meta": 1, "keep": 2, "ranges": 3, "tasks": 4, "projected": 5, "shading_queue": 6,
    "lut": 7, "page_table": 8, "page_meta": 9, "cell_counts": 10, "cell_offsets": 11, "refs": 12,
    "active_tiles": 13, "work_items": 14, "camera_runs": 15, "shadow_runs": 16, "dom_depth": 17,
    "dom_opacity": 18, "atlas": 19, "history": 20, "image": 21, "depth": 22, "composite": 23, "stamp": 24,
    "telemetry": 25, "opaque_tiles": 26, "image_rgba8": 27, "composite_rgba8": 28, "dom_requests": 29,
    "dom_page_table": 30,
}


def _u32(*names):
    return np.dtype([(n, "<u4") for n in names])


# ----------------------------------------------------------- numpy dtypes
strand_meta_dtype = _u32("count", "offset", "material_idx", "pad")
aabb_dtype = np.dtype([("min", "<f4", 3), ("pad1", "<f4"), ("max", "<f4", 3), ("pad2", "<f4")])
strand_geo_dtype = np.dtype([("strand_count", "<u4"), ("max_segments_in_strand", "<u4"), ("offset", "<u4"),
                             ("pad", "<u4"), ("aabb", aabb_dtype), ("pad3", "<u4", 4)])
strand_instance_dtype = np.dtype([("vertex_id", "<u4"), ("index_id", "<u4"), ("meta_id", "<u4"), ("geo_id", "<u4"),
                                  ("material_id", "<u4"), ("pad0", "<u4"), ("pad1", "<u4"), ("pad2", "<u4"),
                                  ("world_from_local", "<f4", 16), ("local_from_world", "<f4", 16)])
strand_material_dtype = np.dtype([("absorption_color", "<f4", 4), ("specular_color", "<f4", 4),
                                  ("ambient_factor", "<f4"), ("ao_factor", "<f4"), ("eta", "<f4"), ("beta", "<f4"),
                                  ("alpha", "<f4"), ("shift", "<f4"), ("min_radius_pixels", "<f4"),
                                  ("max_radius_pixels", "<f4")])
frustum_desc_dtype = _u32("screen_width", "screen_height", "froxel_size_x", "froxel_size_y", "depth_slices",
                          "bucket_base", "bucket_count", "kind", "coarse_depth_tile_base", "coarse_depth_tile_count",
                          "coarse_tiles_x", "coarse_tiles_y", "cascade_index", "fine_depth_tile_base", "pad1", "pad2")
view_dtype = np.dtype([("clip_from_world", "<f4", 16), ("world_position", "<f4", 4)])
light_dtype = np.dtype([("direction_to_light", "<f4", 4), ("color", "<f4", 4)])
binning_task_dtype = _u32("id_info", "chunk_id", "seg_idx", "packed_field")
projected_segment_dtype = _u32("p0_xy", "p1_xy", "depths")
raster_work_item_dtype = _u32("seg_ref_base", "seg_ref_count", "frustum_id", "screen_tile_id", "depth_key",
                              "load_score", "page_idx", "cell_idx")
raster_tile_run_dtype = _u32("work_base", "work_count", "frustum_id", "screen_tile_id", "load_score", "flags",
                             "reserved_a", "reserved_b")
fine_seg_ref_dtype = _u32("inst_id", "seg_id", "material_idx")
fine_page_meta_dtype = _u32("seg_ref_base", "seg_ref_count", "work_base", "work_count", "coarse_tile_id",
                            "coarse_z", "first_depth_key", "flags")
coarse_asset_range_dtype = _u32("geo_id", "frustum_id", "min_x", "max_x", "min_y", "max_y", "z_min_q", "z_max_q")
coarse_interval_ref_dtype = _u32("next", "range_id", "z_min_q", "z_max_q")
coarse_depth_lut_entry_dtype = _u32("z_min_q", "z_max_q", "virtual_start_q", "virtual_count_q")
broad_instance_meta_dtype = np.dtype([("strand_count", "<u4"), ("visible", "<u4"),
                                      ("camera_keep_probability", "<f4"), ("shadow_keep_probability", "<f4")])
frame_counters_dtype = _u32("binning_tasks", "camera_tasks", "allocated_pages", "fine_refs", "active_tiles",
                            "work_items", "camera_runs", "shadow_runs", "overflow_pages", "coarse_tiles",
                            "task_capacity", "ref_capacity", "page_drops", "page_candidates", "capacity_exceeded", "dom_pages")
push_constants_dtype = np.dtype([("workgroup_offset", "<u4"), ("num_elements", "<u4"), ("frustum_count", "<u4"),
                                 ("scan_load_base", "<u4"), ("scan_save_base", "<u4"),
                                 ("stochastic_cull_enabled", "<u4"), ("target_strands_per_pixel", "<f4"),
                                 ("min_keep_probability", "<f4"), ("shadow_keep_probability", "<f4"),
                                 ("telemetry_enabled", "<u4"), ("fine_binning_backend", "<u4")])
device_ptr_dtype = _u32("slab", "offset", "size")
froxel_config_dtype = _u32("screen_width", "screen_height", "froxel_size_x", "froxel_size_y", "depth_slices")

DUMP_DTYPES = {
    "counters": frame_counters_dtype, "broad_meta": broad_instance_meta_dtype, "keep": np.dtype("<f4"),
    "ranges": coarse_asset_range_dtype, "tasks": binning_task_dtype, "projected": projected_segment_dtype,
    "shading_queue": np.dtype("<u4"), "lut": coarse_depth_lut_entry_dtype, "page_table": np.dtype("<u4"),
    "page_meta": fine_page_meta_dtype, "cell_counts": np.dtype("<u4"), "cell_offsets": np.dtype("<u4"),
    "refs": fine_seg_ref_dtype, "active_tiles": np.dtype("<u4"), "work_items": raster_work_item_dtype,
    "camera_runs": raster_tile_run_dtype, "shadow_runs": raster_tile_run_dtype, "dom_depth": np.dtype("<f4"),
    "dom_opacity": np.dtype("<f4"), "atlas": np.dtype("<f4"), "history": np.dtype("<f4"),
    "image": np.dtype("<f4"), "depth": np.dtype("<f4"), "composite": np.dtype("<f4"), "stamp": np.dtype("<f4"),
    "telemetry": np.dtype("<u4"), "opaque_tiles": np.dtype([("covered", "<u4"), ("farthest_q", "<u4")]),
    "image_rgba8": np.dtype("u1"), "composite_rgba8": np.dtype("u1"), "dom_requests": np.dtype("<u4"),
    "dom_page_table": np.dtype("<u4"),
}


# ----------------------------------------------------------- ctypes structs
class FrameSettings(C.Structure):
    _fields_ = [("stochastic_cull_enabled", C.c_uint32), ("target_strands_per_pixel", C.c_float),
                ("min_keep_probability", C.c_float), ("shadow_keep_probability", C.c_float),
                ("telemetry_enabled", C.c_uint32), ("fine_binning_backend", C.c_uint32),
                ("dom_layers", C.c_uint32), ("shadow_history_valid", C.c_uint32),
                ("ambient_color", C.c_float * 4), ("scene_clear_color", C.c_float * 4),
                ("scene_clear_depth", C.c_float), ("stamp_target_width", C.c_uint32),
                ("stamp_target_height", C.c_uint32), ("flags", C.c_uint32),
                ("dom_resident_pages", C.c_uint32), ("coarse_depth_slices", C.c_uint32)]


class FrameDesc(C.Structure):
    _fields_ = [("instances", C.c_void_p), ("instance_count", C.c_uint32), ("frustum_count", C.c_uint32),
                ("frusta", C.c_void_p), ("views", C.c_void_p), ("lights", C.c_void_p),
                ("light_count", C.c_uint32), ("mark_waves", C.c_uint32), ("settings", FrameSettings),
                ("opaque_depth", C.c_void_p), ("opaque_samples", C.c_uint32), ("raster_waves", C.c_uint32)]


class AssetSlabs(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32), ("index_count", C.c_uint32),
                ("indices", C.c_void_p), ("meta", C.c_void_p), ("strand_count", C.c_uint32),
                ("material_count", C.c_uint32), ("geo", C.c_void_p), ("materials", C.c_void_p)]


class DevicePtr(C.Structure):
    _fields_ = [("slab", C.c_uint32), ("offset", C.c_uint32), ("size", C.c_uint32)]


class AssetInfo(C.Structure):
    _fields_ = [("vertex_id", C.c_uint32), ("index_id", C.c_uint32), ("meta_id", C.c_uint32),
                ("geo_id", C.c_uint32), ("material_id", C.c_uint32), ("strand_count", C.c_uint32),
                ("index_count", C.c_uint32), ("max_segments_in_strand", C.c_uint32), ("vertices", DevicePtr),
                ("indices", DevicePtr), ("meta", DevicePtr), ("geo", DevicePtr), ("materials", DevicePtr),
                ("pad", C.c_uint32)]


class DeviceConfig(C.Structure):
    _fields_ = [("device_index", C.c_int32), ("flags", C.c_uint32), ("slab_bytes", C.c_uint64),
                ("hbm_budget_bytes", C.c_uint64)]


class ScalpParams(C.Structure):
    _fields_ = [("strand_count", C.c_uint32), ("points_per_strand", C.c_uint32), ("seed", C.c_uint64),
                ("scalp_radius", C.c_float), ("length_min", C.c_float), ("length_max", C.c_float),
                ("curl_radius_max", C.c_float), ("width", C.c_float), ("color_mean", C.c_float),
                ("color_sigma", C.c_float), ("radius_pixels", C.c_float), ("pad", C.c_float)]


# Every symbol the header declares, with its ctypes signature.
tune_trial_dtype = np.dtype([("fine_binning_backend", "<u4"), ("camera_froxel", "<u4"), ("light_froxel", "<u4"),
                             ("hip_graphs", "<u4"), ("ms_per_frame", "<f4"), ("psnr_db", "<f4"), ("accepted", "<u4"),
                             ("mark_waves", "<u4"), ("raster_waves", "<u4"), ("coarse_depth_slices", "<u4")])
TUNE_MAX_TRIALS = 24
tune_result_dtype = np.dtype([("best", tune_trial_dtype), ("trial_count", "<u4"), ("pad", "<u4"),
                              ("trials", tune_trial_dtype, TUNE_MAX_TRIALS)])

SERVER_MAX_LIGHTS = 4
request_dtype = np.dtype([("instance", strand_instance_dtype), ("camera", froxel_config_dtype),
                          ("camera_view", view_dtype), ("light_count", "<u4"), ("flags", "<u4"),
                          ("light_frusta", froxel_config_dtype, SERVER_MAX_LIGHTS),
                          ("light_views", view_dtype, SERVER_MAX_LIGHTS), ("lights", light_dtype, SERVER_MAX_LIGHTS),
                          ("settings", np.dtype([("stochastic_cull_enabled", "<u4"), ("target_strands_per_pixel", "<f4"),
                                                 ("min_keep_probability", "<f4"), ("shadow_keep_probability", "<f4"),
                                                 ("telemetry_enabled", "<u4"), ("fine_binning_backend", "<u4"),
                                                 ("dom_layers", "<u4"), ("shadow_history_valid", "<u4"),
                                                 ("ambient_color", "<f4", 4), ("scene_clear_color", "<f4", 4),
                                                 ("scene_clear_depth", "<f4"), ("stamp_target_width", "<u4"),
                                                 ("stamp_target_height", "<u4"), ("flags", "<u4"),
                                                 ("dom_resident_pages", "<u4"), ("coarse_depth_slices", "<u4")])),
                          ("out_rgba", "<u8"), ("out_depth", "<u8"), ("tag", "<u8")], align=True)
request_result_dtype = np.dtype([("ticket", "<u8"), ("tag", "<u8"), ("status", "<i4"), ("frame_views", "<u4"),
                                 ("submit_ms", "<f8"), ("launch_ms", "<f8"), ("done_ms", "<f8"),
                                 ("counters", frame_counters_dtype)], align=True)
server_config_dtype = np.dtype([("slots", "<u4"), ("max_views_per_frame", "<u4"), ("hbm_budget_bytes", "<u8")])
server_stats_dtype = np.dtype([("submitted", "<u8"), ("completed", "<u8"), ("failed", "<u8"), ("frames", "<u8"),
                               ("peak_held_bytes", "<u8"), ("budget_waits", "<u8"), ("groups_split", "<u8"),
                               ("max_frame_views", "<u4"),
                               ("max_concurrent_frames", "<u4")])

strand_cache_info_dtype = np.dtype([("version", "<u4"), ("scale", "<f4"), ("vertex_count", "<u8"),
                                    ("strand_count", "<u8"), ("index_count", "<u8"),
                                    ("max_segments_in_strand", "<u4"), ("format", "<u4"), ("aabb", aabb_dtype)])

_SIGNATURES = {
    "sr_version": (C.c_char_p, []),
    "sr_create": (C.c_int, [C.POINTER(DeviceConfig), C.POINTER(C.c_void_p)]),
    "sr_destroy": (C.c_int, [C.c_void_p]),
    "sr_asset_upload": (C.c_int, [C.c_void_p, C.POINTER(AssetSlabs), C.POINTER(AssetInfo)]),
    "sr_asset_release": (C.c_int, [C.c_void_p, C.c_uint32]),
    "sr_frame_submit": (C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_void_p]),
    "sr_frame_submit_noalloc": (C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "sr_frame_wait": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sr_frame_dump": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sr_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "sr_stage_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_int32), C.c_char_p,
                                 C.c_int32]),
    "sr_last_error": (C.c_char_p, [C.c_void_p]),
    "sr_scalp_sizes": (C.c_int, [C.POINTER(ScalpParams)] + [C.POINTER(C.c_uint32)] * 4),
    "sr_scalp_generate": (C.c_int, [C.POINTER(ScalpParams)] + [C.c_void_p] * 5),
    "sr_build_frustum_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "sr_camera_view": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]),
    "sr_light_view": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]),
    "sr_strand_cache_parse": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_char_p, C.c_uint64]),
    "sr_autotune": (C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_uint32, C.c_float, C.c_void_p]),
    "sr_server_start": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sr_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "sr_server_poll": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_int32]),
    "sr_server_stats_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sr_server_stop": (C.c_int, [C.c_void_p]),
}

_lib = None


def load_library(path: Path | None = None) -> C.CDLL:
    """Loads libstrand_rt.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"native strand runtime not built: {p} is missing (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def ptr(a: np.ndarray) -> int:
    """Address of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


def check(lib, rc: int, ctx=None, what: str = "") -> None:
    if rc != SR_OK:
        msg = lib.sr_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise RuntimeError(f"{what}: {STATUS_NAMES.get(rc, rc)}: {msg}")
