"""Synthetic strand workloads (the benchmark configurations) and frame tables.

An asset is a set of host slabs in the reference layout (vertices float4 with
w = 1, identity index slab, StrandMeta, one StrandGeo, StrandMaterial per
strand).  A workload adds instances, frusta in frustum-table order (cameras
are kind 0, directional-light cascades kind 1), one view per frustum, the
directional lights and the frame settings.
"""
from __future__ import annotations

import ctypes as C
import math
import struct
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Asset:
    vertices: np.ndarray   # float32 (V, 4)
    indices: np.ndarray    # uint32 (I,)
    meta: np.ndarray       # strand_meta_dtype (N,)
    geo: np.ndarray        # strand_geo_dtype (1,)
    materials: np.ndarray  # strand_material_dtype (M,)

    @property
    def strand_count(self) -> int:
        return int(self.meta.shape[0])

    @property
    def segment_count(self) -> int:
        return int(np.maximum(self.meta["count"].astype(np.int64) - 1, 0).sum())

    def slabs(self) -> abi.AssetSlabs:
        s = abi.AssetSlabs()
        s.vertices = abi.ptr(self.vertices)
        s.vertex_count = self.vertices.shape[0]
        s.indices = abi.ptr(self.indices)
        s.index_count = self.indices.shape[0]
        s.meta = abi.ptr(self.meta)
        s.strand_count = self.meta.shape[0]
        s.geo = abi.ptr(self.geo)
        s.materials = abi.ptr(self.materials)
        s.material_count = self.materials.shape[0]
        return s


@dataclass
class Workload:
    name: str
    assets: list
    instances: np.ndarray
    frusta: np.ndarray
    views: np.ndarray
    lights: np.ndarray
    settings: dict = field(default_factory=dict)
    # opaque scene depth: float32 (camera pixels in frustum order, samples),
    # reverse-Z, 0 = nothing drawn; None = no opaque geometry
    opaque: np.ndarray | None = None

    @property
    def camera_count(self) -> int:
        return int((self.frusta["kind"] == 0).sum())

    @property
    def input_segments(self) -> int:
        return sum(self.assets[int(i["geo_id"])].segment_count for i in self.instances)

    def frame_settings(self) -> abi.FrameSettings:
        st = abi.FrameSettings()
        d = default_settings()
        d.update(self.settings)
        st.stochastic_cull_enabled = int(d["stochastic_cull_enabled"])
        st.target_strands_per_pixel = d["target_strands_per_pixel"]
        st.min_keep_probability = d["min_keep_probability"]
        st.shadow_keep_probability = d["shadow_keep_probability"]
        st.telemetry_enabled = int(d["telemetry_enabled"])
        st.fine_binning_backend = int(d["fine_binning_backend"])
        st.dom_layers = int(d["dom_layers"])
        st.shadow_history_valid = int(d["shadow_history_valid"])
        for i in range(4):
            st.ambient_color[i] = d["ambient_color"][i]
            st.scene_clear_color[i] = d["scene_clear_color"][i]
        st.scene_clear_depth = d["scene_clear_depth"]
        st.stamp_target_width = int(d["stamp_target_width"])
        st.stamp_target_height = int(d["stamp_target_height"])
        st.flags = int(d["frame_flags"])
        st.dom_resident_pages = int(d["dom_resident_pages"])
        st.coarse_depth_slices = int(d.get("coarse_depth_slices", 0))
        return st

    def frame_desc(self, keep: list | None = None) -> abi.FrameDesc:
        """Builds a sr_frame_desc pointing at this workload's arrays."""
        for a in (self.instances, self.frusta, self.views, self.lights):
            assert a.flags["C_CONTIGUOUS"]
        fd = abi.FrameDesc()
        fd.instances = abi.ptr(self.instances) if len(self.instances) else None
        fd.instance_count = len(self.instances)
        fd.frustum_count = len(self.frusta)
        fd.frusta = abi.ptr(self.frusta)
        fd.views = abi.ptr(self.views)
        fd.lights = abi.ptr(self.lights) if len(self.lights) else None
        fd.light_count = len(self.lights)
        fd.settings = self.frame_settings()
        fd.mark_waves = int(self.settings.get("mark_waves", 0))  # launch shape (0 = defaults)
        fd.raster_waves = int(self.settings.get("raster_waves", 0))
        if self.opaque is not None:
            assert self.opaque.dtype == np.float32 and self.opaque.flags["C_CONTIGUOUS"] and self.opaque.ndim == 2
            fd.opaque_depth = abi.ptr(self.opaque)
            fd.opaque_samples = self.opaque.shape[1]
        return fd


def default_settings() -> dict:
    """StochasticCullSettings defaults (src/resources.rs:95-104) and the
    headless scene parameters used by the benchmark workloads."""
    return {
        "stochastic_cull_enabled": 1,
        "target_strands_per_pixel": 2.0,
        "min_keep_probability": 0.02,
        "shadow_keep_probability": 1.0,
        "telemetry_enabled": 0,
        "fine_binning_backend": 0,
        "dom_layers": 4,
        "shadow_history_valid": 0,
        "ambient_color": (80.0, 80.0, 80.0, 1.0),
        # ClearColor srgb(0.025, 0.025, 0.03) in linear RGB
        "scene_clear_color": (0.0019353, 0.0019353, 0.0025950, 1.0),
        "scene_clear_depth": 0.0,
        "stamp_target_width": 0,
        "stamp_target_height": 0,
        "frame_flags": 0,  # SR_FRAME_* bits
        "dom_resident_pages": 0,  # SR_FRAME_VIRTUAL_DOMS: resident pages per light, 0 = all requested
        "coarse_depth_slices": 0,  # LUT slices per coarse tile: 8, 16 or 32 (0 = 16)
    }


# ----------------------------------------------------------- generators
def scalp_asset(strands: int, points: int, seed: int, width_m: float, radius_px: float,
                curl_max: float = 0.01, scalp_radius: float = 0.1, length=(0.05, 0.25),
                color=(0.3, 0.05)) -> Asset:
    """Seeded scalp hair: roots uniform on a hemisphere, gravity-drooping
    Catmull-Rom strands with helical curl, per-strand colour."""
    lib = abi.load_library()
    p = abi.ScalpParams()
    p.strand_count, p.points_per_strand, p.seed = strands, points, seed
    p.scalp_radius, p.length_min, p.length_max = scalp_radius, length[0], length[1]
    p.curl_radius_max, p.width = curl_max, width_m
    p.color_mean, p.color_sigma, p.radius_pixels = color[0], color[1], radius_px
    nv, ni, ns, nm = (C.c_uint32() for _ in range(4))
    abi.check(lib, lib.sr_scalp_sizes(C.byref(p), C.byref(nv), C.byref(ni), C.byref(ns), C.byref(nm)),
              None, "sr_scalp_sizes")
    a = Asset(vertices=np.zeros((nv.value, 4), np.float32), indices=np.zeros(ni.value, np.uint32),
              meta=np.zeros(ns.value, abi.strand_meta_dtype), geo=np.zeros(1, abi.strand_geo_dtype),
              materials=np.zeros(nm.value, abi.strand_material_dtype))
    abi.check(lib, lib.sr_scalp_generate(C.byref(p), abi.ptr(a.vertices), abi.ptr(a.indices), abi.ptr(a.meta),
                                         abi.ptr(a.geo), abi.ptr(a.materials)), None, "sr_scalp_generate")
    return a


def asset_from_strands(strands: list, material: dict | None = None, radius_px: float = 1.0) -> Asset:
    """Packs explicit polylines like pack_strands (src/plugin.rs:1315-1355):
    strands with fewer than 2 points are dropped, vertices become (x, y, z, 1),
    StrandMeta{count, offset, material}."""
    pts, meta = [], []
    for s in strands:
        s = np.asarray(s, np.float32)
        if s.shape[0] < 2:
            continue
        meta.append((s.shape[0], sum(len(p) for p in pts), len(meta), 0))
        pts.append(s)
    v = np.concatenate(pts).astype(np.float32)
    verts = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
    a = Asset(vertices=np.ascontiguousarray(verts), indices=np.arange(v.shape[0], dtype=np.uint32),
              meta=np.array(meta, dtype=abi.strand_meta_dtype), geo=np.zeros(1, abi.strand_geo_dtype),
              materials=np.zeros(len(meta), abi.strand_material_dtype))
    g = a.geo[0]
    g["strand_count"] = len(meta)
    g["max_segments_in_strand"] = max(m[0] for m in meta)
    g["aabb"]["min"] = v.min(axis=0)
    g["aabb"]["max"] = v.max(axis=0)
    a.geo[0] = g
    m = default_material()
    if material:
        m.update(material)
    m["min_radius_pixels"] = m.get("min_radius_pixels", radius_px)
    m["max_radius_pixels"] = m.get("max_radius_pixels", radius_px)
    for k, val in m.items():
        a.materials[k] = val
    return a


class StrandCacheError(ValueError):
    """A .strands cache that the loader rejects (message as the reference's)."""


def parse_strand_cache(data: bytes):
    """Parses + validates a .strands cache (binary or JSON) through the
    native loader (StrandCacheAssetLoader, src/strand_cache.rs:223-232).
    Returns (info record, vertices float4[V] scaled with w = 1,
    StrandMeta[N], indices u32[I])."""
    lib = abi.load_library()
    buf = C.create_string_buffer(bytes(data), len(data))
    info = np.zeros(1, abi.strand_cache_info_dtype)
    err = C.create_string_buffer(512)
    rc = lib.sr_strand_cache_parse(buf, len(data), abi.ptr(info), None, None, None, err, 512)
    if rc != abi.SR_OK:
        raise StrandCacheError(err.value.decode())
    i = info[0]
    verts = np.zeros((int(i["vertex_count"]), 4), np.float32)
    meta = np.zeros(int(i["strand_count"]), abi.strand_meta_dtype)
    idx = np.zeros(int(i["index_count"]), np.uint32)
    rc = lib.sr_strand_cache_parse(buf, len(data), abi.ptr(info), abi.ptr(verts), abi.ptr(meta),
                                   abi.ptr(idx) if idx.size else None, err, 512)
    if rc != abi.SR_OK:
        raise StrandCacheError(err.value.decode())
    return info[0], verts, meta, idx


def asset_from_cache(data: bytes, material: dict | None = None, radius_px: float = 1.0) -> Asset:
    """set_cached_strand_geometry (src/plugin.rs:1468-1490): the parsed cache
    packed into slabs with one material (every StrandMeta.material_idx = 0)."""
    info, verts, meta, idx = parse_strand_cache(data)
    a = Asset(vertices=np.ascontiguousarray(verts), indices=idx, meta=meta,
              geo=np.zeros(1, abi.strand_geo_dtype), materials=np.zeros(1, abi.strand_material_dtype))
    g = a.geo[0]
    g["strand_count"] = len(meta)
    g["max_segments_in_strand"] = int(info["max_segments_in_strand"])
    g["aabb"] = info["aabb"]
    a.geo[0] = g
    m = default_material()
    if material:
        m.update(material)
    m["min_radius_pixels"] = m.get("min_radius_pixels", radius_px)
    m["max_radius_pixels"] = m.get("max_radius_pixels", radius_px)
    for k, val in m.items():
        a.materials[k] = val
    return a


def strand_cache_bytes(vertices, strands_count_offset, indices, scale: float = 1.0) -> bytes:
    """The binary cache layout written by the reference's converter
    (tools/alembic_to_strands.py:146-162): magic, '<IfQQQ' header, then the
    vertex, (count, offset) and index payloads, little-endian."""
    v = np.ascontiguousarray(vertices, dtype="<f4").reshape(-1, 3)
    so = np.ascontiguousarray(strands_count_offset, dtype="<u4").reshape(-1, 2)
    ix = np.ascontiguousarray(indices, dtype="<u4").reshape(-1)
    head = b"SSRSTRD\0" + struct.pack("<IfQQQ", 1, scale, v.shape[0], so.shape[0], ix.shape[0])
    return head + v.tobytes() + so.tobytes() + ix.tobytes()


def default_material() -> dict:
    """benchmark_material of the reference bench (examples/strand_bench.rs:489-502)."""
    return {"absorption_color": (0.56, 0.25, 0.14, 0.45), "specular_color": (0.9, 0.72, 0.52, 0.2),
            "ambient_factor": 0.15, "ao_factor": 0.1, "eta": 1.55, "beta": 0.85, "alpha": 0.35, "shift": 0.1}


# ----------------------------------------------------------- frame tables
def identity_instance(asset_id: int) -> np.ndarray:
    inst = np.zeros(1, abi.strand_instance_dtype)
    for k in ("vertex_id", "index_id", "meta_id", "geo_id", "material_id"):
        inst[k] = asset_id
    eye = np.eye(4, dtype=np.float32).reshape(-1)
    inst["world_from_local"] = eye
    inst["local_from_world"] = eye
    return inst


def frustum_table(configs: list, kinds: list, cascades: list | None = None) -> np.ndarray:
    """FroxelConfig list -> GpuFrustumDesc table (src/plugin.rs:386-458)."""
    lib = abi.load_library()
    cfg = np.array([tuple(c) for c in configs], dtype=abi.froxel_config_dtype)
    kinds_a = np.array(kinds, np.uint32)
    casc = np.array(cascades if cascades else [0] * len(configs), np.uint32)
    out = np.zeros(len(configs), abi.frustum_desc_dtype)
    abi.check(lib, lib.sr_build_frustum_table(abi.ptr(cfg), abi.ptr(kinds_a), abi.ptr(casc), len(configs),
                                              abi.ptr(out)), None, "sr_build_frustum_table")
    return out


def camera_view(eye, target, fov_y_deg: float, width: int, height: int, near: float = 0.1) -> np.ndarray:
    lib = abi.load_library()
    v = np.zeros(1, abi.view_dtype)
    e = np.array(eye, np.float32)
    t = np.array(target, np.float32)
    up = np.array((0.0, 1.0, 0.0), np.float32)
    abi.check(lib, lib.sr_camera_view(abi.ptr(e), abi.ptr(t), abi.ptr(up), math.radians(fov_y_deg),
                                      width / height, near, abi.ptr(v)), None, "sr_camera_view")
    return v


def light_view(direction_to_light, center, radius: float) -> np.ndarray:
    lib = abi.load_library()
    v = np.zeros(1, abi.view_dtype)
    d = np.array(direction_to_light, np.float32)
    c = np.array(center, np.float32)
    abi.check(lib, lib.sr_light_view(abi.ptr(d), abi.ptr(c), radius, abi.ptr(v)), None, "sr_light_view")
    return v


def make_light(direction_to_light, color=(255.0, 245.0, 230.0)) -> np.ndarray:
    l = np.zeros(1, abi.light_dtype)
    d = np.array(direction_to_light, np.float64)
    d = d / np.linalg.norm(d)
    l["direction_to_light"] = (d[0], d[1], d[2], 0.0)
    l["color"] = (*color, 1.0)
    return l


LIGHT_DIRS = [(0.35, 0.85, 0.4), (-0.6, 0.7, 0.3), (0.2, 0.6, -0.75), (-0.3, 0.9, -0.3)]
HEAD_CENTER = (0.0, -0.06, 0.0)
FOV_Y = 45.0


def radius_px_for(width_m: float, height_px: int, distance: float, fov_y_deg: float = FOV_Y) -> float:
    focal = (height_px / 2.0) / math.tan(math.radians(fov_y_deg) / 2.0)
    return 0.5 * width_m * focal / distance


def bounding_sphere(asset: Asset):
    g = asset.geo[0]["aabb"]
    mn, mx = np.asarray(g["min"], np.float64), np.asarray(g["max"], np.float64)
    return tuple((mn + mx) / 2.0), float(np.linalg.norm(mx - mn) / 2.0) * 1.05


def orbit_eye(i: int, n: int, distance: float, center=HEAD_CENTER, elev_deg: float = 15.0):
    az = 2.0 * math.pi * i / max(n, 1)
    el = math.radians(elev_deg * math.sin(3.0 * az))
    return (center[0] + distance * math.cos(el) * math.sin(az), center[1] + distance * math.sin(el),
            center[2] + distance * math.cos(el) * math.cos(az))


def build_workload(name: str, asset: Asset, cameras: list, light_count: int, light_res: int, dom_layers: int,
                   froxel: int = 8, light_froxel: int = 8, settings: dict | None = None) -> Workload:
    """cameras: list of (eye, width, height).  Frusta are ordered cameras first,
    then one cascade per light."""
    center, radius = bounding_sphere(asset)
    cfgs, kinds, views = [], [], []
    for eye, w, h in cameras:
        cfgs.append((w, h, froxel, froxel, 16))
        kinds.append(0)
        views.append(camera_view(eye, HEAD_CENTER, FOV_Y, w, h))
    lights = []
    for li in range(light_count):
        d = LIGHT_DIRS[li % len(LIGHT_DIRS)]
        cfgs.append((light_res, light_res, light_froxel, light_froxel, 16))
        kinds.append(1)
        views.append(light_view(d, center, radius))
        lights.append(make_light(d))
    st = {"dom_layers": dom_layers}
    if settings:
        st.update(settings)
    return Workload(name=name, assets=[asset], instances=identity_instance(0), frusta=frustum_table(cfgs, kinds),
                    views=np.ascontiguousarray(np.concatenate(views)),
                    lights=np.ascontiguousarray(np.concatenate(lights)) if lights else np.zeros(0, abi.light_dtype),
                    settings=st)


def retile(w: Workload, camera_froxel: int = 0, light_froxel: int = 0, backend: int | None = None) -> Workload:
    """The workload with its frustum table rebuilt for other froxel sizes (0
    keeps a frustum's size) and, optionally, another binning backend: how an
    autotuner result is applied."""
    cfgs, kinds = [], []
    for f in w.frusta:
        size = int(light_froxel if f["kind"] == 1 else camera_froxel)
        fx, fy = (size, size) if size else (int(f["froxel_size_x"]), int(f["froxel_size_y"]))
        cfgs.append((int(f["screen_width"]), int(f["screen_height"]), fx, fy, int(f["depth_slices"])))
        kinds.append(int(f["kind"]))
    st = dict(w.settings)
    if backend is not None:
        st["fine_binning_backend"] = int(backend)
    return Workload(name=w.name, assets=w.assets, instances=w.instances, frusta=frustum_table(cfgs, kinds),
                    views=w.views, lights=w.lights, settings=st, opaque=w.opaque)


def apply_tuning(w: Workload, best: dict) -> Workload:
    """The workload with an autotuner result applied: froxel sizes and
    binning backend (retile), depth-warp slices per coarse tile and the
    launch shape (marking waves per workgroup, raster waves per CU)."""
    t = retile(w, best["camera_froxel"], best["light_froxel"], best["fine_binning_backend"])
    t.settings["coarse_depth_slices"] = int(best.get("coarse_depth_slices", 0))
    t.settings["mark_waves"] = int(best.get("mark_waves", 0))
    t.settings["raster_waves"] = int(best.get("raster_waves", 0))
    return t


def select_views(w: Workload, cameras: list, name: str | None = None) -> Workload:
    """The workload with only the camera frusta `cameras` (indices into its
    frustum table, kept in the order given) and every light frustum after
    them, the frustum table rebuilt with the same froxel sizes: one view of a
    batched frame drawn on its own."""
    fr = w.frusta
    keep = list(cameras) + [i for i in range(len(fr)) if int(fr[i]["kind"]) == 1]
    for i in cameras:
        if int(fr[i]["kind"]) != 0:
            raise ValueError(f"frustum {i} is not a camera")
    cfgs = [(int(fr[i]["screen_width"]), int(fr[i]["screen_height"]), int(fr[i]["froxel_size_x"]),
             int(fr[i]["froxel_size_y"]), int(fr[i]["depth_slices"])) for i in keep]
    kinds = [int(fr[i]["kind"]) for i in keep]
    return Workload(name=name or f"{w.name}-views{'-'.join(map(str, cameras))}", assets=w.assets,
                    instances=w.instances, frusta=frustum_table(cfgs, kinds),
                    views=np.ascontiguousarray(w.views[keep]), lights=w.lights, settings=dict(w.settings),
                    opaque=None)


def camera_pixel_offsets(w: Workload) -> list:
    """(first pixel, width, height) of every camera frustum in the frame's
    concatenated image / composite / depth outputs, in frustum order."""
    out, off = [], 0
    for f in w.frusta:
        if int(f["kind"]) == 0:
            cw, ch = int(f["screen_width"]), int(f["screen_height"])
            out.append((off, cw, ch))
            off += cw * ch
    return out


def opaque_disc(w: Workload, depth: float, radius_frac: float = 0.3, samples: int = 4, seed: int = 1) -> np.ndarray:
    """A synthetic opaque scene depth for every camera of `w`: a disc (an
    occluder in front of or behind the hair) of constant reverse-Z `depth`
    around the image centre, jittered per MSAA sample near its rim, zero
    elsewhere.  Shape (camera pixels, samples), float32."""
    rng = np.random.default_rng(seed)
    out = []
    for f in w.frusta:
        if f["kind"] != 0:
            continue
        W, H = int(f["screen_width"]), int(f["screen_height"])
        yy, xx = np.mgrid[0:H, 0:W]
        r = np.hypot(xx + 0.5 - W / 2, yy + 0.5 - H / 2) / (radius_frac * min(W, H))
        d = np.zeros((H * W, samples), np.float32)
        for k in range(samples):
            inside = (r + rng.uniform(-0.02, 0.02, size=r.shape)).reshape(-1) < 1.0
            d[inside, k] = depth * (1.0 + 0.01 * k)
        out.append(d)
    return np.ascontiguousarray(np.concatenate(out))


def with_asset_id(w: Workload, asset_id: int) -> Workload:
    """The same workload drawing the uploaded asset `asset_id`."""
    inst = w.instances.copy()
    for k in ("vertex_id", "index_id", "meta_id", "geo_id", "material_id"):
        inst[k] = asset_id
    return Workload(name=w.name, assets=w.assets, instances=inst, frusta=w.frusta, views=w.views, lights=w.lights,
                    settings=dict(w.settings), opaque=w.opaque)


def request_from_workload(w: Workload, tag: int = 0) -> np.ndarray:
    """A serving request (sr_request) equal to a one-camera workload: its
    camera frustum (resolution, froxel size), view, lights and settings, and
    its single instance (asset ids as in the workload)."""
    cams = np.nonzero(w.frusta["kind"] == 0)[0]
    lights = np.nonzero(w.frusta["kind"] == 1)[0]
    if len(cams) != 1 or len(w.instances) != 1 or len(lights) > abi.SERVER_MAX_LIGHTS:
        raise ValueError("a request holds one camera, one instance and at most 4 lights")
    r = np.zeros(1, abi.request_dtype)
    r["instance"] = w.instances[0]
    f = w.frusta[cams[0]]
    for k in ("screen_width", "screen_height", "froxel_size_x", "froxel_size_y", "depth_slices"):
        r["camera"][k] = f[k]
    r["camera_view"] = w.views[cams[0]]
    r["light_count"] = len(lights)
    for j, li in enumerate(lights):
        for k in ("screen_width", "screen_height", "froxel_size_x", "froxel_size_y", "depth_slices"):
            r["light_frusta"][0, j][k] = w.frusta[li][k]
        r["light_views"][0, j] = w.views[li]
        r["lights"][0, j] = w.lights[j]
    st = w.frame_settings()
    for name, _ in st._fields_:
        v = getattr(st, name)
        r["settings"][name] = list(v) if not isinstance(v, (int, float)) else v
    r["tag"] = tag
    return r


def config_small(froxel: int = 16, seed: int = 0x5EEDCAFE) -> Workload:
    """2,000 strands x 16 points, width 0.0008 m, 512x512 camera at 0.6 m,
    one directional light with a 512^2 deep shadow map of 4 layers."""
    dist = 0.6
    asset = scalp_asset(2000, 16, seed, 0.0008, radius_px_for(0.0008, 512, dist))
    eye = (HEAD_CENTER[0], HEAD_CENTER[1] + 0.05, HEAD_CENTER[2] + dist)
    return build_workload(f"scalp-2k-x16-512sq-f{froxel}", asset, [(eye, 512, 512)], 1, 512, 4, froxel, froxel)


def with_virtual_doms(w: Workload, budget: int = 0) -> Workload:
    """The workload drawn with virtual deep shadow maps (SR_FRAME_VIRTUAL_DOMS)
    holding at most `budget` resident pages per light (0: every requested page)."""
    st = dict(w.settings)
    st["frame_flags"] = int(st.get("frame_flags", 0)) | abi.FRAME_VIRTUAL_DOMS
    st["dom_resident_pages"] = int(budget)
    return dataclasses.replace(w, settings=st)


def config_virtual_doms(light_res: int = 2048, spread: float = 2.5, budget: int = 0, froxel: int = 8,
                        seed: int = 0x5EEDCAFE) -> Workload:
    """config_small's 2,000 x 16 scalp and 512x512 camera with one light whose
    light_res^2 deep shadow map (pages of 256^2 texels) frames a sphere `spread`
    times the asset's bounding sphere, so the hair's shadow tiles request only
    the pages around the middle; drawn with virtual DOMs of at most `budget`
    resident pages (0: every requested page)."""
    dist = 0.6
    asset = scalp_asset(2000, 16, seed, 0.0008, radius_px_for(0.0008, 512, dist))
    eye = (HEAD_CENTER[0], HEAD_CENTER[1] + 0.05, HEAD_CENTER[2] + dist)
    w = build_workload(f"scalp-2k-x16-512sq-light{light_res}-x{spread:g}-vdom{budget}", asset, [(eye, 512, 512)],
                       1, light_res, 4, froxel, froxel)
    center, radius = bounding_sphere(asset)
    views = w.views.copy()
    views[1] = light_view(LIGHT_DIRS[0], center, radius * spread)[0]
    return with_virtual_doms(dataclasses.replace(w, views=np.ascontiguousarray(views)), budget)


def config_named(seed: int = 0x5EEDCAFE, strands: int = 100_000, points: int = 32) -> Workload:
    """100k strands x 32 points, width 0.0006 m, 1920x1080 camera, one
    directional light with a 2048^2 deep shadow map of 4 layers, 8x8 froxels."""
    dist = 0.6
    asset = scalp_asset(strands, points, seed, 0.0006, radius_px_for(0.0006, 1080, dist))
    eye = (HEAD_CENTER[0], HEAD_CENTER[1] + 0.05, HEAD_CENTER[2] + dist)
    return build_workload(f"scalp-{strands // 1000}k-x{points}-1080p", asset, [(eye, 1920, 1080)], 1, 2048, 4)


def config_batched(views: int = 64, seed: int = 0x5EEDCAFE, strands: int = 2000, points: int = 16,
                   lights: int = 4, width_m: float = 0.0008, resolutions=((1280, 720), (1920, 1080), (2560, 1440)),
                   light_res: int = 2048, froxel: int = 8) -> Workload:
    """Many views in one frame: a scalp asset seen from `views` orbit cameras
    with resolutions drawn (seeded) from `resolutions`, and `lights`
    light_res^2 x 4-layer lights shared by every view.  The defaults are the
    batching configuration: the 2,000 x 16 scalp, 64 cameras over {720p,
    1080p, 1440p}, 4 lights of 2048^2."""
    dist = 0.6
    asset = scalp_asset(strands, points, seed, width_m, radius_px_for(width_m, 1080, dist))
    rng = np.random.default_rng(seed)
    cams = []
    for i in range(views):
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        cams.append((orbit_eye(i, views, dist), w, h))
    return build_workload(f"scalp-{strands}x{points}-{views}views-{lights}lights", asset, cams, lights, light_res, 4,
                          froxel, froxel)


def config_batched_named(views: int = 8, seed: int = 0x5EEDCAFE) -> Workload:
    """The named asset (100k x 32, width 0.0006 m) from `views` 1920x1080
    orbit cameras sharing one 2048^2 x 4-layer light."""
    return config_batched(views=views, seed=seed, strands=100_000, points=32, lights=1, width_m=0.0006,
                          resolutions=((1920, 1080),))


def config_stress_chunks(views: int = 16, views_per_frame: int = 4, first_frame_views: int = 2,
                         seed: int = 0x5EEDCAFE, strands: int = 1_000_000, points: int = 64, lights: int = 4,
                         light_res: int = 2048, dom_layers: int = 8, width: int = 3840,
                         height: int = 2160) -> list:
    """The stress configuration: 1M strands x 64 points (63M segments), curl
    radius up to 0.02 m, `views` 3840x2160 orbit cameras and 4 lights with
    2048^2 x 8-layer deep shadow maps, as a batch of frames that share one
    build of the lights.  One frame of every view would need more fine
    references than 32-bit offsets address, so the first frame draws
    `first_frame_views` cameras and builds the lights, and every later frame
    draws `views_per_frame` cameras with the shadow maps the first one left
    (SR_FRAME_REUSE_LIGHT_DOMS; the same lights, instance and settings).  All
    frames draw the same asset (id 0, upload it once)."""
    dist = 0.6
    asset = scalp_asset(strands, points, seed, 0.0006, radius_px_for(0.0006, height, dist), curl_max=0.02)
    cams = [(orbit_eye(i, views, dist), width, height) for i in range(views)]
    bounds = [0, min(first_frame_views, views)]
    while bounds[-1] < views:
        bounds.append(min(bounds[-1] + views_per_frame, views))
    frames = []
    for k in range(len(bounds) - 1):
        a, b = bounds[k], bounds[k + 1]
        w = build_workload(f"stress-{strands // 1000}k-x{points}-views{a}-{b - 1}", asset, cams[a:b], lights,
                           light_res, dom_layers)
        if k > 0:
            w.settings["frame_flags"] = abi.FRAME_REUSE_LIGHT_DOMS
        frames.append(w)
    return frames


SERVING_SIZES = (10_000, 25_000, 50_000, 100_000, 150_000, 200_000, 300_000, 500_000)
SERVING_RESOLUTIONS = ((256, 256), (512, 512), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))


def serving_assets(sizes=SERVING_SIZES, points: int = 32, seed: int = 0x5EEDCAFE) -> list:
    """The serving configuration's device-resident assets: scalps of `sizes`
    strands x 32 points (width 0.0006 m), one seed each."""
    return [scalp_asset(n, points, seed + i, 0.0006, radius_px_for(0.0006, 1080, 0.6)) for i, n in enumerate(sizes)]


def serving_request(assets: list, asset_ids: list, rng: np.random.Generator, tag: int = 0, backend: int = 0):
    """One request of the serving stream: a uniform-random asset, a random
    orbit camera and a resolution drawn from SERVING_RESOLUTIONS, lit by one
    2048^2 x 4-layer light (shared by requests of the same asset), binned
    with fine-binning `backend` (0 segment scatter, 1 page CSR).  Returns
    (request record, workload of the same frame drawn alone)."""
    k = int(rng.integers(0, len(assets)))
    w, h = SERVING_RESOLUTIONS[int(rng.integers(0, len(SERVING_RESOLUTIONS)))]
    eye = orbit_eye(int(rng.integers(0, 360)), 360, 0.6)
    wl = with_asset_id(build_workload(f"serve-{k}-{w}x{h}", assets[k], [(eye, w, h)], 1, 2048, 4, 8, 8,
                                      settings={"fine_binning_backend": backend}), asset_ids[k])
    return request_from_workload(wl, tag=tag), wl


def config_two_strands(light_offset: float = 0.0, dom_layers: int = 4, res: int = 64,
                       near=(0.9, 0.1, 0.1, 0.6), far=(0.1, 0.2, 0.9, 0.7)) -> Workload:
    """Two straight one-segment strands on pixel row 20 of a res x res view,
    seen by a camera and a light whose clip_from_world is the identity
    (screen x = (x / 2 + 1 / 2) res, y = (1 / 2 - y / 2) res, depth = z): the
    near strand at depth 0.8 from x = -0.5 to 0.5, the far one at depth 0.5
    from x = -0.4 to 0.4, both 1.5 px in radius, so pixel (res / 2, 20) is
    covered fully by both.  ambient_factor 1 with no ambient light makes the
    shaded colour each strand's absorption colour.  light_offset shifts the
    light's view sideways (clip x + offset): an offset of 4 puts both strands
    outside the light's map, so nothing shadows them.  Stochastic culling is
    off."""
    y = 1.0 - 2.0 * 20.5 / res  # pixel-centre row 20
    asset = asset_from_strands([[(-0.5, y, 0.8), (0.5, y, 0.8)], [(-0.4, y, 0.5), (0.4, y, 0.5)]],
                               material={"ambient_factor": 1.0}, radius_px=1.5)
    asset.materials["absorption_color"][0] = near
    asset.materials["absorption_color"][1] = far
    eye = np.eye(4, dtype=np.float32)
    views = np.zeros(2, abi.view_dtype)
    views["clip_from_world"][0] = eye.T.reshape(-1)
    views["world_position"][0] = (0.0, 0.0, 5.0, 1.0)
    light = eye.copy()
    light[0, 3] = light_offset
    views["clip_from_world"][1] = light.T.reshape(-1)
    views["world_position"][1] = (0.0, 0.0, 5.0, 1.0)
    frusta = frustum_table([(res, res, 8, 8, 16), (res, res, 8, 8, 16)], [0, 1])
    return Workload(name=f"two-strands-{res}-light{light_offset:g}-L{dom_layers}", assets=[asset],
                    instances=identity_instance(0), frusta=frusta, views=views,
                    lights=np.ascontiguousarray(make_light((0.0, 0.0, 1.0))),
                    settings={"dom_layers": dom_layers, "stochastic_cull_enabled": 0,
                              "ambient_color": (0.0, 0.0, 0.0, 1.0)})


def config_one_cell_pair(swap: bool = False, res: int = 64) -> Workload:
    """Two straight one-segment strands on pixel row 20 of a res x res view
    (the identity view of config_two_strands, no shadowing light), one at
    depth 0.6 (colour (0.9, 0.1, 0.1), alpha 0.6) and one a single float step
    nearer (colour (0.1, 0.2, 0.9), alpha 0.7): both depths quantise to the
    same 24-bit value, so every pixel they share holds both fragments in one
    cell, where the (inst, seg) order and the depth order disagree.  swap
    lists the nearer strand first."""
    y = 1.0 - 2.0 * 20.5 / res
    z0 = np.float32(0.6)
    z1 = np.nextafter(z0, np.float32(1.0))
    strands = [[(-0.5, y, float(z0)), (0.5, y, float(z0))], [(-0.5, y, float(z1)), (0.5, y, float(z1))]]
    colors = [(0.9, 0.1, 0.1, 0.6), (0.1, 0.2, 0.9, 0.7)]
    if swap:
        strands, colors = strands[::-1], colors[::-1]
    w = config_two_strands(light_offset=4.0, res=res)
    asset = asset_from_strands(strands, material={"ambient_factor": 1.0}, radius_px=1.5)
    for i, c in enumerate(colors):
        asset.materials["absorption_color"][i] = c
    return dataclasses.replace(w, name=f"one-cell-pair-{res}{'-swapped' if swap else ''}", assets=[asset])


def config_tiny(strands: int = 24, points: int = 6, res: int = 64, froxel: int = 8, seed: int = 7,
                light_res: int = 64, dom_layers: int = 4, lights: int = 1) -> Workload:
    """Tiny scalp for exhaustive parity checks."""
    dist = 0.6
    asset = scalp_asset(strands, points, seed, 0.004, radius_px_for(0.004, res, dist))
    eye = (HEAD_CENTER[0], HEAD_CENTER[1] + 0.05, HEAD_CENTER[2] + dist)
    return build_workload(f"tiny-{strands}x{points}-{res}", asset, [(eye, res, res)], lights, light_res,
                          dom_layers, froxel, froxel)
