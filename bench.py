"""Benchmark of the MI355X strand rendering path.

One step = one frame of the plain per-frame path (prepass/culling,
projection, froxel binning, scans and compaction, raster-work emission,
deep-shadow-map build and stampback, hair shading, strand raster,
composite) over the named workload: 100k strands x 32 points (3.1M
segments), 1920x1080 camera, one directional light with a 2048^2 x 4-layer
deep shadow map, 8x8 froxels.  Inputs are resident in HBM before timing.

Prints one JSON line.  value = camera frames per second over K pipelined
frames (whole job, one GPU); Msegments/s, p50/p99 frame latency, per-stage
times, the roofline of the dominant kernel and the CPU baseline ride along,
and so do the two system patterns:
  plain_tuned -- the named frame at the autotuned configuration (froxel
             sizes, depth-warp slices, launch shape), replayed from a graph;
  batched -- many views and lights in one frame replayed from a HIP graph
             with the autotuned configuration: the 2,000 x 16 scalp from 64
             cameras over {720p, 1080p, 1440p} with 4 lights of 2048^2, and
             the named asset from 8 and from 16 1080p cameras sharing its
             light;
  serving -- the request loop over 8 resident assets (10k..500k strands x
             32 points) under a Poisson stream, p50/p99 latency and frames/s;
  stress  -- 1M strands x 64 points from 16 4K views with 4 lights of
             2048^2 x 8 layers, in frames that share one build of the lights.
The CPU baseline is the oracle (the C++ restatement) on the host cores over
full frames of the named workload and of the 2,000 x 16 scalp.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def percentile(samples, f):
    """Nearest rank on sorted samples, index round((n-1) f) rounded half away
    from zero, as the reference bench's summarize
    (examples/strand_bench.rs:931-947)."""
    s = sorted(samples)
    if not s:
        return None
    return s[int(math.floor((len(s) - 1) * f + 0.5))]


def refs_by_kind(r, w):
    """Fine references of camera and of light frusta: page reference counts
    summed over each frustum's coarse tiles (page_meta dump)."""
    import numpy as np

    pm = r.dump("page_meta")[: int(r.dump("counters")[0]["allocated_pages"])]
    out = {0: 0, 1: 0}
    for f in w.frusta:
        lo = int(f["coarse_depth_tile_base"])
        sel = (pm["coarse_tile_id"] >= lo) & (pm["coarse_tile_id"] < lo + int(f["coarse_depth_tile_count"]))
        out[int(f["kind"])] += int(pm["seg_ref_count"][sel].astype(np.int64).sum())
    return out


def stage_bytes(ctr, w, refs=None):
    """Algorithmic bytes per stage: every ABI record read or written once.
    T tasks (valid slots), C = 256 P cells, H references, N sorted hits,
    W work items, R runs.  Projection and marking: per valid slot its
    float4 vertex and index, the 12-byte projected segment, the two float4
    raw screen endpoints and the hit count; per hit its 4-byte cell record.
    Raster stages: per reference its 12-byte record and the segment's two
    float4 screen endpoints (+ two atlas texels and a history texel for the
    camera), plus the framebuffers written once."""
    T, P, H = ctr["binning_tasks"], ctr["allocated_pages"], ctr["fine_refs"]
    C = 256 * P
    W = ctr["work_items"]
    R = ctr["camera_runs"] + ctr["shadow_runs"]
    tiles16 = 16 * ctr["coarse_tiles"]
    N = ctr.get("page_candidates", 0)
    ptab = 16 * ctr["coarse_tiles"]  # coarse tiles x depth slices
    page_cap = min(ptab, max(4096, ptab // 2))  # the runtime's first-frame page capacity
    passes = (min(max(page_cap.bit_length(), 1), 24) + 7) // 8  # radix passes over the page bits
    out = {
        "project_mark": T * (16 + 4 + 12 + 32 + 4) + 4 * N,
        "count": 28 * T + 4 * C,
        # page CSR: (key, slot) pairs; each radix pass over the page bits
        # reads keys for the digit histogram, key + slot for the scatter and
        # writes key + slot
        "sort_hits": passes * 20 * N,
        # per page: keys for the cell histogram, key + slot + the slot's
        # material for the stable cell sort, the 12-byte references, cell
        # counts / offsets, page totals (the page's bounds are a few probes)
        "page_cells": 4 * N + 12 * N + 12 * N + 8 * C + 8 * P,
        "prefix": 8 * C + 32 * P,
        "fill": 28 * T + 4 * C + 12 * H,
        "sort_cells": 24 * H,
        "compact": 8 * tiles16,
        "emit": 4 * C + 32 * W + 32 * R,
        # per camera task: task record, four vertex words, meta, atlas and
        # history texels written
        "shading": ctr["camera_tasks"] * (16 + 2 * 4 + 2 * 16 + 16 + 16 + 8),
    }
    if refs is not None:
        L = int(w.settings["dom_layers"])
        light_px = sum(int(f["screen_width"]) * int(f["screen_height"]) for f in w.frusta if f["kind"] == 1)
        cam_px = sum(int(f["screen_width"]) * int(f["screen_height"]) for f in w.frusta if f["kind"] == 0)
        out["shadow"] = refs[1] * (12 + 32) + light_px * 4 * (1 + L)
        out["raster"] = refs[0] * (12 + 32 + 32 + 8) + cam_px * (16 + 4)
    return out


# Prepass stages that bin (marking through the last page pass) per backend,
# as named by the library's stage timer.
BINNING_STAGES = {
    "csr": ("project_mark", "project_compact", "allocate", "hit_offsets", "project_write", "sort_hits",
            "page_cells", "page_bases"),
    "scatter": ("project_mark", "project_compact", "allocate", "project_write", "count", "prefix", "fill",
                "sort_cells"),
}


def binning_headline(stage_acc, ctr, backend: str):
    """Count + scan + fill algorithmic bytes (every ABI record read or written
    once: count 28 T + 4 C, scan 8 C + 32 P, fill 28 T + 4 C + 12 H, with T
    binning tasks, P pages, C = 256 P cells, H fine references) over the summed
    time of every binning stage of the backend, marking and writing included,
    against the 8 TB/s HBM peak; and the same with the projection's bytes
    (per task a float4 vertex, an index, the projected segment and the two raw
    screen endpoints, fused into the marking and write passes) added."""
    T, P, H = ctr["binning_tasks"], ctr["allocated_pages"], ctr["fine_refs"]
    C = 256 * P
    count, scan, fill = 28 * T + 4 * C, 8 * C + 32 * P, 28 * T + 4 * C + 12 * H
    proj = T * (16 + 4 + 12 + 32)
    stages = [k for k in BINNING_STAGES[backend] if k in stage_acc]
    ms = sum(stage_acc[k] for k in stages)
    b = count + scan + fill
    gbs = b / (ms * 1e-3) / 1e9
    gbs_p = (b + proj) / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_frame": int(b),
            "bytes": {"count": int(count), "scan": int(scan), "fill": int(fill)}, "ms": round(ms, 4),
            "stages": {k: round(stage_acc[k], 4) for k in stages},
            "with_projection": {"bytes_per_frame": int(b + proj), "achieved": round(gbs_p, 1),
                                "frac": round(gbs_p / HBM_PEAK_GBS, 4)}}


def profile_row(kernel: str):
    """(HBM bytes, VALU issue, lane use) per launch of `kernel` from the
    committed rocprofv3 summary of the named workload, or (None, None, None).
    VALU issue is non-transcendental wave64 VALU instructions against one per
    4 cycles per SIMD at the measured clock (0.95 is what a pure FMA stream
    reaches); lane use is the fraction of a wave's lanes active per VALU
    instruction."""
    f = Path(__file__).resolve().parent / "profiles" / "named_1080p_csr_kernels.md"
    if not f.exists():
        return None, None, None
    for line in f.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 10 and cells[0].strip("`").startswith(kernel):
            try:
                return int((float(cells[3]) + float(cells[4])) * 1e6), float(cells[7]), float(cells[9])
            except ValueError:
                return None, None, None
    return None, None, None


STAGE_KERNEL = {"shadow": "k_shadow_raster", "raster": "k_camera_raster", "shading": "k_shading",
                "project_mark": "k_project_mark", "project_write": "k_project_write"}
STAGE_WORK = {"shadow": "exact per-pixel coverage arithmetic", "raster": "exact per-pixel coverage arithmetic",
              "shading": "hair BSDF per camera segment",
              "project_mark": "exact projection, clipping and cell walks with correctly rounded divisions"}


def roofline_for(stage, ms, ctr, w, refs=None):
    b = stage_bytes(ctr, w, refs).get(stage)
    if b is None or ms <= 0:
        return None
    gbs = b / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": stage,
           "bytes_per_launch": int(b), "avg_ms": round(ms, 4)}
    kern = STAGE_KERNEL.get(stage)
    if kern:
        traffic, valu, lanes = profile_row(kern)
        out["kernel"] = kern
        out["traffic"] = traffic
        if valu is not None:
            out["valu_issue"] = valu
            out["valu_lane_use"] = lanes
            if valu > 0.6:
                out["note"] = (f"VALU-issue bound ({STAGE_WORK.get(stage, 'exact arithmetic')}): {valu:.2f} of one "
                               "wave64 instruction per 4 cycles per SIMD, against 0.95 for a pure FMA stream, "
                               f"with {lanes:.2f} of the lanes active; traffic and VALU figures from profiles/")
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_frames(w, threads: int, warmup: int, samples: int) -> dict:
    """Whole frames of workload `w` through the oracle (the C++ restatement,
    `threads` worker threads): per-frame wall times in the reference bench's
    summary (examples/strand_bench.rs:931-947) after `warmup` untimed frames."""
    from bevy_strand_rasterizer_amd import report
    from oracle import oracle

    for _ in range(warmup):
        oracle.render(w, threads=threads).close()
    ms = []
    for _ in range(samples):
        t0 = time.perf_counter()
        f = oracle.render(w, threads=threads)
        ms.append((time.perf_counter() - t0) * 1e3)
        f.close()
    st = report.summarize(ms)
    return {"workload": w.name, "segments": w.input_segments, "views": w.camera_count,
            "warmup_frames": warmup, "sample_frames": samples, "frame_ms": {k: round(v, 3) if isinstance(v, float)
                                                                            else v for k, v in st.items()},
            "frames_per_s": round(w.camera_count * 1e3 / st["median"], 5),
            "msegments_per_s": round(w.input_segments * w.camera_count / st["median"] / 1e3, 4)}


def cpu_baseline(threads: int, named_samples: int = 3, small_samples: int = 20):
    """The oracle (C++ restatement of the reference's pipeline) on the box's
    host cores: the full named frame (1 warmup + `named_samples` timed frames)
    and the small configuration (2,000 x 16 at 512^2 with its 512^2 light,
    3 warmup + `small_samples` timed frames).  `value` is the named frame's
    frames/s from the median frame time."""
    from bevy_strand_rasterizer_amd import scene

    named = cpu_frames(scene.config_named(), threads, 1, named_samples)
    small = cpu_frames(scene.config_small(froxel=16), threads, 3, small_samples)
    return {"value": named["frames_per_s"], "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{named_samples} full frames of the named workload after 1 warmup frame (median), "
                      f"and {small_samples} frames of the 2,000 x 16 scalp after 3",
            "cpu": cpu_model(), "named": named, "small": small}


def percentile_ms(xs):
    return {"p50": round(percentile(xs, 0.5), 4), "p99": round(percentile(xs, 0.99), 4)}


def time_frames(r, w, steps: int, warmup: int):
    """Latency (one frame at a time) and throughput (`steps` pipelined
    frames) of workload `w` on renderer `r`."""
    import torch

    for _ in range(max(1, warmup)):
        r.submit(w)
    r.wait()
    lat = []
    for _ in range(steps):
        t0 = time.perf_counter()
        r.submit(w)
        r.wait()
        lat.append((time.perf_counter() - t0) * 1e3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.submit(w)
    ctr = r.wait()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return ctr, lat, dt


def write_report(args, r, w):
    """The reference bench's report for this workload: per-stage GPU times and
    telemetry over sampled frames (examples/strand_bench.rs:859-909)."""
    import copy

    import torch

    from bevy_strand_rasterizer_amd import report

    wt = copy.copy(w)
    wt.settings = dict(w.settings, telemetry_enabled=1)
    r.set_profiling(True)
    for _ in range(2):
        r.render(wt)
    stages, tele = {}, []
    for _ in range(args.report_frames):
        r.render(wt)
        for k, v in r.stage_times().items():
            stages.setdefault(k, []).append(v)
        tele.append(report.telemetry_samples(r.dump("telemetry"), len(wt.frusta)))
    r.set_profiling(False)
    cam = wt.frusta[0]
    a = wt.assets[0]
    rep = report.make_report(
        wt.name, int(cam["screen_width"]), int(cam["screen_height"]), 2, args.report_frames, True, args.backend,
        0x5EEDCAFE,
        {"geometry_count": len(wt.assets), "strand_count": int(a.strand_count), "vertex_count": int(len(a.vertices)),
         "clusters": None, "strands_per_cluster": None, "points_per_strand": int(a.meta[0]["count"]),
         "asset": "synthetic scalp"},
        {"name": torch.cuda.get_device_name(0), "backend": "HIP", "driver": torch.version.hip, "driver_info": ""},
        stages, tele)
    with open(args.report, "w") as f:
        json.dump(rep, f, indent=1)


def plain_views_fps(w, repeats: int = 2) -> float:
    """Frames/s of the plain per-frame path over the same views as the batched
    frame `w`: every camera drawn alone in a frame with all of w's lights
    (which that frame renders again), no graph, w's own froxel sizes and
    binning backend."""
    import torch

    from bevy_strand_rasterizer_amd import scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    r = StrandRenderer(device=0)
    r.upload_workload(w)
    ones = [scene.select_views(w, [i]) for i in range(len(w.frusta)) if int(w.frusta[i]["kind"]) == 0]
    for one in ones:  # each shape once, so buffers have grown before timing
        r.render(one)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        for one in ones:
            r.submit(one)
    r.wait()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    r.close()
    return len(ones) * repeats / dt


def run_batched(w, steps: int, warmup: int, plain_fps: float | None, tune: bool = True):
    """One multi-view frame per step, replayed from a HIP graph with the
    autotuned backend and froxel sizes."""
    from bevy_strand_rasterizer_amd import abi, scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    r = StrandRenderer(device=0, flags=abi.CONFIG_HIP_GRAPHS)
    r.upload_workload(w)
    best, trials = None, []
    if tune:
        res = r.autotune(w, frames_per_trial=2, min_psnr_db=50.0)
        best, trials = res["best"], res["trials"]
        w = scene.apply_tuning(w, best)
    ctr, lat, dt = time_frames(r, w, steps, warmup)
    r.close()
    views = w.camera_count
    fps = views * steps / dt
    out = {"workload": w.name, "views": views, "lights": int(len(w.lights)),
           "resolutions": sorted({f"{int(f['screen_width'])}x{int(f['screen_height'])}"
                                  for f in w.frusta if f["kind"] == 0}),
           "frames_per_s": round(fps, 3), "msegments_per_s": round(w.input_segments * views * steps / dt / 1e6, 2),
           "ms_per_step": round(dt * 1e3 / steps, 4), "latency_ms": percentile_ms(lat),
           "tuned": best, "trials": trials,
           "counters": {k: v for k, v in ctr.items() if not k.startswith("pad")}}
    if plain_fps:
        out["vs_plain"] = round(fps / plain_fps, 3)
    return out


def run_stress(views: int, views_per_frame: int, first_frame_views: int, backend: int, repeats: int = 2,
               tune: bool = True):
    """The stress configuration as batches of `views` 4K views: the first
    frame of a batch draws `first_frame_views` cameras and builds the 4
    lights' 8-layer shadow maps, every later frame draws `views_per_frame`
    cameras with those maps (SR_FRAME_REUSE_LIGHT_DOMS).  As every batched
    configuration, the froxel sizes and launch shape come from the autotuner
    (run on the light-building frame, keeping the binning backend: a segment
    scatter trial of a 1.85G-reference frame is not worth its time), applied
    to every frame of the batch.  Every frame shape first renders until its
    capacities fit (the runtime grows them after a frame reports
    capacity_exceeded), then `repeats` batches are timed, each building its
    lights again."""
    import torch

    from bevy_strand_rasterizer_amd import abi, scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    ws = [scene.retile(w, 0, 0, backend)
          for w in scene.config_stress_chunks(views=views, views_per_frame=views_per_frame,
                                              first_frame_views=first_frame_views)]
    r = StrandRenderer(device=0)
    r.upload(ws[0].assets[0])
    best, trials = None, []
    if tune:
        first = ws[0]
        saved = int(first.settings.get("frame_flags", 0))
        first.settings["frame_flags"] = saved | abi.FRAME_TUNE_KEEP_BACKEND
        res = r.autotune(first, frames_per_trial=1, min_psnr_db=50.0)
        first.settings["frame_flags"] = saved
        best, trials = res["best"], res["trials"]
        ws = [scene.apply_tuning(w, best) for w in ws]
        print(f"bench: stress tuned {best}", file=sys.stderr, flush=True)
    grow = 0
    per_frame = []
    for w in ws:
        for _ in range(6):
            ctr = r.render(w)
            print(f"bench: stress warm {w.name}: refs {ctr['fine_refs']} / {ctr['ref_capacity']}, "
                  f"exceeded {ctr['capacity_exceeded']}", file=sys.stderr, flush=True)
            if not ctr["capacity_exceeded"]:
                break
            grow += 1
        if ctr["capacity_exceeded"]:
            r.close()
            return {"error": "capacity not reached", "frame": w.name, "counters": ctr}
        per_frame.append({"frame": w.name, "cameras": w.camera_count, "fine_refs": ctr["fine_refs"],
                          "binning_tasks": ctr["binning_tasks"], "shadow_runs": ctr["shadow_runs"]})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    exceeded = 0
    for _ in range(repeats):
        for w in ws:
            r.submit(w)
            c = r.wait()
            exceeded += int(c["capacity_exceeded"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    r.close()
    segs = ws[0].input_segments
    return {"workload": f"stress-1000k-x64-4k-{views}views-4lights-8layers", "views": views,
            "frames_per_batch": len(ws), "views_per_frame": [w.camera_count for w in ws],
            "lights_built_per_batch": 1, "batches": repeats,
            "fine_binning_backend": ["scatter", "csr"][backend],
            "frames_per_s": round(views * repeats / dt, 3), "ms_per_view": round(dt * 1e3 / (views * repeats), 3),
            "ms_per_batch": round(dt * 1e3 / repeats, 2),
            "msegments_per_s": round(segs * views * repeats / dt / 1e6, 2), "warm_growths": grow,
            "frames_over_capacity": exceeded, "frames": per_frame, "tuned": best, "trials": trials}


class PinnedPool:
    """Reusable pinned host buffers for served outputs, per resolution: an
    rgba8unorm composite (H x W x 4 bytes) and an FP32 depth (H x W)."""

    def __init__(self):
        self.free = {}

    def take(self, w: int, h: int):
        import torch

        lst = self.free.setdefault((w, h), [])
        if lst:
            return lst.pop()
        return (torch.empty((h, w, 4), dtype=torch.uint8, pin_memory=True).numpy(),
                torch.empty((h, w), dtype=torch.float32, pin_memory=True).numpy())

    def give(self, out) -> None:
        rgba, depth = out
        self.free.setdefault((depth.shape[1], depth.shape[0]), []).append(out)


def run_virtual_doms(backend: int, steps: int = 10):
    """Dense against virtual deep shadow maps (SR_FRAME_VIRTUAL_DOMS, every
    requested page resident) on the named frame and on the 2,000-strand scalp
    under a 2048^2 light framed 2.5x wider than the asset: frames/s of both,
    the pages made resident and the DOM bytes each holds."""
    import dataclasses

    from bevy_strand_rasterizer_amd import abi, scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    out = {}
    for name, wv in (("named", scene.with_virtual_doms(scene.config_named())),
                     ("scalp_2k_wide_light", scene.config_virtual_doms())):
        wv.settings["fine_binning_backend"] = backend
        wd = dataclasses.replace(wv, settings={k: v for k, v in wv.settings.items()
                                               if k not in ("frame_flags", "dom_resident_pages")})
        L = int(wv.settings["dom_layers"])
        lights = [f for f in wv.frusta if int(f["kind"]) == 1]
        dense_px = sum(int(f["screen_width"]) * int(f["screen_height"]) for f in lights)
        rec = {"lights": len(lights), "light_pages": sum(-(-int(f["screen_width"]) // abi.DOM_PAGE_XY) *
                                                         -(-int(f["screen_height"]) // abi.DOM_PAGE_XY) for f in lights)}
        with StrandRenderer(device=0) as r:
            r.upload_workload(wv)
            for label, w in (("dense", wd), ("virtual", wv)):
                ctr, lat, dt = time_frames(r, w, steps, 3)
                rec[f"{label}_frames_per_s"] = round(w.camera_count * steps / dt, 3)
                if label == "virtual":
                    rec["resident_pages"] = int(ctr["dom_pages"])
        rec["dom_bytes_dense"] = 4 * dense_px * (1 + L)
        rec["dom_bytes_virtual"] = 4 * rec["resident_pages"] * abi.DOM_PAGE_XY ** 2 * (1 + L)
        out[name] = rec
    return out


def run_sorted_blend(backend: int, steps: int = 10):
    """The named frame with depth-sorted blending inside each work item
    (SR_FRAME_DEPTH_SORTED_BLEND) against the default (inst, seg) order:
    frames/s and the camera raster's stage time of each."""
    from bevy_strand_rasterizer_amd import abi, scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    w = scene.config_named()
    w.settings["fine_binning_backend"] = backend
    out = {"workload": w.name, "fragments_held_per_pixel": abi.BLEND_SORT_FRAGMENTS}
    with StrandRenderer(device=0) as r:
        r.upload_workload(w)
        for label, flags in (("default", 0), ("sorted", abi.FRAME_DEPTH_SORTED_BLEND)):
            w.settings["frame_flags"] = flags
            ctr, lat, dt = time_frames(r, w, steps, 3)
            out[f"{label}_frames_per_s"] = round(steps / dt, 3)
            r.set_profiling(True)
            r.render(w)
            out[f"{label}_raster_ms"] = round(r.stage_times().get("raster", 0.0), 4)
            r.set_profiling(False)
    return out


def run_serving(seconds: float, rate: float, slots: int, max_views: int, budget_gib: float, seed: int = 7,
                backend: int = 1):
    """Poisson arrivals at `rate` req/s for `seconds` over the 8 resident
    serving assets.  Every request's outputs are delivered to the host -- its
    composite in the reference's rgba8unorm framebuffer format and its FP32
    depth, copied by the frame straight into pinned buffers from a reusable
    pool -- and its latency is submit -> outputs in host memory on the
    server clock."""
    import numpy as np

    from bevy_strand_rasterizer_amd import scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer, StrandServer

    assets = scene.serving_assets()
    r = StrandRenderer(device=0)
    ids = [r.upload(a).geo_id for a in assets]
    rng = np.random.default_rng(seed)
    n = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = [scene.serving_request(assets, ids, rng, tag=i, backend=backend) for i in range(n)]
    segs = [assets[int(np.nonzero(np.array(ids) == int(q[0]["instance"]["geo_id"][0]))[0][0])].segment_count
            for q in reqs]
    size = [(int(q[0]["camera"]["screen_width"][0]), int(q[0]["camera"]["screen_height"][0])) for q in reqs]
    pool = PinnedPool()
    held = {}

    def submit(srv, k):
        out = pool.take(*size[k])
        t = srv.submit(reqs[k][0], rgba8=True, out=out, pinned=True)
        held[t] = out

    def collect(got):
        for x in got:  # the buffers go back to the pool; results keep timing only
            pool.give(held.pop(x["ticket"]))
            x["rgba"] = x["depth"] = None
        return got

    res = []
    with StrandServer(r, slots=slots, max_views=max_views, hbm_budget=int(budget_gib * (1 << 30))) as srv:
        for k in range(min(16, n)):  # warm: every slot grows its buffers once
            submit(srv, k)
        collect(srv.drain(min(16, n), timeout_s=300))
        t0 = time.perf_counter()
        for i in range(n):
            while time.perf_counter() - t0 < arrivals[i]:
                res += collect(srv.poll(wait_ms=0))
                time.sleep(0.0002)
            submit(srv, i)
        res += collect(srv.drain(n - len(res), timeout_s=600))
        st = srv.stats()
    r.close()
    ok = [x for x in res if x["status"] == 0]
    lat = [x["latency_ms"] for x in ok]
    span = (max(x["done_ms"] for x in ok) - min(x["submit_ms"] for x in ok)) / 1e3 if ok else 0.0
    return {"workload": "8 resident scalps {10k..500k} x 32, Poisson arrivals",
            "offered_req_per_s": rate, "seconds": seconds, "requests": n, "completed": len(ok),
            "failed": len(res) - len(ok), "frames_per_s": round(len(ok) / span, 3) if span else None,
            "msegments_per_s": round(sum(segs[x["tag"]] for x in ok) / span / 1e6, 2) if span else None,
            "latency_ms": percentile_ms(lat) if lat else None,
            "mean_views_per_frame": round(float(np.mean([x["frame_views"] for x in ok])), 3) if ok else None,
            "outputs": "rgba8unorm composite + FP32 depth per request, copied into pinned host buffers",
            "fine_binning_backend": ["scatter", "csr"][backend],
            "slots": slots, "max_views_per_frame": max_views, "hbm_budget_gib": budget_gib,
            "server": st, "resolutions": [f"{w}x{h}" for w, h in scene.SERVING_RESOLUTIONS]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="named", choices=["named", "small"])
    ap.add_argument("--backend", default="csr", choices=["scatter", "csr"],
                    help="fine binning backend (FineBinningBackend::SegmentScatter / PageCsr)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-virtual-doms", action="store_true")
    ap.add_argument("--no-sorted-blend", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-peak-rate", type=float, default=400.0,
                    help="a second serving run at this offered load, above what one GPU completes (0: none)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-max-views", type=int, default=16)
    ap.add_argument("--serve-budget-gib", type=float, default=96.0)
    ap.add_argument("--report", default="", help="also write the reference bench's report (schema_version 1) here, "
                                                  "from --report-frames frames with telemetry on")
    ap.add_argument("--report-frames", type=int, default=10)
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the stress configuration (1M x 64, 16 4K views, 4 lights x 8 layers)")
    ap.add_argument("--stress-views-per-frame", type=int, default=4)
    ap.add_argument("--stress-first-frame-views", type=int, default=2)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a plain copy kernel of 1 GiB read + 1 GiB written (torch.neg), the reference "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent frames and views shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch  # initialise the device through torch before the library does

    torch.cuda.init()

    from bevy_strand_rasterizer_amd import scene
    from bevy_strand_rasterizer_amd.runtime import StrandRenderer

    w = scene.config_named() if args.config == "named" else scene.config_small()
    w.settings["fine_binning_backend"] = 1 if args.backend == "csr" else 0
    views = w.camera_count
    segs = w.input_segments

    r = StrandRenderer(device=0)
    r.upload_workload(w)
    for _ in range(max(1, args.warmup)):
        r.submit(w)
    ctr = r.wait()
    print(f"bench: plain warm ({w.name})", file=sys.stderr, flush=True)

    # stage timing (HIP events on the library stream), a few frames
    r.set_profiling(True)
    stage_acc = {}
    nprof = 3
    for _ in range(nprof):
        r.submit(w)
        r.wait()
        for k, v in r.stage_times().items():
            stage_acc[k] = stage_acc.get(k, 0.0) + v / nprof
    r.set_profiling(False)

    ctr, lat, dt = time_frames(r, w, args.steps, 0)
    ms_per_step = dt * 1e3 / args.steps
    fps = views * args.steps / dt
    refs = refs_by_kind(r, w)
    if args.report:
        write_report(args, r, w)
    r.close()
    print(f"bench: plain {fps:.2f} frames/s", file=sys.stderr, flush=True)

    dom = max((k for k in stage_acc if k not in ("end",)), key=lambda k: stage_acc[k])
    roof = roofline_for(dom, stage_acc[dom], ctr, w, refs)
    binning = binning_headline(stage_acc, ctr, args.backend)
    # the sort kernels against the bytes they move themselves (radix passes,
    # per-page cell sort), beside the algorithmic headline above
    bstages = ("sort_hits", "page_cells") if args.backend == "csr" else ("count", "prefix", "fill")
    binning_kernels = {k: roofline_for(k, stage_acc.get(k, 0.0), ctr, w, refs) for k in bstages}
    if roof is None:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": dom, "avg_ms": round(stage_acc[dom], 4)}

    batched = plain_tuned = None
    if not args.no_batched:
        # the named plain frame at the autotuned configuration (froxel sizes,
        # depth-warp slices, launch shape; the image within 50 dB of the
        # reference configuration's): what the tuner alone buys, the fair
        # denominator for what batching adds
        wt = scene.config_named() if args.config == "named" else scene.config_small()
        wt.settings["fine_binning_backend"] = 1 if args.backend == "csr" else 0
        plain_tuned = run_batched(wt, steps=args.steps, warmup=2, plain_fps=fps)
        plain_tuned["plain"] = "the named plain frame at the reference configuration (value)"
        print(f"bench: plain tuned {plain_tuned['frames_per_s']} frames/s", file=sys.stderr, flush=True)
        batched = {}
        for name, wb in (("64views_4lights_2kx16", scene.config_batched()),
                         ("named_8views_1light", scene.config_batched_named(8)),
                         ("named_16views_1light", scene.config_batched_named(16))):
            wb.settings["fine_binning_backend"] = 1 if args.backend == "csr" else 0
            if name.startswith("named_"):  # each view alone is the named plain frame
                plain, plain_what = fps, "the named plain frame (one view, its light rendered per frame)"
            else:
                plain = plain_views_fps(wb)
                plain_what = "each view alone with the 4 lights, one plain frame per view"
            batched[name] = run_batched(wb, steps=max(3, args.steps // 4), warmup=2, plain_fps=plain)
            batched[name]["plain_frames_per_s"] = round(plain, 3)
            batched[name]["plain"] = plain_what
            if name.startswith("named_"):
                batched[name]["vs_plain_tuned"] = round(batched[name]["frames_per_s"] / plain_tuned["frames_per_s"], 3)
            print(f"bench: batched {name} {batched[name]['frames_per_s']} frames/s", file=sys.stderr, flush=True)

    vdoms = None
    if not args.no_virtual_doms:
        vdoms = run_virtual_doms(1 if args.backend == "csr" else 0)
        print(f"bench: virtual DOMs {vdoms}", file=sys.stderr, flush=True)

    sorted_blend = None
    if not args.no_sorted_blend:
        sorted_blend = run_sorted_blend(1 if args.backend == "csr" else 0)
        print(f"bench: depth-sorted blending {sorted_blend}", file=sys.stderr, flush=True)

    serving = None
    if not args.no_serving:
        bk = 1 if args.backend == "csr" else 0
        serving = run_serving(args.serve_seconds, args.serve_rate, args.serve_slots, args.serve_max_views,
                              args.serve_budget_gib, backend=bk)
        print(f"bench: serving {serving['frames_per_s']} frames/s", file=sys.stderr, flush=True)
        if args.serve_peak_rate > 0:  # an offered load above saturation: the most one GPU completes
            serving["saturated"] = run_serving(args.serve_seconds, args.serve_peak_rate, args.serve_slots,
                                               args.serve_max_views, args.serve_budget_gib, seed=11, backend=bk)
            print(f"bench: serving at {args.serve_peak_rate} req/s offered: "
                  f"{serving['saturated']['frames_per_s']} frames/s", file=sys.stderr, flush=True)

    stress = None
    if not args.no_stress:
        stress = run_stress(16, args.stress_views_per_frame, args.stress_first_frame_views,
                            1 if args.backend == "csr" else 0)
        print(f"bench: stress {stress.get('frames_per_s')} frames/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(threads)
        except Exception as e:  # the baseline never blocks the GPU line
            cpu = {"error": str(e)}

    out = {
        "metric": "frames/s + Msegments/s (100k x 32 strands, 1080p), batched vs plain; p99 latency",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scalp generator)",
        "config": {"workload": w.name, "mode": "plain", "strands": int(w.assets[0].strand_count),
                   "points_per_strand": int(w.assets[0].meta[0]["count"]), "segments": segs, "views": views,
                   "lights": int(len(w.lights)), "dom_layers": int(w.settings["dom_layers"]),
                   "froxel": int(w.frusta[0]["froxel_size_x"]), "fine_binning_backend": args.backend},
        "msegments_per_s": round(segs * views * args.steps / dt / 1e6, 2),
        "latency_ms": percentile_ms(lat),
        "stage_ms": {k: round(v, 4) for k, v in stage_acc.items()},
        "counters": {k: v for k, v in ctr.items() if not k.startswith("pad")},
        "refs_by_kind": {"camera": refs[0], "light": refs[1]},
        "roofline": roof,
        "binning_roofline": binning,
        "binning_kernel_traffic": binning_kernels,
        "cpu_baseline": cpu,
        "plain_tuned": plain_tuned,
        "batched": batched,
        "virtual_doms": vdoms,
        "depth_sorted_blend": sorted_blend,
        "serving": serving,
        "stress": stress,
    }
    print(json.dumps(out))
    if args.calibrate_copy:
        x = torch.ones(1 << 28, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.neg(x, out=y)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
