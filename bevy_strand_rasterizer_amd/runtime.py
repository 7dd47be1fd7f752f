"""Python front end of the native strand runtime (libstrand_rt.so).

``StrandRenderer`` owns one ``sr_context`` on one device: it uploads assets
into device paging slabs, submits frames and reads back dumps.  There is no
CPU fallback: constructing a renderer without the built library or without a
HIP device raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class StrandRenderer:
    def __init__(self, device: int = 0, slab_bytes: int = 0, hbm_budget: int = 0, profiling: bool = False,
                 flags: int = 0):
        self.lib = abi.load_library()
        cfg = abi.DeviceConfig()
        cfg.device_index = device
        cfg.flags = flags
        cfg.slab_bytes = slab_bytes
        cfg.hbm_budget_bytes = hbm_budget
        h = C.c_void_p()
        abi.check(self.lib, self.lib.sr_create(C.byref(cfg), C.byref(h)), None, "sr_create")
        self.ctx = h
        self._assets = {}
        self._keep = None
        if profiling:
            self.set_profiling(True)

    # ------------------------------------------------------------ assets
    def upload(self, asset) -> abi.AssetInfo:
        info = abi.AssetInfo()
        slabs = asset.slabs()
        abi.check(self.lib, self.lib.sr_asset_upload(self.ctx, C.byref(slabs), C.byref(info)), self.ctx,
                  "sr_asset_upload")
        self._assets[info.geo_id] = info
        return info

    def release(self, asset_id: int) -> None:
        abi.check(self.lib, self.lib.sr_asset_release(self.ctx, asset_id), self.ctx, "sr_asset_release")
        self._assets.pop(asset_id, None)

    def upload_workload(self, workload) -> list:
        """Uploads every asset of a workload; asset i must receive id i."""
        infos = [self.upload(a) for a in workload.assets]
        for i, info in enumerate(infos):
            if info.geo_id != i:
                raise RuntimeError("workload assets must be uploaded into a fresh renderer")
        return infos

    # ------------------------------------------------------------ frames
    def submit(self, workload, stream: int | None = None, history_valid: bool | None = None) -> None:
        fd = workload.frame_desc()
        if history_valid is not None:
            fd.settings.shadow_history_valid = int(history_valid)
        self._keep = (fd, workload)
        abi.check(self.lib, self.lib.sr_frame_submit(self.ctx, C.byref(fd), stream), self.ctx, "sr_frame_submit")

    def wait(self) -> dict:
        c = np.zeros(1, abi.frame_counters_dtype)
        abi.check(self.lib, self.lib.sr_frame_wait(self.ctx, abi.ptr(c)), self.ctx, "sr_frame_wait")
        return {k: int(c[0][k]) for k in c.dtype.names}

    def render(self, workload) -> dict:
        self.submit(workload)
        return self.wait()

    def dump(self, name: str) -> np.ndarray:
        kind = abi.DUMP[name]
        nbytes = C.c_uint64()
        abi.check(self.lib, self.lib.sr_frame_dump(self.ctx, kind, None, 0, C.byref(nbytes)), self.ctx,
                  "sr_frame_dump")
        dt = abi.DUMP_DTYPES[name]
        out = np.zeros(nbytes.value // dt.itemsize, dt)
        if nbytes.value:
            abi.check(self.lib, self.lib.sr_frame_dump(self.ctx, kind, abi.ptr(out), nbytes.value, C.byref(nbytes)),
                      self.ctx, "sr_frame_dump")
        return out

    def autotune(self, workload, frames_per_trial: int = 3, min_psnr_db: float = 50.0) -> dict:
        """sr_autotune on this workload's frame: every trial and the best."""
        fd = workload.frame_desc()
        res = np.zeros(1, abi.tune_result_dtype)
        abi.check(self.lib, self.lib.sr_autotune(self.ctx, C.byref(fd), frames_per_trial, min_psnr_db,
                                                 abi.ptr(res)), self.ctx, "sr_autotune")
        names = abi.tune_trial_dtype.names
        trials = [{k: (float(t[k]) if k in ("ms_per_frame", "psnr_db") else int(t[k])) for k in names if k != "pad"}
                  for t in res[0]["trials"][: int(res[0]["trial_count"])]]
        best = {k: (float(res[0]["best"][k]) if k in ("ms_per_frame", "psnr_db") else int(res[0]["best"][k]))
                for k in names if k != "pad"}
        return {"best": best, "trials": trials}

    def set_profiling(self, enabled: bool) -> None:
        abi.check(self.lib, self.lib.sr_set_profiling(self.ctx, int(enabled)), self.ctx, "sr_set_profiling")

    def stage_times(self) -> dict:
        ms = (C.c_float * 128)()
        n = C.c_int32()
        names = C.create_string_buffer(8192)
        abi.check(self.lib, self.lib.sr_stage_times(self.ctx, ms, 128, C.byref(n), names, 8192), self.ctx,
                  "sr_stage_times")
        keys = names.value.decode().split("\n")[: n.value]
        out = {}
        for i, k in enumerate(keys):
            out[k] = out.get(k, 0.0) + ms[i]
        return out

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.sr_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StrandServer:
    """The native serving loop (sr_server_*) over a renderer's resident assets.

    ``submit`` copies a request record (``scene.request_from_workload``) and
    returns a ticket; the composite image (float32 H x W x 4) and depth
    (H x W) land in arrays owned here and come back from ``poll`` with the
    request's timing and status."""

    def __init__(self, renderer: StrandRenderer, slots: int = 2, max_views: int = 16, hbm_budget: int = 0):
        self.r = renderer
        self.lib = renderer.lib
        cfg = np.zeros(1, abi.server_config_dtype)
        cfg["slots"], cfg["max_views_per_frame"], cfg["hbm_budget_bytes"] = slots, max_views, hbm_budget
        h = C.c_void_p()
        abi.check(self.lib, self.lib.sr_server_start(renderer.ctx, abi.ptr(cfg), C.byref(h)), renderer.ctx,
                  "sr_server_start")
        self.srv = h
        self._pending = {}

    def submit(self, request: np.ndarray, outputs: bool = True, rgba8: bool = False, out=None,
               pinned: bool = False) -> int:
        """Queues one request.  outputs=False asks for none; otherwise the
        composite (float32 H x W x 4, or with rgba8 uint8 H x W x 4 in the
        reference's rgba8unorm format) and depth (float32 H x W) land in
        arrays allocated here, or in `out` = (rgba, depth) given by the
        caller; pinned=True says those are pinned host memory (the frame
        copies straight into them)."""
        req = np.array(request, dtype=abi.request_dtype).reshape(1)
        w, h = int(req["camera"]["screen_width"][0]), int(req["camera"]["screen_height"][0])
        if out is not None:
            rgba, depth = out
            assert rgba.shape == (h, w, 4) and rgba.dtype == (np.uint8 if rgba8 else np.float32)
            assert depth.shape == (h, w) and depth.dtype == np.float32
            assert rgba.flags["C_CONTIGUOUS"] and depth.flags["C_CONTIGUOUS"]
        elif outputs:
            rgba = np.empty((h, w, 4), np.uint8 if rgba8 else np.float32)
            depth = np.empty((h, w), np.float32)
        else:
            rgba = depth = None
        req["flags"] = (abi.REQUEST_RGBA8 if rgba8 else 0) | (abi.REQUEST_PINNED_OUTPUTS if pinned and rgba is not None
                                                              else 0)
        req["out_rgba"] = rgba.ctypes.data if rgba is not None else 0
        req["out_depth"] = depth.ctypes.data if depth is not None else 0
        t = C.c_uint64()
        abi.check(self.lib, self.lib.sr_server_submit(self.srv, abi.ptr(req), C.byref(t)), self.r.ctx,
                  "sr_server_submit")
        self._pending[t.value] = (rgba, depth)
        return t.value

    def poll(self, wait_ms: int = 0, capacity: int = 256) -> list:
        res = np.zeros(capacity, abi.request_result_dtype)
        n = C.c_uint32()
        abi.check(self.lib, self.lib.sr_server_poll(self.srv, abi.ptr(res), capacity, C.byref(n), wait_ms),
                  self.r.ctx, "sr_server_poll")
        out = []
        for rec in res[: n.value]:
            rgba, depth = self._pending.pop(int(rec["ticket"]), (None, None))
            out.append({"ticket": int(rec["ticket"]), "tag": int(rec["tag"]), "status": int(rec["status"]),
                        "frame_views": int(rec["frame_views"]), "submit_ms": float(rec["submit_ms"]),
                        "launch_ms": float(rec["launch_ms"]), "done_ms": float(rec["done_ms"]),
                        "latency_ms": float(rec["done_ms"] - rec["submit_ms"]),
                        "counters": {k: int(rec["counters"][k]) for k in abi.frame_counters_dtype.names},
                        "rgba": rgba, "depth": depth})
        return out

    def drain(self, count: int, timeout_s: float = 120.0) -> list:
        """Polls until `count` results have arrived."""
        import time

        got, t0 = [], time.monotonic()
        while len(got) < count:
            if time.monotonic() - t0 > timeout_s:
                raise TimeoutError(f"{len(got)} of {count} results after {timeout_s} s")
            got += self.poll(wait_ms=50)
        return got

    def stats(self) -> dict:
        st = np.zeros(1, abi.server_stats_dtype)
        abi.check(self.lib, self.lib.sr_server_stats_get(self.srv, abi.ptr(st)), self.r.ctx, "sr_server_stats_get")
        return {k: int(st[0][k]) for k in abi.server_stats_dtype.names}

    def stop(self) -> None:
        if getattr(self, "srv", None):
            self.lib.sr_server_stop(self.srv)
            self.srv = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
