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
