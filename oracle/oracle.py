"""ctypes wrapper of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (as the
reference result) and bench.py's cpu_baseline leg.  The shipped library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from bevy_strand_rasterizer_amd import abi

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        from oracle.build import build_oracle

        build_oracle()
    lib = C.CDLL(str(LIB_PATH))
    lib.oracle_render.restype = C.c_int
    lib.oracle_render.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(abi.FrameDesc), C.c_int, C.c_uint32, C.c_int,
                                  C.POINTER(C.c_void_p)]
    lib.oracle_render_history.restype = C.c_int
    lib.oracle_render_history.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(abi.FrameDesc), C.c_int, C.c_uint32,
                                          C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]
    lib.oracle_get.restype = C.c_int
    lib.oracle_get.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.oracle_stage_times.restype = C.c_int
    lib.oracle_stage_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_char_p, C.c_int]
    lib.oracle_free.restype = None
    lib.oracle_free.argtypes = [C.c_void_p]
    for name, res, args in [
        ("oracle_wang_hash", C.c_uint32, [C.c_uint32]),
        ("oracle_hash_to_unit_float", C.c_float, [C.c_uint32]),
        ("oracle_pack2x16unorm", C.c_uint32, [C.c_float, C.c_float]),
        ("oracle_quantize_depth01", C.c_uint32, [C.c_float]),
        ("oracle_normalize_depth01", C.c_float, [C.c_float]),
        ("oracle_build_lut", None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
        ("oracle_clip_segment", C.c_int, [C.c_void_p] * 5),
        ("oracle_gaussian_detector", C.c_float, [C.c_float, C.c_float, C.c_uint32]),
        ("oracle_fresnel", C.c_float, [C.c_float, C.c_float, C.c_float]),
        ("oracle_bessel", C.c_float, [C.c_float, C.c_uint32]),
        ("oracle_mp", C.c_float, [C.c_float] * 4),
        ("oracle_np", None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_float,
                             C.c_void_p]),
        ("oracle_phi", C.c_float, [C.c_uint32, C.c_float, C.c_float]),
        ("oracle_a", None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_float, C.c_void_p]),
        ("oracle_marschner", None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def gpu_ref_capacity(w) -> int:
    """The reference-slab capacity the runtime uses for a workload's first
    frame (8 x task capacity, at least 65536), so overflow rules agree."""
    pos = sum(int(w.assets[int(i["index_id"])].indices.shape[0]) for i in w.instances)
    slots = max(64, (pos + 63) // 64 * 64)
    task_cap = len(w.frusta) * slots
    return max(8 * task_cap, 1 << 16)


class OracleFrame:
    def __init__(self, lib, handle, keep):
        self._lib, self._h, self._keep = lib, handle, keep

    def get(self, name: str) -> np.ndarray:
        data, nbytes = C.c_void_p(), C.c_uint64()
        rc = self._lib.oracle_get(self._h, abi.DUMP[name], C.byref(data), C.byref(nbytes))
        if rc != 0:
            raise KeyError(name)
        dt = abi.DUMP_DTYPES[name]
        n = nbytes.value // dt.itemsize
        if n == 0:
            return np.zeros(0, dt)
        buf = (C.c_char * nbytes.value).from_address(data.value)
        return np.frombuffer(bytes(buf), dtype=dt, count=n)

    def counters(self) -> dict:
        c = self.get("counters")[0]
        return {k: int(c[k]) for k in c.dtype.names}

    def stage_times(self) -> dict:
        ms = (C.c_double * 64)()
        names = C.create_string_buffer(4096)
        n = self._lib.oracle_stage_times(self._h, ms, 64, names, 4096)
        keys = names.value.decode().split("\n")[:n]
        return {k: ms[i] for i, k in enumerate(keys)}

    def close(self):
        if self._h:
            self._lib.oracle_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(workload, threads: int | None = None, ref_capacity: int | None = None,
           binning_only: bool = False, prev_history: np.ndarray | None = None,
           history_valid: bool | None = None) -> OracleFrame:
    """One frame; `prev_history` (the previous frame's "history" dump, float2
    per atlas texel) is read as shadow_history_prev when history_valid."""
    lib = load()
    slabs = (abi.AssetSlabs * len(workload.assets))(*[a.slabs() for a in workload.assets])
    fd = workload.frame_desc()
    if history_valid is not None:
        fd.settings.shadow_history_valid = int(history_valid)
    out = C.c_void_p()
    cap = ref_capacity if ref_capacity is not None else gpu_ref_capacity(workload)
    nthreads = threads if threads is not None else min(16, os.cpu_count() or 1)
    hist = None if prev_history is None else np.ascontiguousarray(prev_history, np.float32)
    rc = lib.oracle_render_history(slabs, len(workload.assets), C.byref(fd), nthreads, cap, int(binning_only),
                                   None if hist is None else hist.ctypes.data,
                                   0 if hist is None else hist.size // 2, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle_render failed ({rc})")
    return OracleFrame(lib, out.value, (slabs, fd, workload, hist))
