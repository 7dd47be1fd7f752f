"""Failure handling at the C ABI: every rejected asset, frame, wait, dump or
release returns its status without touching the device, and the context
stays usable: a valid frame afterwards matches the oracle.  (The checks are
host-side; a context needs a device, so the tests are marked gpu.)"""
import copy
import dataclasses

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from oracle import oracle

from test_gpu_parity import EXACT, FLOAT, compare

pytestmark = pytest.mark.gpu


def status_of(fn):
    """The sr_status name a failing call reports through abi.check."""
    with pytest.raises(RuntimeError) as e:
        fn()
    return str(e.value).split(":")[1].strip()


def bad_asset(kind: str):
    a = copy.deepcopy(scene.config_tiny().assets[0])
    if kind == "index-past-vertices":
        a.indices = a.indices.copy()
        a.indices[5] = a.vertices.shape[0]
    elif kind == "material-out-of-range":
        a.meta = a.meta.copy()
        a.meta[3]["material_idx"] = a.materials.shape[0]
    elif kind == "strand-past-indices":
        a.meta = a.meta.copy()
        a.meta[-1]["count"] += 4
    return a


@pytest.mark.parametrize("kind", ["index-past-vertices", "material-out-of-range", "strand-past-indices"])
def test_invalid_assets_are_rejected(kind):
    with StrandRenderer() as r:
        assert status_of(lambda: r.upload(bad_asset(kind))) == "SR_ERR_INVALID_ASSET"
        assert status_of(lambda: r.release(7)) == "SR_ERR_INVALID_ARGUMENT"


def invalid_frames(w):
    """(name, workload, expected status) for frames sr_frame_submit refuses."""
    out = []

    def with_settings(name, status, **kv):
        t = dataclasses.replace(w, settings={**w.settings, **kv})
        out.append((name, t, status))

    with_settings("unknown-frame-flag", "SR_ERR_INVALID_ARGUMENT", frame_flags=1 << 10)
    with_settings("zero-dom-layers", "SR_ERR_INVALID_ARGUMENT", dom_layers=0)
    with_settings("17-dom-layers", "SR_ERR_INVALID_ARGUMENT", dom_layers=17)
    with_settings("unknown-backend", "SR_ERR_INVALID_ARGUMENT", fine_binning_backend=2)
    with_settings("3-marking-waves", "SR_ERR_INVALID_ARGUMENT", mark_waves=3)
    with_settings("33-raster-waves", "SR_ERR_INVALID_ARGUMENT", raster_waves=33)
    with_settings("12-depth-slices", "SR_ERR_INVALID_ARGUMENT", coarse_depth_slices=12)
    with_settings("64-depth-slices", "SR_ERR_INVALID_ARGUMENT", coarse_depth_slices=64)
    big = scene.retile(w, 32, 0)  # 32 x 32 px raster tiles: 1024 pixels
    out.append(("froxel-tiles-above-256-px", big, "SR_ERR_UNSUPPORTED"))
    inst = w.instances.copy()
    for k in ("vertex_id", "index_id", "meta_id", "geo_id", "material_id"):
        inst[k] = 9
    out.append(("unknown-asset", dataclasses.replace(w, instances=inst), "SR_ERR_INVALID_ARGUMENT"))
    fr = w.frusta.copy()
    fr["screen_width"][0] = 0
    out.append(("zero-extent", dataclasses.replace(w, frusta=fr), "SR_ERR_INVALID_ARGUMENT"))
    return out


def test_rejected_frames_leave_the_context_usable():
    w = scene.config_tiny()
    with StrandRenderer() as r:
        r.upload_workload(w)
        assert status_of(lambda: r.wait()) == "SR_ERR_NOT_READY"  # nothing submitted yet
        for name, bad, status in invalid_frames(w):
            print(f"[errors] {name}", flush=True)
            assert status_of(lambda: r.submit(bad)) == status, name
        ctr = r.render(w)
        gpu = {k: r.dump(k) for k in EXACT + FLOAT}
        # a dump buffer that is too small, and a dump kind that does not exist
        nbytes = gpu["refs"].nbytes
        small = np.zeros(max(nbytes // 4 - 1, 0), np.uint32)
        import ctypes as C
        got = C.c_uint64()
        rc = r.lib.sr_frame_dump(r.ctx, abi.DUMP["refs"], abi.ptr(small), small.nbytes, C.byref(got))
        assert abi.STATUS_NAMES[rc] == "SR_ERR_CAPACITY"
        rc = r.lib.sr_frame_dump(r.ctx, 999, None, 0, C.byref(got))
        assert abi.STATUS_NAMES[rc] == "SR_ERR_INVALID_ARGUMENT"
    compare(w, ctr, gpu, oracle.render(w))
