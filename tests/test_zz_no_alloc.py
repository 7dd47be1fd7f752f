"""sr_frame_submit_noalloc, the entry point that allocates nothing: its
composite, depth and counters equal sr_frame_submit's for the same frame,
byte for byte, on two calls in a row (the second frame smaller, so it runs
over what the first left in the scratch array); frames that do not fit the
array, or need a staging upload, are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from bufcmp import assert_same

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def camera_pixels(w) -> int:
    cams = w.frusta[w.frusta["kind"] == 0]
    return int((cams["screen_width"].astype(np.int64) * cams["screen_height"]).sum())


def noalloc(r, w):
    """One frame through sr_frame_submit_noalloc on torch's current stream:
    (status, composite, depth, counters).  The outputs start as 7.0 so that
    a pixel the frame leaves unwritten shows."""
    px = camera_pixels(w)
    comp = torch.full((px * 4,), 7.0, dtype=torch.float32, device="cuda")
    depth = torch.full((px,), 7.0, dtype=torch.float32, device="cuda")
    ctr = torch.zeros(16, dtype=torch.int32, device="cuda")
    fd = w.frame_desc()
    stream = torch.cuda.current_stream().cuda_stream
    rc = r.lib.sr_frame_submit_noalloc(r.ctx, C.byref(fd), comp.data_ptr(), depth.data_ptr(), ctr.data_ptr(),
                                       stream)
    torch.cuda.synchronize()
    if rc != 0:
        return rc, None, None, None
    c = ctr.cpu().numpy().view(abi.frame_counters_dtype)[0]
    return rc, comp.cpu().numpy(), depth.cpu().numpy(), {k: int(c[k]) for k in abi.frame_counters_dtype.names}


def plain(r, w):
    ctr = r.render(w)
    return r.dump("composite").reshape(-1), r.dump("depth"), ctr


@pytest.mark.parametrize("backend,slices", [(0, (16, 16)), (1, (16, 16)), (1, (32, 8))])
def test_noalloc_matches_frame_submit(backend, slices):
    """Two calls in a row, the second with a smaller frame (and, in the last
    case, 32 and then 8 depth-warp slices per coarse tile)."""
    frames = [scene.config_small(froxel=8), scene.config_tiny(strands=60, res=80, light_res=72)]
    with StrandRenderer() as r:
        for w, sc in zip(frames, slices):
            w.settings["fine_binning_backend"] = backend
            w.settings["coarse_depth_slices"] = sc
        ids = [r.upload(w.assets[0]).geo_id for w in frames]
        frames = [scene.with_asset_id(w, i) for w, i in zip(frames, ids)]
        for k, w in enumerate(frames):
            print(f"[noalloc] call {k}: {w.name} backend {backend} slices {slices[k]}", flush=True)
            rc, comp, depth, ctr = noalloc(r, w)
            assert rc == 0, r.lib.sr_last_error(r.ctx)
            want_comp, want_depth, want_ctr = plain(r, w)
            assert ctr == want_ctr
            assert ctr["capacity_exceeded"] == 0 and ctr["fine_refs"] > 0
            assert_same(comp, want_comp, f"call {k}: composite differs")
            assert_same(depth, want_depth, f"call {k}: depth differs")


def test_noalloc_refuses_what_it_cannot_hold():
    """A 4096^2 camera needs more than the scratch array; opaque depth needs
    a staging upload.  Both are refused on the host, before any launch."""
    big = scene.config_tiny(strands=24, res=4096, light_res=64)
    with StrandRenderer() as r:
        r.upload_workload(big)
        rc, *_ = noalloc(r, big)
        assert abi.STATUS_NAMES[rc] == "SR_ERR_CAPACITY"
        assert b"SR_NOALLOC_SCRATCH_BYTES" in r.lib.sr_last_error(r.ctx)
        occluded = scene.config_tiny()
        occluded.opaque = scene.opaque_disc(occluded, 0.5)
        rc, *_ = noalloc(r, occluded)
        assert abi.STATUS_NAMES[rc] == "SR_ERR_UNSUPPORTED"
        # the context still renders after the refusals
        small = scene.config_tiny()
        rc, comp, depth, ctr = noalloc(r, small)
        assert rc == 0 and ctr["capacity_exceeded"] == 0
        want_comp, want_depth, _ = plain(r, small)
        assert_same(comp, want_comp, "composite")
        assert_same(depth, want_depth, "depth")
