"""The serving loop (sr_server_*) on the GPU: each request's composite and
depth equal the same request rendered alone through sr_frame_submit, in a
mix of assets, resolutions and merged multi-view frames; admission keeps the
held HBM within the budget and rejects a request that cannot fit alone."""
import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer, StrandServer
from bufcmp import assert_same

pytestmark = pytest.mark.gpu

# torch (pinned host buffers) is imported before any test loads the native
# library, so that its device runtime comes up in this process
torch = pytest.importorskip("torch")


def assets():
    a = scene.scalp_asset(300, 8, 11, 0.002, 2.0)
    b = scene.scalp_asset(700, 12, 23, 0.0015, 1.5)
    return [a, b]


def request_workload(asset, asset_id, i, n, w, h, lights=1):
    eye = scene.orbit_eye(i, n, 0.6)
    wl = scene.build_workload(f"req-{asset_id}-{i}-{w}x{h}", asset, [(eye, w, h)], lights, 96, 4, 8, 8)
    return scene.with_asset_id(wl, asset_id)


def alone(r, wl):
    """The same request through the plain frame path."""
    r.render(wl)
    comp = r.dump("composite").reshape(-1, 4)
    depth = r.dump("depth")
    return comp, depth


MIX = [(0, 0, 64, 64), (0, 1, 96, 80), (1, 2, 80, 96), (0, 3, 64, 64), (1, 4, 128, 96), (0, 5, 112, 64),
       (1, 6, 64, 64), (0, 7, 96, 96), (1, 8, 72, 72), (0, 9, 64, 128)]


def test_alone_vs_mix():
    a = assets()
    with StrandRenderer() as r:
        for x in a:
            r.upload(x)
        wls = [request_workload(a[aid], aid, i, 10, w, h) for aid, i, w, h in MIX]
        ref = [alone(r, wl) for wl in wls]
        print(f"[serving] alone renders done ({len(wls)})", flush=True)
        with StrandServer(r, slots=3, max_views=4) as srv:
            tickets = {srv.submit(scene.request_from_workload(wl, tag=k)): k for k, wl in enumerate(wls)}
            res = srv.drain(len(wls))
            st = srv.stats()
    print(f"[serving] stats {st}", flush=True)
    assert st["completed"] == len(wls) and st["failed"] == 0
    assert st["max_frame_views"] > 1, "same-asset requests waiting together are merged into one frame"
    for rec in res:
        k = rec["tag"]
        assert rec["status"] == 0 and tickets[rec["ticket"]] == k
        comp, depth = ref[k]
        assert_same(rec["rgba"].reshape(-1, 4), comp, f"request {k}: composite differs from alone")
        assert_same(rec["depth"].reshape(-1), depth, f"request {k}: depth differs from alone")
        assert rec["latency_ms"] >= 0.0 and rec["done_ms"] >= rec["launch_ms"] >= rec["submit_ms"]


def test_one_slot_one_view_equals_frame_path():
    a = assets()
    with StrandRenderer() as r:
        r.upload(a[0])
        wl = request_workload(a[0], 0, 2, 8, 96, 64)
        comp, depth = alone(r, wl)
        with StrandServer(r, slots=1, max_views=1) as srv:
            srv.submit(scene.request_from_workload(wl))
            (rec,) = srv.drain(1)
    assert rec["frame_views"] == 1
    assert_same(rec["rgba"].reshape(-1, 4), comp, "composite")
    assert_same(rec["depth"].reshape(-1), depth, "depth")


def test_budget_admission():
    a = assets()
    with StrandRenderer() as r:
        for x in a:
            r.upload(x)
        wls = [request_workload(a[aid], aid, i, 6, w, h) for aid, i, w, h in MIX[:6]]
        ref = [alone(r, wl) for wl in wls]
        # measure what one frame holds, then allow assets + a bit more than one frame
        with StrandServer(r, slots=1, max_views=1) as probe:
            probe.submit(scene.request_from_workload(wls[4]))
            probe.drain(1)
            one = probe.stats()["peak_held_bytes"]
        with StrandServer(r, slots=3, max_views=1, hbm_budget=int(one * 1.3)) as srv:
            for k, wl in enumerate(wls):
                srv.submit(scene.request_from_workload(wl, tag=k))
            res = srv.drain(len(wls))
            st = srv.stats()
        print(f"[serving] budget {int(one * 1.3)} stats {st}", flush=True)
        assert st["completed"] == len(wls)
        assert st["peak_held_bytes"] <= int(one * 1.3)
        for rec in res:
            comp, depth = ref[rec["tag"]]
            assert_same(rec["rgba"].reshape(-1, 4), comp, "composite")
        # a budget below what the assets alone hold rejects the request
        with StrandServer(r, slots=1, max_views=1, hbm_budget=1 << 20) as srv:
            srv.submit(scene.request_from_workload(wls[0]))
            (rec,) = srv.drain(1)
        assert rec["status"] == abi.STATUS_CODES["SR_ERR_CAPACITY"]


def test_rgba8_pinned_outputs_equal_alone():
    """Requests asking for the reference's rgba8unorm composite, copied
    straight into pinned host buffers, equal the same requests drawn alone
    (their FP32 composite rounded to 8 bits) and their depth is exact."""
    torch.cuda.init()
    a = assets()
    mix = MIX[:6]
    with StrandRenderer() as r:
        for x in a:
            r.upload(x)
        wls = [request_workload(a[aid], aid, i, 10, w, h) for aid, i, w, h in mix]
        ref = [alone(r, wl) for wl in wls]
        outs = []
        for aid, i, w, h in mix:
            rgba = torch.empty((h, w, 4), dtype=torch.uint8, pin_memory=True).numpy()
            depth = torch.empty((h, w), dtype=torch.float32, pin_memory=True).numpy()
            outs.append((rgba, depth))
        with StrandServer(r, slots=2, max_views=3) as srv:
            for k, wl in enumerate(wls):
                srv.submit(scene.request_from_workload(wl, tag=k), rgba8=True, out=outs[k], pinned=True)
            res = srv.drain(len(wls))
    for rec in res:
        k = rec["tag"]
        assert rec["status"] == 0
        comp, depth = ref[k]
        want8 = np.rint(np.clip(comp, 0.0, 1.0) * np.float32(255.0)).astype(np.uint8)
        assert_same(outs[k][0].reshape(-1, 4), want8, f"request {k}: rgba8 composite")
        assert_same(outs[k][1].reshape(-1), depth, f"request {k}: depth")
