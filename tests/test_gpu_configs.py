"""Parity of the HIP path at the benchmark configurations' full sizes.

* Named frame (100k x 32 strands, 1920x1080 camera, 2048^2 x 4-layer light,
  30.6M fine references): every canonical dump against the oracle, both
  binning backends (integer dumps byte-equal, float dumps <= 1e-3 max abs,
  images >= 50 dB PSNR).
* Batched frame (the 2,000 x 16 scalp from 64 orbit cameras of 720p, 1080p
  and 1440p, 4 lights of 2048^2 x 4 layers) with the autotuner's
  configuration: every view of the batch equals that view drawn alone, and
  two views (one 1440p) match the oracle.
* The named asset from 16 views at the tuned configuration (8 depth-warp
  slices): every view equals that view drawn alone, one matches the oracle.
* Serving mix on the 8 resident serving assets (10k .. 500k strands):
  requests including the 500k asset and a 3840x2160 view equal the same
  requests drawn alone.
* More than 2^16 pages (the named asset under an 8192^2 light): the page-CSR
  backend's hit sort then runs three radix passes, and every dump equals the
  segment-scatter backend's, which sorts no hits.
"""
import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer, StrandServer
from oracle import oracle

from test_gpu_parity import EXACT, FLOAT, compare
from bufcmp import assert_same

pytestmark = pytest.mark.gpu


def render_until_fits(r, w, tries=3):
    """A frame whose capacities overflowed is drawn again at the grown size."""
    for _ in range(tries):
        ctr = r.render(w)
        if not ctr["capacity_exceeded"]:
            return ctr
    raise AssertionError(f"{w.name}: capacities still exceeded after {tries} frames")


@pytest.fixture(scope="module")
def named():
    w = scene.config_named()
    print(f"[configs] oracle {w.name}", flush=True)
    ref = oracle.render(w)
    print("[configs] oracle done", flush=True)
    return w, ref


@pytest.mark.parametrize("backend", [0, 1])
def test_named_frame_matches_oracle(named, backend):
    w0, ref = named
    w = scene.retile(w0, backend=backend)
    print(f"[configs] named frame backend {backend}", flush=True)
    with StrandRenderer() as r:
        r.upload_workload(w)
        ctr = render_until_fits(r, w)
        gpu = {k: r.dump(k) for k in EXACT + FLOAT}
    assert ctr["fine_refs"] > 30_000_000 and ctr["allocated_pages"] > 10_000
    assert ctr["page_drops"] == 0
    compare(w, ctr, gpu, ref)


def test_many_pages_backends_agree():
    base = scene.config_named()
    eye = (scene.HEAD_CENTER[0], scene.HEAD_CENTER[1] + 0.05, scene.HEAD_CENTER[2] + 0.6)
    w = scene.build_workload("scalp-100k-x32-1080p-light8k", base.assets[0], [(eye, 1920, 1080)], 1, 8192, 4)
    out = {}
    for backend in (1, 0):
        print(f"[configs] 8192^2 light, backend {backend}", flush=True)
        with StrandRenderer() as r:
            r.upload_workload(w)
            ctr = render_until_fits(r, scene.retile(w, backend=backend))
            out[backend] = (ctr, {k: r.dump(k) for k in EXACT + FLOAT})
    ctr = out[1][0]
    print(f"[configs] pages {ctr['allocated_pages']}, refs {ctr['fine_refs']}", flush=True)
    assert ctr["allocated_pages"] > 1 << 16 and ctr["page_drops"] == 0  # the third radix pass runs
    assert out[0][0]["fine_refs"] == ctr["fine_refs"]
    for k in EXACT + FLOAT:
        assert_same(out[1][1][k], out[0][1][k], k)


def split_cameras(w, flat, channels):
    """Per camera slices of a concatenated per-pixel output."""
    return [flat[off * channels:(off + cw * ch) * channels] for off, cw, ch in scene.camera_pixel_offsets(w)]


def test_named_views_at_tuned_configuration():
    """The named asset from 16 1080p views at the configuration the tuner
    picks for it (16-px camera and 8-px light froxels, 8 depth-warp slices,
    page CSR, graph replay): every view of the batch equals that view drawn
    alone, and the first matches the oracle at the same configuration."""
    w = scene.retile(scene.config_batched_named(16), 16, 8, 1)
    w.settings["coarse_depth_slices"] = 8
    with StrandRenderer(flags=abi.CONFIG_HIP_GRAPHS) as r:
        r.upload_workload(w)
        render_until_fits(r, w)
        ctr = r.render(w)
        assert ctr["capacity_exceeded"] == 0 and ctr["page_drops"] == 0
        comp = split_cameras(w, r.dump("composite"), 4)
        depth = split_cameras(w, r.dump("depth"), 1)
    cams = [i for i in range(len(w.frusta)) if int(w.frusta[i]["kind"]) == 0]
    assert len(cams) == 16
    with StrandRenderer() as r:
        r.upload_workload(w)
        for k, fi in enumerate(cams):
            one = scene.select_views(w, [fi])
            ctr = render_until_fits(r, one)
            assert_same(r.dump("composite"), comp[k], f"view {k}: composite differs from alone")
            assert_same(r.dump("depth"), depth[k], f"view {k}: depth differs from alone")
            if k == 0:
                print(f"[configs] oracle named view 0 at 8 slices", flush=True)
                gpu = {n: r.dump(n) for n in EXACT + FLOAT}
                compare(one, ctr, gpu, oracle.render(one))
    print("[configs] 16 named views equal alone", flush=True)


def test_batched_views_equal_alone_and_oracle():
    base = scene.config_batched()
    with StrandRenderer() as r:
        r.upload_workload(base)
        tuned = r.autotune(base, frames_per_trial=1, min_psnr_db=50.0)["best"]
    print(f"[configs] tuned {tuned}", flush=True)
    w = scene.apply_tuning(base, tuned)
    flags = abi.CONFIG_HIP_GRAPHS if tuned["hip_graphs"] else 0
    with StrandRenderer(flags=flags) as r:
        r.upload_workload(w)
        render_until_fits(r, w)
        ctr = r.render(w)  # a replay of the captured frame when graphs are on
        assert ctr["capacity_exceeded"] == 0
        comp = split_cameras(w, r.dump("composite"), 4)
        depth = split_cameras(w, r.dump("depth"), 1)
    cams = [i for i in range(len(w.frusta)) if int(w.frusta[i]["kind"]) == 0]
    assert len(cams) == 64 and int((w.frusta["kind"] == 1).sum()) == 4
    heights = [int(w.frusta[i]["screen_height"]) for i in cams]
    picks = [heights.index(1440), heights.index(720)]
    with StrandRenderer() as r:
        r.upload_workload(w)
        for k, fi in enumerate(cams):
            one = scene.select_views(w, [fi])
            render_until_fits(r, one)
            assert_same(r.dump("composite"), comp[k], f"view {k}: composite differs from alone")
            assert_same(r.dump("depth"), depth[k], f"view {k}: depth differs from alone")
            if k in picks:
                print(f"[configs] oracle view {k} ({heights[k]}p)", flush=True)
                ctr = r.render(one)
                gpu = {n: r.dump(n) for n in EXACT + FLOAT}
                compare(one, ctr, gpu, oracle.render(one))
    print("[configs] 64 views equal alone", flush=True)


def test_serving_mix_on_serving_assets():
    assets = scene.serving_assets()
    rng = np.random.default_rng(2024)
    with StrandRenderer() as r:
        ids = [r.upload(a).geo_id for a in assets]
        # both binning backends (served frames of one backend are merged)
        reqs = [scene.serving_request(assets, ids, rng, tag=k, backend=k % 2) for k in range(12)]
        # the 500k-strand asset at 3840x2160, and the 500k asset again at 1080p
        for k, (asset, res) in enumerate(((7, (3840, 2160)), (7, (1920, 1080)), (0, (3840, 2160)))):
            eye = scene.orbit_eye(40 * k, 360, 0.6)
            wl = scene.with_asset_id(scene.build_workload(f"serve-{asset}-{res[0]}x{res[1]}", assets[asset],
                                                          [(eye, *res)], 1, 2048, 4, 8, 8,
                                                          settings={"fine_binning_backend": 1}), ids[asset])
            reqs.append((scene.request_from_workload(wl, tag=len(reqs)), wl))
        want = []
        for req, wl in reqs:
            render_until_fits(r, wl)
            want.append((r.dump("composite").tobytes(), r.dump("depth").tobytes()))
        print(f"[configs] {len(reqs)} alone renders done", flush=True)
        with StrandServer(r, slots=3, max_views=4) as srv:
            for req, _ in reqs:
                srv.submit(req)
            res = srv.drain(len(reqs), timeout_s=100.0)
            st = srv.stats()
    print(f"[configs] server {st}", flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for rec in res:
        k = rec["tag"]
        assert rec["status"] == 0
        assert_same(rec["rgba"], want[k][0], f"request {k}: composite differs from alone")
        assert_same(rec["depth"], want[k][1], f"request {k}: depth differs from alone")
