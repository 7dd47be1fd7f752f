"""Virtual deep shadow maps (SR_FRAME_VIRTUAL_DOMS) on the HIP path.

* With every requested page resident, a frame equals the same frame with
  dense DOMs byte for byte (the 2,000-strand scalp with a 2048^2 light that
  requests 20 of 64 pages, and the named 100k x 32 frame).
* Against the oracle, with and without a page budget: requests and page
  table byte-equal, every other dump to the parity bar.
* The page pool starts small under SR_CONFIG_TIGHT_CAPACITY, reports its
  capacity exceeded, grows, and the next frame equals the dense one.
* Graph replays and frames that reuse the lights' maps keep the pages.
"""
import dataclasses

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from oracle import oracle

from test_gpu_parity import EXACT, FLOAT, compare
from bufcmp import assert_same

pytestmark = pytest.mark.gpu

OUTPUTS = ["dom_depth", "dom_opacity", "stamp", "atlas", "history", "image", "depth", "composite"]


def render_until_fits(r, w, tries=4):
    for _ in range(tries):
        ctr = r.render(w)
        if not ctr["capacity_exceeded"]:
            return ctr
    raise AssertionError(f"{w.name}: capacities still exceeded after {tries} frames")


def dense(w):
    st = {k: v for k, v in w.settings.items() if k not in ("frame_flags", "dom_resident_pages")}
    return dataclasses.replace(w, name=w.name + "-dense", settings=st)


def outputs(r):
    return {k: r.dump(k).tobytes() for k in OUTPUTS}


@pytest.mark.parametrize("which", ["spread", "named"])
def test_every_requested_page_resident_equals_dense(which):
    if which == "spread":
        w = scene.config_virtual_doms()
    else:
        w = scene.with_virtual_doms(scene.retile(scene.config_named(), backend=1))
    print(f"[vsms] {w.name}", flush=True)
    with StrandRenderer() as r:
        r.upload_workload(w)
        render_until_fits(r, dense(w))
        want = outputs(r)
        ctr = render_until_fits(r, w)
        got = outputs(r)
        req = r.dump("dom_requests")
        table = r.dump("dom_page_table")
    for k in OUTPUTS:
        assert_same(got[k], want[k], f"{w.name}: {k} differs from dense DOMs")
    assert ctr["dom_pages"] == int(req.sum()) == int((table != abi.INVALID_PTR).sum())
    if which == "spread":
        assert ctr["dom_pages"] == 20


@pytest.mark.parametrize("budget,slices", [(0, 16), (7, 16), (7, 8), (7, 32)])
def test_virtual_doms_match_oracle(budget, slices):
    w = scene.config_virtual_doms(budget=budget)
    w.settings["coarse_depth_slices"] = slices
    print(f"[vsms] oracle {w.name} at {slices} depth slices", flush=True)
    ref = oracle.render(w)
    with StrandRenderer() as r:
        r.upload_workload(w)
        ctr = render_until_fits(r, w)
        gpu = {k: r.dump(k) for k in EXACT + FLOAT}
        for k in ("dom_requests", "dom_page_table"):
            assert_same(r.dump(k), ref.get(k), k)
    assert ctr["dom_pages"] == ref.counters()["dom_pages"]
    compare(w, ctr, gpu, ref)


def test_page_pool_grows_from_tight():
    w = scene.config_virtual_doms()
    with StrandRenderer() as r:
        r.upload_workload(w)
        render_until_fits(r, dense(w))
        want = outputs(r)
    with StrandRenderer(flags=abi.CONFIG_TIGHT_CAPACITY) as r:
        r.upload_workload(w)
        first = r.render(w)
        assert first["capacity_exceeded"] == 1 and first["dom_pages"] <= 1, "the pool starts at one page"
        render_until_fits(r, w, tries=8)  # every tight capacity grows
        got = outputs(r)
    for k in OUTPUTS:
        assert_same(got[k], want[k], k)


def test_graph_replays_keep_pages():
    w = scene.config_virtual_doms(budget=7)
    with StrandRenderer() as r:
        r.upload_workload(w)
        render_until_fits(r, w)
        want = outputs(r)
    with StrandRenderer(flags=abi.CONFIG_HIP_GRAPHS) as r:
        r.upload_workload(w)
        render_until_fits(r, w)
        for k in range(3):
            r.render(dense(w))  # eager work of another shape between the replays
            ctr = r.render(w)
            assert ctr["capacity_exceeded"] == 0
            got = outputs(r)
            for n in OUTPUTS:
                assert_same(got[n], want[n], f"replay {k}: {n}")


def test_reused_virtual_maps_equal_built_ones():
    base = scene.config_virtual_doms(budget=7)
    cams = [(scene.orbit_eye(i, 5, 0.6), 320 + 64 * i, 256) for i in range(2)]
    w = scene.build_workload("vsms-2cam", base.assets[0], cams, 1, 2048, 4)
    w = dataclasses.replace(w, views=np.ascontiguousarray(np.concatenate([w.views[:2], base.views[1:]])),
                            settings=dict(base.settings))
    with StrandRenderer() as r:
        r.upload_workload(w)
        render_until_fits(r, scene.select_views(w, [1]))
        want = (r.dump("composite").tobytes(), r.dump("depth").tobytes(), r.dump("atlas").tobytes())
        render_until_fits(r, scene.select_views(w, [0]))
        maps = (r.dump("dom_depth").tobytes(), r.dump("dom_page_table").tobytes())
        later = scene.select_views(w, [1])
        later.settings["frame_flags"] |= abi.FRAME_REUSE_LIGHT_DOMS
        ctr = render_until_fits(r, later)
        assert ctr["shadow_runs"] == 0
        assert (r.dump("dom_depth").tobytes(), r.dump("dom_page_table").tobytes()) == maps
        assert (r.dump("composite").tobytes(), r.dump("depth").tobytes(), r.dump("atlas").tobytes()) == want
