"""Virtual deep shadow maps (SR_FRAME_VIRTUAL_DOMS) in the CPU restatement,
pinned by what the paging rules fix independently of its code:

* requests: the pages holding any pixel of a shadow tile run, dilated by one
  page in x and y (request_shadow_dom_pages, strand_prepass.wgsl:2537-2579),
  computed here with numpy from the run dump;
* residency: the requested pages of each light in page order, the first
  dom_resident_pages of them, numbered consecutively;
* with every requested page resident the frame equals the dense frame byte
  for byte (no shadow tile lies outside a requested page);
* with fewer resident pages the DOM is the dense DOM cleared outside them,
  and shading only brightens (a page that is not resident lets all light
  through, strand_shading.wgsl:238-257).

The residency policy of the reference's un-vendored bevy_vsms crate is not
available; the page-order policy above is this restatement's own, so the
choice of which pages stay resident under a budget is parity unpinned
against the reference (every other rule above follows its WGSL).
"""
import dataclasses

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from oracle import oracle
from bufcmp import assert_same

P = abi.DOM_PAGE_XY


@pytest.fixture(scope="module")
def frames():
    out = {}
    for budget in (0, 7, 1):
        w = scene.config_virtual_doms(budget=budget)
        out[budget] = (w, oracle.render(w, threads=4))
    dense = dataclasses.replace(out[0][0], settings={k: v for k, v in out[0][0].settings.items()
                                                     if k not in ("frame_flags", "dom_resident_pages")})
    out["dense"] = (dense, oracle.render(dense, threads=4))
    return out


def requested_pages(w, runs):
    """numpy: pages touched by the light's shadow tile runs, then a 3 x 3 max filter."""
    li = int(np.nonzero(w.frusta["kind"] == 1)[0][0])
    f = w.frusta[li]
    W, H, fx, fy = (int(f[k]) for k in ("screen_width", "screen_height", "froxel_size_x", "froxel_size_y"))
    px, py = -(-W // P), -(-H // P)
    ftx = -(-W // fx)
    touched = np.zeros((py + 2, px + 2), bool)  # one page of padding on each side
    r = runs[(runs["frustum_id"] == li) & (runs["work_count"] > 0)]
    tx, ty = r["screen_tile_id"] % ftx, r["screen_tile_id"] // ftx
    x0, y0 = tx * fx, ty * fy
    x1, y1 = np.minimum(x0 + fx - 1, W - 1), np.minimum(y0 + fy - 1, H - 1)
    for a, b, c, d in zip(x0 // P, y0 // P, x1 // P, y1 // P):
        touched[b + 1:d + 2, a + 1:c + 2] = True
    dil = np.zeros_like(touched)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            dil[1:-1, 1:-1] |= touched[1 + dy:py + 1 + dy, 1 + dx:px + 1 + dx]
    return dil[1:-1, 1:-1].reshape(-1).astype(np.uint32)


def test_requests_are_the_dilated_pages_of_shadow_runs(frames):
    w, ref = frames[0]
    want = requested_pages(w, ref.get("shadow_runs"))
    got = ref.get("dom_requests")
    assert got.tolist() == want.tolist()
    assert 0 < got.sum() < got.size, "the workload requests some pages but not all"


@pytest.mark.parametrize("budget", [0, 7, 1])
def test_residency_is_page_order_up_to_the_budget(frames, budget):
    w, ref = frames[budget]
    req, table = ref.get("dom_requests"), ref.get("dom_page_table")
    ids = np.nonzero(req)[0]
    keep = ids if budget == 0 else ids[:budget]
    want = np.full(req.size, abi.INVALID_PTR, np.uint32)
    want[keep] = np.arange(keep.size, dtype=np.uint32)
    assert table.tolist() == want.tolist()
    assert ref.counters()["dom_pages"] == keep.size


def test_every_requested_page_resident_equals_dense(frames):
    _, v = frames[0]
    _, d = frames["dense"]
    for k in ("dom_depth", "dom_opacity", "stamp", "atlas", "history", "image", "depth", "composite"):
        assert_same(v.get(k), d.get(k), k)


@pytest.mark.parametrize("budget", [7, 1])
def test_pages_not_resident_read_empty(frames, budget):
    w, v = frames[budget]
    _, d = frames["dense"]
    f = w.frusta[int(np.nonzero(w.frusta["kind"] == 1)[0][0])]
    W, H = int(f["screen_width"]), int(f["screen_height"])
    table = v.get("dom_page_table").reshape(-(-H // P), -(-W // P))
    resident = np.repeat(np.repeat(table != abi.INVALID_PTR, P, 0), P, 1)[:H, :W]
    L = int(w.settings["dom_layers"])
    dd, vd = d.get("dom_depth").reshape(H, W), v.get("dom_depth").reshape(H, W)
    assert (vd[resident] == dd[resident]).all() and (vd[~resident] == 0).all()
    assert (dd[~resident] != 0).any(), "the budget drops pages that hold shadow"
    do, vo = d.get("dom_opacity").reshape(L, H, W), v.get("dom_opacity").reshape(L, H, W)
    assert (vo[:, resident] == do[:, resident]).all() and (vo[:, ~resident] == 0).all()
    # shading: visibility 1 replaces the DOM's attenuation, never the reverse
    va, da = v.get("atlas").reshape(-1, 4)[:, :3], d.get("atlas").reshape(-1, 4)[:, :3]
    assert (va >= da - 1e-6).all() and (va > da + 1e-4).any()
