"""Frames that reuse the lights' deep shadow maps (SR_FRAME_REUSE_LIGHT_DOMS),
and the stress configuration at its full size (1M strands x 64 points, 4K
views, 4 lights of 2048^2 x 8 layers).

Small frames are checked exactly: a camera drawn with reused shadow maps
equals the same camera drawn in a frame that builds them, and frames whose
lights differ from the ones the maps were built from are refused.  At the
stress size the checks are size-independent properties: counters agree with
the dumps, no capacity is exceeded, two draws are byte-identical and a view
drawn alone equals the view drawn in its batch frame.
"""
import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from bufcmp import assert_same

pytestmark = pytest.mark.gpu


def render_until_fits(r, w, tries=4):
    for _ in range(tries):
        ctr = r.render(w)
        if not ctr["capacity_exceeded"]:
            return ctr
    raise AssertionError(f"{w.name}: capacities still exceeded after {tries} frames")


def reuse(w):
    w.settings["frame_flags"] = abi.FRAME_REUSE_LIGHT_DOMS
    w.name += "-reuse"
    return w


def three_cameras(backend):
    base = scene.config_small(froxel=8)
    cams = [(scene.orbit_eye(i, 5, 0.6), 320 + 64 * i, 256) for i in range(3)]
    w = scene.build_workload("small-3cam-2light", base.assets[0], cams, 2, 384, 4)
    w.settings["fine_binning_backend"] = backend
    return w


@pytest.mark.parametrize("backend", [0, 1])
def test_reused_shadow_maps_equal_built_ones(backend):
    w = three_cameras(backend)
    build = scene.select_views(w, [0])
    later = reuse(scene.select_views(w, [1, 2]))
    with StrandRenderer() as r:
        r.upload_workload(w)
        # reference: each camera in a frame that builds both lights
        want = []
        for i in range(3):
            render_until_fits(r, scene.select_views(w, [i]))
            want.append((r.dump("composite").tobytes(), r.dump("depth").tobytes(), r.dump("atlas").tobytes()))
        render_until_fits(r, build)
        dom = (r.dump("dom_depth").tobytes(), r.dump("dom_opacity").tobytes(), r.dump("stamp").tobytes())
        ctr = render_until_fits(r, later)
        assert ctr["shadow_runs"] == 0, "reused lights are not rasterised again"
        assert (r.dump("dom_depth").tobytes(), r.dump("dom_opacity").tobytes(), r.dump("stamp").tobytes()) == dom
        tasks = r.dump("tasks")
        assert ((tasks["packed_field"] >> abi.BINNING_FRUSTUM_SHIFT) < 2).all(), "no light frustum is binned"
        comp, depth = r.dump("composite"), r.dump("depth")
        atlas = r.dump("atlas").reshape(2, -1)
        offs = scene.camera_pixel_offsets(later)
        for k, (off, cw, ch) in enumerate(offs):
            assert_same(comp[off * 4:(off + cw * ch) * 4], want[k + 1][0], f"camera {k + 1}: composite")
            assert_same(depth[off:off + cw * ch], want[k + 1][1], f"camera {k + 1}: depth")
            assert_same(atlas[k], want[k + 1][2], f"camera {k + 1}: shading atlas")


def test_reuse_refused_for_other_lights():
    w = three_cameras(1)
    with StrandRenderer() as r:
        r.upload_workload(w)
        fresh = reuse(scene.select_views(w, [1]))
        with pytest.raises(RuntimeError, match="light-building frame"):
            r.render(fresh)
        render_until_fits(r, scene.select_views(w, [0]))
        moved = reuse(scene.select_views(w, [1]))
        moved.views[-1:] = scene.light_view((0.0, 1.0, 0.0), *scene.bounding_sphere(w.assets[0]))
        with pytest.raises(RuntimeError, match="other than those"):
            r.render(moved)
        fewer = reuse(scene.build_workload("one-light", w.assets[0], [(scene.orbit_eye(1, 5, 0.6), 384, 256)],
                                           1, 384, 4))
        with pytest.raises(RuntimeError, match="other than those"):
            r.render(fewer)
        ok = reuse(scene.select_views(w, [1, 2]))
        assert render_until_fits(r, ok)["shadow_runs"] == 0


def check_counters(r, ctr):
    """Counters against the frame's own dumps."""
    pm = r.dump("page_meta")
    counts = r.dump("cell_counts").reshape(len(pm), -1) if len(pm) else np.zeros((0, 256), np.uint32)
    assert ctr["allocated_pages"] == len(pm) and ctr["page_drops"] == 0 and ctr["overflow_pages"] == 0
    per_page = counts.sum(axis=1, dtype=np.uint64)
    assert (per_page == pm["seg_ref_count"]).all(), "page totals are the sums of their cells"
    assert int(per_page.sum()) == ctr["fine_refs"]
    bases = pm["seg_ref_base"].astype(np.uint64)
    assert (bases[1:] == bases[:-1] + pm["seg_ref_count"][:-1]).all(), "page bases are a scan of page totals"
    offs = r.dump("cell_offsets").reshape(len(pm), -1).astype(np.uint64)
    excl = np.cumsum(counts, axis=1, dtype=np.uint64) - counts
    assert (offs == excl).all(), "cell offsets are page-local exclusive scans"
    assert int((counts > 0).sum()) == ctr["work_items"], "one work item per non-empty cell"


def test_stress_batch_properties():
    frames = scene.config_stress_chunks()
    views = sum(w.camera_count for w in frames)
    assert views == 16 and all(w.input_segments == 63_000_000 for w in frames)
    build, batch = frames[0], frames[1]
    with StrandRenderer() as r:
        r.upload(build.assets[0])
        print(f"[stress] {build.name}", flush=True)
        ctr = render_until_fits(r, build)
        print(f"[stress] built: {ctr}", flush=True)
        assert ctr["shadow_runs"] > 0 and ctr["fine_refs"] > 1_000_000_000
        check_counters(r, ctr)
        print(f"[stress] {batch.name}", flush=True)
        ctr = render_until_fits(r, batch)
        print(f"[stress] batch: {ctr}", flush=True)
        assert ctr["shadow_runs"] == 0 and ctr["camera_runs"] > 0
        check_counters(r, ctr)
        first = (r.dump("composite").tobytes(), r.dump("depth").tobytes(), r.dump("page_meta").tobytes())
        again = r.render(batch)
        assert again == ctr
        assert (r.dump("composite").tobytes(), r.dump("depth").tobytes(), r.dump("page_meta").tobytes()) == first
        cams = [i for i in range(len(batch.frusta)) if int(batch.frusta[i]["kind"]) == 0]
        k = 2
        alone = reuse(scene.select_views(batch, [cams[k]]))
        render_until_fits(r, alone)
        off, cw, ch = scene.camera_pixel_offsets(batch)[k]
        comp = np.frombuffer(first[0], np.float32)
        assert_same(r.dump("composite"), comp[off * 4:(off + cw * ch) * 4], "view alone vs batch")
        img = comp[off * 4:(off + cw * ch) * 4].reshape(-1, 4)
        assert np.isfinite(img).all() and (img[:, 3] > 0).mean() > 0.05, "the view shows hair"
