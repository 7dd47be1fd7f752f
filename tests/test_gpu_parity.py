"""Parity of the HIP path (through the C ABI) against the CPU restatement.

Bar: binning, scans, compaction and emission are bit-exact (every record of
the canonical dumps is compared byte for byte); deep-shadow layers, shading
atlas and images are within 1e-3 max abs in linear FP32, images reach
PSNR >= 50 dB.
"""
import dataclasses

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from oracle import oracle
from bufcmp import assert_same

pytestmark = pytest.mark.gpu

EXACT = ["broad_meta", "keep", "ranges", "tasks", "projected", "shading_queue", "lut", "page_table", "page_meta",
         "cell_counts", "cell_offsets", "refs", "active_tiles", "work_items", "camera_runs", "shadow_runs"]
FLOAT = ["dom_depth", "dom_opacity", "stamp", "atlas", "history", "image", "depth", "composite"]
COUNTERS = ["binning_tasks", "camera_tasks", "allocated_pages", "fine_refs", "active_tiles", "work_items",
            "camera_runs", "shadow_runs", "overflow_pages"]
TOL = 1e-3  # max abs, linear FP32
PSNR_MIN = 50.0


def gpu_frame(w, frames=1, flags=0):
    with StrandRenderer(flags=flags) as r:
        r.upload_workload(w)
        for _ in range(frames):
            ctr = r.render(w)
        out = {k: r.dump(k) for k in EXACT + FLOAT}
    return ctr, out


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)


def assert_exact(name, got, want):
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    g = np.ascontiguousarray(got).view(np.uint8).reshape(len(got), -1) if len(got) else got
    e = np.ascontiguousarray(want).view(np.uint8).reshape(len(want), -1) if len(want) else want
    if len(got):
        bad = np.nonzero((g != e).any(axis=1))[0]
        assert bad.size == 0, f"{name}: {bad.size} records differ, first at {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"


def compare(w, ctr, gpu, ref):
    rc = ref.counters()
    for k in COUNTERS:
        assert ctr[k] == rc[k], f"counter {k}: gpu {ctr[k]} oracle {rc[k]}"
    assert ctr["overflow_pages"] == 0
    for k in EXACT:
        assert_exact(k, gpu[k], ref.get(k))
    for k in FLOAT:
        g, e = gpu[k], ref.get(k)
        assert g.shape == e.shape, f"{k}: shape {g.shape} != {e.shape}"
        assert np.isfinite(g).all() == np.isfinite(e).all(), f"{k}: finiteness differs"
        if g.size:
            d = float(np.nanmax(np.abs(g.astype(np.float64) - e)))
            assert d <= TOL, f"{k}: max abs diff {d}"
    for k in ("image", "composite"):
        p = psnr(gpu[k], ref.get(k))
        assert p >= PSNR_MIN, f"{k}: PSNR {p:.1f} dB"


CASES = {
    "tiny-f8": lambda: scene.config_tiny(),
    "tiny-f16": lambda: scene.config_tiny(froxel=16, res=96, light_res=80),
    # 12-px froxels: tile edges that are not powers of two (divided, not
    # multiplied by a reciprocal, in the walks) and 144-pixel raster tiles
    "tiny-f12": lambda: scene.config_tiny(froxel=12, res=96, light_res=84),
    "tiny-2lights-8layers": lambda: scene.config_tiny(strands=40, lights=2, dom_layers=8, res=72),
    "small-f16": lambda: scene.config_small(froxel=16),
    "small-f8": lambda: scene.config_small(froxel=8),
}


def csr(w):
    """The same workload on the page-CSR fine-binning backend."""
    w.settings["fine_binning_backend"] = 1
    w.name += "-csr"
    return w


def cache_workload():
    """A scalp written as a .strands cache and loaded back (one material)."""
    base = scene.scalp_asset(300, 8, 11, 0.002, 2.0)
    so = np.stack([base.meta["count"], base.meta["offset"]], axis=1)
    data = scene.strand_cache_bytes(base.vertices[:, :3], so, base.indices, scale=1.0)
    a = scene.asset_from_cache(data, radius_px=2.0)
    eye = (scene.HEAD_CENTER[0], scene.HEAD_CENTER[1] + 0.05, scene.HEAD_CENTER[2] + 0.6)
    return scene.build_workload("cache-300x8-96", a, [(eye, 96, 96)], 1, 96, 4, 8, 8)


CASES["cache-asset"] = cache_workload


def depth_stack(n: int = 10, step: float = 0.5):
    """n copies of the tiny scalp, each step metres further along the view
    axis: n disjoint depth intervals over the same coarse tiles (more than 8,
    so an 8-slice LUT merges intervals into their nearest slices)."""
    w = scene.config_tiny(strands=30, points=6, res=64, light_res=64)
    extra = []
    for k in range(1, n):
        inst = scene.identity_instance(0)
        m = np.eye(4)
        m[:3, 3] = (0.0, 0.0, -step * k)
        inst["world_from_local"] = m.T.reshape(-1).astype(np.float32)
        extra.append(inst)
    return dataclasses.replace(w, name=f"tiny-stack-{n}", instances=np.concatenate([w.instances] + extra))


CASES["depth-stack"] = depth_stack


def edge_workload(res: int = 64, far_away: bool = False):
    """Strands at the edges of the pipeline's arithmetic, under the identity
    camera and light of config_two_strands (screen x = (x / 2 + 1 / 2) res,
    y = (1 / 2 - y / 2) res, depth = z): repeated points (zero-length
    segments), a one-point strand (dropped when packing, src/plugin.rs:1334),
    an exactly vertical and an exactly horizontal strand (a zero direction
    component in both DDAs), a strand pointing at the camera (segments of
    zero screen length, depth only), one on a fine-tile edge, a diagonal
    through tile corners, one leaving the screen and one leaving the depth
    range.  far_away moves every strand off both views: a frame with no
    binning task at all."""
    px = lambda sx: 2.0 * sx / res - 1.0  # screen x (pixels) -> clip x
    py = lambda sy: 1.0 - 2.0 * sy / res
    strands = [
        [(px(10.5), py(12.5), 0.7), (px(10.5), py(12.5), 0.7), (px(30.5), py(12.5), 0.7), (px(30.5), py(12.5), 0.7)],
        [(px(40.0), py(40.0), 0.5)],
        [(px(20.0), py(4.0), 0.6), (px(20.0), py(60.0), 0.6)],
        [(px(4.0), py(50.0), 0.45), (px(60.0), py(50.0), 0.45)],
        [(px(44.3), py(20.7), 0.3), (px(44.3), py(20.7), 0.9), (px(44.3), py(20.7), 0.95)],
        [(px(16.0), py(2.0), 0.55), (px(16.0), py(30.0), 0.65)],
        [(px(8.0), py(8.0), 0.2), (px(32.0), py(32.0), 0.6), (px(56.0), py(56.0), 0.95)],
        [(px(-30.0), py(36.0), 0.5), (px(34.0), py(36.5), 0.5)],
        [(px(24.0), py(44.0), -0.3), (px(40.0), py(46.0), 0.35), (px(48.0), py(47.0), 1.4)],
    ]
    if far_away:
        strands = [[(x + 5.0, y, z) for (x, y, z) in st] for st in strands]
    w = scene.config_two_strands(res=res)
    asset = scene.asset_from_strands(strands, radius_px=1.5)
    for i in range(asset.materials.shape[0]):
        asset.materials["absorption_color"][i] = (0.2 + 0.08 * i, 0.5, 0.8 - 0.07 * i, 0.55)
    return dataclasses.replace(w, name=f"edge-strands-{res}{'-far' if far_away else ''}", assets=[asset])


CASES["edge-strands"] = lambda: edge_workload()
CASES["empty-view"] = lambda: edge_workload(far_away=True)


def rect_froxels(w, fx: int, fy: int):
    """The workload with every frustum's froxels fx x fy pixels."""
    cfgs = [(int(f["screen_width"]), int(f["screen_height"]), fx, fy, int(f["depth_slices"])) for f in w.frusta]
    return dataclasses.replace(w, name=f"{w.name}-f{fx}x{fy}", frusta=scene.frustum_table(cfgs, list(w.frusta["kind"])),
                               settings=dict(w.settings))


# non-square froxels: 16 x 8 pixel tiles
CASES["tiny-f16x8"] = lambda: rect_froxels(scene.config_tiny(strands=60, res=96, light_res=80), 16, 8)
# two strands whose DOM layers and blended pixel have closed forms
# (tests/test_oracle_closed_forms.py): lit, and with the light moved off them
CASES["two-strands"] = lambda: scene.config_two_strands()
CASES["two-strands-unlit"] = lambda: scene.config_two_strands(light_offset=4.0)

CSR_CASES = {
    "tiny-f8-csr": lambda: csr(scene.config_tiny()),
    "tiny-2lights-8layers-csr": lambda: csr(scene.config_tiny(strands=40, lights=2, dom_layers=8, res=72)),
    "small-f16-csr": lambda: csr(scene.config_small(froxel=16)),
    "small-f8-csr": lambda: csr(scene.config_small(froxel=8)),
    "tiny-f12-csr": lambda: csr(scene.config_tiny(froxel=12, res=96, light_res=84)),
    "tiny-f16x8-csr": lambda: csr(rect_froxels(scene.config_tiny(strands=60, res=96, light_res=80), 16, 8)),
    # thousands of strands over one coarse tile: pages of thousands of hits
    # (several steps of the per-page cell sort)
    "dense-one-tile-csr": lambda: csr(scene.config_tiny(strands=3000, points=8, res=32, light_res=32)),
    # a 4096^2 light: page capacity above 2^16, three radix passes over the page bits
    "wide-light-csr": lambda: csr(scene.config_tiny(res=64, light_res=4096)),
    "edge-strands-csr": lambda: csr(edge_workload()),
    "empty-view-csr": lambda: csr(edge_workload(far_away=True)),
    "depth-stack-csr": lambda: csr(depth_stack()),
}


CSR_FLAGS = {"few-hits": abi.CONFIG_CSR_FEW_HITS, "few-hits+graphs": abi.CONFIG_CSR_FEW_HITS | abi.CONFIG_HIP_GRAPHS}


@pytest.mark.parametrize("flags", list(CSR_FLAGS))
@pytest.mark.parametrize("case", ["tiny-f8-csr", "small-f8-csr", "dense-one-tile-csr", "edge-strands-csr"])
def test_parity_csr_paths(case, flags):
    """Most segments through the re-trace path (2 recorded cells per
    segment), eager and replayed from a HIP graph."""
    w = CSR_CASES[case]()
    print(f"[parity] {case} ({flags}): {w.name}", flush=True)
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w, flags=CSR_FLAGS[flags])
    compare(w, ctr, gpu, ref)


@pytest.mark.parametrize("slices", [8, 32])
@pytest.mark.parametrize("case", ["tiny-f8", "small-f8", "depth-stack", "tiny-f8-csr", "small-f8-csr", "depth-stack-csr",
                                  "dense-one-tile-csr", "edge-strands-csr", "tiny-2lights-8layers-csr"])
def test_parity_depth_slices(case, slices):
    """8 and 32 depth-warp LUT slices per coarse tile (the oracle bins with
    the same count): LUT, page table, pages, references, work and images.
    At 32 slices a fine tile has 512 depth lanes, and the tiny and stacked
    scenes have tile runs of more than 256 work items."""
    w = {**CASES, **CSR_CASES}[case]()
    w.settings["coarse_depth_slices"] = slices
    print(f"[parity] {case} at {slices} slices: {w.name}", flush=True)
    ref = oracle.render(w)
    assert ref.get("lut").size == slices * int(w.frusta["coarse_depth_tile_count"].sum())
    if slices == 32 and case.startswith(("tiny-f8", "depth-stack")):
        runs = np.concatenate([ref.get("camera_runs"), ref.get("shadow_runs")])
        assert runs["work_count"].max() > 256
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)


@pytest.mark.parametrize("flags", list(CSR_FLAGS))
def test_parity_depth_slices_csr_paths(flags):
    """32 slices through the re-trace path, eager and from a HIP graph."""
    w = csr(depth_stack())
    w.settings["coarse_depth_slices"] = 32
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w, flags=CSR_FLAGS[flags])
    compare(w, ctr, gpu, ref)


@pytest.mark.parametrize("case", ["small-f8", "small-f8-csr"])
def test_capacity_growth(case):
    """A slot starting at minimal reference, work and page capacities: the
    first frame overflows and says so, later frames of the same shape grow
    to what they needed and then match the oracle exactly."""
    w = {**CASES, **CSR_CASES}[case]()
    print(f"[parity] capacity growth {case}: {w.name}", flush=True)
    ref = oracle.render(w)
    with StrandRenderer(flags=abi.CONFIG_TIGHT_CAPACITY) as r:
        r.upload_workload(w)
        exceeded = [r.render(w)["capacity_exceeded"]]
        while exceeded[-1] and len(exceeded) < 5:
            exceeded.append(r.render(w)["capacity_exceeded"])
        ctr = r.render(w)
        gpu = {k: r.dump(k) for k in EXACT + FLOAT}
    assert exceeded[0] == 1 and exceeded[-1] == 0, exceeded
    assert ctr["capacity_exceeded"] == 0
    compare(w, ctr, gpu, ref)


@pytest.mark.parametrize("case", list(CASES) + list(CSR_CASES))
def test_parity(case):
    w = {**CASES, **CSR_CASES}[case]()
    print(f"[parity] {case}: {w.name}", flush=True)
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)
    if w.name.endswith(("-far", "-far-csr")):
        assert ctr["binning_tasks"] == 0 and ctr["fine_refs"] == 0
    elif w.settings.get("fine_binning_backend", 0):
        assert ctr["page_candidates"] > 0


@pytest.mark.parametrize("backend", [0, 1])
def test_two_strands_closed_forms(backend):
    """The HIP path gives the hand-derived values of the two-strand scene:
    DOM layers 0.3, 0.4225, 0.5235625, 0.5235625 at depth 0.8, and the near
    strand over the far one, unpremultiplied."""
    lit, unlit = scene.config_two_strands(), scene.config_two_strands(light_offset=4.0)
    for w in (lit, unlit):
        w.settings["fine_binning_backend"] = backend
    _, g = gpu_frame(lit)
    np.testing.assert_allclose(g["dom_opacity"].reshape(4, 64, 64)[:, 20, 32], [0.3, 0.4225, 0.5235625, 0.5235625],
                               atol=1e-6)
    assert g["dom_depth"].reshape(64, 64)[20, 32] == np.float32(0.8)
    _, g = gpu_frame(unlit)
    near, far = np.array([0.9, 0.1, 0.1, 0.6]), np.array([0.1, 0.2, 0.9, 0.7])
    a = near[3] + far[3] * (1 - near[3])
    rgb = (near[:3] * near[3] + far[:3] * far[3] * (1 - near[3])) / a
    np.testing.assert_allclose(g["image"].reshape(64, 64, 4)[20, 32], np.r_[rgb, a], atol=1e-6)


def unorm8(x):
    """rgba8unorm of float channels: clamp, x 255, round half to even."""
    return np.rint(np.clip(x, 0.0, 1.0) * np.float32(255.0)).astype(np.uint8)


@pytest.mark.parametrize("backend", [0, 1])
def test_rgba8_outputs(backend):
    """SR_FRAME_RGBA8_OUTPUTS: the image and composite in the reference's
    rgba8unorm framebuffer format (src/pipelines/raster.rs:500-513) are the
    FP32 outputs rounded to 8 bits, and within one step (1/255) of the
    oracle's FP32 image."""
    w = scene.config_small(froxel=8)
    w.settings["fine_binning_backend"] = backend
    w.settings["frame_flags"] = abi.FRAME_RGBA8_OUTPUTS
    ref = oracle.render(w)
    with StrandRenderer() as r:
        r.upload_workload(w)
        r.render(w)
        img, comp = r.dump("image"), r.dump("composite")
        img8, comp8 = r.dump("image_rgba8"), r.dump("composite_rgba8")
        w.settings["frame_flags"] = 0
        r.render(w)
        assert r.dump("image_rgba8").size == 0, "rgba8 outputs only when asked for"
        assert_same(r.dump("image"), img, "float image with rgba8 outputs on")
    assert_same(img8, unorm8(img), "rgba8 image")
    assert_same(comp8, unorm8(comp), "rgba8 composite")
    for got, want in ((img8, ref.get("image")), (comp8, ref.get("composite"))):
        d = np.abs(got.astype(np.float64) / 255.0 - np.clip(want, 0.0, 1.0)).max()
        assert d <= 1.0 / 255.0 + 1e-6, d
    assert (img8.reshape(-1, 4)[:, 3] > 0).sum() > 1000


def test_parity_multiview():
    """Two cameras of different resolution and two lights in one frame."""
    base = scene.config_small(froxel=8)
    asset = base.assets[0]
    cams = [(scene.orbit_eye(0, 4, 0.6), 320, 240), (scene.orbit_eye(1, 4, 0.6), 256, 384)]
    w = scene.build_workload("small-2cam-2light", asset, cams, 2, 384, 4)
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)


def instance_transform(kind: str) -> np.ndarray:
    """world_from_local (column-major) turning the scalp 0.3 rad about its
    centre, scaled by 1.1 and shifted; "projective" adds 0.05 x to w, so world
    points have w != 1, which the rasterizers keep when they project
    (strand_rasterizer.wgsl:289-303)."""
    c = np.array(scene.HEAD_CENTER, np.float64)
    a = 0.3
    rot = np.array([[np.cos(a), 0.0, np.sin(a)], [0.0, 1.0, 0.0], [-np.sin(a), 0.0, np.cos(a)]]) * 1.1
    m = np.eye(4)
    m[:3, :3] = rot
    m[:3, 3] = c - rot @ c + np.array([0.01, -0.005, 0.02])
    if kind == "projective":
        m[3, 0] = 0.05
    return m.T.reshape(-1).astype(np.float32)


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("kind", ["affine", "projective"])
def test_parity_instance_transform(kind, backend):
    """A transformed instance (affine, and with a projective row)."""
    w = scene.config_tiny(strands=60, res=80, light_res=72)
    w.instances["world_from_local"][0] = instance_transform(kind)
    w.settings["fine_binning_backend"] = backend
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)


def three_instances():
    """Three instances in one frame: the tiny scalp, a second scalp asset
    shifted right, and the first asset again shifted left (per-instance
    index positions, asset ranges and materials)."""
    w = scene.config_tiny(strands=40, points=6, res=80, light_res=72)
    second = scene.scalp_asset(30, 5, 19, 0.004, 2.0)
    extra = []
    for asset_id, shift in ((1, (0.03, 0.0, 0.01)), (0, (-0.03, 0.01, -0.01))):
        inst = scene.identity_instance(asset_id)
        m = np.eye(4)
        m[:3, 3] = shift
        inst["world_from_local"] = m.T.reshape(-1).astype(np.float32)
        extra.append(inst)
    return dataclasses.replace(w, name="tiny-3-instances", assets=[w.assets[0], second],
                               instances=np.concatenate([w.instances] + extra))


@pytest.mark.parametrize("backend", [0, 1])
def test_parity_three_instances(backend):
    w = three_instances()
    w.settings["fine_binning_backend"] = backend
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)


CULL_SETTINGS = {
    "no-stochastic-cull": {"stochastic_cull_enabled": 0},
    "half-shadow-keep": {"shadow_keep_probability": 0.5, "target_strands_per_pixel": 0.5},
}


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("cull", list(CULL_SETTINGS))
def test_parity_cull_settings(cull, backend):
    """Stochastic culling off, and a shadow keep probability below one with
    fewer strands per pixel (camera and shadow slots keep different strands)."""
    w = scene.config_tiny(strands=60, res=80, light_res=72)
    w.settings.update(CULL_SETTINGS[cull])
    w.settings["fine_binning_backend"] = backend
    ref = oracle.render(w)
    ctr, gpu = gpu_frame(w)
    compare(w, ctr, gpu, ref)


def test_backends_agree_multiframe():
    """Both fine-binning backends give byte-identical dumps, frame after
    frame, on the medium workload."""
    w0 = scene.config_small(froxel=8)
    w1 = csr(scene.config_small(froxel=8))
    _, a = gpu_frame(w0, frames=2)
    ctr, b = gpu_frame(w1, frames=2)
    assert ctr["overflow_pages"] == 0
    for k in EXACT + FLOAT:
        assert_same(a[k], b[k], k)


def test_deterministic_repeat():
    """Two frames of the same workload give byte-identical dumps."""
    w = scene.config_small(froxel=8)
    with StrandRenderer() as r:
        r.upload_workload(w)
        r.render(w)
        a = {k: r.dump(k) for k in EXACT + FLOAT}
        r.render(w)
        b = {k: r.dump(k) for k in EXACT + FLOAT}
    for k in EXACT + FLOAT:
        assert_same(a[k], b[k], k)


def test_graph_replays_match_eager():
    """Frames replayed from a captured HIP graph equal eager frames, for the
    frame that was captured and for another frame of the same shape (new
    per-frame tables), replayed several times with eager work in between."""
    w1 = scene.config_small(froxel=8)
    w2 = scene.build_workload("small-orbit", w1.assets[0], [(scene.orbit_eye(3, 8, 0.6), 512, 512)], 1, 512, 4, 8, 8)
    with StrandRenderer() as eager, StrandRenderer(flags=abi.CONFIG_HIP_GRAPHS) as graph:
        eager.upload_workload(w1)
        graph.upload_workload(w1)
        want = {}
        for name, w in (("w1", w1), ("w2", w2)):
            eager.render(w)
            want[name] = {k: eager.dump(k) for k in EXACT + FLOAT}
        for rep in range(3):
            for name, w in (("w1", w1), ("w2", w2)):
                graph.render(w)
                got = {k: graph.dump(k) for k in EXACT + FLOAT}
                eager.render(w2 if name == "w1" else w1)  # eager work between replays
                for k in EXACT + FLOAT:
                    assert_same(got[k], want[name][k], f"replay {rep} {name}: {k} differs")


def test_autotune_picks_an_accepted_configuration():
    w = scene.config_small(froxel=8)
    with StrandRenderer() as r:
        r.upload_workload(w)
        res = r.autotune(w, frames_per_trial=2, min_psnr_db=40.0)
    print(f"[autotune] best {res['best']}", flush=True)
    trials = res["trials"]
    # reference, 8 backend x froxel candidates, 2 other slice counts, 8 launch shapes of the best
    assert len(trials) == 19
    assert trials[0]["hip_graphs"] == 0 and trials[0]["psnr_db"] > 900  # the reference itself
    assert res["best"]["accepted"] == 1 and res["best"]["ms_per_frame"] > 0
    assert res["best"]["ms_per_frame"] == min(t["ms_per_frame"] for t in trials if t["accepted"])
    # same froxel sizes and backend as the reference: the graph replay is exact
    same = [t for t in trials[1:9] if t["camera_froxel"] == 8 and t["light_froxel"] == 8
            and t["fine_binning_backend"] == trials[0]["fine_binning_backend"]]
    assert same and same[0]["psnr_db"] > 900
    assert all(t["coarse_depth_slices"] == 16 for t in trials[:9])
    # the slice counts 8 and 32 of the stage-one winner's configuration
    first = min((t for t in trials[1:9] if t["accepted"]), key=lambda t: t["ms_per_frame"])
    key = lambda t: (t["fine_binning_backend"], t["camera_froxel"], t["light_froxel"])
    assert sorted(t["coarse_depth_slices"] for t in trials[9:11]) == [8, 32]
    assert all(key(t) == key(first) for t in trials[9:11])
    # the launch shapes keep the stage-two winner's configuration and change no result
    second = min((t for t in [first] + trials[9:11] if t["accepted"]), key=lambda t: t["ms_per_frame"])
    shapes = trials[11:]
    assert {(t["mark_waves"], t["raster_waves"]) for t in shapes} == {
        (m, r) for m in (1, 2, 4) for r in (8, 16, 24)} - {(4, 16)}
    for t in shapes:
        assert key(t) == key(second) and t["coarse_depth_slices"] == second["coarse_depth_slices"]
        assert t["psnr_db"] == second["psnr_db"]
    tuned = scene.apply_tuning(w, res["best"])
    ref = oracle.render(tuned)
    ctr, gpu = gpu_frame(tuned)
    compare(tuned, ctr, gpu, ref)


@pytest.mark.parametrize("backend", [0, 1])
def test_autotune_can_keep_the_backend(backend):
    """SR_FRAME_TUNE_KEEP_BACKEND: every trial keeps the frame's binning
    backend; the froxel sizes, slice counts and launch shapes are still
    searched."""
    w = scene.config_small(froxel=8)
    w.settings["fine_binning_backend"] = backend
    w.settings["frame_flags"] = abi.FRAME_TUNE_KEEP_BACKEND
    with StrandRenderer() as r:
        r.upload_workload(w)
        res = r.autotune(w, frames_per_trial=1, min_psnr_db=40.0)
    trials = res["trials"]
    print(f"[autotune keep backend {backend}] {len(trials)} trials, best {res['best']}", flush=True)
    assert len(trials) == 15  # reference, 4 froxel candidates, 2 other slice counts, 8 launch shapes of the best
    assert all(t["fine_binning_backend"] == backend for t in trials)
    assert {(t["camera_froxel"], t["light_froxel"]) for t in trials[1:5]} == {(c, l) for c in (8, 16) for l in (8, 16)}
    assert res["best"]["accepted"] == 1 and res["best"]["fine_binning_backend"] == backend


def telemetry_from_oracle(w, ref, page_cap):
    """finalize_telemetry's words (strand_prepass.wgsl:2145-2188) and the
    per-frustum words (:736-737, 1085, 1136, 2830-2832), restated from the
    oracle's canonical dumps: a page's candidates are the distinct segments
    among its references (each (task, page) visit writes at least one)."""
    lib = oracle.load()
    F = len(w.frusta)
    fr = w.frusta
    ctr = ref.counters()
    pm, refs, tasks = ref.get("page_meta"), ref.get("refs"), ref.get("tasks")
    t = np.zeros(40 + page_cap + 11 * F, np.uint64)

    def frustum_of_tile(tile):
        for fi in range(F):
            b = int(fr[fi]["coarse_depth_tile_base"])
            if b <= tile < b + int(fr[fi]["coarse_depth_tile_count"]):
                return fi
        return None

    def fw(fi, field):
        return 40 + page_cap + 11 * fi + field

    P = len(pm)
    t[0], t[1], t[3] = P, ctr["binning_tasks"], ctr["fine_refs"]
    for p in range(P):
        b, n = int(pm["seg_ref_base"][p]), int(pm["seg_ref_count"][p])
        cand = len({(int(i), int(s)) for i, s in zip(refs["inst_id"][b:b + n], refs["seg_id"][b:b + n])})
        t[40 + p] = cand
        fi = frustum_of_tile(int(pm["coarse_tile_id"][p]))
        if fi is not None:
            t[fw(fi, 4)] += 1
        if cand == 0:
            continue
        t[2] += cand
        t[4] += 1
        t[5] = max(t[5], cand)
        if fi is not None:
            t[fw(fi, 5)] += cand
            t[fw(fi, 6)] += n
            t[fw(fi, 9)] = max(t[fw(fi, 9)], cand)
        t[8 + min(cand.bit_length() - 1, 31)] += 1
    for run in np.concatenate([ref.get("camera_runs"), ref.get("shadow_runs")]):
        fi = int(run["frustum_id"])
        t[fw(fi, 7)] += 1
        t[fw(fi, 8)] += int(run["load_score"])
        t[fw(fi, 10)] = max(t[fw(fi, 10)], int(run["load_score"]))
    task_f = tasks["packed_field"] >> abi.BINNING_FRUSTUM_SHIFT
    for fi in range(F):
        t[fw(fi, 3)] = int((task_f == fi).sum())
    broad, keep, ranges = ref.get("broad_meta"), ref.get("keep"), ref.get("ranges")
    light_layer, nl = {}, len(w.lights)
    for fi in range(F):
        if fr[fi]["kind"] == 1:
            light_layer[fi] = len(light_layer)
    for ii in range(len(broad)):
        b = broad[ii]
        if not b["visible"]:
            continue
        sc = int(b["strand_count"])
        rnd = np.array([lib.oracle_hash_to_unit_float(lib.oracle_wang_hash(((ii * 16777619) & 0xFFFFFFFF) ^ s))
                        for s in range(sc)], np.float32)
        mx = max(float(b["camera_keep_probability"]), float(b["shadow_keep_probability"]))
        for fi in range(F):
            if int(ranges[ii * F + fi]["geo_id"]) == ii:
                t[fw(fi, 0)] += 1
                t[fw(fi, 1)] += sc
            if fr[fi]["kind"] > 1 or (fr[fi]["kind"] == 1 and light_layer[fi] >= nl):
                continue
            t[fw(fi, 2)] += int(((rnd <= np.float32(mx)) & (rnd <= keep[ii * F + fi])).sum())
    return t


@pytest.mark.parametrize("backend", [0, 1])
def test_telemetry_matches_oracle(backend):
    w = scene.config_small(froxel=8)
    w.settings["telemetry_enabled"] = 1
    w.settings["fine_binning_backend"] = backend
    ref = oracle.render(w)
    with StrandRenderer() as r:
        r.upload_workload(w)
        r.render(w)
        tel = r.dump("telemetry")
    page_cap = len(tel) - 40 - 11 * len(w.frusta)
    want = telemetry_from_oracle(w, ref, page_cap)
    assert tel.shape == want.shape
    bad = np.nonzero(tel.astype(np.uint64) != want)[0]
    assert bad.size == 0, f"telemetry words differ at {bad[:10]}: gpu {tel[bad[:10]]} oracle {want[bad[:10]]}"
    assert tel[2] > 0 and tel[5] >= tel[2] / max(tel[4], 1)


def test_shadow_history_across_frames():
    """Three frames under a moving light with shadow history on: every frame
    matches the oracle fed the oracle's previous history
    (strand_shading.wgsl:660-679; history ping-pong, plugin.rs:1120-1140)."""
    base = scene.config_small(froxel=8)
    asset = base.assets[0]
    center, radius = scene.bounding_sphere(asset)
    dirs = [(0.35, 0.85, 0.4), (-0.5, 0.75, 0.45), (0.1, 0.6, -0.8)]
    frames = []
    for k, d in enumerate(dirs):
        w = scene.config_small(froxel=8)
        d = np.asarray(d) / np.linalg.norm(d)
        w.lights["direction_to_light"][0, :3] = d
        w.views[1] = scene.light_view(d, center, radius)
        w.settings["shadow_history_valid"] = 1
        w.name += f"-light{k}"
        frames.append(w)
    prev = None
    with StrandRenderer() as r:
        r.upload_workload(frames[0])
        for k, w in enumerate(frames):
            print(f"[parity] history frame {k}", flush=True)
            ctr = r.render(w)
            gpu = {n: r.dump(n) for n in EXACT + FLOAT}
            ref = oracle.render(w, history_valid=True, prev_history=prev)
            compare(w, ctr, gpu, ref)
            prev = ref.get("history")
            if k > 0:
                h = gpu["history"].reshape(-1, 2)
                assert (np.abs(h[:, 0] - h[:, 1]) > 1e-4).any(), "the filter blends in the previous frame"


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("depth", [0.9, 0.2])
def test_parity_opaque_occlusion(depth, backend):
    """Opaque depth reduce (K1) and the fine-tile occlusion of every binning
    walk, against the oracle: an occluder in front of all strands and one
    cutting through them."""
    w = scene.config_small(froxel=8)
    w.opaque = scene.opaque_disc(w, depth)
    w.settings["fine_binning_backend"] = backend
    w.name += f"-opaque{depth}-b{backend}"
    print(f"[parity] {w.name}", flush=True)
    ref = oracle.render(w)
    with StrandRenderer() as r:
        r.upload_workload(w)
        ctr = r.render(w)
        gpu = {k: r.dump(k) for k in EXACT + FLOAT + ["opaque_tiles"]}
    compare(w, ctr, gpu, ref)
    assert_exact("opaque_tiles", gpu["opaque_tiles"], ref.get("opaque_tiles"))
    assert ctr["fine_refs"] < oracle.render(scene.config_small(froxel=8)).counters()["fine_refs"]
