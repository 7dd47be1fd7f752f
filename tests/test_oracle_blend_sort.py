"""Depth-sorted blending inside a work item (SR_FRAME_DEPTH_SORTED_BLEND) in
the CPU restatement, pinned by closed forms on two strands that share every
cell they cover (depths 0.6 and one float step nearer, which quantise alike):

* the default order, (inst, seg), blends the listed-first (farther) strand
  first; the sorted order blends the nearer one first; each pixel equals the
  front-to-back "over" of the two fragments in that order, computed here
  from the strands' colours and alphas;
* the sorted frame equals the default frame of the same strands listed
  nearest first, byte for byte, and does not depend on the listing order.
"""
import dataclasses

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi, scene
from bufcmp import assert_same
from oracle import oracle

RED = np.array([0.9, 0.1, 0.1]), 0.6   # depth 0.6
BLUE = np.array([0.1, 0.2, 0.9]), 0.7  # one float step nearer


def sorted_blend(w):
    st = dict(w.settings)
    st["frame_flags"] = int(st.get("frame_flags", 0)) | abi.FRAME_DEPTH_SORTED_BLEND
    return dataclasses.replace(w, name=w.name + "-sorted", settings=st)


def over(first, second):
    """Front-to-back premultiplied over of two fully covering fragments, unpremultiplied."""
    (c1, a1), (c2, a2) = first, second
    a = a1 + a2 * (1.0 - a1)
    return (c1 * a1 + c2 * a2 * (1.0 - a1)) / a, a


@pytest.fixture(scope="module")
def frames():
    a, b = scene.config_one_cell_pair(False), scene.config_one_cell_pair(True)
    return {k: oracle.render(w, threads=2) for k, w in
            (("a", a), ("b", b), ("a_sorted", sorted_blend(a)), ("b_sorted", sorted_blend(b)))}


def pixels(ref):
    # pixels of row 20 well inside both strands (x = -0.5 .. 0.5 covers columns 16..47)
    return ref.get("image").reshape(64, 64, 4)[20, 24:40]


def test_pair_shares_cells(frames):
    cc = frames["a"].get("cell_counts")
    assert cc.max() == 2 and (cc == 2).sum() >= 2


@pytest.mark.parametrize("key,order", [("a", (RED, BLUE)), ("b", (BLUE, RED)), ("a_sorted", (BLUE, RED)),
                                       ("b_sorted", (BLUE, RED))])
def test_pixels_are_the_closed_form_over(frames, key, order):
    rgb, a = over(*order)
    px = pixels(frames[key])
    assert np.abs(px[:, :3] - rgb).max() < 2e-6 and np.abs(px[:, 3] - a).max() < 2e-6


def test_sorted_equals_default_listed_nearest_first(frames):
    for k in ("image", "depth", "composite"):
        assert_same(frames["a_sorted"].get(k), frames["b"].get(k), f"sorted vs nearest-first default: {k}")
        assert_same(frames["a_sorted"].get(k), frames["b_sorted"].get(k), f"sorted, either listing: {k}")
    assert np.abs(frames["a"].get("image") - frames["b"].get("image")).max() > 0.1
