"""Depth-sorted blending inside a work item (SR_FRAME_DEPTH_SORTED_BLEND) on
the HIP path.

* Two strands that share every cell they cover: the sorted frame equals the
  oracle's and, byte for byte, the default frame of the same strands listed
  nearest first.
* Scalp frames (8x8 and 16x16 tiles, and the named 100k x 32 frame, where
  many pixels hold more fragments per cell than the sorted list and take the
  list's overflow rule) and ten stacked scalps at 32 depth-warp slices (tile
  runs above 256 work items) against the oracle at the parity bar.
"""
import numpy as np
import pytest

from bevy_strand_rasterizer_amd import scene
from bevy_strand_rasterizer_amd.runtime import StrandRenderer
from bufcmp import assert_same
from oracle import oracle
from test_gpu_parity import EXACT, FLOAT, compare
from test_oracle_blend_sort import sorted_blend

pytestmark = pytest.mark.gpu


def render(w, tries=4):
    with StrandRenderer() as r:
        r.upload_workload(w)
        for _ in range(tries):
            ctr = r.render(w)
            if not ctr["capacity_exceeded"]:
                break
        assert ctr["capacity_exceeded"] == 0
        return ctr, {k: r.dump(k) for k in EXACT + FLOAT}


def test_pair_sorted_matches_oracle_and_nearest_first_default():
    a = sorted_blend(scene.config_one_cell_pair(False))
    ctr, gpu = render(a)
    compare(a, ctr, gpu, oracle.render(a))
    _, nearest_first = render(scene.config_one_cell_pair(True))
    for k in ("image", "depth", "composite"):
        assert_same(gpu[k], nearest_first[k], f"sorted vs nearest-first default: {k}")


def depth_stack_32():
    """Ten stacked scalps at 32 depth-warp slices: tile runs of more than
    256 work items."""
    from test_gpu_parity import depth_stack

    w = scene.retile(depth_stack(), backend=1)
    w.settings["coarse_depth_slices"] = 32
    return w


@pytest.mark.parametrize("which", ["small8", "small16", "named", "stack32"])
def test_sorted_blend_matches_oracle(which):
    w = {"small8": lambda: scene.retile(scene.config_small(froxel=8), backend=1),
         "small16": lambda: scene.config_small(froxel=16),
         "named": lambda: scene.retile(scene.config_named(), backend=1),
         "stack32": depth_stack_32}[which]()
    w = sorted_blend(w)
    print(f"[sorted] {w.name}", flush=True)
    ctr, gpu = render(w)
    ref = oracle.render(w)
    compare(w, ctr, gpu, ref)
