"""The .strands cache loader (src/strand_cache.rs) through the C ABI: the
reference's own known-answer bytes, its error messages, the JSON form, the
converter's layout (tools/alembic_to_strands.py:146-162) and the packing of
set_cached_strand_geometry (src/plugin.rs:1468-1490).  Host code only."""
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import scene
from bufcmp import assert_same

GOLDEN = Path(__file__).resolve().parent / "golden"


def kat():
    return (GOLDEN / "strand_cache_kat.strands").read_bytes()


def test_reference_known_answer():
    """parses_binary_cache (src/strand_cache.rs:254-281)."""
    data = kat()
    want = json.loads((GOLDEN / "strand_cache_kat.expected.json").read_text())
    assert len(data) == 8 + 4 + 4 + 24 + 3 * 12 + 8 + 12
    info, verts, meta, idx = scene.parse_strand_cache(data)
    assert int(info["version"]) == 1 and int(info["format"]) == 0
    assert int(info["vertex_count"]) == 3 and int(info["strand_count"]) == 1 and int(info["index_count"]) == 3
    assert verts[:, :3].tolist() == want["vertices"]
    assert (verts[:, 3] == 1.0).all()
    assert [[int(m["count"]), int(m["offset"])] for m in meta] == want["strands"]
    assert idx.tolist() == want["indices"]
    assert int(meta[0]["material_idx"]) == 0
    assert int(info["max_segments_in_strand"]) == 3
    assert list(info["aabb"]["min"]) == [0.0, 0.0, 0.0] and list(info["aabb"]["max"]) == [1.0, 1.0, 0.0]


def test_writer_layout_matches_known_answer():
    assert scene.strand_cache_bytes([[0, 0, 0], [1, 0, 0], [1, 1, 0]], [[3, 0]], [0, 1, 2]) == kat()


@pytest.mark.parametrize("cut", [0, 4, 8, 11, 20, 40, len(kat()) - 1])
def test_truncated(cut):
    data = kat()[:cut]
    with pytest.raises(scene.StrandCacheError) as e:
        scene.parse_strand_cache(data)
    if cut >= 8:  # with the magic present the binary parser reports the short read
        assert "unexpected end of strand cache" in str(e.value)


def test_trailing_bytes():
    with pytest.raises(scene.StrandCacheError, match="strand cache has 3 trailing bytes after payload"):
        scene.parse_strand_cache(kat() + b"\0\0\0")


def test_version_and_magic():
    bad = bytearray(kat())
    bad[8:12] = struct.pack("<I", 2)
    with pytest.raises(scene.StrandCacheError, match="unsupported strand cache version 2"):
        scene.parse_strand_cache(bytes(bad))
    # without the magic the loader reads JSON (parse_cache, :165-171)
    with pytest.raises(scene.StrandCacheError, match="JSON"):
        scene.parse_strand_cache(b"SSRSTRX\0" + kat()[8:])


def test_validation_messages():
    v = [[0, 0, 0], [1, 0, 0], [1, 1, 0]]
    with pytest.raises(scene.StrandCacheError, match="strand 0 has fewer than two vertices"):
        scene.parse_strand_cache(scene.strand_cache_bytes(v, [[1, 0]], [0]))
    with pytest.raises(scene.StrandCacheError, match=r"strand 0 range 1\.\.4 exceeds 3 cache indices"):
        scene.parse_strand_cache(scene.strand_cache_bytes(v, [[3, 1]], [0, 1, 2]))
    with pytest.raises(scene.StrandCacheError, match="strand 1 references vertex 7, but cache has 3 vertices"):
        scene.parse_strand_cache(scene.strand_cache_bytes(v, [[2, 0], [2, 1]], [0, 1, 7]))
    with pytest.raises(scene.StrandCacheError, match="strand 0 index range overflows"):
        scene.parse_strand_cache(scene.strand_cache_bytes(v, [[2, 0xFFFFFFFF]], [0, 1]))


def test_json_equals_binary():
    doc = {"version": 1, "source": "kat.abc", "scale": 1.0, "extra": {"ignored": [1, 2]},
           "vertices": [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]], "strands": [[0, 1, 2]]}
    jb = json.dumps(doc).encode()
    a = scene.parse_strand_cache(jb)
    b = scene.parse_strand_cache(kat())
    assert int(a[0]["format"]) == 1
    for x, y in zip(a[1:], b[1:]):
        assert_same(x, y, "JSON and binary cache slabs")
    # scale defaults to 1 (unit_scale, :42-44); source may be null
    doc.pop("scale")
    doc["source"] = None
    assert_same(scene.parse_strand_cache(json.dumps(doc).encode())[1], b[1], "vertices without scale or source")
    with pytest.raises(scene.StrandCacheError, match="missing field `strands`"):
        scene.parse_strand_cache(b'{"version": 1, "vertices": []}')


def test_scale_and_aabb_and_packing():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(40, 3)).astype(np.float32)
    strands = [[5, 0], [7, 5], [2, 12]]          # (count, offset) over an index list
    idx = rng.integers(0, 40, size=14).astype(np.uint32)
    scale = 0.01
    info, verts, meta, out_idx = scene.parse_strand_cache(scene.strand_cache_bytes(v, strands, idx, scale))
    # Vec3 * scale in f32, (x, y, z, 1)
    np.testing.assert_array_equal(verts[:, :3], v * np.float32(scale))
    # Aabb3d::from_point_cloud over every vertex, referenced or not
    np.testing.assert_array_equal(info["aabb"]["min"], (v * np.float32(scale)).min(axis=0))
    np.testing.assert_array_equal(info["aabb"]["max"], (v * np.float32(scale)).max(axis=0))
    assert int(info["max_segments_in_strand"]) == 7
    assert meta["count"].tolist() == [5, 7, 2] and meta["offset"].tolist() == [0, 5, 12]
    assert out_idx.tolist() == idx.tolist()


def test_asset_from_cache_has_one_material():
    v = [[0, 0, 0], [0, 0.1, 0], [0, 0.2, 0.01], [0.05, 0, 0], [0.05, 0.1, 0]]
    a = scene.asset_from_cache(scene.strand_cache_bytes(v, [[3, 0], [2, 3]], [0, 1, 2, 3, 4]),
                               radius_px=1.5)
    assert a.materials.shape == (1,) and (a.meta["material_idx"] == 0).all()
    assert float(a.materials[0]["min_radius_pixels"]) == 1.5
    assert int(a.geo[0]["strand_count"]) == 2 and int(a.geo[0]["max_segments_in_strand"]) == 3
    assert a.segment_count == 3
