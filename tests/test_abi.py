"""ABI: every struct of include/strand_rt.h has the reference's byte layout,
the numpy/ctypes mirrors agree with the C compiler, and the built library
exports every function the header declares (no compute calls)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from bevy_strand_rasterizer_amd import abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "strand_rt.h"

# (C struct, numpy dtype, reference size in bytes, source of the size)
STRUCTS = [
    ("sr_push_constants", abi.push_constants_dtype, 44, "src/shader_types.rs:9-22"),
    ("sr_strand_meta", abi.strand_meta_dtype, 16, "src/shader_types.rs:26-31"),
    ("sr_aabb", abi.aabb_dtype, 32, "src/shader_types.rs:46-51"),
    ("sr_strand_geo", abi.strand_geo_dtype, 64, "src/shader_types.rs:66-72"),
    ("sr_strand_instance", abi.strand_instance_dtype, 160, "src/shader_types.rs:77-88"),
    ("sr_strand_material", abi.strand_material_dtype, 64, "src/components.rs:61-72"),
    ("sr_device_ptr", abi.device_ptr_dtype, 12, "src/shaders/types.wgsl:23-27"),
    ("sr_froxel_config", abi.froxel_config_dtype, 20, "src/components.rs:93-100"),
    ("sr_frustum_desc", abi.frustum_desc_dtype, 64, "src/plugin.rs:1545-1562"),
    ("sr_binning_task", abi.binning_task_dtype, 16, "src/pipelines/task_contract.rs:70-75"),
    ("sr_projected_segment", abi.projected_segment_dtype, 12, "src/pipelines/task_contract.rs:86"),
    ("sr_raster_work_item", abi.raster_work_item_dtype, 32, "src/pipelines/task_contract.rs:90-99"),
    ("sr_raster_tile_run", abi.raster_tile_run_dtype, 32, "src/pipelines/task_contract.rs:103-112"),
    ("sr_fine_seg_ref", abi.fine_seg_ref_dtype, 12, "src/pipelines/task_contract.rs:122"),
    ("sr_fine_page_meta", abi.fine_page_meta_dtype, 32, "src/pipelines/task_contract.rs:126-135"),
    ("sr_coarse_asset_range", abi.coarse_asset_range_dtype, 32, "src/plugin.rs:1575-1584"),
    ("sr_coarse_interval_ref", abi.coarse_interval_ref_dtype, 16, "src/plugin.rs:1588-1593"),
    ("sr_coarse_depth_lut_entry", abi.coarse_depth_lut_entry_dtype, 16, "src/plugin.rs:1566-1571"),
    ("sr_broad_instance_meta", abi.broad_instance_meta_dtype, 16, "src/shaders/strand_prepass.wgsl:193-198"),
    ("sr_view", abi.view_dtype, 80, "per-frustum view (new)"),
    ("sr_directional_light", abi.light_dtype, 32, "directional light (new)"),
    ("sr_frame_counters", abi.frame_counters_dtype, 64, "frame counters (new)"),
    ("sr_strand_cache_info", abi.strand_cache_info_dtype, 72, "strand cache parse result (new)"),
    ("sr_request", abi.request_dtype, 912, "serving request (new)"),
    ("sr_request_result", abi.request_result_dtype, 112, "serving result (new)"),
    ("sr_server_config", abi.server_config_dtype, 16, "serving config (new)"),
    ("sr_server_stats", abi.server_stats_dtype, 64, "serving stats (new)"),
    ("sr_tune_trial", abi.tune_trial_dtype, 40, "autotuner trial (new)"),
    ("sr_tune_result", abi.tune_result_dtype, 40 + 8 + 24 * 40, "autotuner result (new)"),
]


# numpy field -> C member expression where the C struct uses an array
C_MEMBER = {}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """Offsets and sizes reported by gcc for every field of every struct."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, dt, _, _ in STRUCTS:
        lines.append(f'  printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for f in dt.names:
            member = C_MEMBER.get((cname, f), f)
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {member}));')
    for f, *_ in abi.FrameDesc._fields_:
        lines.append(f'  printf("sr_frame_desc {f} %zu\\n", offsetof(sr_frame_desc, {f}));')
    lines.append('  printf("sr_frame_desc __sizeof__ %zu\\n", sizeof(sr_frame_desc));')
    lines.append("  return 0; }")
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for ln in out.splitlines():
        name, field, val = ln.split()
        table[(name, field)] = int(val)
    return table


@pytest.mark.parametrize("cname,dt,size,src", STRUCTS, ids=[s[0] for s in STRUCTS])
def test_struct_layout(c_layout, cname, dt, size, src):
    assert c_layout[(cname, "__sizeof__")] == size, f"{cname} must be {size} bytes ({src})"
    assert dt.itemsize == size
    for f in dt.names:
        assert c_layout[(cname, f)] == dt.fields[f][1], f"{cname}.{f} offset"


def test_ctypes_mirrors_match_header(c_layout):
    import ctypes as C

    assert C.sizeof(abi.FrameSettings) == 88
    assert C.sizeof(abi.DevicePtr) == 12
    assert C.sizeof(abi.FrameDesc) == c_layout[("sr_frame_desc", "__sizeof__")]
    for name, *_ in abi.FrameDesc._fields_:
        assert getattr(abi.FrameDesc, name).offset == c_layout[("sr_frame_desc", name)], name


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    decl = declared_functions()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), f"libstrand_rt.so does not export {name}"
    assert set(decl) == set(abi.exported_symbols()), "abi.py signatures out of sync with the header"


def test_version_string():
    lib = abi.load_library()
    assert b"gfx950" in lib.sr_version()


def test_flag_constants_match_header():
    """The SR_CONFIG_*, SR_FRAME_* and SR_REQUEST_* bits of abi.py are the
    header's, and no two flags of one family share a bit."""
    text = HEADER.read_text()
    for family in ("CONFIG", "FRAME", "REQUEST"):
        defs = {m.group(1): int(m.group(2)) for m in re.finditer(rf"#define SR_{family}_(\w+) (\d+)u\b", text)}
        assert defs, family
        for name, value in defs.items():
            assert getattr(abi, f"{family}_{name}") == value, f"SR_{family}_{name}"
        assert len(set(defs.values())) == len(defs), family


def test_binning_field_packing():
    # task_contract.rs:137-157: level bits 0-1, shadow bit 2, frustum bits 3-31
    packed = (0 & 3) | (1 << 2) | (37 << 3)
    assert packed >> abi.BINNING_FRUSTUM_SHIFT == 37
    assert (packed >> 2) & 1 == 1


def test_frustum_table_builder():
    from bevy_strand_rasterizer_amd import scene

    t = scene.frustum_table([(1920, 1080, 8, 8, 16), (2048, 2048, 8, 8, 16)], [0, 1])
    # plugin.rs:386-458: fine tiles ceil(W/8), coarse = ceil(fine/4), running bases
    assert t[0]["coarse_tiles_x"] == 60 and t[0]["coarse_tiles_y"] == 34
    assert t[0]["coarse_depth_tile_count"] == 2040
    assert t[1]["coarse_depth_tile_base"] == 2040 and t[1]["coarse_depth_tile_count"] == 4096
    assert t[1]["fine_depth_tile_base"] == 240 * 135
    assert t[1]["bucket_base"] == 240 * 135 * 16
    assert t[1]["kind"] == 1 and t[0]["kind"] == 0


def dump_kinds():
    """sr_dump_kind enumerators of the header (name -> value)."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef enum sr_dump_kind \{(.*?)\} sr_dump_kind;", text, flags=re.S).group(1)
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"SR_DUMP_([A-Z0-9_]+)\s*=\s*(\d+)", body)}


def test_dump_kind_count_exceeds_every_kind():
    kinds = dump_kinds()
    count = kinds.pop("KIND_COUNT")
    assert count == max(kinds.values()) + 1
    assert sorted(kinds.values()) == list(range(count)), "dump kinds are dense and listed once"
    assert list(kinds.values()) == sorted(kinds.values()), "enumerators in numeric order"
    assert {k.lower(): v for k, v in kinds.items()} == abi.DUMP


def test_integration_binding_declares_every_entry_point():
    """The Rust extern block a maintainer would add names
    every function the header exports."""
    text = (ROOT / "INTEGRATION.md").read_text()
    block = re.search(r'#\[link\(name = "strand_rt"\)\]\s*extern "C" \{(.*?)\n\}', text, flags=re.S).group(1)
    bound = set(re.findall(r"pub fn (sr_[a-z_0-9]+)\(", block))
    assert bound == set(declared_functions())
