"""The rasterizers' conservative pixel spans never drop a covered pixel.

Builds tests/span_check.cpp for the host (hipcc, no GPU needed) and runs it
over random and adversarial segments: zero, sub-pixel and very long
lengths, axis-aligned and nearly axis-aligned directions, endpoints and
radii on pixel centres, tiles near 16k pixels, 4x4 / 8x8 / 16x8 / 16x16
tiles.  Every pixel that the shadow or camera exact coverage test accepts
must be in the span mask (a miss would change rendered results).  Also
checks the kernels' exact 16-bit unorm decode (tests/unorm_check.cpp).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def span_check(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("spans") / "span_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "span_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def test_spans_keep_every_covered_pixel(span_check):
    out = subprocess.run([str(span_check), "400000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    cases, misses, kept, covered = map(int, out.stdout.split()[-4:])
    # exact reciprocal / square root, and both moved 2 ulp down and up (the
    # kernels use the hardware's, within 1 ulp)
    assert cases == 3 * 700000 and misses == 0
    assert covered > 3_000_000
    # the spans are tight: at most twice the covered pixels are kept
    assert kept < 2 * covered


def test_unorm16_matches_division(tmp_path):
    """The kernels' i / 65535 and q / 16777215 (FMA-corrected products) equal
    the correctly rounded quotients for all 65,536 values of a 16-bit unorm
    and all 2^24 quantised depths."""
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "unorm_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "unorm_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["0"], out.stdout

def test_even_slice_matches_the_slice_scan(tmp_path):
    """The marking pass's arithmetic LUT slice of a depth in an evenly split
    single-interval tile equals the first of build_depth_warp_lut's 8, 16 or 32
    slices
    holding it, for reciprocals of the span within 3 ulp
    (tests/even_slice_check.cpp)."""
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "even_slice_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "even_slice_check.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["0"], out.stdout
