"""Report restatements of examples/strand_bench.rs (summarize :931-947,
histogram_percentile_upper :911-929, derived telemetry :757-788) against
hand-computed answers, and the bench's percentile helper."""
import numpy as np

import bench
from bevy_strand_rasterizer_amd import report


def test_summarize_nearest_rank_rounds_half_away_from_zero():
    # n = 18: (n - 1) * 0.5 = 8.5 -> index 9 (Rust f64::round), not 8 (banker's)
    v = list(range(18, 0, -1))
    s = report.summarize(v)
    assert s["median"] == 10.0 and s["min"] == 1.0 and s["max"] == 18.0 and s["samples"] == 18
    assert s["p90"] == 16.0   # 15.3 -> 15 -> value 16
    assert s["p95"] == 17.0   # 16.15 -> 16 -> 17
    assert s["p99"] == 18.0   # 16.83 -> 17 -> 18
    assert abs(s["mean"] - 9.5) < 1e-12
    assert bench.percentile(v, 0.5) == 10


def test_histogram_percentile_upper():
    h = [0] * 32
    h[0], h[3], h[5] = 10, 80, 10       # 100 pages: 1, 8..15, 32..63 candidates
    assert report.histogram_percentile_upper(h, 0.05) == 1
    assert report.histogram_percentile_upper(h, 0.50) == 15
    assert report.histogram_percentile_upper(h, 0.90) == 15
    assert report.histogram_percentile_upper(h, 0.91) == 63
    assert report.histogram_percentile_upper([0] * 32, 0.5) is None
    h31 = [0] * 32
    h31[31] = 1
    assert report.histogram_percentile_upper(h31, 0.5) == 0xFFFFFFFF


def test_telemetry_samples_and_report_schema():
    F, page_cap = 2, 5
    tel = np.zeros(40 + page_cap + 11 * F, np.uint32)
    tel[:6] = [4, 100, 30, 200, 3, 15]
    tel[8 + 3] = 2   # two pages with 8..15 candidates
    tel[8 + 2] = 1   # one page with 4..7
    base = 40 + page_cap
    tel[base + 1], tel[base + 2] = 1000, 250          # frustum 0: visible / retained strands
    tel[base + 11 + 1], tel[base + 11 + 2] = 0, 0     # frustum 1: nothing visible
    s = report.telemetry_samples(tel, F)
    assert s["telemetry"]["page_candidates"] == 30 and s["telemetry"]["frustum/0/retained_strands"] == 250
    assert s["derived"]["candidates_per_active_page/mean"] == 10.0
    assert s["derived"]["candidates_per_active_page/p50_upper_bound"] == 15.0
    assert s["derived"]["frustum/0/retention_ratio"] == 0.25 and "frustum/1/retention_ratio" not in s["derived"]
    r = report.make_report("named", 1920, 1080, 5, 3, True, "csr", 7, {"strand_count": 1}, {"name": "x"},
                           {"raster": [1.0, 2.0, 3.0]}, [s, s, s])
    assert r["schema_version"] == 1 and r["fine_binning_backend"] == "page-csr" and r["mode"] == "headless"
    assert r["timings_ms"]["raster/elapsed_gpu"]["median"] == 2.0
    assert r["telemetry"]["candidates_per_active_page/mean"]["samples"] == 3
