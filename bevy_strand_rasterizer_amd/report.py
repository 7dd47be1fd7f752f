"""Benchmark reports in the reference bench's JSON schema.

Restates the report side of examples/strand_bench.rs: ``summarize``
(:931-947, nearest-rank percentiles at index round((n - 1) f) with Rust's
round-half-away-from-zero), ``histogram_percentile_upper`` (:911-929), the
telemetry samples and their derived values (:700-788) and ``make_report``
(:859-909, schema_version 1).  Telemetry words come from the runtime's
telemetry dump (finalize_telemetry layout).
"""
from __future__ import annotations

import math

import numpy as np

TELEMETRY_FIELDS = ("allocated_pages", "binning_tasks", "page_candidates", "fine_cell_hits",
                    "active_candidate_pages", "max_candidates_per_page")  # prepass.rs:1345-1353
FRUSTUM_FIELDS = ("visible_instances", "visible_strands", "retained_strands", "emitted_segments", "allocated_pages",
                  "page_candidates", "fine_refs", "raster_runs", "raster_load", "max_page_candidates",
                  "max_raster_load")  # prepass.rs:1362-1374
HIST_BINS, HIST_OFF, PAGE_OFF, FRUSTUM_STRIDE = 32, 8, 40, 11


def nearest_rank(sorted_values, fraction: float):
    """sorted[round((n - 1) * fraction)] with round half away from zero."""
    return sorted_values[int(math.floor((len(sorted_values) - 1) * fraction + 0.5))]


def summarize(values) -> dict:
    s = sorted(float(v) for v in values)
    return {"samples": len(s), "min": s[0], "mean": sum(s) / len(s), "median": nearest_rank(s, 0.5),
            "p90": nearest_rank(s, 0.9), "p95": nearest_rank(s, 0.95), "p99": nearest_rank(s, 0.99), "max": s[-1]}


def histogram_percentile_upper(histogram, fraction: float):
    """Upper bound of the log2 bin holding the `fraction` quantile."""
    total = int(sum(int(c) for c in histogram))
    if total == 0:
        return None
    target = max(int(math.ceil(total * fraction)), 1)
    cumulative = 0
    for b, c in enumerate(histogram):
        cumulative += int(c)
        if cumulative >= target:
            return 0xFFFFFFFF if b == 31 else (1 << (b + 1)) - 1
    return None


def telemetry_samples(tel: np.ndarray, frustum_count: int) -> dict:
    """One frame's telemetry diagnostics and derived values, by name."""
    tel = np.asarray(tel, np.uint64)
    page_cap = len(tel) - PAGE_OFF - FRUSTUM_STRIDE * frustum_count
    out = {name: float(tel[i]) for i, name in enumerate(TELEMETRY_FIELDS)}
    hist = [int(x) for x in tel[HIST_OFF:HIST_OFF + HIST_BINS]]
    for b in range(HIST_BINS):
        out[f"candidate_histogram/{b}"] = float(hist[b])
    base = PAGE_OFF + page_cap
    for fi in range(frustum_count):
        for k, name in enumerate(FRUSTUM_FIELDS):
            out[f"frustum/{fi}/{name}"] = float(tel[base + fi * FRUSTUM_STRIDE + k])
    derived = {}
    if out["active_candidate_pages"] > 0:
        derived["candidates_per_active_page/mean"] = out["page_candidates"] / out["active_candidate_pages"]
    for name, f in (("p50", 0.50), ("p90", 0.90), ("p95", 0.95), ("p99", 0.99)):
        ub = histogram_percentile_upper(hist, f)
        if ub is not None:
            derived[f"candidates_per_active_page/{name}_upper_bound"] = float(ub)
    for fi in range(frustum_count):
        vis, ret = out[f"frustum/{fi}/visible_strands"], out[f"frustum/{fi}/retained_strands"]
        if vis > 0:
            derived[f"frustum/{fi}/retention_ratio"] = ret / vis
    return {"telemetry": out, "derived": derived}


def make_report(case: str, width: int, height: int, warmup_frames: int, sample_frames: int, telemetry_enabled: bool,
                backend: str, seed: int, workload: dict, adapter: dict, timing_samples: dict,
                telemetry_frames: list) -> dict:
    """The reference bench's report: timings_ms from per-stage GPU samples,
    telemetry (and derived) summaries over the sampled frames."""
    timings = {f"{k}/elapsed_gpu": summarize(v) for k, v in sorted(timing_samples.items()) if v}
    tele: dict = {}
    for frame in telemetry_frames:
        for group in ("telemetry", "derived"):
            for k, v in frame[group].items():
                tele.setdefault(k, []).append(v)
    return {
        "schema_version": 1,
        "case": case,
        "mode": "headless",
        "resolution": {"width": width, "height": height},
        "warmup_frames": warmup_frames,
        "sample_frames": sample_frames,
        "telemetry_enabled": telemetry_enabled,
        "fine_binning_backend": {"scatter": "segment", "csr": "page-csr"}[backend],
        "seed": seed,
        "workload": workload,
        "adapter": adapter,
        "timings_ms": timings,
        "telemetry": {k: summarize(v) for k, v in sorted(tele.items())},
    }
