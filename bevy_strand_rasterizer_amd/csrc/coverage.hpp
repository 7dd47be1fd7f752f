// coverage.hpp -- conservative pixel spans of a strand segment over a tile.
//
// The rasterizers' exact coverage tests (strand_rasterizer.wgsl:138-147,
// 553-566, strand_shadow_raster.wgsl) accept a pixel centre only when its
// projection parameter t = dot(p - p0, p1 - p0) / |p1 - p0|^2 lies in
// [0, 1] and its distance to the point at t is below the radius, which is
// at most rmax.  Such a pixel lies in the rectangle
//   dot(p - p0, ab)   in [0, |ab|^2]          (along the segment)
//   cross(p - p0, ab) in [-rmax |ab|, rmax |ab|]  (across it)
// inside the segment's box grown by rmax.  For each pixel row of a tile the
// rectangle is one interval of pixel centres; the masks below set every
// pixel of those intervals.  All limits are widened by a pad that exceeds
// the f32 rounding of both the exact tests and this computation, so a
// pixel outside the mask is never covered -- skipping it changes no result.
#pragma once

#include "common.hpp"

namespace srk {

// Largest radius the exact tests can use along a segment: they evaluate
// r = mix(minr, maxr, clamp(z, 0, 1)) at a depth z between the endpoint
// depths, and r grows with z (minr <= maxr), so the deeper endpoint bounds
// it (rounding is absorbed by the span pad).
SR_HD float span_radius(float minr, float maxr, float zdeep01) {
  return wmix(minr, maxr, wclamp(zdeep01, 0.0f, 1.0f));
}

// The largest normalize_depth01 over raw depths in [za, zb] (camera
// segments): it is the identity on [0, 1] and z / 2 + 1/2 clamped outside,
// so below 0 it approaches 1/2.
SR_HD float span_depth01_max(float za, float zb) {
  const float lo = fminf(za, zb), hi = fmaxf(za, zb);
  const float n = normalize_depth01(hi);
  return lo < 0.0f ? fmaxf(n, 0.5f) : n;
}

struct SpanRef {
  float x0, y0, abx, aby;
  float bu0, bu1, bv0, bv1;  // box of p - p0, grown by the padded radius
  float dlo, dhi, c;         // padded along / across limits
  float ia, ib;              // 1 / abx, 1 / aby (unused when zero)
  float pad;
};

SR_HD SpanRef span_setup(float x0, float y0, float abx, float aby, float rmax) {
  SpanRef s;
  const float l2 = abx * abx + aby * aby;
  const float len = sqrtf(l2);
  s.pad = 0.03f + 4e-6f * (fabsf(x0) + fabsf(y0) + len + rmax);
  const float rp = rmax + s.pad;
  s.x0 = x0;
  s.y0 = y0;
  s.abx = abx;
  s.aby = aby;
  s.bu0 = fminf(0.0f, abx) - rp;
  s.bu1 = fmaxf(0.0f, abx) + rp;
  s.bv0 = fminf(0.0f, aby) - rp;
  s.bv1 = fmaxf(0.0f, aby) + rp;
  s.dlo = -s.pad * len;
  s.dhi = l2 + s.pad * len;
  s.c = rp * len;
  s.ia = abx != 0.0f ? 1.0f / abx : 0.0f;
  s.ib = aby != 0.0f ? 1.0f / aby : 0.0f;
  return s;
}

// Interval [*lo, *hi] of u = px - x0 on the row v = py - y0; false when the
// row misses the rectangle.
SR_HD bool span_row(const SpanRef& s, float v, float* lo, float* hi) {
  if (!(v >= s.bv0 && v <= s.bv1)) return false;
  float a = s.bu0, b = s.bu1;
  const float va = v * s.aby;  // along: u * abx in [dlo - v aby, dhi - v aby]
  if (s.abx != 0.0f) {
    const float e0 = (s.dlo - va) * s.ia, e1 = (s.dhi - va) * s.ia;
    a = fmaxf(a, fminf(e0, e1));
    b = fminf(b, fmaxf(e0, e1));
  } else if (!(va >= s.dlo && va <= s.dhi)) {
    return false;
  }
  const float vc = v * s.abx;  // across: u * aby in [v abx - c, v abx + c]
  if (s.aby != 0.0f) {
    const float e0 = (vc - s.c) * s.ib, e1 = (vc + s.c) * s.ib;
    a = fmaxf(a, fminf(e0, e1));
    b = fminf(b, fmaxf(e0, e1));
  } else if (!(fabsf(vc) <= s.c)) {
    return false;
  }
  *lo = a;
  *hi = b;
  return a <= b;
}

// Pixel mask of a fx x fy tile at (tx0, ty0) (pixel y * fx + x, W words of
// 64 pixels): the pixels whose centres fall in the segment's row spans.
template <int W>
SR_HD void span_mask(const SpanRef& s, uint32_t tx0, uint32_t ty0, uint32_t fx, uint32_t fy, uint64_t (&m)[W]) {
  for (int j = 0; j < W; ++j) m[j] = 0ull;
  const float c0 = (float)tx0 + 0.5f - s.x0;  // u of pixel x is x + c0
  for (uint32_t y = 0; y < fy; ++y) {
    const float v = ((float)(ty0 + y) + 0.5f) - s.y0;
    float lo, hi;
    if (!span_row(s, v, &lo, &hi)) continue;
    const float xl = lo - c0, xh = hi - c0;
    const float fl = ceilf(xl - (0.01f + 1e-6f * (fabsf(lo) + fabsf(c0))));
    const float fh = floorf(xh + (0.01f + 1e-6f * (fabsf(hi) + fabsf(c0))));
    const float cl = fmaxf(fl, 0.0f), ch = fminf(fh, (float)(fx - 1u));
    if (!(cl <= ch)) continue;
    const uint32_t b0 = y * fx + (uint32_t)cl, b1 = y * fx + (uint32_t)ch + 1u;  // bits [b0, b1)
    for (int j = 0; j < W; ++j) {
      const uint32_t lo_b = max(b0, 64u * j), hi_b = min(b1, 64u * j + 64u);
      if (lo_b >= hi_b) continue;
      const uint32_t n = hi_b - lo_b, sh = lo_b - 64u * j;
      const uint64_t bits = n >= 64u ? ~0ull : ((1ull << n) - 1ull);
      m[j] |= bits << sh;
    }
  }
}

}  // namespace srk
