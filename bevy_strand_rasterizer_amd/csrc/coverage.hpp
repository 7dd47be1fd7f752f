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

// Reciprocal and square root of the span setup.  The spans only need to be
// conservative, and their pad (0.03 px plus 4e-6 and 1e-6 relative terms)
// exceeds a few ulp of error in either, so the rasterisers use the
// hardware's (within 1 ulp) and the host check uses the correctly rounded
// ones and ones perturbed by a few ulp (tests/span_check.cpp).
struct ExactMath {
  SR_HD static float rcp(float x) { return 1.0f / x; }
  SR_HD static float sqrt(float x) { return sqrtf(x); }
};

// The rectangle's interval on pixel row y of a tile is
//   x in [max(box_lo, al0 + y al1, cl0 + y cl1), min(box_hi, ah0 + y al1, ch0 + y cl1)]
// (x, y relative to the tile's first pixel centre): the along and across
// limits are linear in y, so the coefficients are set up once per segment
// and every row costs a few fused multiply-adds.  A zero component of ab
// leaves that limit independent of x; it then bounds the rows instead.
struct SpanRef {
  float box_lo, box_hi;  // x range of the padded box
  float al0, ah0, al1;   // along limits: low / high intercepts, common slope
  float cl0, ch0, cl1;   // across limits
  float eps;             // absolute rounding bound of the x limits
  int32_t y0, y1;        // rows that can meet the rectangle
};

template <class M = ExactMath>
SR_HD SpanRef span_setup(float x0, float y0, float abx, float aby, float rmax, uint32_t tx0, uint32_t ty0,
                         uint32_t fy) {
  SpanRef s;
  const float l2 = abx * abx + aby * aby;
  const float len = M::sqrt(l2);
  const float pad = 0.03f + 4e-6f * (fabsf(x0) + fabsf(y0) + len + rmax);
  const float rp = rmax + pad;
  const float dlo = -pad * len, dhi = l2 + pad * len, c = rp * len;
  const float c0 = (float)tx0 + 0.5f - x0;  // u = px - x0 of pixel x is x + c0
  const float d0 = (float)ty0 + 0.5f - y0;  // v = py - y0 of row y is y + d0
  s.box_lo = (fminf(0.0f, abx) - rp) - c0;
  s.box_hi = (fmaxf(0.0f, abx) + rp) - c0;
  float vlo = fminf(0.0f, aby) - rp, vhi = fmaxf(0.0f, aby) + rp;  // rows, in v
  const float inf = __builtin_huge_valf();
  float m = fabsf(c0) + fabsf(s.box_lo) + fabsf(s.box_hi);
  if (abx != 0.0f) {  // u abx in [dlo - v aby, dhi - v aby]
    const float ia = M::rcp(abx);
    const float e0 = (dlo - d0 * aby) * ia - c0, e1 = (dhi - d0 * aby) * ia - c0;
    s.al0 = fminf(e0, e1);
    s.ah0 = fmaxf(e0, e1);
    s.al1 = -aby * ia;
    m += fabsf(e0) + fabsf(e1) + fabsf((dlo - d0 * aby) * ia) + fabsf((dhi - d0 * aby) * ia) +
         (float)fy * fabsf(s.al1);
  } else {  // v aby in [dlo, dhi]: a row range when aby != 0, else every row (l2 = 0)
    s.al0 = -inf;
    s.ah0 = inf;
    s.al1 = 0.0f;
    if (aby != 0.0f) {
      const float iy = M::rcp(aby);
      const float w0 = dlo * iy, w1 = dhi * iy;
      vlo = fmaxf(vlo, fminf(w0, w1));
      vhi = fminf(vhi, fmaxf(w0, w1));
    }
  }
  if (aby != 0.0f) {  // u aby in [v abx - c, v abx + c]
    const float ib = M::rcp(aby);
    const float e0 = (d0 * abx - c) * ib - c0, e1 = (d0 * abx + c) * ib - c0;
    s.cl0 = fminf(e0, e1);
    s.ch0 = fmaxf(e0, e1);
    s.cl1 = abx * ib;
    m += fabsf(e0) + fabsf(e1) + fabsf((d0 * abx - c) * ib) + fabsf((d0 * abx + c) * ib) +
         (float)fy * fabsf(s.cl1);
  } else {  // |v abx| <= c
    s.cl0 = -inf;
    s.ch0 = inf;
    s.cl1 = 0.0f;
    if (abx != 0.0f) {
      const float w = c * M::rcp(fabsf(abx));
      vlo = fmaxf(vlo, -w);
      vhi = fminf(vhi, w);
    }
  }
  s.eps = 0.01f + 1e-6f * m;
  // rows y with y + d0 in [vlo, vhi], widened like the x limits
  const float ey = 0.01f + 1e-6f * (fabsf(d0) + fabsf(vlo) + fabsf(vhi));
  const float fy0 = ceilf(vlo - d0 - ey), fy1 = floorf(vhi - d0 + ey);
  s.y0 = (int32_t)fmaxf(fy0, 0.0f);
  s.y1 = (int32_t)fminf(fy1, (float)fy - 1.0f);
  if (!(fy0 <= fy1)) s.y1 = -1;
  return s;
}

// Pixel mask of a fx x fy tile (pixel y * fx + x, W words of 64 pixels):
// the pixels whose centres fall in the segment's row spans.
template <int W>
SR_HD void span_mask(const SpanRef& s, uint32_t fx, uint64_t (&m)[W]) {
  for (int j = 0; j < W; ++j) m[j] = 0ull;
  // every lane walks all rows of the tile (a row outside [y0, y1] or with an
  // empty interval adds no bits), so a wave's lanes stay converged
  const int32_t rows = (int32_t)((64u * W) / fx);
  for (int32_t y = 0; y < rows; ++y) {
    const float fy = (float)y;
    const float lo = fmaxf(s.box_lo, fmaxf(s.al0 + fy * s.al1, s.cl0 + fy * s.cl1));
    const float hi = fminf(s.box_hi, fminf(s.ah0 + fy * s.al1, s.ch0 + fy * s.cl1));
    const float cl = fmaxf(ceilf(lo - s.eps), 0.0f), ch = fminf(floorf(hi + s.eps), (float)(fx - 1u));
    const bool row = y >= s.y0 && y <= s.y1 && cl <= ch;
    const uint32_t b0 = (uint32_t)y * fx + (row ? (uint32_t)cl : 0u);
    const uint32_t b1 = row ? (uint32_t)y * fx + (uint32_t)ch + 1u : b0;  // bits [b0, b1)
    for (int j = 0; j < W; ++j) {
      const uint32_t lo_b = max(b0, 64u * j), hi_b = min(b1, 64u * j + 64u);
      const uint32_t n = hi_b > lo_b ? hi_b - lo_b : 0u, sh = (lo_b - 64u * j) & 63u;
      const uint64_t bits = n >= 64u ? ~0ull : ((1ull << n) - 1ull);
      m[j] |= bits << sh;
    }
  }
}

}  // namespace srk
