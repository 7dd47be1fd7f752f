// Host check of the rasterizers' conservative pixel spans (coverage.hpp):
// over many random and adversarial segments, every pixel centre that the
// shadow or camera exact coverage test accepts must be set in the tile mask.
// The exact tests below restate the kernels' per-pixel arithmetic
// (k_shadow_raster / k_camera_raster; strand_rasterizer.wgsl:138-147,
// 553-566).  Prints "cases misses kept covered"; exits 1 on a miss.
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>

#include "../bevy_strand_rasterizer_amd/csrc/coverage.hpp"

using namespace srk;

// The kernels' span setup takes the hardware reciprocal and square root
// (within 1 ulp); the check runs it with results moved 2 ulp down and up.
template <int ULPS>
struct UlpMath {
  static float move(float x) {
    for (int i = 0; i < (ULPS < 0 ? -ULPS : ULPS); ++i) x = nextafterf(x, ULPS < 0 ? -INFINITY : INFINITY);
    return x;
  }
  static float rcp(float x) { return move(1.0f / x); }
  static float sqrt(float x) { return move(sqrtf(x)); }
};

template <class M>
static SpanRef setup(float x0, float y0, float abx, float aby, float rmax, uint32_t tx0, uint32_t ty0, uint32_t fy) {
  return span_setup<M>(x0, y0, abx, aby, rmax, tx0, ty0, fy);
}

static bool shadow_exact(float x0, float y0, float z0, float x1, float y1, float z1, float minr, float maxr, float px,
                         float py) {
  const float abx = x1 - x0, aby = y1 - y0, l2 = abx * abx + aby * aby;
  float t = 0.0f;
  if (!(l2 == 0.0f)) t = ((px - x0) * abx + (py - y0) * aby) / l2;
  if (t < 0.0f || t > 1.0f) return false;
  const float qx = wmix(x0, x1, t), qy = wmix(y0, y1, t), qz = wmix(z0, z1, t);
  const float r = wmix(minr, maxr, wclamp(qz, 0.0f, 1.0f));
  const float dx = px - qx, dy = py - qy;
  return wclamp(1.0f - sqrtf(dx * dx + dy * dy) / r, 0.0f, 1.0f) > 0.0f;
}

static bool camera_exact(float x0, float y0, float z0, float x1, float y1, float z1, float rmin, float rmax, float cx,
                         float cy) {
  const float minr0 = fmaxf(rmin, 1e-4f), maxr0 = fmaxf(rmax, minr0);
  const float abx = x1 - x0, aby = y1 - y0;
  const float inv_l2 = 1.0f / (abx * abx + aby * aby);
  if (cx < fminf(x0, x1) - maxr0 || cy < fminf(y0, y1) - maxr0 || cx > fmaxf(x0, x1) + maxr0 ||
      cy > fmaxf(y0, y1) + maxr0)
    return false;
  const float tt = ((cx - x0) * abx + (cy - y0) * aby) * inv_l2;
  if (!(tt >= 0.0f && tt <= 1.0f)) return false;
  const float qx = x0 + abx * tt, qy = y0 + aby * tt;
  const float qz = wmix(z0, z1, tt);
  const float r = wmix(minr0, maxr0, wclamp(normalize_depth01(qz), 0.0f, 1.0f));
  const float dx = cx - qx, dy = cy - qy;
  return dx * dx + dy * dy < r * r;
}

template <int W, class M>
static void run(std::mt19937& g, long cases, uint32_t fx, uint32_t fy, long* misses, long* kept, long* covered) {
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  for (long i = 0; i < cases; ++i) {
    const int kind = (int)(U(g) * 8.0f);
    const float scale = kind == 7 ? 16000.0f : 2048.0f;
    const uint32_t tx0 = (uint32_t)(U(g) * (scale / fx - 2.0f)) * fx, ty0 = (uint32_t)(U(g) * (scale / fy - 2.0f)) * fy;
    // endpoints near the tile; lengths from zero to thousands of pixels
    float len;
    switch (kind) {
      case 0: len = 0.0f; break;
      case 1: len = powf(10.0f, -7.0f + 5.0f * U(g)); break;
      case 2: len = 20.0f * U(g); break;
      default: len = powf(10.0f, -1.0f + 4.5f * U(g)); break;
    }
    float x0 = (float)tx0 + (U(g) * 1.6f - 0.3f) * fx, y0 = (float)ty0 + (U(g) * 1.6f - 0.3f) * fy;
    if (U(g) < 0.3f) {  // endpoint on a pixel centre or corner
      x0 = floorf(x0) + (U(g) < 0.5f ? 0.5f : 0.0f);
      y0 = floorf(y0) + (U(g) < 0.5f ? 0.5f : 0.0f);
    }
    float ang = U(g) * 6.2831853f;
    const float sel = U(g);
    if (sel < 0.1f) ang = 0.0f;
    else if (sel < 0.2f) ang = 1.5707964f;
    else if (sel < 0.25f) ang = 1e-6f * U(g);
    else if (sel < 0.3f) ang = 3.1415927f;
    float x1 = x0 + len * cosf(ang), y1 = y0 + len * sinf(ang);
    if (sel < 0.1f) y1 = y0;
    if (sel >= 0.1f && sel < 0.2f) x1 = x0;
    if (U(g) < 0.5f) {  // p0 side alternates
      const float tx = x0; x0 = x1; x1 = tx;
      const float ty = y0; y0 = y1; y1 = ty;
    }
    const float minr = powf(10.0f, -4.0f + 5.0f * U(g)) * (U(g) < 0.2f ? 0.0f : 1.0f);
    float maxr = minr + powf(10.0f, -3.0f + 4.0f * U(g));
    if (U(g) < 0.2f) maxr = floorf(maxr) + (U(g) < 0.5f ? 0.5f : 1.0f);  // radius reaching pixel centres
    const float z0 = -0.5f + 2.0f * U(g), z1 = -0.5f + 2.0f * U(g);
    const float minr_s = fmaxf(minr, 1e-4f), maxr_s = fmaxf(maxr, minr_s);
    const float abx = x1 - x0, aby = y1 - y0;
    uint64_t ms[W], mc[W];
    // the kernels' span radii: the exact radius at the deeper endpoint
    span_mask<W>(setup<M>(x0, y0, abx, aby, span_radius(minr_s, maxr_s, fmaxf(z0, z1)), tx0, ty0, fy), fx, ms);
    const float minr_c = fmaxf(minr, 1e-4f), maxr_c = fmaxf(maxr, minr_c);
    const float rcam = span_radius(minr_c, maxr_c, span_depth01_max(z0, z1));
    span_mask<W>(setup<M>(x0, y0, abx, aby, rcam, tx0, ty0, fy), fx, mc);
    for (uint32_t y = 0; y < fy; ++y)
      for (uint32_t x = 0; x < fx; ++x) {
        const uint32_t p = y * fx + x;
        const bool bs = (ms[p / 64] >> (p % 64)) & 1ull, bc = (mc[p / 64] >> (p % 64)) & 1ull;
        const float px = (float)(tx0 + x) + 0.5f, py = (float)(ty0 + y) + 0.5f;
        const bool es = shadow_exact(x0, y0, z0, x1, y1, z1, minr_s, maxr_s, px, py);
        const bool ec = camera_exact(x0, y0, z0, x1, y1, z1, minr, maxr, px, py);
        *kept += bs + bc;
        *covered += es + ec;
        if ((es && !bs) || (ec && !bc)) {
          if (*misses < 5)
            std::printf("miss: tile (%u,%u) px (%u,%u) p0 (%.9g,%.9g) p1 (%.9g,%.9g) r %.9g..%.9g shadow %d/%d camera %d/%d\n",
                        tx0, ty0, x, y, x0, y0, x1, y1, minr, maxr, es, bs, ec, bc);
          ++*misses;
        }
      }
  }
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 200000;
  std::mt19937 g(12345);
  long misses = 0, kept = 0, covered = 0;
  long cases = 0;
  auto all = [&](auto tag) {
    using M = decltype(tag);
    run<1, M>(g, n, 8, 8, &misses, &kept, &covered);
    run<1, M>(g, n / 4, 4, 4, &misses, &kept, &covered);
    run<4, M>(g, n / 4, 16, 16, &misses, &kept, &covered);
    run<2, M>(g, n / 4, 16, 8, &misses, &kept, &covered);
    cases += n + 3 * (n / 4);
  };
  all(ExactMath{});
  all(UlpMath<-2>{});
  all(UlpMath<2>{});
  std::printf("%ld %ld %ld %ld\n", cases, misses, kept, covered);
  return misses ? 1 : 0;
}
