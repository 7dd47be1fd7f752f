// Host check of even_slice (common.hpp): for a tile whose one depth interval
// [zmin, zmax] spans at least S quantised depths (S = 8, 16 or 32 slices),
// the arithmetic slice equals the first of the S slices of
// build_depth_warp_lut (strand_prepass.wgsl:984-1013, restated below) whose
// range holds q, for the correctly rounded reciprocal of the span and for
// reciprocals a few ulp away (the hardware's is within one).  Spans: every
// one from S to 4096, then random ones up to 2^24, with every depth of small
// spans and the slice boundaries and random depths of large ones.  Prints the
// mismatches; exits 1 on any.
#include <cmath>
#include <cstdio>
#include <random>

#include "../bevy_strand_rasterizer_amd/csrc/common.hpp"

static uint32_t scan_slice(uint32_t zmin, uint32_t zmax, uint32_t q, uint32_t S) {
  const uint32_t span = zmax - zmin + 1u;
  for (uint32_t ls = 0; ls < S; ++ls) {
    const uint32_t start = zmin + (span * ls) / S, ex_end = zmin + (span * (ls + 1u)) / S;
    const uint32_t end = std::min(ex_end > start ? ex_end - 1u : start, zmax);
    if (q >= start && q <= end) return ls;
  }
  return 99u;
}

static unsigned bad = 0;

static void check(uint32_t zmin, uint32_t zmax, uint32_t q, uint32_t shift) {
  const uint32_t want = scan_slice(zmin, zmax, q, 1u << shift);
  const float r = 1.0f / (float)(zmax - zmin + 1u);
  for (int k = -3; k <= 3; ++k) {
    float rk = r;
    for (int i = 0; i < (k < 0 ? -k : k); ++i) rk = std::nextafter(rk, k < 0 ? 0.0f : 1.0f);
    const uint32_t got = srk::even_slice(zmin, zmax, q, rk, shift);
    if (got != want && bad++ < 10)
      std::printf("S %u zmin %u zmax %u q %u: %u != %u (rcp %+d ulp)\n", 1u << shift, zmin, zmax, q, got, want, k);
  }
}

int main() {
  std::mt19937 rng(7);
  for (uint32_t shift = 3; shift <= 5; ++shift) {
    const uint32_t S = 1u << shift;
    for (uint32_t span = S; span <= 4096; ++span) {
      const uint32_t zmin = rng() % (16777216u - span);
      for (uint32_t d = 0; d < span; ++d) check(zmin, zmin + span - 1u, zmin + d, shift);
    }
    for (int t = 0; t < 20000; ++t) {
      const uint32_t span = S + rng() % (16777216u - S);
      const uint32_t zmin = rng() % (16777216u - span + 1u), zmax = zmin + span - 1u;
      for (uint32_t ls = 0; ls <= S; ++ls) {
        const uint32_t s = zmin + (span * ls) / S;
        for (int o = -2; o <= 2; ++o) {
          const int64_t q = (int64_t)s + o;
          if (q >= zmin && q <= zmax) check(zmin, zmax, (uint32_t)q, shift);
        }
      }
      for (int i = 0; i < 16; ++i) check(zmin, zmax, zmin + rng() % span, shift);
    }
  }
  std::printf("%u\n", bad);
  return bad ? 1 : 0;
}
