// Host check of unorm16_to_float and dequantize_depth01 (common.hpp): for
// every 16-bit value the FMA-corrected product equals the correctly rounded
// quotient i / 65535, and for every 24-bit value q / 16777215.
// Prints the number of mismatches; exits 1 on any.
#include <cstdio>

#include "../bevy_strand_rasterizer_amd/csrc/common.hpp"

int main() {
  unsigned bad = 0;
  for (uint32_t i = 0; i <= 0xFFFFu; ++i) {
    const float want = (float)i / 65535.0f;
    if (srk::unorm16_to_float(i) != want) ++bad;
    const srk::V2 v = srk::unpack2x16unorm(i | ((0xFFFFu - i) << 16));
    if (v.x != want || v.y != (float)(0xFFFFu - i) / 65535.0f) ++bad;
  }
  for (uint32_t q = 0; q < (1u << 24); ++q)
    if (srk::dequantize_depth01(q) != (float)q / 16777215.0f) ++bad;
  std::printf("%u\n", bad);
  return bad ? 1 : 0;
}
