// scan.hpp -- device-wide exclusive scan for wave64 / CDNA4.
//
// Reduce-then-scan in three launches, no inter-workgroup spin waits:
//   1. scan_reduce_k : each 2048-item tile reduces to one partial;
//   2. scan_spine_k  : one workgroup scans the tile partials in place and
//                      publishes the grand total;
//   3. scan_down_k   : each tile rescans with its offset and calls the
//                      consumer's store(i, exclusive, value).
// Items are read striped (item k of lane t at tile + k*256 + t) so every
// load is coalesced; the element count may live in device memory so the
// launch shape is fixed by a capacity and never needs a host round trip.
#pragma once

#include "common.hpp"

namespace srk {

constexpr int kScanWg = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanWg * kScanItems;

struct U4 {
  uint32_t x, y, z, w;
};
__device__ __forceinline__ U4 operator+(U4 a, U4 b) { return U4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ U4 operator-(U4 a, U4 b) { return U4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <class T> __device__ __forceinline__ T zero_v() { return T{}; }

// Inclusive scan across one wave64 with every lane active, in the VALU: DPP
// row_shr:1, 2, 4, 8 scan each 16-lane row (a lane with no source lane in its
// row adds 0), then row_bcast:15 adds row r's last lane to row r + 1 (rows 1
// and 3) and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_u32(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
  return v;
}

template <class T> __device__ __forceinline__ T wave_inclusive(T v);
template <> __device__ __forceinline__ uint32_t wave_inclusive<uint32_t>(uint32_t v) { return wave_inclusive_u32(v); }
template <> __device__ __forceinline__ U4 wave_inclusive<U4>(U4 v) {
  return U4{wave_inclusive_u32(v.x), wave_inclusive_u32(v.y), wave_inclusive_u32(v.z), wave_inclusive_u32(v.w)};
}

// Exclusive scan across the 256-thread block; returns the block total too.
template <class T>
__device__ __forceinline__ T block_exclusive(T v, T* total, T* lds4) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T inc = wave_inclusive(v);
  if (lane == 63) lds4[wave] = inc;
  __syncthreads();
  T wave_base = zero_v<T>(), tot = zero_v<T>();
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    T s = lds4[w];
    if (w < wave) wave_base = wave_base + s;
    tot = tot + s;
  }
  __syncthreads();
  *total = tot;
  return wave_base + (inc - v);  // exclusive = inclusive - v (modular, exact)
}

__device__ __forceinline__ uint32_t scan_count(uint32_t cap, const uint32_t* n_dev) {
  return n_dev ? min(cap, *n_dev) : cap;
}

template <class T, class Src>
__global__ __launch_bounds__(kScanWg) void scan_reduce_k(Src src, uint32_t cap, const uint32_t* n_dev, T* partials) {
  __shared__ T lds4[4];
  const uint32_t n = scan_count(cap, n_dev);
  const uint32_t tile = blockIdx.x * (uint32_t)kScanTile;
  if (tile >= n) return;
  T acc = zero_v<T>();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = tile + k * kScanWg + threadIdx.x;
    if (i < n) acc = acc + src.load(i);
  }
  T total;
  (void)block_exclusive(acc, &total, lds4);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

template <class T, class Dst>
__global__ __launch_bounds__(kScanWg) void scan_spine_k(T* partials, uint32_t cap, const uint32_t* n_dev, Dst dst) {
  __shared__ T lds4[4];
  const uint32_t n = scan_count(cap, n_dev);
  const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  T carry = zero_v<T>();
  for (uint32_t b = 0; b < tiles; b += kScanWg) {
    uint32_t i = b + threadIdx.x;
    T v = i < tiles ? partials[i] : zero_v<T>();
    T total;
    T ex = block_exclusive(v, &total, lds4);
    if (i < tiles) partials[i] = carry + ex;
    carry = carry + total;
  }
  if (threadIdx.x == 0) dst.total(carry);
}

template <class T, class Src, class Dst>
__global__ __launch_bounds__(kScanWg) void scan_down_k(Src src, Dst dst, uint32_t cap, const uint32_t* n_dev,
                                                       const T* partials) {
  __shared__ T lds4[4];
  const uint32_t n = scan_count(cap, n_dev);
  const uint32_t tile = blockIdx.x * (uint32_t)kScanTile;
  if (tile >= n) return;
  T carry = partials[blockIdx.x];
  // all loads first, then the block scans, then every store: no barrier
  // sits between a store and the next item, so consumers whose store reads
  // more memory overlap their latencies
  T v[kScanItems], ex[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = tile + k * kScanWg + threadIdx.x;
    v[k] = i < n ? src.load(i) : zero_v<T>();
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    T total;
    ex[k] = carry + block_exclusive(v[k], &total, lds4);
    carry = carry + total;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t i = tile + k * kScanWg + threadIdx.x;
    if (i < n) dst.store(i, ex[k], v[k]);
  }
}

// Host-side launcher: partials must hold ceil(cap / 2048) items.
template <class T, class Src, class Dst>
inline void launch_scan(Src src, Dst dst, uint32_t cap, const uint32_t* n_dev, T* partials, hipStream_t s) {
  if (cap == 0) {
    hipLaunchKernelGGL((scan_spine_k<T, Dst>), dim3(1), dim3(kScanWg), 0, s, partials, 0u, nullptr, dst);
    return;
  }
  uint32_t tiles = (cap + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((scan_reduce_k<T, Src>), dim3(tiles), dim3(kScanWg), 0, s, src, cap, n_dev, partials);
  hipLaunchKernelGGL((scan_spine_k<T, Dst>), dim3(1), dim3(kScanWg), 0, s, partials, cap, n_dev, dst);
  hipLaunchKernelGGL((scan_down_k<T, Src, Dst>), dim3(tiles), dim3(kScanWg), 0, s, src, dst, cap, n_dev, partials);
}

}  // namespace srk
