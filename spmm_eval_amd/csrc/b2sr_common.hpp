// Pieces shared by the B2SR translation units (b2sr.hip: conversion, T <= 8
// kernels, narrow and sign-packed kernels, entry points; b2sr_wide.hip: the
// T >= 16 block-row kernel).
#pragma once
#include "csr_core.hpp"

#include <atomic>
#include <type_traits>

namespace spmm {

template <int T> struct Word;
template <> struct Word<2> { using type = uint8_t; };
template <> struct Word<4> { using type = uint8_t; };
template <> struct Word<8> { using type = uint8_t; };
template <> struct Word<16> { using type = uint16_t; };
template <> struct Word<32> { using type = uint32_t; };
template <> struct Word<64> { using type = uint64_t; };

// A tile row's bits as a register: 64-bit only for T = 64.
template <int T> using Bits = typename std::conditional<(T > 32), uint64_t, uint32_t>::type;
template <typename U> __device__ __forceinline__ int popcount(U x) {
  return sizeof(U) == 8 ? __popcll(static_cast<unsigned long long>(x))
                        : __popc(static_cast<uint32_t>(x));
}
template <typename U> __device__ __forceinline__ int top_bit(U x) {  // x != 0
  return sizeof(U) == 8 ? 63 - __clzll(static_cast<unsigned long long>(x))
                        : 31 - __clz(static_cast<uint32_t>(x));
}

constexpr int kB2Threads = 256;
constexpr int kB2Waves = kB2Threads / kWave;
constexpr int kLongChunks = 2;  // block rows above this many tile chunks go to b2sr_long_kernel
// plan-free wide tiles: block rows of at least this many tiles go first
// (on a power-law graph: T = 16 at 256 tiles 105 us against 124 at 512 and
// 173 at 128; T = 32 and 64 at 512 tiles 145 and 264 us against 208 and
// 380 at 256)
constexpr int wide_heavy(int T) { return T <= 16 ? 256 : 512; }
// ints per tile in a wide-tile plan's list buffer (round lists written by
// the decode kernel, read by the gather kernel)
constexpr int kWideSlots = 8;
// tiles per round of the wide-tile kernels (8 / 16 / 32 KB of words for
// T = 16 / 32 / 64)
constexpr int wide_round_tiles(int T) { return T == 16 ? 256 : (T == 32 ? 128 : 64); }

template <int EPL>
__device__ __forceinline__ void store_row(float *__restrict__ C, int ldc, int row, int col0,
                                          const float (&v)[EPL]) {
  float *Crow = C + static_cast<size_t>(row) * ldc + col0;
#pragma unroll
  for (int q = 0; q < EPL; q += 4)
    *reinterpret_cast<float4 *>(Crow + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
}

// store_row, or (kAcc: a column slice after the first) C row += v.
template <int EPL, bool kAcc>
__device__ __forceinline__ void put_c_row(float *__restrict__ C, int ldc, int row, int col0,
                                          const float (&v)[EPL]) {
  if constexpr (!kAcc) {
    store_row<EPL>(C, ldc, row, col0, v);
  } else {
    float *Crow = C + static_cast<size_t>(row) * ldc + col0;
#pragma unroll
    for (int q = 0; q < EPL; q += 4) {
      float4 c = *reinterpret_cast<const float4 *>(Crow + q);
      c.x += v[q];
      c.y += v[q + 1];
      c.z += v[q + 2];
      c.w += v[q + 3];
      *reinterpret_cast<float4 *>(Crow + q) = c;
    }
  }
}

// Workgroups of `kernel` resident on the chip at once (per-CU count cached).
inline int resident_wgs(const void *kernel, int cus, std::atomic<int> *cache) {
  int per_cu = cache->load(std::memory_order_relaxed);
  if (per_cu == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kB2Threads, 0) !=
            hipSuccess ||
        per_cu <= 0)
      per_cu = 4;
    cache->store(per_cu, std::memory_order_relaxed);
  }
  return per_cu * cus;
}

// The T = 16 / 32 / 64 FBF launch (b2sr_wide.hip): G lanes per row over
// `tiles` column tiles; p == nullptr is the plan-free form, over the block
// rows sorted by tile count when `sorted` is given, else listing the heavy
// block rows on `state` first.
template <typename TB>
void b2sr_wide_launch(int T, int G, int tiles, const spmm_csr_plan_s *p, const int *brp,
                      int *state, const int *sorted, int cus, int M, int K, const int *bci,
                      const void *val, const TB *B, int ldb, float *C, int ldc, hipStream_t s);

}  // namespace spmm
