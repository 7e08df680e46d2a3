// Shared internals of libspmm_mi355x: status plumbing, the handle and its
// grow-only device workspace, and small device helpers used by every kernel
// file. gfx950 only: wave64, 16-byte global loads, 64-bit ballot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "../../include/spmm_mi355x.h"

namespace spmm {

constexpr int kWave = 64;

// Thread-local message of the last failing call (spmm_last_error()).
void set_error(const std::string &msg);

#define SPMM_FAIL(code, msg)                                                   \
  do {                                                                         \
    ::spmm::set_error(std::string(__func__) + ": " + (msg));                   \
    return (code);                                                             \
  } while (0)

#define SPMM_CHECK_ARG(cond, msg)                                              \
  do {                                                                         \
    if (!(cond)) SPMM_FAIL(SPMM_ERR_INVALID_ARG, msg);                         \
  } while (0)

#define SPMM_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e__ = (call);                                                   \
    if (e__ != hipSuccess)                                                     \
      SPMM_FAIL(SPMM_ERR_HIP,                                                  \
                std::string(#call) + " -> " + hipGetErrorString(e__));         \
  } while (0)

// Checks the launch just made (configuration errors surface here).
#define SPMM_LAUNCHED() SPMM_HIP(hipGetLastError())

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
inline long long ceil_div_ll(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace spmm

// Handle: owns the grow-only device buffers of the calls made through it
// (the workspace of the synchronous helper paths -- analysis, conversions,
// transposes --, the plan-free list state, the plan-free transposed
// structure). Not thread-safe; use one handle per host thread, and order the
// plan-free calls made through one handle on one stream.
struct spmm_handle_s {
  int device = 0;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  void *pinned = nullptr;  // small pinned host buffer for count read-backs
  size_t pinned_bytes = 0;
  // plan-free list state (B2SR and CSR drop-ins): kStateHead counters, then
  // two row / block-row lists; counters start at 0 and are reset by the
  // kernel that drains the lists
  int *list_state = nullptr;
  size_t list_cap = 0;  // ints
  // the plan-free transposed SpMM's structure: trowptr | tcol | perm
  int *tstruct = nullptr;
  size_t tstruct_cap = 0;  // ints
  // the plan-free column-sliced SpMM's per-row slice offsets, (S + 1) x M
  int *slice_off = nullptr;
  size_t slice_off_cap = 0;  // ints
  // buffers replaced by larger ones, freed by spmm_destroy: work queued
  // earlier, or a graph captured from an earlier call, may still use them
  std::vector<void *> retired;
};

namespace spmm {
// Counters at the head of the plan-free list state (zero between calls; each
// kernel that counts resets those it used), then the lists.
constexpr int kStateHead = 8;
// Returns a workspace of at least `bytes` (device) or nullptr with the
// error already recorded. A buffer that grows is replaced, not freed (the
// old one is retired until spmm_destroy); with `s` given, growing while s
// is being captured into a graph is refused.
void *workspace(spmm_handle_s *h, size_t bytes, hipStream_t s = nullptr);
void *pinned(spmm_handle_s *h, size_t bytes);
// Grow-only zeroed int buffer of at least `ints` elements in *buf / *cap,
// same retirement and capture rules as the workspace.
int grow_ints(spmm_handle_s *h, int **buf, size_t *cap, size_t ints, hipStream_t s);
// The plan-free list state for lists of up to n entries each.
inline int list_state(spmm_handle_s *h, int n, hipStream_t s, int **out) {
  const int st = grow_ints(h, &h->list_state, &h->list_cap,
                           kStateHead + 2 * static_cast<size_t>(n), s);
  *out = h->list_state;
  return st;
}
}  // namespace spmm

// ------------------------------------------------------------------ device
namespace spmm {

__device__ __forceinline__ float bf16_lo(uint32_t u) {
  return __uint_as_float(u << 16);
}
__device__ __forceinline__ float bf16_hi(uint32_t u) {
  return __uint_as_float(u & 0xffff0000u);
}

// 16-byte operand fetch: four floats or eight bf16 widened to float. The
// raw load and the widening are separate so a batch of loads can be issued
// back to back and waited for once (widening right after each load would
// force a wait per load).
template <typename TB> struct Vec16;
template <> struct Vec16<float> {
  static constexpr int kElems = 4;
  using Raw = float4;
  __device__ __forceinline__ static Raw raw(const float *p) {
    return *reinterpret_cast<const float4 *>(p);
  }
  __device__ __forceinline__ static void widen(const Raw &v, float *out) {
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
  }
  __device__ __forceinline__ static void load(const float *p, float *out) { widen(raw(p), out); }
};
template <> struct Vec16<uint16_t> {
  static constexpr int kElems = 8;
  using Raw = uint4;
  __device__ __forceinline__ static Raw raw(const uint16_t *p) {
    return *reinterpret_cast<const uint4 *>(p);
  }
  __device__ __forceinline__ static void widen(const Raw &v, float *out) {
    out[0] = bf16_lo(v.x); out[1] = bf16_hi(v.x);
    out[2] = bf16_lo(v.y); out[3] = bf16_hi(v.y);
    out[4] = bf16_lo(v.z); out[5] = bf16_hi(v.z);
    out[6] = bf16_lo(v.w); out[7] = bf16_hi(v.w);
  }
  __device__ __forceinline__ static void load(const uint16_t *p, float *out) { widen(raw(p), out); }
};

// Gathers of 16-byte row segments of a dense operand. RowLoader<TB, false>
// uses flat 64-bit addresses; RowLoader<TB, true> goes through a buffer
// resource with 32-bit byte offsets (one VALU op and one VGPR per address
// instead of 64-bit arithmetic and two), usable when the operand spans
// less than 2 GiB; the hardware also range-checks every such load (reads
// past num_records return zeros).
constexpr int kBufferWord3 = 0x00020000;  // gfx9-family resource: 32-bit data format

template <typename TB, bool kBuf> struct RowLoader;
template <typename TB> struct RowLoader<TB, false> {
  const TB *base;  // already offset to the lane's first column
  int ld;          // elements per row
  __device__ __forceinline__ RowLoader(const TB *B, int ldb, int col, int) : base(B + col), ld(ldb) {}
  __device__ __forceinline__ typename Vec16<TB>::Raw operator()(int row) const {
    return Vec16<TB>::raw(base + static_cast<size_t>(row) * ld);
  }
};
template <typename TB> struct RowLoader<TB, true> {
  __amdgpu_buffer_rsrc_t rsrc;
  int ld_bytes, col_bytes;
  __device__ __forceinline__ RowLoader(const TB *B, int ldb, int col, int bytes)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<TB *>(B), static_cast<short>(0), bytes,
                                               kBufferWord3)),
        ld_bytes(ldb * static_cast<int>(sizeof(TB))),
        col_bytes(col * static_cast<int>(sizeof(TB))) {}
  __device__ __forceinline__ typename Vec16<TB>::Raw operator()(int row) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, row * ld_bytes + col_bytes, 0, 0);
    return __builtin_bit_cast(typename Vec16<TB>::Raw, v);
  }
};

// Byte span of rows [0, rows) of a [rows][ld] operand if a buffer resource
// can address it (< 2^31), else 0.
inline int buffer_span(long long rows, long long ld, int elem_bytes) {
  const long long b = rows * ld * elem_bytes;
  return (b > 0 && b < (1ll << 31)) ? static_cast<int>(b) : 0;
}

template <typename TB>
__device__ __forceinline__ float load_scalar(const TB *p);
template <> __device__ __forceinline__ float load_scalar<float>(const float *p) {
  return *p;
}
template <>
__device__ __forceinline__ float load_scalar<uint16_t>(const uint16_t *p) {
  return __uint_as_float(static_cast<uint32_t>(*p) << 16);
}

}  // namespace spmm

namespace spmm {
// scan.hip: in-place exclusive prefix sum of d[0..n) on stream s;
// `part` must hold scan_workspace_bytes(n) bytes of device memory.
size_t scan_workspace_bytes(int n);
hipError_t exclusive_scan(int *d, int n, int *part, hipStream_t s);
}  // namespace spmm
