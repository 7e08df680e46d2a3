// Device exclusive prefix sum over int32 (reduce-then-scan, three kernels),
// used to turn per-row counts into row pointers (B2SR block rows, the
// transpose's column pointers). Workspace: one int per 2048 elements.
#include "common.hpp"

namespace spmm {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanChunk = kScanThreads * kScanPer;  // 2048

__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread across the block; returns the
// block total through *total.
__device__ int block_exclusive_scan(int v, int *total) {
  __shared__ int wsum[kScanThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) wsum[wave] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    if (w < wave) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(kScanThreads)
scan_reduce_kernel(const int *__restrict__ d, int n, int *__restrict__ part) {
  const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) s += base + i < n ? d[base + i] : 0;
  int tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads)
scan_partials_kernel(int *__restrict__ part, int np) {
  int carry = 0;
  for (int c0 = 0; c0 < np; c0 += kScanThreads) {
    int i = c0 + threadIdx.x;
    int v = i < np ? part[i] : 0;
    int tot;
    int ex = block_exclusive_scan(v, &tot);
    if (i < np) part[i] = carry + ex;
    carry += tot;
  }
}

__global__ void __launch_bounds__(kScanThreads)
scan_apply_kernel(int *__restrict__ d, int n, const int *__restrict__ part) {
  const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int v[kScanPer];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = base + i < n ? d[base + i] : 0;
    s += v[i];
  }
  int tot;
  int ex = block_exclusive_scan(s, &tot) + part[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < n) d[base + i] = ex;
    ex += v[i];
  }
}

size_t scan_workspace_bytes(int n) {
  return sizeof(int) * static_cast<size_t>(ceil_div(n > 0 ? n : 1, kScanChunk));
}

// In-place exclusive scan of d[0..n). `part` needs scan_workspace_bytes(n).
hipError_t exclusive_scan(int *d, int n, int *part, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int nb = ceil_div(n, kScanChunk);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, s, d, n, part);
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, s, part, nb);
  hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanThreads), 0, s, d, n, part);
  return hipGetLastError();
}

}  // namespace spmm
