// Adjoint of Y = A X for CSR A: the stable device transpose (so dX = A^T dY
// runs through the same CSR SpMM kernels) and SDDMM for the value gradient
// dval[e] = <dY[i], X[j]> at each stored (i, j).
//
// Transpose (the device form of csr2csc, backend/readMtx.hpp:319-345):
// column counts by atomics, device scan to column pointers, scatter of the
// source positions through atomic per-column cursors, then every column's
// segment of source positions is sorted ascending (one workgroup per
// column, bitonic in LDS, or in global memory for segments longer than the
// LDS tile). Sorting restores the stable order, so the output is unique and
// identical to the sequential counting sort. Row indices and values are then
// gathered through the sorted positions.
//
// SDDMM: entry-chunked (each subgroup owns 32 consecutive stored entries,
// so hub rows are split across subgroups); per row the dY segment stays in
// registers, X row segments are gathered eight at a time and the partial
// dots are reduce-scattered across the G lanes (about one shuffle per
// entry). The kernel body lives in csr_core.hpp (shared with the fused
// batched launch).
#include "csr_core.hpp"

#include <string>

#include <algorithm>
#include <climits>

namespace spmm {

constexpr int kAdjThreads = 256;
constexpr int kSortTile = 4096;  // LDS bitonic tile (ints)

__global__ void __launch_bounds__(kAdjThreads)
col_count_kernel(int nnz, const int *__restrict__ colind, int *__restrict__ cnt) {
  for (int e = blockIdx.x * kAdjThreads + threadIdx.x; e < nnz; e += gridDim.x * kAdjThreads)
    atomicAdd(&cnt[colind[e]], 1);
}

__global__ void __launch_bounds__(kAdjThreads)
row_of_entry_kernel(int M, const int *__restrict__ rowptr, int *__restrict__ rowidx) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + threadIdx.x / 64;
  if (row >= M) return;
  for (int e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) rowidx[e] = row;
}

__global__ void __launch_bounds__(kAdjThreads)
scatter_kernel(int nnz, const int *__restrict__ colind, int *__restrict__ cursor,
               int *__restrict__ pos) {
  for (int e = blockIdx.x * kAdjThreads + threadIdx.x; e < nnz; e += gridDim.x * kAdjThreads)
    pos[atomicAdd(&cursor[colind[e]], 1)] = e;
}

// One step of the all-ascending bitonic network ("flip" form): comparator
// (i, i ^ mask) puts the smaller value at the lower index. Partners at or
// past `len` stand for +infinity and never move into range, so skipping
// them sorts any length without padding.
template <typename Seg>
__device__ __forceinline__ void bitonic_step(Seg seg, int len, int n2, int mask) {
  for (int i = threadIdx.x; i < n2; i += kAdjThreads) {
    const int p = i ^ mask;
    if (p > i && p < len) {
      const int a = seg[i], b = seg[p];
      if (a > b) { seg[i] = b; seg[p] = a; }
    }
  }
  __syncthreads();
}

template <typename Seg>
__device__ void bitonic_sort(Seg seg, int len) {
  int n2 = 1;
  while (n2 < len) n2 <<= 1;
  for (int k = 2; k <= n2; k <<= 1) {
    bitonic_step(seg, len, n2, k - 1);  // mirror compare within blocks of k
    for (int j = k >> 2; j > 0; j >>= 1) bitonic_step(seg, len, n2, j);
  }
}

// Sorts pos[colptr[c] .. colptr[c+1]) ascending; one workgroup per column.
__global__ void __launch_bounds__(kAdjThreads)
segment_sort_kernel(const int *__restrict__ colptr, int *__restrict__ pos) {
  __shared__ int tile[kSortTile];
  const int c = blockIdx.x;
  const int b = colptr[c], len = colptr[c + 1] - b;
  if (len <= 1) return;
  int *seg = pos + b;
  if (len <= 32) {
    if (threadIdx.x == 0) {  // insertion sort
      for (int i = 1; i < len; ++i) {
        int v = seg[i], j = i - 1;
        while (j >= 0 && seg[j] > v) { seg[j + 1] = seg[j]; --j; }
        seg[j + 1] = v;
      }
    }
    return;
  }
  if (len <= kSortTile) {
    for (int i = threadIdx.x; i < len; i += kAdjThreads) tile[i] = seg[i];
    __syncthreads();
    bitonic_sort<int *>(tile, len);
    for (int i = threadIdx.x; i < len; i += kAdjThreads) seg[i] = tile[i];
    return;
  }
  bitonic_sort<int *>(seg, len);  // long segment: same network in global memory
}

__global__ void __launch_bounds__(kAdjThreads)
gather_kernel(int nnz, const int *__restrict__ pos, const int *__restrict__ rowidx,
              const float *__restrict__ val, int *__restrict__ tcol, float *__restrict__ tval,
              int *__restrict__ perm) {
  for (int d = blockIdx.x * kAdjThreads + threadIdx.x; d < nnz; d += gridDim.x * kAdjThreads) {
    const int e = pos[d];
    tcol[d] = rowidx[e];
    if (tval) tval[d] = val[e];
    if (perm) perm[d] = e;
  }
}

// Set / copy int arrays with kernels rather than memset / memcpy nodes, so
// the plan-free transposed SpMM can be captured into a graph.
__global__ void __launch_bounds__(kAdjThreads)
fill_ints_kernel(int *__restrict__ d, int n, int v) {
  for (int i = blockIdx.x * kAdjThreads + threadIdx.x; i < n; i += gridDim.x * kAdjThreads) d[i] = v;
}

__global__ void __launch_bounds__(kAdjThreads)
copy_ints_kernel(const int *__restrict__ a, int *__restrict__ b, int n) {
  for (int i = blockIdx.x * kAdjThreads + threadIdx.x; i < n; i += gridDim.x * kAdjThreads) b[i] = a[i];
}

template <typename TB, int G>
__global__ void __launch_bounds__(kAdjThreads)
sddmm_kernel(int M, int nnz, const int *__restrict__ rowptr, const int *__restrict__ colind,
             const TB *__restrict__ dY, int lddy, const TB *__restrict__ X, int ldx,
             float *__restrict__ out, int ntile) {
  sddmm_wg<TB, G>(blockIdx.x, M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile);
}

// Narrow or unaligned N: one lane per stored entry, 64 consecutive entries a
// wave. The wave finds the rows of its first and last entries by a uniform
// binary search over rowptr; when those are fewer than 64 rows apart each
// lane holds one of the rows' end offsets and every lane finds its row in
// six shuffle steps, else each lane searches rowptr itself. The dot product
// takes 8 columns at a time, all 16 loads in flight, summed in column order.
template <typename TB>
__global__ void __launch_bounds__(kAdjThreads)
sddmm_narrow_kernel(int M, int N, int nnz, const int *__restrict__ rowptr,
                    const int *__restrict__ colind, const TB *__restrict__ dY, int lddy,
                    const TB *__restrict__ X, int ldx, float *__restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long first = (static_cast<long long>(blockIdx.x) * kAdjThreads + threadIdx.x) -
                          lane;
  if (first >= nnz) return;  // whole wave
  const int e0 = static_cast<int>(first);
  const int e = e0 + lane;
  const int elast = min(e0 + kWave - 1, nnz - 1);
  // the row holding entry x: the last r with rowptr[r] <= x
  auto row_of = [&](int x) {
    int lo = 0, hi = M - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rowptr[mid] <= x) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  const int rlo = row_of(e0), rhi = row_of(elast);
  int row;
  if (rhi - rlo < kWave) {
    const int n = rhi - rlo;                           // ends of rows rlo .. rhi-1
    const int end = rowptr[min(rlo + 1 + lane, M)];    // lane l: end of row rlo + l
    int k = 0;                                         // rows of the window ending at or before e
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const int v = __shfl(end, min(k + step - 1, kWave - 1), kWave);
      if (k + step <= n && v <= e) k += step;
    }
    row = rlo + k;
  } else {
    row = e < nnz ? row_of(e) : rhi;
  }
  if (e >= nnz) return;
  const TB *y = dY + static_cast<size_t>(row) * lddy;
  const TB *x = X + static_cast<size_t>(colind[e]) * ldx;
  float acc = 0.f;
  for (int k0 = 0; k0 < N; k0 += 8) {
    float a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = min(k0 + u, N - 1);
      a[u] = load_scalar<TB>(y + k);
      b[u] = load_scalar<TB>(x + k);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u < N) acc = fmaf(a[u], b[u], acc);
  }
  out[e] = acc;
}

template <typename TB, int G>
static void launch_sddmm(int M, int nnz, const int *rowptr, const int *colind, const TB *dY,
                         int lddy, const TB *X, int ldx, float *out, int ntile, hipStream_t s) {
  const long long per_wg = static_cast<long long>(kAdjThreads / G) * kSddmmChunk;
  hipLaunchKernelGGL((sddmm_kernel<TB, G>), dim3(static_cast<unsigned>(ceil_div_ll(nnz, per_wg))),
                     dim3(kAdjThreads), 0, s, M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile);
}

template <typename TB, int G>
__global__ void __launch_bounds__(kAdjThreads)
sddmm_row_kernel(LaunchTable tab, const int *__restrict__ rows, const int2 *__restrict__ span,
                 const int *__restrict__ colind, const TB *__restrict__ dY, int lddy,
                 const TB *__restrict__ X, int ldx, float *__restrict__ out, int ntile) {
  sddmm_row_wg<TB, G>(tab, blockIdx.x, rows, span, colind, dY, lddy, X, ldx, out, ntile);
}

template <typename TB, int G>
static void launch_sddmm_rows(const spmm_csr_plan_s *p, const int *rowptr, const int *colind,
                              const TB *dY, int lddy, const TB *X, int ldx, float *out,
                              int ntile, hipStream_t s) {
  LaunchTable t;
  build_table(p->counts, p->offs, G, 1, &t);
  if (t.total_wg == 0) return;
  hipLaunchKernelGGL((sddmm_row_kernel<TB, G>), dim3(t.total_wg), dim3(kAdjThreads), 0, s, t,
                     p->d_rows, p->d_span, colind, dY, lddy, X, ldx, out, ntile);
}

// plan == nullptr: entry-chunked (no schedule needed); else rows of the plan.
template <typename TB>
int run_sddmm(const spmm_csr_plan_s *plan, int M, int K, int N, int nnz,
                     const int *rowptr, const int *colind, const TB *dY, int lddy, const TB *X,
                     int ldx, float *out, void *stream) {
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  if (plan) {
    if (plan->tile != 0) SPMM_FAIL(SPMM_ERR_MISMATCH, "a B2SR plan was given to a CSR call");
    if (plan->M != M || plan->nnz != nnz)
      SPMM_FAIL(SPMM_ERR_MISMATCH, "plan was built for M=" + std::to_string(plan->M) +
                                       " nnz=" + std::to_string(plan->nnz));
  }
  if (M == 0 || nnz == 0) return SPMM_OK;
  SPMM_CHECK_ARG(rowptr && colind && out, "null rowptr, colind or out");
  SPMM_CHECK_ARG(N == 0 || (dY && X && lddy >= N && ldx >= N), "null operand or ld < N");
  hipStream_t s = as_stream(stream);
  if (N == 0) {
    SPMM_HIP(hipMemsetAsync(out, 0, sizeof(float) * static_cast<size_t>(nnz), s));
    return SPMM_OK;
  }
  const int e16 = 16 / static_cast<int>(sizeof(TB));
  const bool vec = aligned16(dY) && aligned16(X) && lddy % e16 == 0 && ldx % e16 == 0 &&
                   N % e16 == 0;
  int G = 0;
  if (vec) {
    const int lanes = N / e16;
    for (int t = 32; t >= 1; t >>= 1)  // the reduce-scatter keeps G partials per lane
      if (t <= lanes && lanes % t == 0) { G = t; break; }
  }
  if (G == 0) {
    hipLaunchKernelGGL((sddmm_narrow_kernel<TB>),
                       dim3(static_cast<unsigned>(ceil_div_ll(nnz, kAdjThreads))),
                       dim3(kAdjThreads), 0, s, M, N, nnz, rowptr, colind, dY, lddy, X, ldx, out);
  } else if (plan) {
    const int ntile = N / (G * e16);
    switch (G) {
      case 1: launch_sddmm_rows<TB, 1>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 2: launch_sddmm_rows<TB, 2>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 4: launch_sddmm_rows<TB, 4>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 8: launch_sddmm_rows<TB, 8>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 16: launch_sddmm_rows<TB, 16>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      default: launch_sddmm_rows<TB, 32>(plan, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
    }
  } else {
    const int ntile = N / (G * e16);
    switch (G) {
      case 1: launch_sddmm<TB, 1>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 2: launch_sddmm<TB, 2>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 4: launch_sddmm<TB, 4>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 8: launch_sddmm<TB, 8>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      case 16: launch_sddmm<TB, 16>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
      default: launch_sddmm<TB, 32>(M, nnz, rowptr, colind, dY, lddy, X, ldx, out, ntile, s); break;
    }
  }
  SPMM_LAUNCHED();
  return SPMM_OK;
}

template int run_sddmm<float>(const spmm_csr_plan_s *, int, int, int, int, const int *,
                              const int *, const float *, int, const float *, int, float *,
                              void *);
template int run_sddmm<uint16_t>(const spmm_csr_plan_s *, int, int, int, int, const int *,
                                 const int *, const uint16_t *, int, const uint16_t *, int,
                                 float *, void *);

// The stable transpose (csr2csc order, readMtx.hpp:319-345) without the
// final synchronisation: atomic column counts -> scan -> scatter of source
// positions -> per-column sort of those positions -> gather. Work is queued
// on `s` only (kernels, no memset or memcpy nodes), so it can be captured
// once the handle's workspace has grown to this size.
int transpose_async(spmm_handle_s *h, int M, int K, int nnz, const int *rowptr, const int *colind,
                    const float *val, int *tRowPtr, int *tColInd, float *tVal, int *perm,
                    hipStream_t s) {
  SPMM_CHECK_ARG(h && tRowPtr, "null handle or tRowPtr");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && nnz >= 0, "negative size");
  SPMM_CHECK_ARG(M == 0 || rowptr, "null rowptr");
  SPMM_CHECK_ARG(nnz == 0 || (colind && tColInd), "null colind or tColInd");
  SPMM_CHECK_ARG(!(tVal && !val && nnz), "tVal requested without val");
  // Workspace: cursor[K+1] | scan partials | pos[nnz] | rowidx[nnz]
  const size_t a0 = 0;
  const size_t a1 = (a0 + sizeof(int) * (static_cast<size_t>(K) + 1) + 255) & ~size_t(255);
  const size_t a2 = (a1 + scan_workspace_bytes(K + 1) + 255) & ~size_t(255);
  const size_t a3 = (a2 + sizeof(int) * static_cast<size_t>(nnz) + 255) & ~size_t(255);
  const size_t total = a3 + sizeof(int) * static_cast<size_t>(nnz) + 256;
  char *ws = static_cast<char *>(workspace(h, total, s));
  if (!ws) return SPMM_ERR_ALLOC;
  int *cursor = reinterpret_cast<int *>(ws + a0);
  int *part = reinterpret_cast<int *>(ws + a1);
  int *pos = reinterpret_cast<int *>(ws + a2);
  int *rowidx = reinterpret_cast<int *>(ws + a3);

  const int kgrid = std::max(1, std::min(ceil_div(K + 1, kAdjThreads), 1024));
  hipLaunchKernelGGL(fill_ints_kernel, dim3(kgrid), dim3(kAdjThreads), 0, s, tRowPtr, K + 1, 0);
  SPMM_LAUNCHED();
  const int grid = std::max(1, std::min(ceil_div(nnz, kAdjThreads), 4096));
  if (nnz > 0) {
    hipLaunchKernelGGL(col_count_kernel, dim3(grid), dim3(kAdjThreads), 0, s, nnz, colind,
                       tRowPtr);
    SPMM_LAUNCHED();
  }
  SPMM_HIP(exclusive_scan(tRowPtr, K + 1, part, s));
  if (nnz == 0) return SPMM_OK;
  hipLaunchKernelGGL(copy_ints_kernel, dim3(kgrid), dim3(kAdjThreads), 0, s, tRowPtr, cursor,
                     K + 1);
  SPMM_LAUNCHED();
  hipLaunchKernelGGL(scatter_kernel, dim3(grid), dim3(kAdjThreads), 0, s, nnz, colind, cursor, pos);
  SPMM_LAUNCHED();
  if (K > 0) {
    hipLaunchKernelGGL(segment_sort_kernel, dim3(K), dim3(kAdjThreads), 0, s, tRowPtr, pos);
    SPMM_LAUNCHED();
  }
  hipLaunchKernelGGL(row_of_entry_kernel, dim3(ceil_div(M, 4)), dim3(kAdjThreads), 0, s, M, rowptr,
                     rowidx);
  SPMM_LAUNCHED();
  hipLaunchKernelGGL(gather_kernel, dim3(grid), dim3(kAdjThreads), 0, s, nnz, pos, rowidx, val,
                     tColInd, tVal, perm);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // namespace spmm

using namespace spmm;

extern "C" {

int csr_transpose(spmm_handle_t h, int M, int K, int nnz, const int *rowptr, const int *colind,
                  const float *val, int *tRowPtr, int *tColInd, float *tVal, int *perm,
                  void *stream) {
  const int st = transpose_async(h, M, K, nnz, rowptr, colind, val, tRowPtr, tColInd, tVal, perm,
                                 as_stream(stream));
  if (st) return st;
  // the workspace is reused by the next call on this handle, whatever stream
  // that call uses
  SPMM_HIP(hipStreamSynchronize(as_stream(stream)));
  return SPMM_OK;
}

int sddmm_csr_f32(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                  const int *colind, const float *dY, int lddy, const float *X, int ldx,
                  float *dval, void *stream) {
  (void)h;
  return run_sddmm<float>(nullptr, M, K, N, nnz, rowptr, colind, dY, lddy, X, ldx, dval, stream);
}

int sddmm_csr_bf16(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                   const int *colind, const uint16_t *dY, int lddy, const uint16_t *X, int ldx,
                   float *dval, void *stream) {
  (void)h;
  return run_sddmm<uint16_t>(nullptr, M, K, N, nnz, rowptr, colind, dY, lddy, X, ldx, dval,
                             stream);
}

int sddmm_csr_f32_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz, const int *rowptr,
                          const int *colind, const float *dY, int lddy, const float *X, int ldx,
                          float *dval, void *stream) {
  SPMM_CHECK_ARG(plan != nullptr, "plan is null");
  return run_sddmm<float>(plan, M, K, N, nnz, rowptr, colind, dY, lddy, X, ldx, dval, stream);
}

int sddmm_csr_bf16_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                           const int *rowptr, const int *colind, const uint16_t *dY, int lddy,
                           const uint16_t *X, int ldx, float *dval, void *stream) {
  SPMM_CHECK_ARG(plan != nullptr, "plan is null");
  return run_sddmm<uint16_t>(plan, M, K, N, nnz, rowptr, colind, dY, lddy, X, ldx, dval, stream);
}

}  // extern "C"
