// Aligned B2SR-8 / B2SR-16 and the binary SpMV that reads it.
//
// Aligned layout (the reference's packUchar2AlignedUnsigned /
// packUshort2AlignedUnsigned, csr2bsr_batch_bsrbmv.cu:607-647, 858-888):
// each block row's tiles are taken in groups of GS = 32 / T (4 for T = 8,
// 2 for T = 16) and padded with empty tiles of column 0 to a whole group.
// newRowPtr counts padded tiles, newColInd holds their block columns, and
// newVal[T * g + j] (g a global group index) is the uint32 whose bits
// [32 - T(k+1), 32 - Tk) are row j of the group's tile k -- first tile in
// the highest bits. One 32-bit AND + popcount then covers GS tiles.
//
// Conversion: a count kernel and the device scan make newRowPtr; the pack
// kernel runs one thread per output word or column index and finds its block
// row by binary search over newRowPtr, so hub block rows cost no more than
// any other (the reference's packers are single-thread loops over all block
// rows).
//
// SpMV y = A_bin * x: L lanes per block row (a multiple of T, up to a whole
// workgroup, chosen from the average group count). Lane (gl, j) keeps row j fixed and
// strides over the block row's groups gl, gl + L/T, ...; consecutive lanes
// read consecutive words. Binary x (bit-packed per block column, the tile
// rows' bit order) gives exact int32 counts and the binarised block-row word
// (bit T-1-j set iff row j's count is non-zero, the vector the reference's
// bin2full decodes, utility.cu:154-178); FP32 x gives FP32 y, summed over the
// set bits in order within a lane, then across lanes.
#include "common.hpp"

#include <climits>

namespace spmm {
namespace {

template <int T> struct AWord;
template <> struct AWord<8> { using type = uint8_t; };
template <> struct AWord<16> { using type = uint16_t; };

constexpr int kAThreads = 256;

// Last block row br with nrp[br] <= p (p < nrp[MB]); empty block rows are
// skipped because their successor starts at the same offset.
__device__ __forceinline__ int find_block_row(const int *__restrict__ nrp, int MB, int p) {
  int lo = 0, hi = MB;  // invariant: nrp[lo] <= p < nrp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (nrp[mid] <= p) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int GS>
__global__ void __launch_bounds__(kAThreads)
aligned_count_kernel(const int *__restrict__ rp, int MB, int *__restrict__ nrp) {
  const int i = blockIdx.x * kAThreads + threadIdx.x;
  if (i < MB) {
    const int n = rp[i + 1] - rp[i];
    nrp[i] = (n + GS - 1) / GS * GS;
  } else if (i == MB) {
    nrp[MB] = 0;  // the exclusive scan turns this slot into the total
  }
}

template <int T>
__global__ void __launch_bounds__(kAThreads)
aligned_pack_kernel(const int *__restrict__ rp, const int *__restrict__ ci,
                    const typename AWord<T>::type *__restrict__ val, const int *__restrict__ nrp,
                    int MB, int64_t nwords, int ntiles, int *__restrict__ nci,
                    uint32_t *__restrict__ nval) {
  constexpr int GS = 32 / T;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kAThreads + threadIdx.x;
  if (i < nwords) {
    const int p = static_cast<int>(i / T) * GS, j = static_cast<int>(i % T);
    const int br = find_block_row(nrp, MB, p);
    const int local = p - nrp[br], base = rp[br], n = rp[br + 1] - base;
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < GS; ++k)
      if (local + k < n)
        w |= static_cast<uint32_t>(val[static_cast<int64_t>(base + local + k) * T + j])
             << (32 - T * (k + 1));
    nval[i] = w;
  } else if (i < nwords + ntiles) {
    const int p = static_cast<int>(i - nwords);
    const int br = find_block_row(nrp, MB, p);
    const int local = p - nrp[br], base = rp[br];
    nci[p] = local < rp[br + 1] - base ? ci[base + local] : 0;
  }
}

// kF32 = false: x is one T-bit word per block column, out = counts / bits.
// kF32 = true:  x is K floats, out = y (FP32).
// L lanes per block row (L <= 64: 64/L block rows per wave; L = 256: one
// block row per workgroup, the waves' partial sums meet in LDS). Each lane
// takes U groups per round with all their word and column loads in flight.
template <int T, int L, bool kF32>
__global__ void __launch_bounds__(kAThreads)
aligned_spmv_kernel(const int *__restrict__ nrp, const int *__restrict__ nci,
                    const uint32_t *__restrict__ nval, const void *__restrict__ xv, int M, int MB,
                    int32_t *__restrict__ y_count, typename AWord<T>::type *__restrict__ y_bin,
                    float *__restrict__ y) {
  static_assert(L <= kWave || L == kAThreads, "a block row spans part of a wave or a workgroup");
  constexpr int GS = 32 / T, GPP = L / T, U = 4;
  constexpr int WL = L < kWave ? L : kWave;  // a subgroup's lanes inside one wave
  using W = typename AWord<T>::type;
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sl = tid % L, j = sl % T, gl = sl / T;
  const int br = blockIdx.x * (kAThreads / L) + tid / L;
  const bool live = br < MB;
  int cnt = 0;
  float acc = 0.f;
  if (live) {
    const int g1 = nrp[br + 1] / GS;
    for (int g = nrp[br] / GS + gl; g < g1; g += GPP * U) {
      // Every slot loads from a valid group (past the end: the block row's
      // last, whose word is then dropped), so no load sits behind a branch
      // and all of them are in flight together.
      uint32_t a[U];
      int col[U][GS];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int gg = min(g + u * GPP, g1 - 1);
        a[u] = nval[static_cast<int64_t>(gg) * T + j];
#pragma unroll
        for (int k = 0; k < GS; ++k) col[u][k] = nci[gg * GS + k];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (g + u * GPP >= g1) a[u] = 0u;
      if constexpr (!kF32) {
        const W *x = static_cast<const W *>(xv);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          uint32_t xw = 0;
#pragma unroll
          for (int k = 0; k < GS; ++k)
            xw |= static_cast<uint32_t>(x[col[u][k]]) << (32 - T * (k + 1));
          cnt += __popc(a[u] & xw);
        }
      } else {
        // Each round takes the highest remaining set bit of every word, so
        // the U gathers of a round go out together; rounds repeat while a
        // word has bits left (most words hold none or one).
        const float *x = static_cast<const float *>(xv);
        bool more = false;
#pragma unroll
        for (int u = 0; u < U; ++u) more = more || a[u] != 0;
        while (more) {
          float v[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int b = a[u] ? 31 - __clz(a[u]) : 31;
            const int k = (31 - b) / T, c = T - 1 - (b % T);
            int ck = col[u][0];
#pragma unroll
            for (int q = 1; q < GS; ++q) ck = k == q ? col[u][q] : ck;
            v[u] = x[static_cast<int64_t>(ck) * T + c];
          }
          more = false;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (a[u]) {
              acc += v[u];
              a[u] &= ~(1u << (31 - __clz(a[u])));
            }
            more = more || a[u] != 0;
          }
        }
      }
    }
  }
  // lanes of one wave that share j differ in the bits at T, 2T, ... WL/2
#pragma unroll
  for (int o = T; o < WL; o <<= 1) {
    if constexpr (!kF32) cnt += __shfl_xor(cnt, o, kWave);
    else acc += __shfl_xor(acc, o, kWave);
  }
  if constexpr (L > kWave) {  // one block row per workgroup: add the waves' sums
    __shared__ int red_i[kAThreads / kWave][T];
    __shared__ float red_f[kAThreads / kWave][T];
    if (lane < T) {
      if constexpr (!kF32) red_i[tid / kWave][lane] = cnt;
      else red_f[tid / kWave][lane] = acc;
    }
    __syncthreads();
    if (tid < T) {
#pragma unroll
      for (int w = 1; w < kAThreads / kWave; ++w) {
        if constexpr (!kF32) cnt += red_i[w][tid];
        else acc += red_f[w][tid];
      }
    }
    if (tid >= kWave) return;  // wave 0 writes the block row
  }
  const int row = br * T + j;
  const bool writer = live && sl < T;  // one lane per row
  if constexpr (!kF32) {
    const uint64_t nz = __ballot(writer && cnt > 0);
    if (writer && y_count && row < M) y_count[row] = cnt;
    if (live && sl == 0 && y_bin) {
      const int shift = lane / WL * WL;
      const uint32_t m = static_cast<uint32_t>(nz >> shift) & ((1u << T) - 1);
      y_bin[br] = static_cast<W>(__brev(m) >> (32 - T));
    }
  } else {
    if (writer && row < M) y[row] = acc;
  }
}

template <int T, bool kF32>
void launch_spmv(int L, const int *nrp, const int *nci, const uint32_t *nval, const void *x, int M,
                 int MB, int32_t *y_count, void *y_bin, float *y, hipStream_t s) {
  const int per_block = kAThreads / L;
  const dim3 grid((MB + per_block - 1) / per_block);
  using W = typename AWord<T>::type;
  auto *yb = static_cast<W *>(y_bin);
#define SPMM_ASPMV(LL)                                                                 \
  aligned_spmv_kernel<T, LL, kF32><<<grid, kAThreads, 0, s>>>(nrp, nci, nval, x, M, MB, \
                                                               y_count, yb, y)
  switch (L) {
    case 8: if constexpr (T <= 8) { SPMM_ASPMV(8); break; } [[fallthrough]];
    case 16: SPMM_ASPMV(16); break;
    case 32: SPMM_ASPMV(32); break;
    case 64: SPMM_ASPMV(64); break;
    default: SPMM_ASPMV(256); break;
  }
#undef SPMM_ASPMV
}

// Lanes per block row. Bit x: enough that U = 4 rounds cover the average
// block row, 128 rounded up to a whole workgroup (measured on a 89k-row
// power-law graph: 8.5 us at T = 8 with 64 lanes, 14.4 us at T = 16 with a
// workgroup). FP32 x: always a workgroup -- its set bits are gathered one
// round each, and the dense words of hub block rows set the critical path.
int lanes_for(int T, int MB, int nnzb, bool f32) {
  if (f32) return kAThreads;
  const int GS = 32 / T;
  const double groups = MB > 0 ? static_cast<double>(nnzb) / GS / MB : 0.0;
  int L = T;
  while (L < kAThreads && 4.0 * L / T < groups) L *= 2;
  return L > kWave ? kAThreads : L;
}

template <bool kF32>
int run_spmv(spmm_handle_t h, int T, int M, int K, int nnzb, const int *nrp, const int *nci,
             const uint32_t *nval, const void *x, int32_t *y_count, void *y_bin, float *y,
             void *stream) {
  SPMM_CHECK_ARG(h, "null handle");
  SPMM_CHECK_ARG(T == 8 || T == 16, "aligned B2SR needs T = 8 or 16");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && nnzb >= 0, "negative size");
  SPMM_CHECK_ARG(nnzb % (32 / T) == 0, "nnzb of an aligned B2SR is a whole number of groups");
  const int MB = (M + T - 1) / T;
  SPMM_CHECK_ARG(MB == 0 || nrp, "null newRowPtr");
  SPMM_CHECK_ARG(nnzb == 0 || (nci && nval && x), "null newColInd, newVal or x");
  SPMM_CHECK_ARG(static_cast<int64_t>(nnzb) / (32 / T) * T < (int64_t{1} << 31),
                 "aligned B2SR above 2^31 words");
  if (MB == 0) return SPMM_OK;
  hipStream_t s = as_stream(stream);
  const int L = lanes_for(T, MB, nnzb, kF32);
  if (T == 8) launch_spmv<8, kF32>(L, nrp, nci, nval, x, M, MB, y_count, y_bin, y, s);
  else launch_spmv<16, kF32>(L, nrp, nci, nval, x, M, MB, y_count, y_bin, y, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // namespace
}  // namespace spmm

using namespace spmm;

extern "C" {

int b2sr_aligned_nnzb(spmm_handle_t h, int T, int MB, const int *bsrRowPtr, int *newRowPtr,
                      int *nnzb_aligned, void *stream) {
  SPMM_CHECK_ARG(h && newRowPtr && nnzb_aligned, "null handle, newRowPtr or nnzb_aligned");
  SPMM_CHECK_ARG(T == 8 || T == 16, "aligned B2SR needs T = 8 or 16");
  SPMM_CHECK_ARG(MB >= 0, "negative size");
  SPMM_CHECK_ARG(bsrRowPtr, "null bsrRowPtr");
  hipStream_t s = as_stream(stream);
  char *ws = static_cast<char *>(workspace(h, scan_workspace_bytes(MB + 1)));
  int *host = static_cast<int *>(pinned(h, sizeof(int)));
  if (!ws || !host) return SPMM_ERR_ALLOC;
  const dim3 grid(MB / kAThreads + 1);
  if (T == 8) aligned_count_kernel<4><<<grid, kAThreads, 0, s>>>(bsrRowPtr, MB, newRowPtr);
  else aligned_count_kernel<2><<<grid, kAThreads, 0, s>>>(bsrRowPtr, MB, newRowPtr);
  SPMM_LAUNCHED();
  SPMM_HIP(exclusive_scan(newRowPtr, MB + 1, reinterpret_cast<int *>(ws), s));
  SPMM_HIP(hipMemcpyAsync(host, newRowPtr + MB, sizeof(int), hipMemcpyDeviceToHost, s));
  SPMM_HIP(hipStreamSynchronize(s));
  *nnzb_aligned = host[0];
  return SPMM_OK;
}

int b2sr_aligned_pack(spmm_handle_t h, int T, int MB, int nnzb_aligned, const int *bsrRowPtr,
                      const int *bsrColInd, const void *bsrVal, const int *newRowPtr,
                      int *newColInd, uint32_t *newVal, void *stream) {
  SPMM_CHECK_ARG(h, "null handle");
  SPMM_CHECK_ARG(T == 8 || T == 16, "aligned B2SR needs T = 8 or 16");
  SPMM_CHECK_ARG(MB >= 0 && nnzb_aligned >= 0, "negative size");
  SPMM_CHECK_ARG(nnzb_aligned % (32 / T) == 0, "nnzb_aligned is a whole number of groups");
  if (MB == 0 || nnzb_aligned == 0) return SPMM_OK;
  SPMM_CHECK_ARG(bsrRowPtr && bsrColInd && bsrVal && newRowPtr && newColInd && newVal,
                 "null input or output");
  const int64_t nwords = static_cast<int64_t>(nnzb_aligned) / (32 / T) * T;
  const int64_t items = nwords + nnzb_aligned;
  SPMM_CHECK_ARG(items / kAThreads < INT32_MAX, "aligned B2SR too large for one launch");
  hipStream_t s = as_stream(stream);
  const dim3 grid(static_cast<unsigned>((items + kAThreads - 1) / kAThreads));
  if (T == 8)
    aligned_pack_kernel<8><<<grid, kAThreads, 0, s>>>(
        bsrRowPtr, bsrColInd, static_cast<const uint8_t *>(bsrVal), newRowPtr, MB, nwords,
        nnzb_aligned, newColInd, newVal);
  else
    aligned_pack_kernel<16><<<grid, kAThreads, 0, s>>>(
        bsrRowPtr, bsrColInd, static_cast<const uint16_t *>(bsrVal), newRowPtr, MB, nwords,
        nnzb_aligned, newColInd, newVal);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

int spmv_b2sr_aligned_bin(spmm_handle_t h, int T, int M, int K, int nnzb_aligned,
                          const int *newRowPtr, const int *newColInd, const uint32_t *newVal,
                          const void *x_bits, int32_t *y_count, void *y_bin, void *stream) {
  SPMM_CHECK_ARG(y_count || y_bin, "no output: y_count and y_bin are both null");
  return run_spmv<false>(h, T, M, K, nnzb_aligned, newRowPtr, newColInd, newVal, x_bits, y_count,
                         y_bin, nullptr, stream);
}

int spmv_b2sr_aligned_f32(spmm_handle_t h, int T, int M, int K, int nnzb_aligned,
                          const int *newRowPtr, const int *newColInd, const uint32_t *newVal,
                          const float *x, float *y, void *stream) {
  SPMM_CHECK_ARG(y || M == 0, "null y");
  return run_spmv<true>(h, T, M, K, nnzb_aligned, newRowPtr, newColInd, newVal, x, nullptr,
                        nullptr, y, stream);
}

}  // extern "C"
