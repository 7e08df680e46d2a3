// B2SR (bit-block sparse row): device CSR -> B2SR-T conversion, the FBF
// SpMM C = A_bin * B (FP32 or BF16 B, FP32 out), the binary x sign-packed
// popcount SpMM (int32 out) and the dense 32-bit packers.
//
// Layout (the reference's, csr2bsr_batch_bsrbmv.cu:207-1070 with
// spmm.cu:165-335): bsrRowPtr[MB+1], bsrColInd[nb] ascending per block row,
// word bsrVal[T*b + r] has bit (T-1-c) set iff A[T*br+r][T*bc+c] is stored.
//
// Conversion: 64/T block rows per wave, one lane per CSR row of the block
// row; the T lanes of a group merge their sorted rows (group minimum of the
// current block column via xor-shuffles), OR their bits for that block
// column and emit one word each. One pass counts, a device scan makes
// bsrRowPtr, a second pass fills. No host loop, no per-block-row transfers
// (the reference does ~10 host<->device round trips per block row,
// csr2bsr_batch_bsrbmv.cu:236-379).
//
// FBF SpMM: one subgroup of G lanes per output row (G lanes x 16-byte
// vectors span the B row segment). For each chunk of tiles (up to 64 for
// T=4; each lane reads the row's word of several consecutive tiles), the
// set bits are counted, prefix summed across the subgroup and decoded into
// an LDS list of B row indices (ascending column order), and the subgroup
// then streams that list with several 16-byte B gathers in flight per lane,
// as the CSR kernel does. Power-law graphs leave ~1 set bit per 4x4 tile,
// so decoding a whole block row per round keeps the gathers deep.
#include "b2sr_common.hpp"

#include <atomic>
#include <climits>
#include <mutex>
#include <type_traits>

namespace spmm {

static bool valid_T(int T) {
  return T == 2 || T == 4 || T == 8 || T == 16 || T == 32 || T == 64;
}

// ------------------------------------------------------------ conversion
// Block rows holding more than hub_entries stored entries are appended to
// hubs (hubs[0] = count, zeroed before the launch) for csr2b2sr_hub_kernel
// (hub_entries = INT_MAX: none are).
template <int T, bool kFill>
__global__ void __launch_bounds__(kB2Threads)
csr2b2sr_kernel(int M, int K, const int *__restrict__ rowptr, const int *__restrict__ colind,
                int *__restrict__ bsrRowPtr, int *__restrict__ bsrColInd,
                typename Word<T>::type *__restrict__ bsrVal, int *__restrict__ err,
                int hub_entries, int *__restrict__ hubs) {
  using W = typename Word<T>::type;
  constexpr int GPW = kWave / T;  // block rows per wave
  const int lane = threadIdx.x & (kWave - 1);
  const int wave_g = (blockIdx.x * kB2Threads + threadIdx.x) / kWave;
  const int br = wave_g * GPW + lane / T;
  const int r = lane % T;
  const int MB = (M + T - 1) / T;
  const bool live = br < MB;
  const int row = br * T + r;
  int pos = 0, end = 0;
  if (live && row < M) {
    pos = rowptr[row];
    end = rowptr[row + 1];
  }
  const bool hub =
      live && rowptr[min(br * T + T, M)] - rowptr[br * T] > hub_entries;  // same in the group
  if (hub) {
    pos = end;
    if (r == 0) hubs[1 + atomicAdd(&hubs[0], 1)] = br;
  }
  int out = 0;
  if (kFill && live) out = bsrRowPtr[br];
  int count = 0, prev = -1, bad = 0;
  while (true) {
    int cur = INT_MAX;
    if (pos < end) {
      const int c = colind[pos];
      cur = (c >= 0 && c < K) ? c / T : INT_MAX - 1;
    }
    int m = cur;
#pragma unroll
    for (int d = 1; d < T; d <<= 1) m = min(m, __shfl_xor(m, d));
    if (__all(m == INT_MAX)) break;
    if (m != INT_MAX) {
      Bits<T> bits = 0;
      while (pos < end) {
        const int c = colind[pos];
        if (c < 0 || c >= K || c < prev) { bad = 1; pos = end; break; }
        if (c / T != m) break;
        bits |= static_cast<Bits<T>>(1) << (T - 1 - (c % T));
        prev = c;
        ++pos;
      }
      if (m == INT_MAX - 1) { bad = 1; pos = end; }
      if (kFill && m < INT_MAX - 1) {
        if (r == 0) bsrColInd[out + count] = m;
        bsrVal[static_cast<size_t>(out + count) * T + r] = static_cast<W>(bits);
      }
      ++count;
    }
  }
  if (!kFill && live && !hub && r == 0) bsrRowPtr[br] = count;
  if (bad) atomicOr(err, 1);
}

// Block rows with more than kHubEntries stored entries (the hub rows of a
// power-law graph, whose merge above is one long serial walk) are listed by
// csr2b2sr_kernel and taken one per workgroup by a persistent grid: the
// block row's block columns are marked in an LDS bitmap over [0, K/T).
// Count pass: the bitmap's popcount is the block row's tile count (column
// indices checked as in the merge). Fill pass: a popcount prefix over the
// bitmap (per-thread segments, then a workgroup scan) ranks every block
// column, so bsrColInd comes out ascending; the block row's words are zeroed
// and each entry ORs its bit in with a 32-bit atomic (bsrVal 4-byte aligned).
constexpr int kHubEntries = 64;
constexpr int kHubBitmapWords = 12288;  // 48 KB of LDS: K / T <= 393,216

template <int T, bool kFill>
__device__ __forceinline__ void hub_block_row(int br, int M, int K, const int *__restrict__ rowptr,
                                              const int *__restrict__ colind,
                                              int *__restrict__ bsrRowPtr,
                                              int *__restrict__ bsrColInd,
                                              typename Word<T>::type *__restrict__ bsrVal,
                                              int *__restrict__ err) {
  __shared__ uint32_t bm[kHubBitmapWords];
  __shared__ int rp[T + 1];
  __shared__ int segpre[kB2Threads];
  __shared__ int wsum[kB2Waves];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int r_end = min(br * T + T, M);
  const int e0 = rowptr[br * T], e1 = rowptr[r_end];
  const int KB = (K + T - 1) / T, KBW = (KB + 31) / 32;
  const int nrows = r_end - br * T;
  for (int i = tid; i <= nrows; i += kB2Threads) rp[i] = rowptr[br * T + i];
  for (int w = tid; w < KBW; w += kB2Threads) bm[w] = 0;
  __syncthreads();
  int bad = 0;
  for (int e = e0 + tid; e < e1; e += kB2Threads) {
    int r = 0;  // this entry's row within the block row
    while (r + 1 < nrows && rp[r + 1] <= e) ++r;
    const int c = colind[e];
    if (c < 0 || c >= K || (e > rp[r] && colind[e - 1] > c)) {  // repeats allowed, as in the merge
      bad = 1;
      continue;
    }
    const int bc = c / T;
    atomicOr(&bm[bc >> 5], 1u << (bc & 31));
  }
  if (!kFill && bad) atomicOr(err, 1);
  __syncthreads();
  // per-thread segments of the bitmap, their popcounts scanned over the block
  const int seg = (KBW + kB2Threads - 1) / kB2Threads;
  const int w0 = tid * seg, w1 = min(w0 + seg, KBW);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bm[w]);
  int inc = mine;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += y;
  }
  if (lane == kWave - 1) wsum[wave] = inc;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kB2Waves; ++w) {
    before += w < wave ? wsum[w] : 0;
    total += wsum[w];
  }
  if constexpr (!kFill) {
    if (tid == 0) bsrRowPtr[br] = total;
  } else {
    segpre[tid] = before + inc - mine;
    const int base = bsrRowPtr[br];
    // block columns in order
    int k = base + before + inc - mine;
    for (int w = w0; w < w1; ++w) {
      uint32_t x = bm[w];
      while (x) {
        const int b = __ffs(x) - 1;
        x &= x - 1;
        bsrColInd[k++] = w * 32 + b;
      }
    }
    using W = typename Word<T>::type;
    for (long long i = tid; i < static_cast<long long>(total) * T; i += kB2Threads)
      bsrVal[static_cast<size_t>(base) * T + i] = static_cast<W>(0);
    __syncthreads();
    uint32_t *val32 = reinterpret_cast<uint32_t *>(bsrVal);
    for (int e = e0 + tid; e < e1; e += kB2Threads) {
      int r = 0;
      while (r + 1 < nrows && rp[r + 1] <= e) ++r;
      const int c = colind[e];
      if (c < 0 || c >= K) continue;  // reported by the count pass
      const int bc = c / T, w = bc >> 5;
      const int sgi = w / seg;
      int rank = segpre[sgi];
      for (int q = sgi * seg; q < w; ++q) rank += __popc(bm[q]);
      rank += __popc(bm[w] & ((1u << (bc & 31)) - 1u));
      const size_t word = static_cast<size_t>(base + rank) * T + r;
      const int p = T - 1 - c % T;  // bit within the tile row's word
      const size_t bit = word * (sizeof(W) * 8) + p;
      atomicOr(&val32[bit >> 5], 1u << (bit & 31));
    }
  }
  __syncthreads();  // the LDS arrays are reused by the workgroup's next block row
}

template <int T, bool kFill>
__global__ void __launch_bounds__(kB2Threads)
csr2b2sr_hub_kernel(int M, int K, const int *__restrict__ rowptr, const int *__restrict__ colind,
                    int *__restrict__ bsrRowPtr, int *__restrict__ bsrColInd,
                    typename Word<T>::type *__restrict__ bsrVal, int *__restrict__ err,
                    const int *__restrict__ hubs) {
  const int n = hubs[0];
  for (int i = blockIdx.x; i < n; i += gridDim.x)
    hub_block_row<T, kFill>(hubs[1 + i], M, K, rowptr, colind, bsrRowPtr, bsrColInd, bsrVal, err);
}


// --------------------------------------------------- decode to LDS list
// LDS list entries per subgroup and tiles decoded per chunk: a chunk of
// TC tiles can hold at most TC*T set bits, which must fit the list.
template <int T, int G> struct Chunk {
  static constexpr int kCap = (16 * G > T) ? 16 * G : T;  // 16 KB per WG (32 KB for G=1,T=32)
  static constexpr int kTiles = kCap / T;                  // TC
  static constexpr int kPerLane = (kTiles + G - 1) / G;    // TPL
};

// The subgroup's tiles [t, min(t + TC, t1)) for row r, TPL consecutive
// tiles per lane: writes the B row indices of the set bits, tile by tile
// and ascending within a tile, to list[0..total).
template <int T, int G>
__device__ __forceinline__ int decode_chunk(int t, int t1, int r, int lane_sub,
                                            const int *__restrict__ bci,
                                            const typename Word<T>::type *__restrict__ words,
                                            int *list) {
  constexpr int TC = Chunk<T, G>::kTiles;
  constexpr int TPL = Chunk<T, G>::kPerLane;
  int bc[TPL];
  Bits<T> w[TPL];
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int ti = lane_sub * TPL + k;
    const int tile = t + ti;
    bc[k] = 0;
    w[k] = 0;
    if (ti < TC && tile < t1) {
      bc[k] = bci[tile];
      w[k] = words[static_cast<size_t>(tile) * T + r];
    }
    cnt += popcount(w[k]);
  }
  int inc = cnt;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const int v = __shfl_up(inc, d, G);
    if (lane_sub >= d) inc += v;
  }
  const int total = __shfl(inc, G - 1, G);
  int off = inc - cnt;
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int base = bc[k] * T + T - 1;
    Bits<T> x = w[k];
    while (x) {
      const int p = top_bit(x);
      x ^= static_cast<Bits<T>>(1) << p;
      list[off++] = base - p;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return total;
}

// Plan-free form when bsrVal lacks the planned kernels' alignment: one row
// per subgroup, Word<T> loads at their natural alignment; block rows of more
// than kLongChunks chunks are listed for b2sr_long_kernel.
template <int T, typename TB, int G, int VEC>
__global__ void __launch_bounds__(kB2Threads)
b2sr_spmm_kernel(int M, const int *__restrict__ brp, const int *__restrict__ bci,
                 const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                 int ldb, float *__restrict__ C, int ldc, int *__restrict__ state) {
  constexpr int EPL = VEC * Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  constexpr int U = EPL >= 16 ? 4 : 8;
  __shared__ int lds[kB2Waves * S][Chunk<T, G>::kCap];

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int gsub = wave * S + lane / G;
  const int row = blockIdx.x * (kB2Waves * S) + gsub;
  if (row >= M) return;
  const int br = row / T, r = row % T;
  const int col0 = blockIdx.y * CW + lane_sub * EPL;
  const TB *Bc = B + col0;
  int *list = lds[gsub];

  const int t0 = brp[br], t1 = brp[br + 1];
  if (t1 - t0 > kLongChunks * Chunk<T, G>::kTiles) {  // left to b2sr_long_kernel
    // one entry per block row: its row 0, first column tile
    if (blockIdx.y == 0 && r == 0 && lane_sub == 0) state[kStateHead + atomicAdd(&state[0], 1)] = br;
    return;
  }
  float acc[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
  for (int t = t0; t < t1; t += Chunk<T, G>::kTiles) {
    const int total = decode_chunk<T, G>(t, t1, r, lane_sub, bci, words, list);
    for (int k = 0; k < total; k += U) {
      // unconditional loads (slots past total re-read entry 0, which
      // exists since total > 0), one wait, predicated accumulation
      typename Vec16<TB>::Raw raw[U][VEC];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const TB *p = Bc + static_cast<size_t>(list[k + u < total ? k + u : 0]) * ldb;
#pragma unroll
        for (int q = 0; q < VEC; ++q) raw[u][q] = Vec16<TB>::raw(p + q * Vec16<TB>::kElems);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k + u < total) {
          float b[EPL];
#pragma unroll
          for (int q = 0; q < VEC; ++q) Vec16<TB>::widen(raw[u][q], &b[q * Vec16<TB>::kElems]);
#pragma unroll
          for (int q = 0; q < EPL; ++q) acc[q] += b[q];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // list is rewritten by the next chunk
  }
  float *Crow = C + static_cast<size_t>(row) * ldc + col0;
#pragma unroll
  for (int q = 0; q < EPL; q += 4)
    *reinterpret_cast<float4 *>(Crow + q) = make_float4(acc[q], acc[q + 1], acc[q + 2], acc[q + 3]);
}

// Block rows with more than kLongChunks chunks of tiles (the hub rows of a
// power-law graph) are appended by b2sr_spmm_kernel to a list that this
// persistent kernel drains one ROW of a listed block row per workgroup: its
// 4S subgroups take the row's tile chunks in turn (decode + gather), a
// butterfly sums each wave's subgroups and the waves' sums meet in LDS in a
// fixed order. The last workgroup to finish resets the list counters, so
// consecutive (or graph-replayed) calls need no memset.
template <int T, typename TB, int G, int VEC>
__global__ void __launch_bounds__(kB2Threads)
b2sr_long_kernel(int M, const int *__restrict__ brp, const int *__restrict__ bci,
                 const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                 int ldb, float *__restrict__ C, int ldc, int *__restrict__ state) {
  constexpr int EPL = VEC * Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  constexpr int U = EPL >= 16 ? 4 : 8;
  constexpr int TC = Chunk<T, G>::kTiles;
  __shared__ int lds[kB2Waves * S][Chunk<T, G>::kCap];
  __shared__ float red[kB2Waves - 1][CW];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int gsub = wave * S + sub;
  const int col0 = blockIdx.y * CW + lane_sub * EPL;
  const TB *Bc = B + col0;
  int *list = lds[gsub];
  const int n_items = state[0] * T;
  for (int i = blockIdx.x; i < n_items; i += gridDim.x) {
    const int br = state[kStateHead + i / T];
    const int r = i % T;
    const int row = br * T + r;
    if (row >= M) continue;  // whole workgroup
    const int t0 = brp[br], t1 = brp[br + 1];
    float acc[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
    for (int t = t0 + gsub * TC; t < t1; t += kB2Waves * S * TC) {
      const int total = decode_chunk<T, G>(t, t1, r, lane_sub, bci, words, list);
      for (int k = 0; k < total; k += U) {
        typename Vec16<TB>::Raw raw[U][VEC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const TB *p = Bc + static_cast<size_t>(list[k + u < total ? k + u : 0]) * ldb;
#pragma unroll
          for (int q = 0; q < VEC; ++q) raw[u][q] = Vec16<TB>::raw(p + q * Vec16<TB>::kElems);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (k + u < total) {
            float b[EPL];
#pragma unroll
            for (int q = 0; q < VEC; ++q) Vec16<TB>::widen(raw[u][q], &b[q * Vec16<TB>::kElems]);
#pragma unroll
            for (int q = 0; q < EPL; ++q) acc[q] += b[q];
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int m = G; m < kWave; m <<= 1) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] += __shfl_xor(acc[q], m);
    }
    if (wave > 0 && sub == 0) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) red[wave - 1][lane_sub * EPL + q] = acc[q];
    }
    __syncthreads();
    if (wave == 0 && sub == 0) {
#pragma unroll
      for (int w = 0; w < kB2Waves - 1; ++w) {
#pragma unroll
        for (int q = 0; q < EPL; ++q) acc[q] += red[w][lane_sub * EPL + q];
      }
      float *Crow = C + static_cast<size_t>(row) * ldc + col0;
#pragma unroll
      for (int q = 0; q < EPL; q += 4)
        *reinterpret_cast<float4 *>(Crow + q) =
            make_float4(acc[q], acc[q + 1], acc[q + 2], acc[q + 3]);
    }
    __syncthreads();  // red[] is reused by the next row
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Every workgroup has read the count before it counts itself done.
    if (atomicAdd(&state[1], 1) == static_cast<int>(gridDim.x * gridDim.y) - 1) {
      atomicExch(&state[0], 0);
      atomicExch(&state[1], 0);
    }
  }
}

// Row r's share of tiles [t, t1) taken in chunks at `stride`: each chunk's
// bits of row r decoded into the subgroup's list, then 8 gathers in flight.
template <int T, typename TB, int G, class LD>
__device__ __forceinline__ void b2sr_range(int t, int t1, int stride, int r, int lane_sub,
                                           const int *__restrict__ bci,
                                           const typename Word<T>::type *__restrict__ words,
                                           int *list, const LD &ld,
                                           float (&acc)[Vec16<TB>::kElems]) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int U = 8;
  for (; t < t1; t += stride) {
    const int total = decode_chunk<T, G>(t, t1, r, lane_sub, bci, words, list);
    for (int k = 0; k < total; k += U) {
      typename Vec16<TB>::Raw raw[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        raw[u] = ld(list[k + u < total ? k + u : 0]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k + u < total) {
          float b[EPL];
          Vec16<TB>::widen(raw[u], b);
#pragma unroll
          for (int q = 0; q < EPL; ++q) acc[q] += b[q];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // list is rewritten by the next chunk
  }
}

// ------------------------------------------- block-row FBF (T = 4 and 8)
// For the small tiles the planned launch schedules BLOCK rows (binned by
// tile count), so each tile is decoded once for all T rows. A tile is one
// 32-bit (T=4) or 64-bit (T=8) word, byte r holding row r. A group of L
// lanes (a subgroup, or a whole wave) loads 4 consecutive tiles per lane;
// the T per-row bit counts go through one group scan as packed 16-bit
// pairs, and the LDS list is written row-segmented (row 0's B rows, then
// row 1's, ...), ascending in column within each row, so every row is
// summed in ascending column order exactly as the CSR kernel sums it. If
// the bits exceed the list (dense tiles) it is filled and drained in
// windows; each window reloads its tiles, so nothing of the decode stays
// live while the gathers run.
//
// Classes by tile count: up to one subgroup chunk (4G tiles, most block
// rows of a power-law graph) -> one subgroup per block row, streamed; up to
// one wave chunk (256 tiles) -> one wave per block row, rows in turn, each
// row's entries dealt round-robin to the subgroups; wider -> one workgroup
// per row of the block row (its 16 subgroups take the tile chunks in turn,
// decoding only that row's bits; the widest block rows set the kernel's
// critical path, so they get the most lanes).
template <int T> struct TileWord;
template <> struct TileWord<2> { using type = uint16_t; };
template <> struct TileWord<4> { using type = uint32_t; };
template <> struct TileWord<8> { using type = uint64_t; };

template <int G> struct BrowChunk {
  static constexpr int kCap = 16 * G;  // list entries per subgroup (16 KB per WG)
  static constexpr int kPerLane = 4;   // tiles per lane per chunk
  static constexpr int kTiles = kPerLane * G;
  static constexpr int kWaveTiles = kPerLane * kWave;  // 256
  static constexpr int kWaveCap = 16 * kWave;          // the wave's subgroup lists, contiguous
};

// Decode tiles [t, min(t + 4L, t1)) by a group of L lanes (lane_l its lane)
// into list window [base, base + CAPL): entry = B row, or with kFlags
// (B row << 4) | (last of its row << 3) | row. Returns the chunk's bit
// count; seg[r] = first position of row r (seg[T] = total).
template <int T, int L, int CAPL, bool kFlags>
__device__ __forceinline__ int decode_rows(int t, int t1, int lane_l, int base,
                                           const int *__restrict__ bci,
                                           const typename TileWord<T>::type *__restrict__ tw,
                                           int *list, int (&seg)[T + 1]) {
  using TWd = typename TileWord<T>::type;
  constexpr int TPL = 4;
  constexpr int NP = T / 2;
  int bc[TPL];
  TWd w[TPL];
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int tile = t + lane_l * TPL + k;
    bc[k] = 0;
    w[k] = 0;
    if (tile < t1) {
      bc[k] = bci[tile];
      w[k] = tw[tile];
    }
  }
  uint32_t pk[NP], inc[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    pk[j] = 0;
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      const uint32_t lo = static_cast<uint32_t>(w[k] >> (16 * j)) & 0xffu;
      const uint32_t hi = static_cast<uint32_t>(w[k] >> (16 * j + 8)) & 0xffu;
      pk[j] += __popc(lo) | (__popc(hi) << 16);
    }
    inc[j] = pk[j];
  }
#pragma unroll
  for (int d = 1; d < L; d <<= 1) {
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const uint32_t v = __shfl_up(inc[j], d, L);
      if (lane_l >= d) inc[j] += v;
    }
  }
  seg[0] = 0;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint32_t tot = __shfl(inc[j], L - 1, L);
    seg[2 * j + 1] = seg[2 * j] + static_cast<int>(tot & 0xffffu);
    seg[2 * j + 2] = seg[2 * j + 1] + static_cast<int>(tot >> 16);
  }
  // this lane's next position per row, two rows per word like the counts
  // (positions stay below 2^16: at most 4L tiles x T*T bits)
  uint32_t Q[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j)
    Q[j] = (inc[j] - pk[j]) + static_cast<uint32_t>(seg[2 * j]) +
           (static_cast<uint32_t>(seg[2 * j + 1]) << 16);
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    TWd x = w[k];
    const int b0 = bc[k] * T + T - 1;
    while (x) {
      // highest set bit first: last row first, lowest column first
      const int q = (sizeof(TWd) == 8 ? 63 - __clzll(static_cast<unsigned long long>(x))
                                      : 31 - __clz(static_cast<uint32_t>(x)));
      x ^= static_cast<TWd>(1) << q;
      const int r = q >> 3;
      const int sh = 16 * (r & 1);
      uint32_t cq = Q[0];
#pragma unroll
      for (int j = 1; j < NP; ++j) cq = (r >> 1) == j ? Q[j] : cq;
      const int pos = static_cast<int>((cq >> sh) & 0xffffu);
      const int e = pos - base;
      if (e >= 0 && e < CAPL) {
        int v = b0 - (q & 7);
        if (kFlags) {
          int end = seg[1];
#pragma unroll
          for (int i = 1; i < T; ++i) end = r == i ? seg[i + 1] : end;
          v = (v << 4) | ((pos + 1 == end) << 3) | r;
        }
        list[e] = v;
      }
#pragma unroll
      for (int j = 0; j < NP; ++j)
        if ((r >> 1) == j) Q[j] += 1u << sh;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return seg[T];
}

// One subgroup, one block row of at most one subgroup chunk: the list is
// streamed flat in groups of 8 gathers into ONE running sum, stored to C at
// each row's last entry (a block row of ~40 bits costs ~5 gather round
// trips). Rows without entries are zeroed.
template <int T, typename TB, int G, class LD, bool kAcc = false>
__device__ __forceinline__ void brow_stream(int t, int t1, int lane_sub,
                                            const int *__restrict__ bci,
                                            const typename TileWord<T>::type *__restrict__ tw,
                                            int *list, const LD &ld, float *__restrict__ C,
                                            int ldc, int row0, int M, int col0) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int CAP = BrowChunk<G>::kCap;
  float cur[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
  int total = 1;
  for (int base = 0; base < total; base += CAP) {
    int seg[T + 1];
    total = decode_rows<T, G, CAP, true>(t, t1, lane_sub, base, bci, tw, list, seg);
    if (base == 0 && !kAcc) {  // (a later column slice adds nothing to such rows)
#pragma unroll
      for (int r = 0; r < T; ++r)
        if (seg[r] == seg[r + 1] && row0 + r < M) store_row<EPL>(C, ldc, row0 + r, col0, cur);
    }
    const int n = min(CAP, total - base);
    for (int k = 0; k < n; k += 8) {
      typename Vec16<TB>::Raw raw[8];
      uint32_t meta = 0;  // low 4 bits (end flag, row) of the 8 entries
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t ent = static_cast<uint32_t>(list[k + u < n ? k + u : k]);
        meta |= (ent & 15u) << (4 * u);
        raw[u] = ld(static_cast<int>(ent >> 4));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (k + u < n) {
          float b[EPL];
          Vec16<TB>::widen(raw[u], b);
#pragma unroll
          for (int q = 0; q < EPL; ++q) cur[q] += b[q];
          const uint32_t m = meta >> (4 * u);
          if (m & 8u) {
            const int row = row0 + static_cast<int>(m & 7u);
            if (row < M) put_c_row<EPL, kAcc>(C, ldc, row, col0, cur);
#pragma unroll
            for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // list is rewritten next
  }
}

// One wave, one block row of at most one wave chunk (256 tiles): decoded
// once by all 64 lanes into the wave's lists; then row by row, the row's
// entries dealt round-robin to the S subgroups (8 gathers in flight each),
// a butterfly over the subgroups in a fixed order, subgroup 0 stores.
template <int T, typename TB, int G, class LD, bool kAcc = false>
__device__ __forceinline__ void brow_wave(int t, int t1, int lane, const int *__restrict__ bci,
                                          const typename TileWord<T>::type *__restrict__ tw,
                                          int *wlist, const LD &ld, float *__restrict__ C,
                                          int ldc, int row0, int M, int col0) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CAPW = BrowChunk<G>::kWaveCap;
  const int sub = lane / G;
  float cur[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
  int total = 1;
  for (int base = 0; base < total; base += CAPW) {
    int seg[T + 1];
    total = decode_rows<T, kWave, CAPW, false>(t, t1, lane, base, bci, tw, wlist, seg);
#pragma unroll
    for (int r = 0; r < T; ++r) {
      if (seg[r + 1] <= base && !(base == 0 && seg[r] == seg[r + 1])) continue;  // done earlier
      if (seg[r] >= base + CAPW) continue;                                     // a later window
      const int hi = min(seg[r + 1], base + CAPW) - base;
      for (int k = max(seg[r], base) - base + sub; k < hi; k += 8 * S) {
        typename Vec16<TB>::Raw raw[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          raw[u] = ld(wlist[k + u * S < hi ? k + u * S : k]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (k + u * S < hi) {
            float b[EPL];
            Vec16<TB>::widen(raw[u], b);
#pragma unroll
            for (int q = 0; q < EPL; ++q) cur[q] += b[q];
          }
        }
      }
      if (seg[r + 1] <= base + CAPW) {  // row complete
#pragma unroll
        for (int m = G; m < kWave; m <<= 1) {
#pragma unroll
          for (int q = 0; q < EPL; ++q) cur[q] += __shfl_xor(cur[q], m);
        }
        if (sub == 0 && row0 + r < M && !(kAcc && seg[r] == seg[r + 1]))
          put_c_row<EPL, kAcc>(C, ldc, row0 + r, col0, cur);
#pragma unroll
        for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
      }
    }
    __builtin_amdgcn_wave_barrier();  // lists are rewritten next
  }
}

// The work of virtual workgroup `vwg` of the block-row launch table. With
// lo != nullptr, plan unit i's tiles are [lo[i], hi[i]) (one column slice of
// the block row: the plan's classes, over spans no longer than the whole
// block row's), and kAcc adds the slice's sums into C.
template <int T, typename TB, int G, bool kBuf, bool kAcc = false>
__device__ __forceinline__ void brow_wg(const LaunchTable &tab, int vwg, int *lds,
                                        const int *__restrict__ brows, int M,
                                        const int2 *__restrict__ span,
                                        const int *__restrict__ lo, const int *__restrict__ hi,
                                        const int *__restrict__ bci,
                                        const typename TileWord<T>::type *__restrict__ tw,
                                        const TB *__restrict__ B, int ldb, int b_bytes,
                                        float *__restrict__ C, int ldc) {
  auto span_of = [&](int pos) { return lo ? make_int2(lo[pos], hi[pos]) : span[pos]; };
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  constexpr int CAP = BrowChunk<G>::kCap;
  const Seg sg = find_seg(tab, vwg);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int gsub = wave * S + sub;
  const int col0 = blockIdx.y * CW + lane_sub * EPL;
  const RowLoader<TB, kBuf> ld(B, ldb, col0, b_bytes);
  const int local_wg = vwg - sg.wg_begin;

  if (sg.kind == kSegA) {
    int *list = lds + gsub * CAP;
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kB2Waves * S;
    const int end = sg.list_off + sg.count;
    int pos = sg.list_off + local_wg * subs_per_wg * R + gsub;
    int br = -1;
    int2 sp = make_int2(0, 0);
    if (pos < end) {  // the block row and its tile span arrive together
      br = brows[pos];
      sp = span_of(pos);
    }
    for (int k = 0; k < R && br >= 0; ++k) {
      const int pos_n = pos + subs_per_wg;
      int br_n = -1;
      int2 sp_n = make_int2(0, 0);
      if (k + 1 < R && pos_n < end) {
        br_n = brows[pos_n];
        sp_n = span_of(pos_n);
      }
      brow_stream<T, TB, G, RowLoader<TB, kBuf>, kAcc>(sp.x, sp.y, lane_sub, bci, tw, list, ld, C,
                                                       ldc, br * T, M, col0);
      br = br_n;
      sp = sp_n;
      pos = pos_n;
    }
    return;
  }
  if (sg.kind == kSegB) {
    const int pos = sg.list_off + local_wg * kB2Waves + wave;
    if (pos >= sg.list_off + sg.count) return;  // whole wave
    const int br = brows[pos];
    const int2 sp = span_of(pos);
    brow_wave<T, TB, G, RowLoader<TB, kBuf>, kAcc>(sp.x, sp.y, lane, bci, tw, lds + wave * S * CAP,
                                                   ld, C, ldc, br * T, M, col0);
    return;
  }
  // class C: one workgroup per (block row, row): its 4S subgroups take the
  // tile chunks in turn decoding only that row's bits, a butterfly sums each
  // wave, and the waves' sums are added in a fixed order through LDS (the
  // list area, free once every subgroup is done).
  const int br = brows[sg.list_off + local_wg / T];
  const int2 sp = span_of(sg.list_off + local_wg / T);
  const int r = local_wg % T;
  const int row = br * T + r;
  if (row >= M) return;  // whole workgroup
  int *list = lds + gsub * CAP;
  const int t0 = sp.x, t1 = sp.y;
  const auto *words = reinterpret_cast<const typename Word<T>::type *>(tw);
  constexpr int RC = Chunk<T, G>::kTiles;
  float acc[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
  b2sr_range<T, TB, G>(t0 + gsub * RC, t1, kB2Waves * S * RC, r, lane_sub, bci, words, list, ld,
                       acc);
#pragma unroll
  for (int m = G; m < kWave; m <<= 1) {
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] += __shfl_xor(acc[q], m);
  }
  static_assert((kB2Waves - 1) * CW <= kB2Waves * S * CAP, "reduction must fit the list area");
  float *red = reinterpret_cast<float *>(lds);
  __syncthreads();
  if (wave > 0 && sub == 0) {
#pragma unroll
    for (int q = 0; q < EPL; ++q) red[(wave - 1) * CW + lane_sub * EPL + q] = acc[q];
  }
  __syncthreads();
  if (wave == 0 && sub == 0) {
#pragma unroll
    for (int w = 0; w < kB2Waves - 1; ++w) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] += red[w * CW + lane_sub * EPL + q];
    }
    put_c_row<EPL, kAcc>(C, ldc, row, col0, acc);
  }
}

// lo == nullptr: the plan's whole block rows; else column slice [lo, hi)
// of each (kAcc: the slices after the first add into C).
template <int T, typename TB, int G, bool kBuf, bool kAcc>
__global__ void __launch_bounds__(kB2Threads)
b2sr_brow_kernel(LaunchTable tab, const int *__restrict__ brows, int M,
                 const int2 *__restrict__ span, const int *__restrict__ lo,
                 const int *__restrict__ hi, const int *__restrict__ bci,
                 const typename TileWord<T>::type *__restrict__ tw, const TB *__restrict__ B,
                 int ldb, int b_bytes, float *__restrict__ C, int ldc) {
  constexpr int S = kWave / G;
  constexpr int CAP = BrowChunk<G>::kCap;
  static_assert(Chunk<T, G>::kCap == CAP, "all paths share the list area");
  __shared__ int lds[kB2Waves * S][CAP];
  brow_wg<T, TB, G, kBuf, kAcc>(tab, blockIdx.x, &lds[0][0], brows, M, span, lo, hi, bci, tw, B,
                                ldb, b_bytes, C, ldc);
}

// ------------------------------------------- plan-free form, T <= 8
// One launch over the classes class_list_kernel listed, with the planned
// form's per-class code and static shares (one counter shared by every
// workgroup would serialise their atomics across the XCDs). The first E
// workgroups (as many as the chip holds at once, or half) take the listed work:
// workgroup x the workgroup-class rows x, x + E, ... (row i = listed block
// row i / T, row i % T), and the wave-class block rows are dealt to their
// waves from the last of them backwards, so they land on workgroups with
// one workgroup-class row fewer. Workgroup E + k then streams block rows
// k SW .. k SW + SW - 1, one subgroup each, skipping the listed ones. The
// last workgroup to finish zeroes the list counts, so consecutive or
// graph-replayed calls need no memset.
//
// kSorted (with a handle): the block rows come sorted by tile-count bin,
// heaviest first (rows; meta[0] workgroup-class and meta[1] wave-class
// block rows at the head, the classes above), from dropin_bin_sort. The
// first E workgroups deal the two listed classes as above, and workgroup
// E + k takes the subgroup-class block rows at sorted positions
// k SW .. k SW + SW - 1 past them, so its subgroups get block rows of
// similar size. No list counts to reset.
template <int T, typename TB, int G, bool kBuf, bool kSorted>
__global__ void __launch_bounds__(kB2Threads)
b2sr_dropin_kernel(int M, int E, const int *__restrict__ brp, const int *__restrict__ bci,
                   const typename TileWord<T>::type *__restrict__ tw, const TB *__restrict__ B,
                   int ldb, int b_bytes, float *__restrict__ C, int ldc,
                   int *__restrict__ state, const int *__restrict__ rows,
                   const int *__restrict__ meta) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int S = kWave / G, SW = kB2Waves * S;
  constexpr int CW = G * EPL;
  constexpr int CAP = BrowChunk<G>::kCap;
  constexpr int RC = Chunk<T, G>::kTiles;
  constexpr int TA = BrowChunk<G>::kTiles;
  static_assert(Chunk<T, G>::kCap == CAP, "all paths share the list area");
  static_assert((kB2Waves - 1) * CW <= SW * CAP, "reduction must fit the list area");
  __shared__ int lds[SW * CAP];
  const int MB = (M + T - 1) / T;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int gsub = wave * S + sub;
  const int col0 = blockIdx.y * CW + lane_sub * EPL;
  const RowLoader<TB, kBuf> ld(B, ldb, col0, b_bytes);
  const int x = blockIdx.x;

  // workgroup-class block row i, wave-class block row j
  auto wg_brow = [&](int i) { return kSorted ? rows[i] : state[kStateHead + i]; };
  auto wave_brow = [&](int j) { return kSorted ? rows[meta[0] + j] : state[kStateHead + MB + j]; };
  if (x < E) {
    const int n_rows = (kSorted ? meta[0] : state[0]) * T;
    const int n_brows = kSorted ? meta[1] : state[2];
    const auto *words = reinterpret_cast<const typename Word<T>::type *>(tw);
    for (int i = x; i < n_rows; i += E) {  // workgroup class
      const int br = wg_brow(i / T);
      const int r = i % T;
      const int row = br * T + r;
      if (row >= M) continue;  // whole workgroup
      float acc[EPL];
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
      b2sr_range<T, TB, G>(brp[br] + gsub * RC, brp[br + 1], SW * RC, r, lane_sub, bci, words,
                           lds + gsub * CAP, ld, acc);
#pragma unroll
      for (int m = G; m < kWave; m <<= 1) {
#pragma unroll
        for (int q = 0; q < EPL; ++q) acc[q] += __shfl_xor(acc[q], m);
      }
      float *red = reinterpret_cast<float *>(lds);
      __syncthreads();
      if (wave > 0 && sub == 0) {
#pragma unroll
        for (int q = 0; q < EPL; ++q) red[(wave - 1) * CW + lane_sub * EPL + q] = acc[q];
      }
      __syncthreads();
      if (wave == 0 && sub == 0) {
#pragma unroll
        for (int w = 0; w < kB2Waves - 1; ++w) {
#pragma unroll
          for (int q = 0; q < EPL; ++q) acc[q] += red[w * CW + lane_sub * EPL + q];
        }
        store_row<EPL>(C, ldc, row, col0, acc);
      }
      __syncthreads();  // red is the next row's list area
    }
    for (int j = (E - 1 - x) * kB2Waves + wave; j < n_brows; j += E * kB2Waves) {  // wave class
      const int br = wave_brow(j);
      brow_wave<T, TB, G>(brp[br], brp[br + 1], lane, bci, tw, lds + wave * S * CAP, ld, C, ldc,
                          br * T, M, col0);
    }
  } else if constexpr (kSorted) {  // subgroup class, sorted
    const int pos = meta[0] + meta[1] + (x - E) * SW + gsub;
    if (pos < MB) {
      const int br = rows[pos];
      brow_stream<T, TB, G>(brp[br], brp[br + 1], lane_sub, bci, tw, lds + gsub * CAP, ld, C, ldc,
                            br * T, M, col0);
    }
  } else {
    const int br = (x - E) * SW + gsub;  // subgroup class
    if (br < MB) {
      const int t0 = brp[br], t1 = brp[br + 1];
      if (t1 - t0 < TA)  // else listed
        brow_stream<T, TB, G>(t0, t1, lane_sub, bci, tw, lds + gsub * CAP, ld, C, ldc, br * T,
                              M, col0);
    }
  }
  if constexpr (kSorted) return;
  __syncthreads();
  if (threadIdx.x == 0) {
    // every workgroup has read the counts and lists before it counts itself
    if (atomicAdd(&state[1], 1) == static_cast<int>(gridDim.x * gridDim.y) - 1) {
      atomicExch(&state[0], 0);
      atomicExch(&state[2], 0);
      atomicExch(&state[1], 0);
    }
  }
}

// Narrow or unaligned N (N not a multiple of one 16-byte vector, or B / C /
// ld not 16-byte aligned; N = 1 is a binary-A SpMV): L lanes per block row
// (64, or a workgroup for T >= 16), lane (slot, j) keeps tile row j and takes
// the block row's tiles slot, slot + L/T, ... four at a time (word and block
// column loads clamped to a valid tile, so they are in flight together).
// Each round takes the highest remaining set bit of each of the four words
// and gathers its NB columns of B together; a butterfly over the slots (and
// LDS across waves) sums each row, slot 0 stores. Wider N runs in passes of
// NB columns.
template <int T, int L, int NB, typename TB>
__global__ void __launch_bounds__(kB2Threads)
b2sr_narrow_kernel(int M, int N, const int *__restrict__ brp, const int *__restrict__ bci,
                   const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                   int ldb, float *__restrict__ C, int ldc) {
  static_assert(L == kWave || L == kB2Threads, "a wave or a workgroup per block row");
  constexpr int SLOTS = L / T, U = 4;
  constexpr int WL = L < kWave ? L : kWave;
  __shared__ float red[kB2Waves][T][NB];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int sl = tid % L, j = sl % T, slot = sl / T;
  const int MB = (M + T - 1) / T;
  const int br = blockIdx.x * (kB2Threads / L) + tid / L;
  const bool live = br < MB;
  int t0 = 0, t1 = 0;
  if (live) {
    t0 = brp[br];
    t1 = brp[br + 1];
  }
  const int row = br * T + j;
  for (int n0 = 0; n0 < N; n0 += NB) {
    float acc[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[n] = 0.f;
    for (int t = t0 + slot; t < t1; t += SLOTS * U) {
      Bits<T> a[U];
      int bc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = min(t + u * SLOTS, t1 - 1);
        a[u] = words[static_cast<size_t>(tt) * T + j];
        bc[u] = bci[tt];
      }
      bool more = false;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (t + u * SLOTS >= t1) a[u] = 0;
        more = more || a[u] != 0;
      }
      while (more) {
        float v[U][NB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int b = a[u] ? top_bit(a[u]) : T - 1;  // empty words read column 0
          const TB *Brow = B + static_cast<size_t>(bc[u] * T + T - 1 - b) * ldb;
#pragma unroll
          for (int n = 0; n < NB; ++n) v[u][n] = load_scalar<TB>(Brow + min(n0 + n, N - 1));
        }
        more = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (a[u]) {
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[n] += v[u][n];
            a[u] ^= static_cast<Bits<T>>(1) << top_bit(a[u]);
          }
          more = more || a[u] != 0;
        }
      }
    }
#pragma unroll
    for (int o = T; o < WL; o <<= 1) {
#pragma unroll
      for (int n = 0; n < NB; ++n) acc[n] += __shfl_xor(acc[n], o, kWave);
    }
    if constexpr (L > kWave) {  // the waves' partial sums meet in LDS
      if (lane < T) {
#pragma unroll
        for (int n = 0; n < NB; ++n) red[tid / kWave][lane][n] = acc[n];
      }
      __syncthreads();
      if (tid < T) {
#pragma unroll
        for (int w = 1; w < kB2Waves; ++w) {
#pragma unroll
          for (int n = 0; n < NB; ++n) acc[n] += red[w][tid][n];
        }
      }
      __syncthreads();  // red[] is reused by the next pass
    }
    if (live && slot == 0 && (L <= kWave || tid < T) && row < M) {
#pragma unroll
      for (int n = 0; n < NB; ++n)
        if (n0 + n < N) C[static_cast<size_t>(row) * ldc + n0 + n] = acc[n];
    }
  }
}

// N <= 8, T <= 8: B staged in LDS, the lane mapping of the reference's
// spmm4_full_full_1_1024_lessthan8 (backend/spmm.cu:417-475) for wave64. A
// wave per block row takes Q = 64 / T tiles a round: lane l stages the N
// values of B row (block column of tile l / T) * T + l % T into the wave's
// slab and loads tile l / T's row l % T word; then output lane o = (r, n)
// (o < T N) adds, tile by tile and column by column (so in ascending column
// order, as the CSR kernels sum), the staged values of its row's set bits.
// The slab is rewritten next round.
template <int T, typename TB>
__global__ void __launch_bounds__(kB2Threads)
b2sr_lds_kernel(int M, int N, const int *__restrict__ brp, const int *__restrict__ bci,
                const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                int ldb, float *__restrict__ C, int ldc) {
  constexpr int Q = kWave / T;
  constexpr int SP = 9;  // slab stride (N <= 8), odd to spread the LDS banks
  __shared__ float slab[kB2Waves][kWave * SP];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int MB = (M + T - 1) / T;
  const int br = blockIdx.x * kB2Waves + wave;
  if (br >= MB) return;  // whole wave; the kernel has no workgroup barrier
  const int t0 = brp[br], t1 = brp[br + 1];
  const bool out = lane < T * N;
  const int r = out ? lane / N : 0, n = out ? lane % N : 0;
  float *sw = slab[wave];
  float acc = 0.f;
  for (int t = t0; t < t1; t += Q) {
    const int tl = t + lane / T, c = lane % T;
    uint32_t w = 0;
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = 0.f;
    if (tl < t1) {
      w = words[static_cast<size_t>(tl) * T + c];  // tile tl, row c
      const TB *Brow = B + static_cast<size_t>(bci[tl] * T + c) * ldb;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < N) v[k] = load_scalar<TB>(Brow + k);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sw[lane * SP + k] = v[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nq = min(Q, t1 - t);
    for (int q = 0; q < nq; ++q) {
      const uint32_t bits = __shfl(w, q * T + r);  // tile q's row r
      if (out) {
#pragma unroll
        for (int cc = 0; cc < T; ++cc)
          if ((bits >> (T - 1 - cc)) & 1u) acc += sw[(q * T + cc) * SP + n];
      }
    }
    __builtin_amdgcn_wave_barrier();  // the slab is rewritten next round
  }
  if (out && br * T + r < M) C[static_cast<size_t>(br * T + r) * ldc + n] = acc;
}

template <int T, typename TB>
static void launch_lds(int M, int N, const int *brp, const int *bci, const void *val,
                       const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  const int MB = (M + T - 1) / T;
  hipLaunchKernelGGL((b2sr_lds_kernel<T, TB>), dim3(ceil_div(MB, kB2Waves)), dim3(kB2Threads), 0,
                     s, M, N, brp, bci, static_cast<const typename Word<T>::type *>(val), B, ldb,
                     C, ldc);
}

// Binary A x sign-packed B -> int32. Lane owns 4 consecutive columns and
// counts the +1 bits; C = 2 * count - (number of stored entries in the row).
template <int T, int G>
__global__ void __launch_bounds__(kB2Threads)
b2sr_sign_kernel(int M, int N, const int *__restrict__ brp, const int *__restrict__ bci,
                 const typename Word<T>::type *__restrict__ words,
                 const uint32_t *__restrict__ Bw, int ldbw, int32_t *__restrict__ C, int ldc) {
  constexpr int S = kWave / G;
  constexpr int U = 8;
  __shared__ int lds[kB2Waves * S][Chunk<T, G>::kCap];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int gsub = wave * S + lane / G;
  const int row = blockIdx.x * (kB2Waves * S) + gsub;
  if (row >= M) return;
  const int br = row / T, r = row % T;
  const int col0 = (blockIdx.y * G + lane_sub) * 4;
  const bool has_cols = col0 < N;
  const int wi = has_cols ? col0 / 32 : 0;
  const int sh = 28 - (col0 % 32);  // bits (31-c%32) .. for 4 columns
  int *list = lds[gsub];
  int cnt[4] = {0, 0, 0, 0};
  int deg = 0;
  const int t1 = brp[br + 1];
  for (int t = brp[br]; t < t1; t += Chunk<T, G>::kTiles) {
    const int total = decode_chunk<T, G>(t, t1, r, lane_sub, bci, words, list);
    deg += total;
    for (int k = 0; k < total; k += U) {
      uint32_t wv[U];
#pragma unroll
      for (int u = 0; u < U; ++u)  // unconditional: past total re-read entry 0
        wv[u] = Bw[static_cast<size_t>(list[k + u < total ? k + u : 0]) * ldbw + wi];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // columns col0..col0+3 at bits 3..0; slots past total count nothing
        const uint32_t nib = (k + u < total && has_cols) ? (wv[u] >> sh) & 0xfu : 0u;
        cnt[0] += (nib >> 3) & 1u;
        cnt[1] += (nib >> 2) & 1u;
        cnt[2] += (nib >> 1) & 1u;
        cnt[3] += nib & 1u;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  int32_t *Crow = C + static_cast<size_t>(row) * ldc;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (col0 + q < N) Crow[col0 + q] = 2 * cnt[q] - deg;
}

// Sign-packed B as +-1 BF16 (+1 where the bit is set). Column c of a row
// is bit 31 - c % 32 of word c / 32 (pack_sign_rows_u32's order). kVec
// (N a multiple of 8): one thread per 8 columns, one 16-byte store, so a
// wave writes 1 KB of consecutive bytes; else one thread per column. The
// item count is below 2^32 (the caller checks).
template <bool kVec>
__global__ void __launch_bounds__(kB2Threads)
sign_expand_kernel(uint32_t items, int N, const uint32_t *__restrict__ Bw, int ldbw,
                   uint16_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * kB2Threads + threadIdx.x;
  if (i >= items) return;
  const uint32_t per_row = kVec ? static_cast<uint32_t>(N) / 8 : static_cast<uint32_t>(N);
  const uint32_t k = i / per_row, j = i - k * per_row;
  const uint32_t c = kVec ? 8 * j : j;  // first column
  const uint32_t w = Bw[static_cast<size_t>(k) * ldbw + c / 32];
  if constexpr (kVec) {
    uint32_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t b = 31 - (c % 32) - 2 * q;  // columns c + 2q, c + 2q + 1
      const uint32_t lo = (w >> b) & 1u ? 0x3f80u : 0xbf80u;
      const uint32_t hi = (w >> (b - 1)) & 1u ? 0x3f80u : 0xbf80u;
      v[q] = lo | (hi << 16);
    }
    *reinterpret_cast<uint4 *>(out + static_cast<size_t>(k) * N + c) =
        make_uint4(v[0], v[1], v[2], v[3]);
  } else {
    out[static_cast<size_t>(k) * N + c] = (w >> (31 - c % 32)) & 1u ? 0x3f80u : 0xbf80u;
  }
}

// C[i][0..N) in place: FP32 values (whole numbers below 2^24) -> int32.
// kVec (N and ldc multiples of 4, C 16-byte aligned): one thread per 16
// bytes; else one per element. The item count is below 2^32.
template <bool kVec>
__global__ void __launch_bounds__(kB2Threads)
f32_to_i32_kernel(uint32_t items, int N, int32_t *__restrict__ C, int ldc) {
  const uint32_t i = blockIdx.x * kB2Threads + threadIdx.x;
  if (i >= items) return;
  const uint32_t per_row = kVec ? static_cast<uint32_t>(N) / 4 : static_cast<uint32_t>(N);
  const uint32_t r = i / per_row, j = i - r * per_row;
  int32_t *p = C + static_cast<size_t>(r) * ldc + (kVec ? 4 * j : j);
  if constexpr (kVec) {
    int4 v = *reinterpret_cast<const int4 *>(p);
    v.x = static_cast<int32_t>(__int_as_float(v.x));
    v.y = static_cast<int32_t>(__int_as_float(v.y));
    v.z = static_cast<int32_t>(__int_as_float(v.z));
    v.w = static_cast<int32_t>(__int_as_float(v.w));
    *reinterpret_cast<int4 *>(p) = v;
  } else {
    *p = static_cast<int32_t>(__int_as_float(*p));
  }
}

// ------------------------------------------------------- dense packers
// One wave per (row, 64-column chunk): a 64-bit ballot gives two words.
// cols == false: out[h][w] (row-major, Wc words per row);
// cols == true : out[w][h] over 32*ceil(H/32) rows (padding rows are 0).
template <bool kSign, bool kCols>
__global__ void __launch_bounds__(kB2Threads)
pack_kernel(int H, int W, const float *__restrict__ X, int ldx, uint32_t *__restrict__ out) {
  const int Wc = (W + 31) / 32;
  const int Hp = kCols ? 32 * ((H + 31) / 32) : H;
  const int chunks = (W + 63) / 64;
  const long long item = (static_cast<long long>(blockIdx.x) * kB2Threads + threadIdx.x) / kWave;
  if (item >= static_cast<long long>(Hp) * chunks) return;
  const int h = static_cast<int>(item / chunks), j = static_cast<int>(item % chunks);
  const int lane = threadIdx.x & (kWave - 1);
  const int col = 64 * j + lane;
  bool pred = false;
  if (h < H && col < W) {
    const float x = X[static_cast<size_t>(h) * ldx + col];
    pred = kSign ? (x >= 0.f) : (x > 0.f);
  }
  const unsigned long long m = __ballot(pred);
  const int w = 2 * j + lane;
  if (lane < 2 && w < Wc) {
    const uint32_t word = __brev(static_cast<uint32_t>(lane == 0 ? m : (m >> 32)));
    if (kCols)
      out[static_cast<size_t>(w) * Hp + h] = word;
    else
      out[static_cast<size_t>(h) * Wc + w] = word;
  }
}

__global__ void __launch_bounds__(kB2Threads)
unpack_cols_kernel(int H, int W, const uint32_t *__restrict__ in, float *__restrict__ X, int ldx) {
  const long long i = static_cast<long long>(blockIdx.x) * kB2Threads + threadIdx.x;
  if (i >= static_cast<long long>(H) * W) return;
  const int h = static_cast<int>(i / W), c = static_cast<int>(i % W);
  const int Hp = 32 * ((H + 31) / 32);
  const uint32_t v = in[static_cast<size_t>(c / 32) * Hp + h];
  X[static_cast<size_t>(h) * ldx + c] = static_cast<float>((v >> (31 - c % 32)) & 1u);
}

// ------------------------------------------------------------------ host
template <int T, bool kFill>
static void launch_convert(int M, int K, const int *rowptr, const int *colind, int *brp,
                           int *bci, void *val, int *err, int *hubs, hipStream_t s) {
  const int MB = (M + T - 1) / T;
  const int waves = ceil_div(MB, kWave / T);
  const int blocks = ceil_div(waves, kB2Waves);
  // hub block rows to the bitmap kernel when K / T fits its LDS bitmap (and,
  // for the fill, bsrVal takes 32-bit atomics); the count and the fill pass
  // of one conversion make the same choice
  const bool use_hubs = MB > 0 && (K + T - 1) / T <= kHubBitmapWords * 32 &&
                        reinterpret_cast<uintptr_t>(val) % 4 == 0;
  const int hub_entries = use_hubs ? kHubEntries : INT_MAX;
  auto *w = static_cast<typename Word<T>::type *>(val);
  hipLaunchKernelGGL((csr2b2sr_kernel<T, kFill>), dim3(blocks), dim3(kB2Threads), 0, s, M, K,
                     rowptr, colind, brp, bci, w, err, hub_entries, hubs);
  if (use_hubs)  // a persistent grid drains the list the merge kernel filled
    hipLaunchKernelGGL((csr2b2sr_hub_kernel<T, kFill>), dim3(min(MB, 1024)), dim3(kB2Threads),
                       0, s, M, K, rowptr, colind, brp, bci, w, err, hubs);
}

// hubs: 1 + MB ints of device workspace, hubs[0] zeroed on s
template <bool kFill>
static void convert(int T, int M, int K, const int *rowptr, const int *colind, int *brp,
                    int *bci, void *val, int *err, int *hubs, hipStream_t s) {
  switch (T) {
    case 2: launch_convert<2, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
    case 4: launch_convert<4, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
    case 8: launch_convert<8, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
    case 16: launch_convert<16, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
    case 32: launch_convert<32, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
    default: launch_convert<64, kFill>(M, K, rowptr, colind, brp, bci, val, err, hubs, s); break;
  }
}

struct Geo {
  int G = 0, VEC = 1, tiles = 1, rows_scale = 1, slices = 1;
};

static Geo geometry(int N, int elem_bytes, const void *B, int ldb, const void *C, int ldc) {
  Geo g;
  const int e16 = 16 / elem_bytes;
  if (!(aligned16(B) && aligned16(C) && ldb % e16 == 0 && ldc % 4 == 0) || N % e16 != 0) return g;
  const int lanes = N / e16;
  for (int t = 64; t >= 1; t >>= 1)
    if (t <= lanes && lanes % t == 0) { g.G = t; break; }
  g.tiles = N / (g.G * e16);
  return g;
}

template <int T, typename TB, int G>
static void launch_fbf(int M, const int *brp, const int *bci, const void *val, const TB *B,
                       int ldb, float *C, int ldc, int tiles, int *state, int long_grid,
                       hipStream_t s) {
  constexpr int S = kWave / G;
  dim3 grid(ceil_div(M, kB2Waves * S), tiles);
  hipLaunchKernelGGL((b2sr_spmm_kernel<T, TB, G, 1>), grid, dim3(kB2Threads), 0, s, M, brp, bci,
                     static_cast<const typename Word<T>::type *>(val), B, ldb, C, ldc, state);
  hipLaunchKernelGGL((b2sr_long_kernel<T, TB, G, 1>), dim3(long_grid, tiles), dim3(kB2Threads), 0, s,
                     M, brp, bci, static_cast<const typename Word<T>::type *>(val), B, ldb, C, ldc,
                     state);
}

template <int T, int L, typename TB>
static void launch_narrow_l(int M, int N, const int *brp, const int *bci, const void *val,
                            const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  const int MB = (M + T - 1) / T;
  const dim3 grid(ceil_div(MB, kB2Threads / L));
  const auto *w = static_cast<const typename Word<T>::type *>(val);
#define SPMM_NW(nb)                                                                           \
  hipLaunchKernelGGL((b2sr_narrow_kernel<T, L, nb, TB>), grid, dim3(kB2Threads), 0, s, M, N, \
                     brp, bci, w, B, ldb, C, ldc)
  if (N <= 1) SPMM_NW(1);
  else if (N <= 4) SPMM_NW(4);
  else SPMM_NW(8);
#undef SPMM_NW
}

// A wave per block row for T <= 8, a workgroup for the larger tiles (their
// block rows hold several times more tiles on power-law graphs).
template <int T, typename TB>
static void launch_narrow(int M, int N, const int *brp, const int *bci, const void *val,
                          const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  if (T <= 8) launch_narrow_l<T, kWave, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s);
  else launch_narrow_l<T, kB2Threads, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s);
}

template <int T, typename TB>
static void dispatch_fbf(const Geo &g, int M, int N, const int *brp, const int *bci,
                         const void *val, const TB *B, int ldb, float *C, int ldc,
                         int *state, int long_grid, hipStream_t s) {
  switch (g.G) {
    case 0: launch_narrow<T, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s); break;
    case 1: launch_fbf<T, TB, 1>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    case 2: launch_fbf<T, TB, 2>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    case 4: launch_fbf<T, TB, 4>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    case 8: launch_fbf<T, TB, 8>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    case 16: launch_fbf<T, TB, 16>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    case 32: launch_fbf<T, TB, 32>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
    default: launch_fbf<T, TB, 64>(M, brp, bci, val, B, ldb, C, ldc, g.tiles, state, long_grid, s); break;
  }
}

// Workgroups of `kernel` the chip holds at once (per-CU occupancy, cached
// per kernel, times the CU count).
// h != nullptr: the block rows sorted by tile-count bin (dropin_bin_sort),
// else listed by class_list_kernel into `state` (the allocation-free form).
template <int T, typename TB, int G>
static int launch_dropin(spmm_handle_s *h, int M, int K, const int *brp, const int *bci,
                         const void *val, const TB *B, int ldb, float *C, int ldc, int tiles,
                         int *state, int cus, hipStream_t s) {
  constexpr int SW = kB2Waves * (kWave / G);
  const int MB = (M + T - 1) / T;
  const int *rows = nullptr, *meta = nullptr;
  if (h) {
    // the classes of class_list_kernel's limits, which are powers of two: a
    // bin b >= 1 holds [2^(b-1), 2^b) tiles
    constexpr int b_sub = 31 - __builtin_clz(BrowChunk<G>::kTiles);
    constexpr int b_wg = 32 - __builtin_clz(BrowChunk<G>::kWaveTiles);
    static_assert((BrowChunk<G>::kTiles & (BrowChunk<G>::kTiles - 1)) == 0 &&
                      (BrowChunk<G>::kWaveTiles & (BrowChunk<G>::kWaveTiles - 1)) == 0,
                  "class limits on bin edges");
    const int st = dropin_bin_sort(h, brp, MB, b_wg, b_sub, s, &rows, &meta);
    if (st != SPMM_OK) return st;
  } else {
    hipLaunchKernelGGL(class_list_kernel<kB2Threads>, dim3(ceil_div(MB, kB2Threads)),
                       dim3(kB2Threads), 0, s, MB, brp, BrowChunk<G>::kTiles,
                       BrowChunk<G>::kWaveTiles, state);
  }
  const auto *tw = static_cast<const typename TileWord<T>::type *>(val);
  const int b_bytes = buffer_span(K, ldb, sizeof(TB));
  auto *const kern =
      h ? (b_bytes ? &b2sr_dropin_kernel<T, TB, G, true, true>
                   : &b2sr_dropin_kernel<T, TB, G, false, true>)
        : (b_bytes ? &b2sr_dropin_kernel<T, TB, G, true, false>
                   : &b2sr_dropin_kernel<T, TB, G, false, false>);
  static std::atomic<int> per_cu[4];
  // listed work on as many workgroups as the chip holds at once, then one
  // workgroup per SW block rows. Sorted, the full count is fastest at every
  // T, N and dtype measured on the 89k-row power-law graph (FP32 N = 64, us
  // per call, full / half / quarter: T = 2 38.6 / 42.5 / 57.6, T = 4 36.4 /
  // 43.6 / 60.9, T = 8 45.9 / 59.5 / 87.0). With class lists, half as many
  // for FP32 B at T <= 4 and BF16 B at T = 2 (full / half: FP32 T = 2 63.6 /
  // 57.0, T = 4 50.0 / 47.0, T = 8 47.3 / 61.9; BF16 T = 2 45.4 / 44.2,
  // T = 4 42.6 / 47.8, T = 8 58.2 / 79.9).
  const int kDiv = h ? 1 : ((T == 2 || (T == 4 && sizeof(TB) == 4)) ? 2 : 1);
  const int E = std::min(std::max(1, resident_wgs(reinterpret_cast<const void *>(kern), cus,
                                                  &per_cu[(b_bytes ? 1 : 0) + (h ? 2 : 0)]) /
                                         kDiv),
                         M);
  const dim3 grid(E + ceil_div(MB, SW), tiles);
  hipLaunchKernelGGL(kern, grid, dim3(kB2Threads), 0, s, M, E, brp, bci, tw, B, ldb, b_bytes, C,
                     ldc, state, rows, meta);
  return SPMM_OK;
}

template <int T, typename TB>
static int dispatch_dropin(spmm_handle_s *h, const Geo &g, int M, int K, const int *brp,
                           const int *bci, const void *val, const TB *B, int ldb, float *C,
                           int ldc, int *state, int cus, hipStream_t s) {
#define SPMM_DI(gg) \
  return launch_dropin<T, TB, gg>(h, M, K, brp, bci, val, B, ldb, C, ldc, g.tiles, state, cus, s)
  switch (g.G) {
    case 1: SPMM_DI(1);
    case 2: SPMM_DI(2);
    case 4: SPMM_DI(4);
    case 8: SPMM_DI(8);
    case 16: SPMM_DI(16);
    case 32: SPMM_DI(32);
    default: SPMM_DI(64);
  }
#undef SPMM_DI
}


template <typename TB>
static int run_fbf(spmm_handle_s *h, int T, int M, int K, int N, const int *brp, const int *bci,
                   const void *val, const TB *B, int ldb, float *C, int ldc, void *stream,
                   int *ext_state = nullptr) {
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0, "negative size");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(brp && C, "null bsrRowPtr or C");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldb >= N), "ldb/ldc smaller than N");
  SPMM_CHECK_ARG(K == 0 || B, "null B");
  SPMM_CHECK_ARG(h != nullptr || ext_state != nullptr, "null handle");
  hipStream_t s = as_stream(stream);
  Geo g = geometry(N, sizeof(TB), B, ldb, C, ldc);
  // the planned form's kernels when bsrVal has their alignment (T bytes for
  // T <= 8, 16 bytes above) and, for T > 8, B rows fit the list entries'
  // 25 bits (above six bits of row within the block row)
  const uintptr_t va = reinterpret_cast<uintptr_t>(val);
  const bool classed = g.G && (T > 8 ? va % 16 == 0 && K < (1 << 25) : va % T == 0);
  int *state = nullptr;
  int long_grid = 0, cus = 0;
  if (g.G) {
    if (ext_state) {
      state = ext_state;
    } else {
      const int st = list_state(h, (M + T - 1) / T, s, &state);
      if (st) return st;
    }
    int dev = 0;
    SPMM_HIP(hipGetDevice(&dev));
    SPMM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    long_grid = 8 * cus;
  }
  if (classed && T > 8) {
    // with a handle: one workgroup per block row, heaviest first after a
    // device-side sort by tile count (no classes: bins above 31 are none)
    const int *sorted = nullptr, *meta = nullptr;
    if (!ext_state) {
      const int st = dropin_bin_sort(h, brp, (M + T - 1) / T, kBins, kBins - 1, s, &sorted, &meta);
      if (st != SPMM_OK) return st;
    }
    b2sr_wide_launch<TB>(T, g.G, g.tiles, nullptr, brp, state, sorted, cus, M, K, bci, val, B, ldb,
                         C, ldc, s);
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  if (classed) {
    // with a handle the block rows are sorted on the device; the
    // allocation-free form (ext_state) lists its classes instead
    spmm_handle_s *hs = ext_state ? nullptr : h;
    int st = SPMM_OK;
    switch (T) {
      case 2: st = dispatch_dropin<2, TB>(hs, g, M, K, brp, bci, val, B, ldb, C, ldc, state, cus, s); break;
      case 4: st = dispatch_dropin<4, TB>(hs, g, M, K, brp, bci, val, B, ldb, C, ldc, state, cus, s); break;
      default: st = dispatch_dropin<8, TB>(hs, g, M, K, brp, bci, val, B, ldb, C, ldc, state, cus, s); break;
    }
    if (st != SPMM_OK) return st;
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
#define SPMM_FBF(t) dispatch_fbf<t, TB>(g, M, N, brp, bci, val, B, ldb, C, ldc, state, long_grid, s)
  switch (T) {
    case 2: SPMM_FBF(2); break;
    case 4: SPMM_FBF(4); break;
    case 8: SPMM_FBF(8); break;
    case 16: SPMM_FBF(16); break;
    case 32: SPMM_FBF(32); break;
    default: SPMM_FBF(64); break;
  }
#undef SPMM_FBF
  SPMM_LAUNCHED();
  return SPMM_OK;
}

// off != nullptr: S column-slice launches over the plan's slice offsets
// (off[s MB + i], plan order), slice s > 0 adding into C.
template <int T, typename TB, int G>
static void launch_brow(const spmm_csr_plan_s *p, int M, int K, const int *brp, const int *bci,
                        const void *val, const TB *B, int ldb, float *C, int ldc, int tiles,
                        int rows_scale, const int *off, int S, hipStream_t s) {
  constexpr int SG = kWave / G;
  constexpr int TC = BrowChunk<G>::kTiles;
  LaunchTable t;
  // bins wholly below one subgroup chunk of tiles: a subgroup per block row
  // (several only below 8 tiles); wholly below one wave chunk: a wave (the
  // wave path decodes exactly one chunk); the rest: a workgroup per row
  build_table_ex(p->counts, p->offs, SG, TC, BrowChunk<G>::kWaveTiles, 8, 16, rows_scale, &t, 1,
                 T);
  if (t.total_wg == 0) return;
  const auto *tw = static_cast<const typename TileWord<T>::type *>(val);
  // B rows are read through a buffer resource when B spans < 2 GiB
  const int b_bytes = buffer_span(K, ldb, sizeof(TB));
  const size_t MB = static_cast<size_t>(ceil_div(M, T));  // plan units (p->M counts rows)
  for (int sl = 0; sl < (off ? S : 1); ++sl) {
    const int *lo = off ? off + sl * MB : nullptr;
    const int *hi = off ? off + (sl + 1) * MB : nullptr;
    auto *const kern =
        b_bytes ? (sl ? &b2sr_brow_kernel<T, TB, G, true, true> : &b2sr_brow_kernel<T, TB, G, true, false>)
                : (sl ? &b2sr_brow_kernel<T, TB, G, false, true>
                      : &b2sr_brow_kernel<T, TB, G, false, false>);
    hipLaunchKernelGGL(kern, dim3(t.total_wg, tiles), dim3(kB2Threads), 0, s, t, p->d_rows, M,
                       p->d_span, lo, hi, bci, tw, B, ldb, b_bytes, C, ldc);
  }
}

template <int T, typename TB>
static void dispatch_brow(const Geo &g, const spmm_csr_plan_s *p, int M, int K, const int *brp,
                          const int *bci, const void *val, const TB *B, int ldb, float *C,
                          int ldc, const int *off, hipStream_t s) {
#define SPMM_BR(gg)   launch_brow<T, TB, gg>(p, M, K, brp, bci, val, B, ldb, C, ldc, g.tiles, g.rows_scale, off, g.slices, s)
  switch (g.G) {
    case 1: SPMM_BR(1); break;
    case 2: SPMM_BR(2); break;
    case 4: SPMM_BR(4); break;
    case 8: SPMM_BR(8); break;
    case 16: SPMM_BR(16); break;
    case 32: SPMM_BR(32); break;
    default: SPMM_BR(64); break;
  }
#undef SPMM_BR
}

// The candidate a planned call runs: `tune` when given, else the plan's
// entry for (N, dtype), else variant 0 at the default lane count (the widest
// power of two dividing N's 16-byte vectors). Variant 0 without the
// alignment of the vector kernels, or variant 2 outside N <= 8 and T <= 8,
// runs variant 1.
static Geo b2sr_candidate(const spmm_csr_plan_s *p, int T, int N, int elem_bytes, const void *B,
                          int ldb, const void *C, int ldc, const spmm_tune_t *tune,
                          int *variant) {
  spmm_tune_t t = {0, 0, 0, 0, 0};
  if (tune) {
    t = *tune;
  } else if (p) {
    for (int i = 0; i < p->ntuned; ++i)
      if (p->tuned[i].N == N && p->tuned[i].elem_bytes == elem_bytes) t = p->tuned[i].tune;
  }
  Geo g = geometry(N, elem_bytes, B, ldb, C, ldc);
  *variant = t.variant;
  if (t.variant == 2 && !(N <= 8 && T <= 8)) *variant = 1;
  if (*variant == 0 && g.G == 0) *variant = 1;
  if (*variant != 0) return Geo{};
  const int lanes = N / (16 / elem_bytes);
  const int G = t.lanes_per_row;
  if (G >= 1 && G <= 64 && (G & (G - 1)) == 0 && lanes % G == 0) {
    g.G = G;
    g.tiles = lanes / G;
  }
  g.rows_scale = t.rows_scale == 2 || t.rows_scale == 4 ? t.rows_scale : 1;
  g.slices = T <= 8 && t.slices >= 2 && t.slices <= 64 ? t.slices : 1;
  return g;
}

template <typename TB>
static int run_planned_fbf(spmm_csr_plan_s *p, int T, int M, int K, int N, const int *brp,
                           const int *bci, const void *val, const TB *B, int ldb, float *C,
                           int ldc, void *stream, const spmm_tune_t *tune = nullptr) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0, "negative size");
  if (p->tile != T || p->M != M)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "plan was not built by spmm_b2sr_analyze for this T and M");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(brp && C, "null bsrRowPtr or C");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldb >= N), "ldb/ldc smaller than N");
  SPMM_CHECK_ARG(K == 0 || B, "null B");
  // T = 2 / 4 / 8 tiles are read as one 16 / 32 / 64-bit word; list entries keep
  // the B row in 27 bits
  SPMM_CHECK_ARG(T > 8 || reinterpret_cast<uintptr_t>(val) % T == 0,
                 "bsrVal must be aligned to T bytes for T = 2, 4 and 8");
  SPMM_CHECK_ARG(T <= 8 || reinterpret_cast<uintptr_t>(val) % 16 == 0,
                 "bsrVal must be 16-byte aligned for T = 16, 32 and 64");
  SPMM_CHECK_ARG(K < (1 << 27), "planned B2SR needs K < 2^27");
  hipStream_t s = as_stream(stream);
  int variant = 0;
  Geo g = b2sr_candidate(p, T, N, sizeof(TB), B, ldb, C, ldc, tune, &variant);
  if (T > 8 && K >= (1 << 25) && variant == 0) {  // wide list entries keep 25 bits of B row
    variant = 1;
    g = Geo{};
  }
  if (variant == 2) {  // B staged in LDS (N <= 8, T <= 8)
    switch (T) {
      case 2: launch_lds<2, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s); break;
      case 4: launch_lds<4, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s); break;
      default: launch_lds<8, TB>(M, N, brp, bci, val, B, ldb, C, ldc, s); break;
    }
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  if (g.G == 0) {  // word lanes (variant 1): no schedule needed
#define SPMM_FBF(t) dispatch_fbf<t, TB>(g, M, N, brp, bci, val, B, ldb, C, ldc, nullptr, 0, s)
    switch (T) {
      case 2: SPMM_FBF(2); break;
      case 4: SPMM_FBF(4); break;
      case 8: SPMM_FBF(8); break;
      case 16: SPMM_FBF(16); break;
      case 32: SPMM_FBF(32); break;
      default: SPMM_FBF(64); break;
    }
#undef SPMM_FBF
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  const int *off = nullptr;  // column slices: the plan's per-block-row tile offsets
  if (T <= 8 && g.slices > 1) {
    // a B2SR plan's units are its ceil(M / T) block rows (p->M counts rows)
    const int st = plan_slice_offsets(p, ceil_div(M, T), g.slices, ceil_div(K, T), bci, s, &off);
    if (st != SPMM_OK) return st;
  }
  switch (T) {
    case 2: dispatch_brow<2, TB>(g, p, M, K, brp, bci, val, B, ldb, C, ldc, off, s); break;
    case 4: dispatch_brow<4, TB>(g, p, M, K, brp, bci, val, B, ldb, C, ldc, off, s); break;
    case 8: dispatch_brow<8, TB>(g, p, M, K, brp, bci, val, B, ldb, C, ldc, off, s); break;
    default:
      SPMM_CHECK_ARG(T == 16 || p->d_wide != nullptr, "wide-tile plan without its round lists");
      b2sr_wide_launch<TB>(T, g.G, g.tiles, p, brp, nullptr, nullptr, 0, M, K, bci, val, B, ldb,
                           C, ldc, s);
      break;
  }
  SPMM_LAUNCHED();
  return SPMM_OK;
}

template <int T, int G>
static void launch_sign(int M, int N, const int *brp, const int *bci, const void *val,
                        const uint32_t *Bw, int ldbw, int32_t *C, int ldc, hipStream_t s) {
  constexpr int S = kWave / G;
  dim3 grid(ceil_div(M, kB2Waves * S), ceil_div(N, 4 * G));
  hipLaunchKernelGGL((b2sr_sign_kernel<T, G>), grid, dim3(kB2Threads), 0, s, M, N, brp, bci,
                     static_cast<const typename Word<T>::type *>(val), Bw, ldbw, C, ldc);
}

template <int T>
static void dispatch_sign(int M, int N, const int *brp, const int *bci, const void *val,
                          const uint32_t *Bw, int ldbw, int32_t *C, int ldc, hipStream_t s) {
  const int lanes = ceil_div(N, 4);
  if (lanes <= 4) launch_sign<T, 4>(M, N, brp, bci, val, Bw, ldbw, C, ldc, s);
  else if (lanes <= 8) launch_sign<T, 8>(M, N, brp, bci, val, Bw, ldbw, C, ldc, s);
  else if (lanes <= 16) launch_sign<T, 16>(M, N, brp, bci, val, Bw, ldbw, C, ldc, s);
  else if (lanes <= 32) launch_sign<T, 32>(M, N, brp, bci, val, Bw, ldbw, C, ldc, s);
  else launch_sign<T, 64>(M, N, brp, bci, val, Bw, ldbw, C, ldc, s);
}

}  // namespace spmm

using namespace spmm;

// Allocation-free form of the plan-free FBF SpMM: the block-row lists live
// in `scratch`, one copy per GPU, held until the process ends. In host code
// `scratch` names a host-side stand-in, so the device address of each GPU's
// copy is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels.
constexpr int kNoAllocMaxBlockRows = 1 << 18;  // 2 MiB of lists + head
constexpr int kMaxDevices = 64;
static __device__ __align__(256) int scratch[kStateHead + 2 * kNoAllocMaxBlockRows];

static int scratch_ptr(int **out) {
  static std::mutex mu;
  static int *ptrs[kMaxDevices] = {};
  int dev = 0;
  SPMM_HIP(hipGetDevice(&dev));
  SPMM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
      SPMM_FAIL(SPMM_ERR_HIP, "hipGetSymbolAddress of the B2SR scratch failed");
    ptrs[dev] = static_cast<int *>(p);
  }
  *out = ptrs[dev];
  return SPMM_OK;
}

// Counters the list kernels expect at zero (a kernel, not a memset node, so
// the call stays correct under graph capture).
__global__ void b2sr_zero_head_kernel(int *state) {
  if (threadIdx.x < kStateHead) state[threadIdx.x] = 0;
}

extern "C" {

int csr2b2sr_nnzb(spmm_handle_t h, int M, int K, int nnz, const int *rowptr, const int *colind,
                  int T, int *bsrRowPtr, int *nnzb, void *stream) {
  SPMM_CHECK_ARG(h && bsrRowPtr && nnzb, "null handle, bsrRowPtr or nnzb");
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && nnz >= 0, "negative size");
  SPMM_CHECK_ARG(M == 0 || rowptr, "null rowptr");
  SPMM_CHECK_ARG(nnz == 0 || colind, "null colind");
  hipStream_t s = as_stream(stream);
  const int MB = (M + T - 1) / T;
  const size_t part_bytes = (scan_workspace_bytes(MB + 1) + 255) / 256 * 256;
  char *ws = static_cast<char *>(workspace(h, 256 + part_bytes + sizeof(int) * (MB + 1)));
  int *host = static_cast<int *>(pinned(h, 2 * sizeof(int)));
  if (!ws || !host) return SPMM_ERR_ALLOC;
  int *err = reinterpret_cast<int *>(ws);
  int *part = reinterpret_cast<int *>(ws + 256);
  int *hubs = reinterpret_cast<int *>(ws + 256 + part_bytes);
  SPMM_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  SPMM_HIP(hipMemsetAsync(hubs, 0, sizeof(int), s));
  SPMM_HIP(hipMemsetAsync(bsrRowPtr + MB, 0, sizeof(int), s));
  if (MB > 0) {
    convert<false>(T, M, K, rowptr, colind, bsrRowPtr, nullptr, nullptr, err, hubs, s);
    SPMM_LAUNCHED();
  }
  SPMM_HIP(exclusive_scan(bsrRowPtr, MB + 1, part, s));
  SPMM_HIP(hipMemcpyAsync(host, bsrRowPtr + MB, sizeof(int), hipMemcpyDeviceToHost, s));
  SPMM_HIP(hipMemcpyAsync(host + 1, err, sizeof(int), hipMemcpyDeviceToHost, s));
  SPMM_HIP(hipStreamSynchronize(s));
  if (host[1])
    SPMM_FAIL(SPMM_ERR_INVALID_ARG,
              "CSR column indices must lie in [0, K) and ascend within each row");
  *nnzb = host[0];
  return SPMM_OK;
}

int csr2b2sr(spmm_handle_t h, int M, int K, int nnz, const int *rowptr, const int *colind, int T,
             const int *bsrRowPtr, int *bsrColInd, void *bsrVal, void *stream) {
  SPMM_CHECK_ARG(h && bsrRowPtr, "null handle or bsrRowPtr");
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && nnz >= 0, "negative size");
  SPMM_CHECK_ARG(M == 0 || rowptr, "null rowptr");
  SPMM_CHECK_ARG(nnz == 0 || (colind && bsrColInd && bsrVal), "null colind or outputs");
  if (M == 0 || nnz == 0) return SPMM_OK;
  hipStream_t s = as_stream(stream);
  const int MB = (M + T - 1) / T;
  char *ws = static_cast<char *>(workspace(h, 256 + sizeof(int) * (MB + 1)));
  if (!ws) return SPMM_ERR_ALLOC;
  int *err = reinterpret_cast<int *>(ws);
  int *hubs = reinterpret_cast<int *>(ws + 256);
  SPMM_HIP(hipMemsetAsync(hubs, 0, sizeof(int), s));
  convert<true>(T, M, K, rowptr, colind, const_cast<int *>(bsrRowPtr), bsrColInd, bsrVal, err,
                hubs, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

int spmm_b2sr_f32(spmm_handle_t h, int T, int M, int K, int N, const int *bsrRowPtr,
                  const int *bsrColInd, const void *bsrVal, const float *B, int ldb, float *C,
                  int ldc, void *stream) {
  return run_fbf<float>(h, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, B, ldb, C, ldc, stream);
}

int spmm_b2sr_bf16(spmm_handle_t h, int T, int M, int K, int N, const int *bsrRowPtr,
                   const int *bsrColInd, const void *bsrVal, const uint16_t *B, int ldb,
                   float *C, int ldc, void *stream) {
  return run_fbf<uint16_t>(h, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, B, ldb, C, ldc, stream);
}

int spmm_b2sr_f32_noalloc(int T, int M, int K, int N, const int *bsrRowPtr,
                          const int *bsrColInd, const void *bsrVal, const float *B, int ldb,
                          float *C, int ldc, void *stream) {
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0, "negative size");
  if ((M + T - 1) / T > kNoAllocMaxBlockRows)
    SPMM_FAIL(SPMM_ERR_UNSUPPORTED,
              "spmm_b2sr_f32_noalloc: more than 262144 block rows (the static scratch limit)");
  int *state = nullptr;
  const int st = scratch_ptr(&state);
  if (st) return st;
  if (M > 0 && N > 0) {
    hipLaunchKernelGGL(b2sr_zero_head_kernel, dim3(1), dim3(kWave), 0, as_stream(stream), state);
    SPMM_LAUNCHED();
  }
  return run_fbf<float>(nullptr, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, B, ldb, C, ldc,
                        stream, state);
}

int spmm_b2sr_analyze(spmm_handle_t h, int T, int M, const int *bsrRowPtr, spmm_csr_plan_t *plan,
                      void *stream) {
  SPMM_CHECK_ARG(h && plan && bsrRowPtr, "null handle, plan or bsrRowPtr");
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M > 0, "M must be positive");
  // block rows binned by tile count (the units of b2sr_brow_kernel and
  // the wide-tile kernels), heaviest first
  const int MB = (M + T - 1) / T;
  const int st = analyze(h, MB, 0, bsrRowPtr, 0, plan, stream);
  if (st != SPMM_OK) return st;
  spmm_csr_plan_s *p = *plan;
  p->M = M;
  p->tile = T;
  if (T > 16) {
    // the wide-tile kernels' rounds (wide_round_tiles(T) tiles of one block
    // row each) numbered on the host, and their list buffer
    hipStream_t s = as_stream(stream);
    std::vector<int> hb(static_cast<size_t>(MB) + 1);
    auto fail = [&](const char *what) {
      spmm_csr_plan_destroy(p);
      *plan = nullptr;
      set_error(std::string("spmm_b2sr_analyze: ") + what);
      return SPMM_ERR_ALLOC;
    };
    std::vector<int> units(static_cast<size_t>(MB));  // the plan's block-row order
    if (hipMemcpyAsync(hb.data(), bsrRowPtr, sizeof(int) * hb.size(), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(units.data(), p->d_rows, sizeof(int) * units.size(),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail("bsrRowPtr read-back failed");
    const int NT = wide_round_tiles(T);
    std::vector<int> first(static_cast<size_t>(MB));
    std::vector<int2> rounds;
    if (hb[0] < 0) return fail("bsrRowPtr[0] is negative");
    for (int b = 0; b < MB; ++b) {
      if (hb[b + 1] < hb[b]) return fail("bsrRowPtr is not monotone");
      first[b] = static_cast<int>(rounds.size());
      for (int t = hb[b]; t < hb[b + 1]; t += NT)
        rounds.push_back(make_int2(t, std::min(t + NT, hb[b + 1])));
    }
    std::vector<int> rbase(static_cast<size_t>(MB));  // by plan unit
    for (int x = 0; x < MB; ++x) rbase[x] = first[units[x]];
    const size_t nb = static_cast<size_t>(hb[MB]);  // lists are placed by absolute tile index
    const size_t R = rounds.size();
    // lists | row counts | round table (2 ints a round) | first rounds
    const size_t n_list = kWideSlots * std::max<size_t>(nb, 1), n_rc = T * std::max<size_t>(R, 1);
    const size_t n = n_list + n_rc + 2 * std::max<size_t>(R, 1) + rbase.size();
    if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device) !=
            hipSuccess ||
        hipMalloc(&p->d_wide, sizeof(int) * n) != hipSuccess)
      return fail("hipMalloc of the round-list buffer failed");
    p->d_wide_rc = p->d_wide + n_list;
    p->d_wide_round = reinterpret_cast<int2 *>(p->d_wide_rc + n_rc);
    p->d_wide_rbase = p->d_wide_rc + n_rc + 2 * std::max<size_t>(R, 1);
    p->wide_rounds = static_cast<int>(R);
    if ((R > 0 && hipMemcpyAsync(p->d_wide_round, rounds.data(), sizeof(int2) * R,
                                 hipMemcpyHostToDevice, s) != hipSuccess) ||
        hipMemcpyAsync(p->d_wide_rbase, rbase.data(), sizeof(int) * rbase.size(),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail("round table upload failed");
  }
  return SPMM_OK;
}

int spmm_b2sr_f32_planned(spmm_csr_plan_t plan, int T, int M, int K, int N, const int *bsrRowPtr,
                          const int *bsrColInd, const void *bsrVal, const float *B, int ldb,
                          float *C, int ldc, void *stream) {
  return run_planned_fbf<float>(plan, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, B, ldb, C, ldc,
                                stream);
}

int spmm_b2sr_bf16_planned(spmm_csr_plan_t plan, int T, int M, int K, int N, const int *bsrRowPtr,
                           const int *bsrColInd, const void *bsrVal, const uint16_t *B, int ldb,
                           float *C, int ldc, void *stream) {
  return run_planned_fbf<uint16_t>(plan, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, B, ldb, C, ldc,
                                   stream);
}

}  // extern "C"

namespace {

// Records `t` in the plan for (N, dtype): the entry's slot, else a free one,
// else the last.
void record_tune(spmm_csr_plan_s *p, int N, int elem_bytes, const spmm_tune_t &t) {
  int slot = -1;
  for (int i = 0; i < p->ntuned; ++i)
    if (p->tuned[i].N == N && p->tuned[i].elem_bytes == elem_bytes) slot = i;
  if (slot < 0) slot = p->ntuned < kTuneSlots ? p->ntuned++ : kTuneSlots - 1;
  p->tuned[slot] = {N, elem_bytes, t};
}

template <typename TB>
int b2sr_autotune(spmm_csr_plan_s *p, int T, int M, int K, int N, const int *brp, const int *bci,
                  const void *val, const TB *B, int ldb, float *C, int ldc,
                  spmm_tune_t *best_out, void *stream) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  if (p->tile != T || p->M != M)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "plan was not built by spmm_b2sr_analyze for this T and M");
  SPMM_CHECK_ARG(N > 0, "N must be positive");
  hipStream_t s = as_stream(stream);
  const int e16 = 16 / static_cast<int>(sizeof(TB));
  spmm_tune_t cands[32];
  int nc = 0;
  if (geometry(N, sizeof(TB), B, ldb, C, ldc).G) {
    const int lanes = N / e16;
    for (int G = 64; G >= 1; G >>= 1) {
      if (lanes % G || G * 8 < std::min(lanes, 64)) continue;  // at most 8 column tiles
      for (int rs = 1; rs <= (T <= 8 ? 4 : 1); rs <<= 1) cands[nc++] = {G, 1, rs, 0, 0};
    }
  }
  cands[nc++] = {0, 0, 1, 0, 1};
  if (N <= 8 && T <= 8) cands[nc++] = {0, 0, 1, 0, 2};
  hipEvent_t e0, e1;
  SPMM_HIP(hipEventCreate(&e0));
  SPMM_HIP(hipEventCreate(&e1));
  spmm_tune_t best = cands[0];
  float best_ms = 1e30f;
  int st = SPMM_OK;
  auto trial = [&](const spmm_tune_t &c) {  // one warm call, then three timed
    st = run_planned_fbf<TB>(p, T, M, K, N, brp, bci, val, B, ldb, C, ldc, stream, &c);
    if (st) return;
    if (hipEventRecord(e0, s) != hipSuccess) { st = SPMM_ERR_HIP; return; }
    for (int r = 0; r < 3 && st == SPMM_OK; ++r)
      st = run_planned_fbf<TB>(p, T, M, K, N, brp, bci, val, B, ldb, C, ldc, stream, &c);
    if (st) return;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      st = SPMM_ERR_HIP;
      return;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  };
  for (int c = 0; c < nc && st == SPMM_OK; ++c) trial(cands[c]);
  // column slices for the block-row kernels when B outgrows two XCDs' L2
  // (the offsets are built once per slice count and kept in the plan)
  if (st == SPMM_OK && best.variant == 0 && T <= 8 && p->sorted != 0 &&
      static_cast<long long>(K) * N * static_cast<long long>(sizeof(TB)) >= (8ll << 20)) {
    const spmm_tune_t base = best;
    for (const int S : {2, 4, 6, 8, 12, 16}) {
      if (st != SPMM_OK) break;
      spmm_tune_t c = base;
      c.slices = S;
      trial(c);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  record_tune(p, N, static_cast<int>(sizeof(TB)), best);
  if (best_out) *best_out = best;
  return SPMM_OK;
}

}  // namespace

extern "C" {

int spmm_b2sr_autotune(spmm_csr_plan_t plan, int T, int M, int K, int N, const int *bsrRowPtr,
                       const int *bsrColInd, const void *bsrVal, const void *B, int bf16, int ldb,
                       float *C, int ldc, spmm_tune_t *best, void *stream) {
  if (bf16)
    return b2sr_autotune<uint16_t>(plan, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal,
                                   static_cast<const uint16_t *>(B), ldb, C, ldc, best, stream);
  return b2sr_autotune<float>(plan, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal,
                              static_cast<const float *>(B), ldb, C, ldc, best, stream);
}

int spmm_b2sr_plan_set(spmm_csr_plan_t plan, int N, int bf16, const spmm_tune_t *tune) {
  SPMM_CHECK_ARG(plan && tune, "null plan or tune");
  SPMM_CHECK_ARG(plan->tile != 0, "not a B2SR plan (spmm_b2sr_analyze)");
  SPMM_CHECK_ARG(N > 0, "N must be positive");
  const int e16 = bf16 ? 8 : 4;
  const int v = tune->variant, G = tune->lanes_per_row, rs = tune->rows_scale;
  SPMM_CHECK_ARG(v >= 0 && v <= 2, "variant must be 0, 1 or 2");
  SPMM_CHECK_ARG(v != 2 || (N <= 8 && plan->tile <= 8), "variant 2 needs N <= 8 and T <= 8");
  SPMM_CHECK_ARG(v != 0 || N % e16 == 0, "variant 0 needs N a multiple of one 16-byte vector");
  SPMM_CHECK_ARG(v != 0 || G == 0 ||
                     (G <= 64 && (G & (G - 1)) == 0 && (N / e16) % G == 0),
                 "lanes_per_row must be a power of two dividing N's 16-byte vectors");
  SPMM_CHECK_ARG(rs == 0 || rs == 1 || rs == 2 || rs == 4, "rows_scale must be 1, 2 or 4");
  SPMM_CHECK_ARG(tune->slices <= 1 || (v == 0 && plan->tile <= 8 && tune->slices <= 64),
                 "slices (2 .. 64) need variant 0 and T <= 8");
  spmm_tune_t t = *tune;
  t.vec = 1;
  t.slices = tune->slices > 1 ? tune->slices : 0;
  record_tune(plan, N, bf16 ? 2 : 4, t);
  return SPMM_OK;
}

int spmm_b2sr_signpacked_i32(spmm_handle_t h, int T, int M, int K, int N, const int *bsrRowPtr,
                             const int *bsrColInd, const void *bsrVal, const uint32_t *Bw,
                             int ldbw, int32_t *C, int ldc, void *stream) {
  SPMM_CHECK_ARG(valid_T(T), "T must be 2, 4, 8, 16, 32 or 64");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0, "negative size");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(bsrRowPtr && C, "null bsrRowPtr or C");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldbw >= (N + 31) / 32), "ldbw/ldc too small");
  SPMM_CHECK_ARG(K == 0 || Bw, "null Bw");
  hipStream_t s = as_stream(stream);
  // The product is the FBF one with B = +-1, exact in BF16, and its sums are
  // whole numbers of magnitude at most K, exact in FP32 below 2^24: B is
  // expanded once into the handle's workspace, the BF16 FBF drop-in writes
  // FP32 into C's memory, and C is converted in place.
  const long long kn = static_cast<long long>(K) * N, mn = static_cast<long long>(M) * N;
  if (h && K > 0 && K < (1 << 24) && kn < (1ll << 32) && mn < (1ll << 32)) {
    auto *Bx = static_cast<uint16_t *>(workspace(h, sizeof(uint16_t) * static_cast<size_t>(kn), s));
    if (!Bx) return SPMM_ERR_ALLOC;
    const bool ev = N % 8 == 0;  // 16-byte rows of Bx (the workspace is 256-byte aligned)
    const uint32_t ei = static_cast<uint32_t>(ev ? kn / 8 : kn);
    hipLaunchKernelGGL(ev ? sign_expand_kernel<true> : sign_expand_kernel<false>,
                       dim3(static_cast<unsigned>(ceil_div_ll(ei, kB2Threads))), dim3(kB2Threads), 0,
                       s, ei, N, Bw, ldbw, Bx);
    SPMM_LAUNCHED();
    const int st = run_fbf<uint16_t>(h, T, M, K, N, bsrRowPtr, bsrColInd, bsrVal, Bx, N,
                                     reinterpret_cast<float *>(C), ldc, stream);
    if (st != SPMM_OK) return st;
    const bool cv = N % 4 == 0 && ldc % 4 == 0 && reinterpret_cast<uintptr_t>(C) % 16 == 0;
    const uint32_t ci = static_cast<uint32_t>(cv ? mn / 4 : mn);
    hipLaunchKernelGGL(cv ? f32_to_i32_kernel<true> : f32_to_i32_kernel<false>,
                       dim3(static_cast<unsigned>(ceil_div_ll(ci, kB2Threads))), dim3(kB2Threads), 0,
                       s, ci, N, C, ldc);
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  switch (T) {
    case 2: dispatch_sign<2>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
    case 4: dispatch_sign<4>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
    case 8: dispatch_sign<8>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
    case 16: dispatch_sign<16>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
    case 32: dispatch_sign<32>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
    default: dispatch_sign<64>(M, N, bsrRowPtr, bsrColInd, bsrVal, Bw, ldbw, C, ldc, s); break;
  }
  SPMM_LAUNCHED();
  return SPMM_OK;
}

static int run_pack(bool sign, bool cols, int H, int W, const float *X, int ldx, uint32_t *out,
                    void *stream) {
  SPMM_CHECK_ARG(H >= 0 && W >= 0, "negative size");
  if (H == 0 || W == 0) return SPMM_OK;
  SPMM_CHECK_ARG(X && out && ldx >= W, "null pointer or ldx < W");
  const int Hp = cols ? 32 * ceil_div(H, 32) : H;
  const long long items = static_cast<long long>(Hp) * ceil_div(W, 64);
  const long long blocks = ceil_div_ll(items, kB2Waves);
  SPMM_CHECK_ARG(blocks < (1ll << 31), "matrix too large");
  hipStream_t s = as_stream(stream);
  dim3 grid(static_cast<unsigned>(blocks));
  if (sign && cols)
    hipLaunchKernelGGL((pack_kernel<true, true>), grid, dim3(kB2Threads), 0, s, H, W, X, ldx, out);
  else if (sign)
    hipLaunchKernelGGL((pack_kernel<true, false>), grid, dim3(kB2Threads), 0, s, H, W, X, ldx, out);
  else
    hipLaunchKernelGGL((pack_kernel<false, false>), grid, dim3(kB2Threads), 0, s, H, W, X, ldx, out);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

int pack_sign_rows_u32(int H, int W, const float *X, int ldx, uint32_t *out, void *stream) {
  return run_pack(true, false, H, W, X, ldx, out, stream);
}

int pack_bin_rows_u32(int H, int W, const float *X, int ldx, uint32_t *out, void *stream) {
  return run_pack(false, false, H, W, X, ldx, out, stream);
}

int pack_sign_cols_u32(int H, int W, const float *X, int ldx, uint32_t *out, void *stream) {
  return run_pack(true, true, H, W, X, ldx, out, stream);
}

int unpack_cols_u32_f32(int H, int W, const uint32_t *in, float *X, int ldx, void *stream) {
  SPMM_CHECK_ARG(H >= 0 && W >= 0, "negative size");
  if (H == 0 || W == 0) return SPMM_OK;
  SPMM_CHECK_ARG(in && X && ldx >= W, "null pointer or ldx < W");
  const long long n = static_cast<long long>(H) * W;
  hipLaunchKernelGGL(unpack_cols_kernel, dim3(static_cast<unsigned>(ceil_div_ll(n, kB2Threads))),
                     dim3(kB2Threads), 0, as_stream(stream), H, W, in, X, ldx);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // extern "C"
