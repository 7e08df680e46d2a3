// CSR SpMM, C[M][N] = A[M][K] * B[K][N] (FP32 values, FP32 or BF16 B,
// FP32 accumulate), the drop-in for the cuSPARSE call of
// evalCSRSpmmFloatCuSPARSE (test_spmm_full.cu:212-342).
//
// Schedule (gfx950, wave64). A row's B gathers are 16-byte loads: G lanes
// cover one B row segment of G*EPL columns (EPL = 4 floats or 8 bf16 per
// 16-byte vector, times `VEC`). A wave therefore holds S = 64/G subgroups.
// Rows are binned by degree at analysis time (bin b: degree in
// [2^(b-1), 2^b), longest bins first) and each bin runs in one of three
// forms inside a single launch of 256-thread workgroups:
//   A  short rows  : one subgroup per row, R rows per subgroup;
//   B  medium rows : one wave per row, its S subgroups split the row;
//   C  long rows   : one workgroup per row, 4*S subgroups split the row,
//                    partial rows reduced through LDS in a fixed order.
// Every row's summation order depends only on its degree, so results are
// bitwise reproducible from run to run. No atomics touch C.
#include "csr_core.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <new>
#include <vector>

namespace spmm {

__device__ __forceinline__ int degree_bin(int deg) {
  return deg <= 0 ? 0 : 32 - __clz(deg);
}

__global__ void __launch_bounds__(256)
bin_count_kernel(const int *__restrict__ rowptr, int M, int shift, int *__restrict__ counts) {
  __shared__ int h[kBins];
  if (threadIdx.x < kBins) h[threadIdx.x] = 0;
  __syncthreads();
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < M;
       r += gridDim.x * blockDim.x) {
    const int q = r >> shift;
    int b = degree_bin(rowptr[q + 1] - rowptr[q]);
    atomicAdd(&h[b], 1);
  }
  __syncthreads();
  if (threadIdx.x < kBins && h[threadIdx.x])
    atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// cursor[b] starts at the bin's offset; each workgroup reserves its rows of
// each bin with one atomic and ranks them through LDS.
__global__ void __launch_bounds__(256)
bin_scatter_kernel(const int *__restrict__ rowptr, int M, int shift, int *__restrict__ cursor,
                   int *__restrict__ rows) {
  __shared__ int h[kBins];
  __shared__ int base[kBins];
  for (int r0 = blockIdx.x * blockDim.x; r0 < M; r0 += gridDim.x * blockDim.x) {
    if (threadIdx.x < kBins) h[threadIdx.x] = 0;
    __syncthreads();
    int r = r0 + threadIdx.x;
    int b = -1, rank = 0;
    if (r < M) {
      const int q = r >> shift;
      b = degree_bin(rowptr[q + 1] - rowptr[q]);
      rank = atomicAdd(&h[b], 1);
    }
    __syncthreads();
    if (threadIdx.x < kBins)
      base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]) : 0;
    __syncthreads();
    if (b >= 0) rows[base[b] + rank] = r;
    __syncthreads();
  }
}

// span[i] = the entry range of unit rows[i] (of block row rows[i] >> shift).
__global__ void __launch_bounds__(256)
span_kernel(const int *__restrict__ rowptr, int M, int shift, const int *__restrict__ rows,
            int2 *__restrict__ span) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
    const int q = rows[i] >> shift;
    span[i] = make_int2(rowptr[q], rowptr[q + 1]);
  }
}

template <typename TB, int G, int VEC>
__global__ void __launch_bounds__(kThreads)
csr_spmm_kernel(LaunchTable tab, const int *__restrict__ rows,
                const int2 *__restrict__ span, const int *__restrict__ colind,
                const float *__restrict__ val, const TB *__restrict__ B, int ldb,
                float *__restrict__ C, int ldc) {
  csr_wg<TB, G, VEC>(tab, blockIdx.x, blockIdx.y, rows, span, colind, val, B, ldb, C, ldc);
}

// C = A^T B on the transposed structure of A, values read through perm.
template <typename TB, int G, int VEC>
__global__ void __launch_bounds__(kThreads)
csr_spmm_perm_kernel(LaunchTable tab, const int *__restrict__ rows,
                     const int2 *__restrict__ span, const int *__restrict__ colind,
                     const float *__restrict__ val, const int *__restrict__ perm,
                     const TB *__restrict__ B, int ldb, float *__restrict__ C, int ldc) {
  csr_wg<TB, G, VEC, true>(tab, blockIdx.x, blockIdx.y, rows, span, colind, val, B, ldb, C,
                           ldc, perm);
}

template <typename TB, int G, int VEC>
__global__ void __launch_bounds__(kThreads)
csr_t_sddmm_kernel(LaunchTable tab, const int *__restrict__ rows,
                   const int2 *__restrict__ span, const int *__restrict__ tcol,
                   const float *__restrict__ val, const int *__restrict__ perm,
                   const TB *__restrict__ dY, int lddy, const TB *__restrict__ X, int ldx,
                   float *__restrict__ dX, int lddx, float *__restrict__ dval) {
  csr_t_sddmm_wg<TB, G, VEC, false>(tab, blockIdx.x, rows, span, tcol, val, perm, dY, lddy, X,
                                    ldx, dX, lddx, dval);
}

// Plan-free CSR SpMM (the drop-in): one launch over the classes that
// class_list_kernel found on the device, with the planned kernel's per-row
// code, so a row's sum is bitwise the planned form's at the same lane
// geometry. Workgroups x < E (the workgroups the chip holds at once) take
// the listed rows: workgroup-class rows (list 0) x, x + E, ..., then the
// wave-class rows (list 1) dealt to their waves from the last workgroup
// backwards, so they land on workgroups with one workgroup-class row fewer.
// Workgroup E + k takes rows k SW R .. (k + 1) SW R - 1, one subgroup per
// row and round (R rounds), skipping the listed ones. Shares are static: one
// counter shared by every workgroup would serialise their atomics across the
// XCDs. list_counts_reset_kernel, launched next, zeroes the list counts, so
// consecutive or graph-replayed calls need no memset. (Before, the last
// workgroup to finish did it, at the cost of one atomic per workgroup on one
// address: a 16k-row graph at N = 32 took 23.7 us a call, then 15.5, against
// 7.1 us planned.)
// R = 1: a 16k-row graph at N = 32 takes 12.8 us a call against 15.8 at
// R = 2 and 23.9 at R = 4 (the 89k-row graph at N = 64: 45.6 / 47.0 / 51.5)
constexpr int kDropinRounds = 1;

template <typename TB, int G, int VEC, bool kPerm>
__global__ void __launch_bounds__(kThreads)
csr_dropin_kernel(int M, int E, int a_limit, const int *__restrict__ rowptr,
                  const int *__restrict__ colind, const float *__restrict__ val,
                  const int *__restrict__ perm, const TB *__restrict__ B, int ldb,
                  float *__restrict__ C, int ldc, int *__restrict__ state) {
  constexpr int E16 = Vec16<TB>::kElems;
  constexpr int S = kWave / G, SW = kWavesPerWG * S;
  constexpr int CW = G * VEC * E16;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1), sub = lane / G;
  const int col0 = blockIdx.y * CW + lane_sub * E16;
  const TB *Bc = B + col0;
  const int x = blockIdx.x;
  if (x < E) {
    const int n_wg = state[0], n_wave = state[2];
    for (int i = x; i < n_wg; i += E) {  // uniform over the workgroup
      const int row = state[kStateHead + i];
      csr_row_wg<TB, G, VEC, kPerm>(row, make_int2(rowptr[row], rowptr[row + 1]), wave, sub,
                                    lane_sub, col0, colind, val, Bc, ldb, C, ldc, perm);
    }
    for (int j = (E - 1 - x) * kWavesPerWG + wave; j < n_wave; j += E * kWavesPerWG) {
      const int row = state[kStateHead + M + j];
      csr_row_wave<TB, G, VEC, kPerm>(row, make_int2(rowptr[row], rowptr[row + 1]), sub, lane_sub,
                                      col0, colind, val, Bc, ldb, C, ldc, perm);
    }
  } else {
    const int base = (x - E) * SW * kDropinRounds + wave * S + sub;
    for (int k = 0; k < kDropinRounds; ++k) {
      const int row = base + k * SW;
      if (row < M) {
        const int2 sp = make_int2(rowptr[row], rowptr[row + 1]);
        if (sp.y - sp.x < a_limit)  // else listed
          csr_row_sub<TB, G, VEC, kPerm>(row, sp, lane_sub, col0, colind, val, Bc, ldb, C, ldc,
                                         perm);
      }
    }
  }
}

// Zeroes the plan-free list counts once the launch that read them is done
// (the next launch on the stream), so every call finds them at zero.
__global__ void __launch_bounds__(kWave) list_counts_reset_kernel(int *__restrict__ state) {
  if (threadIdx.x == 0) {
    state[0] = 0;
    state[2] = 0;
  }
}

// ------------------------------------------------ plan-free column slices
// The plan-free form of the column-sliced product (csr_sliced_kernel's
// launches without a plan): the rows sorted by degree bin and the slice
// offsets built on the device per call, then S launches, slice 0 storing C
// and the others adding into it.
constexpr int kSlicedInFlight = 4;  // B-row loads in flight per lane, sliced launches
constexpr int kUnsortedFlag = 4;  // list-state slot: a row's columns descend

// Slice offsets of every row. A wave takes blocks of kOffRows rows (lane l
// reads row r0 + l's span once); row by row, the lanes stream the row's
// colind 8 x 64 entries a round, the next row's first round requested
// before this row's is scanned. Entry e is the first of slice t for every t
// in (slice(e - 1), slice(e)], slice(c) = min(c / W, S - 1); so
// off[t * M + r] is the first entry of row r in slice t, off[0 * M + r] and
// off[S * M + r] its span, and slices after the row's last entry start at
// its end. Every (t, r) is set once when the columns ascend; a descending
// pair sets state[kUnsortedFlag] instead (the slice launches then run whole
// rows in slice 0). The block's offsets gather in LDS and leave as one
// coalesced store of kOffRows ints per slice (scattered 4-byte stores, one
// per (row, slice), measured 4% slower).
constexpr int kOffChunks = 8;
constexpr int kOffRows = 16;
constexpr int kOffMaxS = 32;  // slices of the plan-free form at most

constexpr int kDppWaveShr1 = 0x138;  // DPP wave_shr:1: lane l reads lane l - 1

// min(c / W, S - 1) for c >= 0 (negatives as 0) by a float reciprocal and a
// one-step correction: below S <= 32 slices the float quotient is within
// 32 * 2^-22 of the exact one, so one step up or down makes it exact, and a
// larger quotient only has to stay >= S - 1.
__device__ __forceinline__ int slice_of(int c, int W, float invW, int S) {
  const unsigned cc = static_cast<unsigned>(max(c, 0));
  int q = static_cast<int>(static_cast<float>(cc) * invW);
  q = min(q, S);  // the correction below stays in 32 bits
  if (static_cast<unsigned>(q) * static_cast<unsigned>(W) > cc) --q;
  else if (static_cast<unsigned>(q + 1) * static_cast<unsigned>(W) <= cc) ++q;
  return min(q, S - 1);
}

__device__ __forceinline__ void off_round(const int *__restrict__ colind, int e, int e1,
                                          int lane, int (&c)[kOffChunks]) {
#pragma unroll
  for (int u = 0; u < kOffChunks; ++u) {
    const int ei = e + u * kWave + lane;
    c[u] = ei < e1 ? __builtin_nontemporal_load(colind + ei) : 0;
  }
}

__global__ void __launch_bounds__(kThreads)
dropin_slice_offsets_kernel(int M, int S, int W, const int *__restrict__ rowptr,
                            const int *__restrict__ colind, int *__restrict__ off,
                            int *__restrict__ state) {
  __shared__ int soff[kWavesPerWG][kOffMaxS + 1][kOffRows];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int waves = gridDim.x * kWavesPerWG;
  const size_t Ms = static_cast<size_t>(M);
  const float invW = 1.0f / static_cast<float>(W);
  int(*const so)[kOffRows] = soff[wave];
  bool bad = false;
  for (int r0 = (blockIdx.x * kThreads + threadIdx.x) / kWave * kOffRows; r0 < M;
       r0 += waves * kOffRows) {
    const int nr = min(kOffRows, M - r0);
    const int my0 = lane < nr ? rowptr[r0 + lane] : 0;
    const int my1 = lane < nr ? rowptr[r0 + lane + 1] : 0;
    int c[kOffChunks], cn[kOffChunks];
    off_round(colind, __shfl(my0, 0), __shfl(my1, 0), lane, cn);
    for (int j = 0; j < nr; ++j) {
      const int e0 = __shfl(my0, j), e1 = __shfl(my1, j);
#pragma unroll
      for (int u = 0; u < kOffChunks; ++u) c[u] = cn[u];
      if (j + 1 < nr) off_round(colind, __shfl(my0, j + 1), __shfl(my1, j + 1), lane, cn);
      if (lane == 0) so[0][j] = e0;
      int prev_c = INT_MIN, prev_s = 0;  // the entry before this chunk's first
      for (int e = e0; e < e1; e += kOffChunks * kWave) {
        if (e != e0) off_round(colind, e, e1, lane, c);  // rows above one round
#pragma unroll
        for (int u = 0; u < kOffChunks; ++u) {
          const int eb = e + u * kWave;
          if (eb >= e1) break;  // uniform
          const int ei = eb + lane;
          const int sc = slice_of(c[u], W, invW, S);
          // lane l - 1's column and slice (lane 0: the previous chunk's last)
          const int cb = __builtin_amdgcn_update_dpp(prev_c, c[u], kDppWaveShr1, 0xf, 0xf, false);
          const int sb = __builtin_amdgcn_update_dpp(prev_s, sc, kDppWaveShr1, 0xf, 0xf, false);
          if (ei < e1) {
            bad |= c[u] < cb;
            for (int t = sb + 1; t <= sc; ++t) so[t][j] = ei;
          }
          const int last = min(kWave - 1, e1 - 1 - eb);  // uniform
          prev_c = __builtin_amdgcn_readlane(c[u], last);
          prev_s = __builtin_amdgcn_readlane(sc, last);
        }
      }
      // slices after the last entry's (all of them for an empty row)
      for (int t = prev_s + 1 + lane; t <= S; t += kWave) so[t][j] = e1;
    }
    // the block's offsets, one store of nr consecutive ints per slice (LDS
    // written and read by this wave only: a wave-scope fence orders them)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int x = lane; x < (S + 1) * kOffRows; x += kWave) {
      const int t = x / kOffRows, j = x % kOffRows;
      if (j < nr) off[t * Ms + r0 + j] = so[t][j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (__ballot(bad) && lane == 0) atomicOr(&state[kUnsortedFlag], 1);
}

constexpr int kDropinHead = 2 * kBins + 16;  // counts | cursors | class sizes, 64-byte aligned

// Descending-bin cursors of the plan-free sliced product's row sort: from
// the per-bin counts of bin_count_kernel, cursor[b] = the rows of bins above
// b, meta[0] = workgroup-class rows (bins >= b_wg), meta[1] = wave-class rows
// (bins b_sub < b < b_wg); the counts are zeroed for the next call.
__global__ void __launch_bounds__(kWave)
dropin_bins_kernel(int *__restrict__ counts, int *__restrict__ cursor, int *__restrict__ meta,
                   int b_wg, int b_sub) {
  __shared__ int cnt[kBins];
  if (threadIdx.x < kBins) {
    cnt[threadIdx.x] = counts[threadIdx.x];
    counts[threadIdx.x] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int off = 0, n_wg = 0, n_wave = 0;
    for (int b = kBins - 1; b >= 0; --b) {
      cursor[b] = off;
      off += cnt[b];
      if (b >= b_wg) n_wg += cnt[b];
      else if (b > b_sub) n_wave += cnt[b];
    }
    meta[0] = n_wg;
    meta[1] = n_wave;
  }
}

// One slice launch of the plan-free sliced product, over the rows sorted by
// degree bin, heaviest first (rows[0, n_wg): workgroup class, then n_wave
// rows of the wave class, then the subgroup class; the classes of the
// planned sliced launch at the same geometry and slice count, so every row
// is summed as that launch sums it), each row's range [lo[r], hi[r]) of
// slice s. The first E workgroups (as many as the chip holds) deal the
// workgroup-class rows (x, x + E, ...) and then the wave-class rows to their
// waves, heaviest first; workgroup E + k takes
// `rounds` subgroup-class rows per subgroup in sorted order. Rows with no
// entry in a later slice are skipped. The last workgroup of each launch
// zeroes the done count; that of the last slice also the unsorted flag.
template <typename TB, int G, int VEC, bool kAcc>
__global__ void __launch_bounds__(kThreads)
csr_dropin_sliced_kernel(int M, int E, int rounds, const int *__restrict__ rows,
                         const int *__restrict__ meta, const int *__restrict__ lo,
                         const int *__restrict__ hi, const int *__restrict__ rowptr,
                         const int *__restrict__ colind, const float *__restrict__ val,
                         const TB *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                         int *__restrict__ state, int last) {
  constexpr int S = kWave / G, SW = kWavesPerWG * S;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1), sub = lane / G;
  const int col0 = lane_sub * Vec16<TB>::kElems;
  const TB *Bc = B + col0;
  const int x = blockIdx.x;
  // descending columns somewhere: slice 0 runs whole rows, the others nothing
  const bool whole = state[kUnsortedFlag] != 0;
  auto span = [&](int row) {
    return whole ? (kAcc ? make_int2(0, 0) : make_int2(rowptr[row], rowptr[row + 1]))
                 : make_int2(lo[row], hi[row]);
  };
  if (!(kAcc && whole)) {
    const int n_wg = meta[0], n_wave = meta[1];
    if (x < E) {
      for (int i = x; i < n_wg; i += E) {  // uniform over the workgroup
        const int row = rows[i];
        const int2 sp = span(row);
        if (kAcc && sp.x == sp.y) continue;
        csr_row_wg<TB, G, VEC, false, kAcc, true, kSlicedInFlight>(row, sp, wave, sub, lane_sub,
                                                                   col0, colind, val, Bc, ldb, C,
                                                                   ldc, nullptr);
      }
      for (int j = (E - 1 - x) * kWavesPerWG + wave; j < n_wave; j += E * kWavesPerWG) {
        const int row = rows[n_wg + j];
        const int2 sp = span(row);
        if (kAcc && sp.x == sp.y) continue;
        csr_row_wave<TB, G, VEC, false, kAcc, true, kSlicedInFlight>(
            row, sp, sub, lane_sub, col0, colind, val, Bc, ldb, C, ldc, nullptr);
      }
    } else {
      const int first = n_wg + n_wave;
      const int base = first + (x - E) * SW * rounds + wave * S + sub;
      for (int k = 0; k < rounds; ++k) {
        const int p = base + k * SW;
        if (p < M) {
          const int row = rows[p];
          const int2 sp = span(row);
          if (!kAcc || sp.x < sp.y)
            csr_row_sub<TB, G, VEC, false, kAcc, true, kSlicedInFlight>(
                row, sp, lane_sub, col0, colind, val, Bc, ldb, C, ldc, nullptr);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // every workgroup has read the flag before it counts itself
    if (atomicAdd(&state[1], 1) == static_cast<int>(gridDim.x) - 1) {
      if (last) atomicExch(&state[kUnsortedFlag], 0);
      atomicExch(&state[1], 0);
    }
  }
}

// ------------------------------------------------- column-sliced launches
// Per-row slice offsets of a plan, one wave per row (plan position i): lane
// s in [1, S) finds by binary search the first entry whose column is at
// least s * W (off[s * M + i]); off[0 * M + i] and off[S * M + i] are the
// row's span. The lanes also check that the row's columns ascend (the
// offsets need it) and flag it otherwise.
__global__ void __launch_bounds__(kThreads)
slice_offsets_kernel(int M, int S, int W, const int2 *__restrict__ span,
                     const int *__restrict__ colind, int *__restrict__ off,
                     int *__restrict__ unsorted) {
  const int lane = threadIdx.x & (kWave - 1);
  const int waves = gridDim.x * kWavesPerWG;
  for (int i = (blockIdx.x * kThreads + threadIdx.x) / kWave; i < M; i += waves) {
    const int2 sp = span[i];
    bool bad = false;
    for (int e = sp.x + lane; e + 1 < sp.y; e += kWave) bad |= colind[e + 1] < colind[e];
    if (__ballot(bad) && lane == 0) atomicOr(unsorted, 1);
    if (lane >= 1 && lane < S) {
      const long long bound = static_cast<long long>(lane) * W;
      int lo = sp.x, hi = sp.y;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (colind[mid] < bound) lo = mid + 1;
        else hi = mid;
      }
      off[static_cast<size_t>(lane) * M + i] = lo;
    }
    if (lane == 0) {
      off[i] = sp.x;
      off[static_cast<size_t>(S) * M + i] = sp.y;
    }
  }
}

// One launch of a column-sliced product: for column tile `ctile`, the
// entries of each row whose column lies in one range of B's rows
// ([lo[i], hi[i]) at plan position i), in the row classes of a launch table
// built on the rows' degrees divided by the slice count. kAcc = false (the
// first slice) stores C, rows with no entry in it storing zeros; kAcc = true
// adds the slice's partial sums into C, so a row's result is
// ((p_0 + p_1) + p_2) + ... in slice order, the same at every call. Launches
// run in slice order on one stream, so no two of them touch C at once.
// colind, val and C are read (and C written) once per launch, with the
// non-temporal policy, so the slice of B keeps the caches. Four B rows are
// in flight per lane (kSlicedInFlight) rather than the unsliced kernel's
// eight. kPerm: entry e's value is val[perm[e]] (the transposed plan, whose
// rows ascend in A's row index, as the stable transpose builds them).
template <typename TB, int G, int VEC, bool kAcc, bool kPerm>
__global__ void __launch_bounds__(kThreads)
csr_sliced_kernel(LaunchTable tab, int ctile, const int *__restrict__ rows,
                  const int *__restrict__ lo, const int *__restrict__ hi,
                  const int *__restrict__ colind, const float *__restrict__ val,
                  const int *__restrict__ perm, const TB *__restrict__ B, int ldb,
                  float *__restrict__ C, int ldc) {
  constexpr int E = Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * VEC * E;
  const Seg sg = find_seg(tab, blockIdx.x);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1), sub = lane / G;
  const int col0 = ctile * CW + lane_sub * E;
  const TB *Bc = B + col0;
  const int local_wg = blockIdx.x - sg.wg_begin;
  if (sg.kind == kSegA) {
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kWavesPerWG * S;
    const int end = sg.list_off + sg.count;
    int pos = sg.list_off + local_wg * subs_per_wg * R + wave * S + sub;
    int row = -1;
    int2 sp = make_int2(0, 0);
    if (pos < end) {
      row = rows[pos];
      sp = make_int2(lo[pos], hi[pos]);
    }
    for (int k = 0; k < R && row >= 0; ++k) {
      // the next row's id and range are requested before this row's gathers
      const int pos_n = pos + subs_per_wg;
      int row_n = -1;
      int2 sp_n = make_int2(0, 0);
      if (k + 1 < R && pos_n < end) {
        row_n = rows[pos_n];
        sp_n = make_int2(lo[pos_n], hi[pos_n]);
      }
      if (!kAcc || sp.x < sp.y)
        csr_row_sub<TB, G, VEC, kPerm, kAcc, true, kSlicedInFlight>(row, sp, lane_sub, col0, colind,
                                                                    val, Bc, ldb, C, ldc, perm);
      row = row_n;
      sp = sp_n;
      pos = pos_n;
    }
  } else if (sg.kind == kSegB) {
    const int pos = sg.list_off + local_wg * kWavesPerWG + wave;
    if (pos >= sg.list_off + sg.count) return;  // whole wave
    const int2 sp = make_int2(lo[pos], hi[pos]);
    if (kAcc && sp.x == sp.y) return;
    csr_row_wave<TB, G, VEC, kPerm, kAcc, true, kSlicedInFlight>(rows[pos], sp, sub, lane_sub, col0,
                                                                 colind, val, Bc, ldb, C, ldc, perm);
  } else {
    const int pos = sg.list_off + local_wg;
    const int2 sp = make_int2(lo[pos], hi[pos]);
    if (kAcc && sp.x == sp.y) return;  // whole workgroup
    csr_row_wg<TB, G, VEC, kPerm, kAcc, true, kSlicedInFlight>(rows[pos], sp, wave, sub, lane_sub,
                                                               col0, colind, val, Bc, ldb, C, ldc,
                                                               perm);
  }
}

// Narrow or unaligned N (N not a multiple of one 16-byte vector, or B / C /
// ld not 16-byte aligned): a row's entries are split across lanes rather
// than its columns. A subgroup is NC column lanes (a power of two covering
// min(N, 64); wider N loops over column chunks) times ES entry slots, and
// the plan's row classes set ES: short rows a subgroup of NC x
// min(4, 64 / NC) lanes each, medium rows a wave (ES = 64 / NC), long rows a
// workgroup (ES = 256 / NC) whose waves meet in LDS. A lane takes
// kNarrowU entries a round, all loads from clamped (in range) indices, so a
// round's gathers are in flight together; slots past the row's end are
// dropped before the FMA. Sums: per lane in entry order, then a butterfly
// over the slots, then the waves in order -- the same for every call.
constexpr int kNarrowU = 4;

template <typename TB, int NC, int ES, bool kPerm>
__device__ __forceinline__ float narrow_range(int e0, int e1, int es, int cc,
                                              const int *__restrict__ colind,
                                              const float *__restrict__ val,
                                              const int *__restrict__ perm,
                                              const TB *__restrict__ B, int ldb) {
  float acc = 0.f;
  for (int e = e0 + es; e < e1; e += ES * kNarrowU) {
    int ci[kNarrowU];
    float v[kNarrowU], b[kNarrowU];
#pragma unroll
    for (int u = 0; u < kNarrowU; ++u) {
      const int eu = min(e + u * ES, e1 - 1);
      ci[u] = colind[eu];
      v[u] = val[kPerm ? perm[eu] : eu];
    }
#pragma unroll
    for (int u = 0; u < kNarrowU; ++u)
      b[u] = load_scalar<TB>(B + static_cast<size_t>(ci[u]) * ldb + cc);
#pragma unroll
    for (int u = 0; u < kNarrowU; ++u)
      if (e + u * ES < e1) acc = fmaf(v[u], b[u], acc);
  }
  return acc;
}

// Entry slots of a class-A subgroup: 4, or all the wave has beside NC.
template <int NC> constexpr int narrow_slots_a() { return kWave / NC < 4 ? kWave / NC : 4; }

template <typename TB, int NC, bool kPerm>
__global__ void __launch_bounds__(kThreads)
csr_narrow_kernel(LaunchTable tab, const int *__restrict__ rows, const int2 *__restrict__ span,
                  const int *__restrict__ colind, const float *__restrict__ val,
                  const int *__restrict__ perm, const TB *__restrict__ B, int ldb,
                  float *__restrict__ C, int ldc, int N) {
  constexpr int ESA = narrow_slots_a<NC>();
  constexpr int GA = NC * ESA, SA = kWave / GA;
  __shared__ float red[kWavesPerWG][NC];
  const Seg sg = find_seg(tab, blockIdx.x);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int local_wg = blockIdx.x - sg.wg_begin;
  const int cl = lane % NC;
  for (int c0 = 0; c0 < N; c0 += NC) {
    const int c = c0 + cl, cc = min(c, N - 1);
    if (sg.kind == kSegA) {
      // R rows per subgroup; every lane runs all R rounds (subgroups past the
      // end take an empty span) so the shuffles stay wave-uniform.
      const int R = sg.rows_per_sub;
      const int sub = lane / GA, es = (lane % GA) / NC;
      const int subs_per_wg = kWavesPerWG * SA;
      const int end = sg.list_off + sg.count;
      int pos = sg.list_off + local_wg * subs_per_wg * R + wave * SA + sub;
      for (int k = 0; k < R; ++k, pos += subs_per_wg) {
        const bool ok = pos < end;
        const int2 sp = ok ? span[pos] : make_int2(0, 0);
        float acc = narrow_range<TB, NC, ESA, kPerm>(sp.x, sp.y, es, cc, colind, val, perm, B, ldb);
#pragma unroll
        for (int o = NC; o < GA; o <<= 1) acc += __shfl_xor(acc, o, kWave);
        if (ok && es == 0 && c < N) C[static_cast<size_t>(rows[pos]) * ldc + c] = acc;
      }
    } else if (sg.kind == kSegB) {
      const int pos = sg.list_off + local_wg * kWavesPerWG + wave;
      if (pos >= sg.list_off + sg.count) return;  // whole wave, no barrier below
      const int2 sp = span[pos];
      const int es = lane / NC;
      float acc = narrow_range<TB, NC, kWave / NC, kPerm>(sp.x, sp.y, es, cc, colind, val, perm,
                                                          B, ldb);
#pragma unroll
      for (int o = NC; o < kWave; o <<= 1) acc += __shfl_xor(acc, o, kWave);
      if (es == 0 && c < N) C[static_cast<size_t>(rows[pos]) * ldc + c] = acc;
    } else {
      const int pos = sg.list_off + local_wg;
      const int2 sp = span[pos];
      const int es = threadIdx.x / NC;
      float acc = narrow_range<TB, NC, kThreads / NC, kPerm>(sp.x, sp.y, es, cc, colind, val,
                                                             perm, B, ldb);
#pragma unroll
      for (int o = NC; o < kWave; o <<= 1) acc += __shfl_xor(acc, o, kWave);
      if (lane < NC) red[wave][lane] = acc;
      __syncthreads();
      if (wave == 0 && lane < NC && c < N) {
#pragma unroll
        for (int w = 1; w < kWavesPerWG; ++w) acc += red[w][lane];
        C[static_cast<size_t>(rows[pos]) * ldc + c] = acc;
      }
      __syncthreads();  // red[] is reused by the next column chunk
    }
  }
}

// ------------------------------------------------------------------- host
static bool tune_valid(int N, int e16, int G, int VEC) {
  if (G < 1 || G > 64 || (G & (G - 1)) || (VEC != 1 && VEC != 2)) return false;
  const int epl = e16 * VEC;
  return N % (G * epl) == 0;
}

// Lane geometry for one call: an explicit tune, else the plan's autotuned
// entry for (N, dtype), else the widest power-of-two subgroup that tiles N.
// G == 0 selects the narrow kernel (unaligned operands or N % e16 != 0).
static Geometry pick_geometry(const spmm_csr_plan_s *p, int N, int elem_bytes, const void *B,
                              int ldb, const void *C, int ldc, const spmm_tune_t *tune) {
  Geometry g;
  const int e16 = 16 / elem_bytes;
  const bool aligned = aligned16(B) && aligned16(C) && (ldb % e16 == 0) && (ldc % 4 == 0);
  if (!aligned || N % e16 != 0) return g;
  spmm_tune_t t = {0, 0, 0};
  if (tune) {
    t = *tune;
  } else if (p) {
    for (int i = 0; i < p->ntuned; ++i)
      if (p->tuned[i].N == N && p->tuned[i].elem_bytes == elem_bytes) t = p->tuned[i].tune;
  }
  g.rows_scale = t.rows_scale > 0 ? t.rows_scale : 1;
  int vec = t.vec == 2 ? 2 : 1;
  if (t.lanes_per_row == 0 && t.vec == 0 && elem_bytes == 2) vec = 2;  // default below
  if (N % (e16 * vec) != 0) vec = 1;
  int G = tune_valid(N, e16, t.lanes_per_row, vec) ? t.lanes_per_row : 0;
  if (G == 0) {
    // Default (no tune, nothing autotuned for this N and dtype): two 16-byte
    // vectors per lane for BF16 B, one for FP32, and the widest power of two
    // up to 4 (BF16) or 8 (FP32) lanes per row that divides the vectors of N;
    // the autotuner's picks on the 232,965-row workload at N = 64 and 128
    // (FP32 8 x 1, BF16 4 x 2) and within a few percent of them elsewhere.
    const int lanes = N / (e16 * vec);
    const int cap = elem_bytes == 2 ? 4 : 8;
    for (int c = cap; c >= 1; c >>= 1)
      if (c <= lanes && lanes % c == 0) { G = c; break; }
  }
  g.G = G;
  g.VEC = vec;
  g.col_tiles = N / (G * e16 * vec);
  const int sl = t.slices;
  g.slices = sl >= 2 && sl <= 64 && G >= 4 && G <= 32 ? sl : 1;
  return g;
}

void build_table_ex(const int *counts, const int *offs, int S, long long a_limit,
                    long long c_limit, int r_num, int r_empty, int rows_scale, LaunchTable *t,
                    int b_items, int c_items) {
  t->nseg = 0;
  int wg = 0;
  for (int b = kBins - 1; b >= 0; --b) {
    const int cnt = counts[b];
    if (cnt == 0) continue;
    const long long lo = b == 0 ? 0 : (1ll << (b - 1));
    const long long hi = b == 0 ? 1 : (1ll << b);  // exclusive upper bound
    Seg s;
    s.list_off = offs[b];
    s.count = cnt;
    s.wg_begin = wg;
    if (hi <= a_limit) {
      s.kind = kSegA;
      int R = b == 0 ? r_empty : std::max(1, static_cast<int>(r_num / hi));
      R = std::max(1, R * rows_scale);
      s.rows_per_sub = R;
      wg += ceil_div(cnt, kWavesPerWG * S * R);
    } else if (lo >= c_limit) {
      s.kind = kSegC;
      s.rows_per_sub = 1;
      wg += cnt * c_items;
    } else {
      s.kind = kSegB;
      s.rows_per_sub = 1;
      wg += ceil_div(cnt * b_items, kWavesPerWG);
    }
    t->seg[t->nseg++] = s;
  }
  t->total_wg = wg;
}

// CSR classes: degree < a_limit -> one subgroup per row (each row costs
// a dependent index -> gather round trip, so few rows per subgroup and many
// workgroups); degree >= 64*S -> one workgroup per row; else one wave.
// a_limit = max(8G, 32) up to 16 lanes (a narrow subgroup keeps rows of up to
// 8 chunks rather than leaving most of a wave idle on a 16-entry row: 16k-row
// N=32 7.9 -> 7.2 us, 89k-row N=32 18.1 -> 16.8 us), 2G for wider subgroups
// (89k-row N=128 51.4 us against 52.7 with 8G).
void build_table(const int *counts, const int *offs, int G, int rows_scale, LaunchTable *t) {
  int a_limit, c_limit;
  csr_class_limits(G, &a_limit, &c_limit);
  build_table_ex(counts, offs, kWave / G, a_limit, c_limit, 8, 16, rows_scale, t);
}

// Launch table of one slice of an S-sliced product: the classes and rows
// per subgroup of build_table, on degrees divided by S (a slice holds about
// 1/S of each row's entries).
static void build_table_sliced(const int *counts, const int *offs, int G, int rows_scale, int S,
                               LaunchTable *t) {
  int a_limit, c_limit;
  csr_class_limits(G, &a_limit, &c_limit);
  build_table_ex(counts, offs, kWave / G, static_cast<long long>(a_limit) * S,
                 static_cast<long long>(c_limit) * S, 8 * S, 16, rows_scale, t);
}

// The plan's slice offsets for (S, K) over its `units` plan units (rows, or
// block rows of a B2SR plan, whose M counts rows), built on first use: one
// allocation, one kernel and a read-back of the sortedness flag
// (synchronises). Up to six (S, K) sets are kept; a seventh replaces the
// last. SPMM_ERR_UNSUPPORTED when colind is not ascending within its rows.
static int slice_offsets(spmm_csr_plan_s *p, int units, int S, int K, const int *colind,
                         hipStream_t s, const int **out) {
  auto &sets = p->slice_sets;
  for (auto &ss : sets)
    if (ss.d_off && ss.S == S && ss.K == K) { *out = ss.d_off; return SPMM_OK; }
  if (p->sorted == 0) SPMM_FAIL(SPMM_ERR_UNSUPPORTED, "column slices need ascending colind in every row");
  int slot = 5;
  for (int i = 0; i < 6; ++i)
    if (!sets[i].d_off) { slot = i; break; }
  if (sets[slot].d_off) {
    SPMM_HIP(hipStreamSynchronize(s));  // launches queued earlier may read it
    (void)hipFree(sets[slot].d_off);
    sets[slot] = {};
  }
  const size_t n = (static_cast<size_t>(S) + 1) * static_cast<size_t>(units);
  int *d = nullptr;
  if (hipMalloc(&d, sizeof(int) * (n + 1)) != hipSuccess)
    SPMM_FAIL(SPMM_ERR_ALLOC, "hipMalloc of the slice offsets failed");
  int *flag = d + n;
  int unsorted = 0;
  const int W = ceil_div(K, S);
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(slice_offsets_kernel, dim3(std::min(ceil_div(units, kWavesPerWG), 8192)),
                       dim3(kThreads), 0, s, units, S, W, p->d_span, colind, d, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&unsorted, flag, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d);
    SPMM_FAIL(SPMM_ERR_HIP, std::string("slice offsets: ") + hipGetErrorString(e));
  }
  p->sorted = unsorted ? 0 : 1;
  if (unsorted) {
    (void)hipFree(d);
    SPMM_FAIL(SPMM_ERR_UNSUPPORTED, "column slices need ascending colind in every row");
  }
  sets[slot].S = S;
  sets[slot].K = K;
  sets[slot].d_off = d;
  *out = d;
  return SPMM_OK;
}

int dropin_bin_sort(spmm_handle_s *h, const int *ptr, int units, int b_wg, int b_sub,
                    hipStream_t s, const int **rows, const int **meta) {
  const int st = grow_ints(h, &h->slice_off, &h->slice_off_cap,
                           static_cast<size_t>(kDropinHead) + static_cast<size_t>(units), s);
  if (st != SPMM_OK) return st;
  int *counts = h->slice_off, *cursor = counts + kBins, *m = counts + 2 * kBins;
  int *r = h->slice_off + kDropinHead;
  const int grid = std::min(ceil_div(units, kThreads), 2048);
  hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(kThreads), 0, s, ptr, units, 0, counts);
  hipLaunchKernelGGL(dropin_bins_kernel, dim3(1), dim3(kWave), 0, s, counts, cursor, m, b_wg,
                     b_sub);
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(kThreads), 0, s, ptr, units, 0, cursor,
                     r);
  *rows = r;
  *meta = m;
  return SPMM_OK;
}

int plan_slice_offsets(spmm_csr_plan_s *p, int units, int S, int K, const int *colind,
                       hipStream_t s, const int **out) {
  return slice_offsets(p, units, S, K, colind, s, out);
}

template <typename TB, int G, int VEC, bool kPerm>
static void launch_sliced(const LaunchTable &t, int col_tiles, int S, const spmm_csr_plan_s *p,
                          const int *off, const int *colind, const float *val, const int *perm,
                          const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  const size_t M = static_cast<size_t>(p->M);
  for (int ct = 0; ct < col_tiles; ++ct)
    for (int sl = 0; sl < S; ++sl) {
      const int *lo = off + sl * M, *hi = off + (sl + 1) * M;
      if (sl == 0)
        hipLaunchKernelGGL((csr_sliced_kernel<TB, G, VEC, false, kPerm>), dim3(t.total_wg),
                           dim3(kThreads), 0, s, t, ct, p->d_rows, lo, hi, colind, val, perm, B,
                           ldb, C, ldc);
      else
        hipLaunchKernelGGL((csr_sliced_kernel<TB, G, VEC, true, kPerm>), dim3(t.total_wg),
                           dim3(kThreads), 0, s, t, ct, p->d_rows, lo, hi, colind, val, perm, B,
                           ldb, C, ldc);
    }
}

template <typename TB>
static void dispatch_sliced(const Geometry &g, const LaunchTable &t, const spmm_csr_plan_s *p,
                            const int *off, const int *colind, const float *val,
                            const int *perm, const TB *B, int ldb, float *C, int ldc,
                            hipStream_t s) {
#define SPMM_SL(gg, vv)                                                                        \
  (perm ? launch_sliced<TB, gg, vv, true>(t, g.col_tiles, g.slices, p, off, colind, val, perm, \
                                          B, ldb, C, ldc, s)                                   \
        : launch_sliced<TB, gg, vv, false>(t, g.col_tiles, g.slices, p, off, colind, val,      \
                                           nullptr, B, ldb, C, ldc, s))
  if (g.VEC == 2) {
    switch (g.G) {
      case 4: SPMM_SL(4, 2); break;
      case 8: SPMM_SL(8, 2); break;
      case 16: SPMM_SL(16, 2); break;
      default: SPMM_SL(32, 2); break;
    }
  } else {
    switch (g.G) {
      case 4: SPMM_SL(4, 1); break;
      case 8: SPMM_SL(8, 1); break;
      case 16: SPMM_SL(16, 1); break;
      default: SPMM_SL(32, 1); break;
    }
  }
#undef SPMM_SL
}

template <typename TB, int G, int VEC>
static void launch_vec(const LaunchTable &t, int col_tiles, const spmm_csr_plan_s *p,
                       const int *rowptr, const int *colind, const float *val,
                       const int *perm, const TB *B, int ldb, float *C, int ldc,
                       hipStream_t s) {
  dim3 grid(t.total_wg, col_tiles);
  if (perm)
    hipLaunchKernelGGL((csr_spmm_perm_kernel<TB, G, VEC>), grid, dim3(kThreads), 0, s, t,
                       p->d_rows, p->d_span, colind, val, perm, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((csr_spmm_kernel<TB, G, VEC>), grid, dim3(kThreads), 0, s, t,
                       p->d_rows, p->d_span, colind, val, B, ldb, C, ldc);
}

template <typename TB, int VEC>
static void dispatch_G(int G, const LaunchTable &t, int col_tiles,
                       const spmm_csr_plan_s *p, const int *rowptr,
                       const int *colind, const float *val, const int *perm, const TB *B,
                       int ldb, float *C, int ldc, hipStream_t s) {
#define SPMM_LV(g) launch_vec<TB, g, VEC>(t, col_tiles, p, rowptr, colind, val, perm, B, ldb, C, ldc, s)
  switch (G) {
    case 1: SPMM_LV(1); break;
    case 2: SPMM_LV(2); break;
    case 4: SPMM_LV(4); break;
    case 8: SPMM_LV(8); break;
    case 16: SPMM_LV(16); break;
    case 32: SPMM_LV(32); break;
    default: SPMM_LV(64); break;
  }
#undef SPMM_LV
}

template <typename TB, int NC>
static void launch_narrow_nc(const spmm_csr_plan_s *p, int N, const int *colind,
                             const float *val, const int *perm, const TB *B, int ldb, float *C,
                             int ldc, hipStream_t s) {
  constexpr int ESA = narrow_slots_a<NC>();
  LaunchTable t;
  // class A: rows that fit one round of a subgroup; class C: rows needing
  // more than 8 rounds of a wave
  build_table_ex(p->counts, p->offs, kWave / (NC * ESA), ESA * kNarrowU,
                 static_cast<long long>(kWave / NC) * kNarrowU * 8, 8, 16, 1, &t);
  if (t.total_wg == 0) return;
  if (perm)
    hipLaunchKernelGGL((csr_narrow_kernel<TB, NC, true>), dim3(t.total_wg), dim3(kThreads), 0, s,
                       t, p->d_rows, p->d_span, colind, val, perm, B, ldb, C, ldc, N);
  else
    hipLaunchKernelGGL((csr_narrow_kernel<TB, NC, false>), dim3(t.total_wg), dim3(kThreads), 0,
                       s, t, p->d_rows, p->d_span, colind, val, perm, B, ldb, C, ldc, N);
}

template <typename TB>
static void launch_narrow(const spmm_csr_plan_s *p, int N, const int *colind, const float *val,
                          const int *perm, const TB *B, int ldb, float *C, int ldc,
                          hipStream_t s) {
#define SPMM_LN(nc) launch_narrow_nc<TB, nc>(p, N, colind, val, perm, B, ldb, C, ldc, s)
  if (N <= 1) SPMM_LN(1);
  else if (N <= 2) SPMM_LN(2);
  else if (N <= 4) SPMM_LN(4);
  else if (N <= 8) SPMM_LN(8);
  else if (N <= 16) SPMM_LN(16);
  else if (N <= 32) SPMM_LN(32);
  else SPMM_LN(64);
#undef SPMM_LN
}

// perm != nullptr: entry e's value is val[perm[e]] (transposed structure).
template <typename TB>
static int run_planned(spmm_csr_plan_s *p, int M, int K, int N, int nnz,
                       const int *rowptr, const int *colind, const float *val,
                       const TB *B, int ldb, float *C, int ldc,
                       const spmm_tune_t *tune, void *stream, const int *perm = nullptr) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  if (p->tile != 0) SPMM_FAIL(SPMM_ERR_MISMATCH, "a B2SR plan was given to a CSR call");
  if (p->M != M || p->nnz != nnz)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "plan was built for M=" + std::to_string(p->M) +
                                     " nnz=" + std::to_string(p->nnz));
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(rowptr && C, "null rowptr or C");
  SPMM_CHECK_ARG(nnz == 0 || (colind && val && B), "null colind, val or B");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldb >= N), "ldb/ldc smaller than N");
  hipStream_t s = as_stream(stream);
  Geometry g = pick_geometry(p, N, sizeof(TB), B, ldb, C, ldc, tune);
  if (g.G == 0) {
    launch_narrow<TB>(p, N, colind, val, perm, B, ldb, C, ldc, s);
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  LaunchTable t;
  if (g.slices > 1 && nnz > 0) {
    const int *off = nullptr;
    const int st = slice_offsets(p, p->M, g.slices, K, colind, s, &off);
    if (st) return st;
    build_table_sliced(p->counts, p->offs, g.G, g.rows_scale, g.slices, &t);
    if (t.total_wg == 0) return SPMM_OK;
    dispatch_sliced<TB>(g, t, p, off, colind, val, perm, B, ldb, C, ldc, s);
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  build_table(p->counts, p->offs, g.G, g.rows_scale, &t);
  if (t.total_wg == 0) return SPMM_OK;
  if (g.VEC == 2)
    dispatch_G<TB, 2>(g.G, t, g.col_tiles, p, rowptr, colind, val, perm, B, ldb, C, ldc, s);
  else
    dispatch_G<TB, 1>(g.G, t, g.col_tiles, p, rowptr, colind, val, perm, B, ldb, C, ldc, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

// Times every valid (lanes per row, vector width, rows per subgroup, column
// slices) on the caller's operands and records the fastest in the plan for
// (N, dtype).
template <typename TB>
static int autotune(spmm_csr_plan_s *p, int M, int K, int N, int nnz, const int *rowptr,
                    const int *colind, const float *val, const TB *B, int ldb, float *C,
                    int ldc, spmm_tune_t *best_out, void *stream, const int *perm = nullptr) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  hipStream_t s = as_stream(stream);
  const int e16 = 16 / static_cast<int>(sizeof(TB));
  spmm_tune_t best = {0, 0, 0, 0};
  float best_ms = 1e30f;
  bool slicing = nnz > 0 && p->sorted != 0;
  hipEvent_t e0, e1;
  SPMM_HIP(hipEventCreate(&e0));
  SPMM_HIP(hipEventCreate(&e1));
  int st = SPMM_OK;
  for (int vec = 1; vec <= 2 && st == SPMM_OK; ++vec) {
    if (N % (e16 * vec)) continue;
    const int lanes = N / (e16 * vec);
    for (int G = 64; G >= 1 && st == SPMM_OK; G >>= 1) {
      if (!tune_valid(N, e16, G, vec) || G * 4 < std::min(lanes, 64)) continue;  // <= 4 column tiles
      // slice counts: none, then 2, 3, 4, 6, ..., 24, 32 (powers of two and
      // the midpoints between them)
      constexpr int kSlices[10] = {0, 2, 3, 4, 6, 8, 12, 16, 24, 32};
      for (int cand_i = 0; cand_i < 4 * 10 && st == SPMM_OK; ++cand_i) {
        const int rs = 1 << (cand_i % 4), sl = kSlices[cand_i / 4];
        // column slices: only where one tile of B outgrows an XCD's L2, only
        // on a plan that can hold them, at rows_scale 1 and 2
        if (sl && (!slicing || G < 4 || G > 32 || rs > 2 ||
                   static_cast<long long>(K) * G * vec * 16 < (4ll << 20)))
          continue;
        spmm_tune_t cand = {G, vec, rs, sl};
        st = run_planned<TB>(p, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, &cand, stream,
                             perm);
        if (sl && st == SPMM_ERR_UNSUPPORTED) {  // colind not ascending: no slices
          slicing = false;
          st = SPMM_OK;
          continue;
        }
        if (st) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = SPMM_ERR_HIP; break; }
        for (int r = 0; r < 3 && st == SPMM_OK; ++r)
          st = run_planned<TB>(p, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, &cand,
                               stream, perm);
        if (st) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          st = SPMM_ERR_HIP;
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best_ms) { best_ms = ms; best = cand; }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  if (best.lanes_per_row == 0) SPMM_FAIL(SPMM_ERR_UNSUPPORTED, "no vector geometry fits this N / alignment");
  int slot = -1;
  for (int i = 0; i < p->ntuned; ++i)
    if (p->tuned[i].N == N && p->tuned[i].elem_bytes == static_cast<int>(sizeof(TB))) slot = i;
  if (slot < 0) slot = p->ntuned < kTuneSlots ? p->ntuned++ : kTuneSlots - 1;
  p->tuned[slot] = {N, static_cast<int>(sizeof(TB)), best};
  if (best_out) *best_out = best;
  return SPMM_OK;
}

// Bins the M rows by degree, where row r's degree is
// rowptr[(r >> shift) + 1] - rowptr[r >> shift] (shift 0: a CSR; shift
// log2(T): the tile count of a B2SR block row).
int analyze(spmm_handle_s *h, int M, int nnz, const int *rowptr, int shift,
            spmm_csr_plan_s **out, void *stream) {
  SPMM_CHECK_ARG(h && out, "null handle or out");
  *out = nullptr;
  SPMM_CHECK_ARG(M >= 0 && nnz >= 0, "negative size");
  SPMM_CHECK_ARG(M == 0 || rowptr, "rowptr is null");
  hipStream_t s = as_stream(stream);
  auto *p = new (std::nothrow) spmm_csr_plan_s();
  if (!p) SPMM_FAIL(SPMM_ERR_ALLOC, "host allocation failed");
  p->M = M;
  p->nnz = nnz;
  p->device = h->device;
  if (M == 0) { *out = p; return SPMM_OK; }
  if (hipMalloc(&p->d_rows, sizeof(int) * static_cast<size_t>(M)) != hipSuccess ||
      hipMalloc(&p->d_span, sizeof(int2) * static_cast<size_t>(M)) != hipSuccess) {
    spmm_csr_plan_destroy(p);
    SPMM_FAIL(SPMM_ERR_ALLOC, "hipMalloc of the row list failed");
  }
  int *ws = static_cast<int *>(workspace(h, sizeof(int) * kBins));
  int *host = static_cast<int *>(pinned(h, sizeof(int) * kBins));
  if (!ws || !host) { spmm_csr_plan_destroy(p); return SPMM_ERR_ALLOC; }
  auto fail = [&](const char *what) {
    spmm_csr_plan_destroy(p);
    set_error(std::string("spmm_csr_analyze: ") + what + " failed");
    return SPMM_ERR_HIP;
  };
  const int grid = std::min(ceil_div(M, kThreads), 2048);
  if (hipMemsetAsync(ws, 0, sizeof(int) * kBins, s) != hipSuccess) return fail("memset");
  hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(kThreads), 0, s, rowptr, M, shift, ws);
  if (hipGetLastError() != hipSuccess) return fail("bin_count_kernel launch");
  if (hipMemcpyAsync(host, ws, sizeof(int) * kBins, hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail("count read-back");
  if (hipStreamSynchronize(s) != hipSuccess) return fail("stream sync");
  int off = 0;
  for (int b = kBins - 1; b >= 0; --b) {
    p->counts[b] = host[b];
    p->offs[b] = off;
    off += host[b];
  }
  if (off != M) {
    spmm_csr_plan_destroy(p);
    SPMM_FAIL(SPMM_ERR_INVALID_ARG, "row degrees do not add up (rowptr not monotone?)");
  }
  for (int b = 0; b < kBins; ++b) host[b] = p->offs[b];
  if (hipMemcpyAsync(ws, host, sizeof(int) * kBins, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail("cursor upload");
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(kThreads), 0, s, rowptr, M, shift, ws,
                     p->d_rows);
  if (hipGetLastError() != hipSuccess) return fail("bin_scatter_kernel launch");
  hipLaunchKernelGGL(span_kernel, dim3(grid), dim3(kThreads), 0, s, rowptr, M, shift, p->d_rows,
                     p->d_span);
  if (hipGetLastError() != hipSuccess) return fail("span_kernel launch");
  // The pinned buffer and workspace are reused by the next call.
  if (hipStreamSynchronize(s) != hipSuccess) return fail("stream sync");
  *out = p;
  return SPMM_OK;
}

// Workgroups of `kernel` the chip holds at once (per-CU occupancy, cached per
// kernel instantiation, times the CU count).
static int resident_wgs(const void *kernel, int dev, std::atomic<int> *cache) {
  int per_cu = cache->load(std::memory_order_relaxed);
  if (per_cu == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess ||
        per_cu <= 0)
      per_cu = 4;
    cache->store(per_cu, std::memory_order_relaxed);
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return per_cu * cus;
}

template <typename TB, int G, int VEC>
static void launch_dropin(int M, int col_tiles, int dev, const int *rowptr, const int *colind,
                          const float *val, const int *perm, const TB *B, int ldb, float *C,
                          int ldc, int *state, hipStream_t s) {
  constexpr int SW = kWavesPerWG * (kWave / G);
  int a_limit, c_limit;
  csr_class_limits(G, &a_limit, &c_limit);
  hipLaunchKernelGGL(class_list_kernel<kThreads>, dim3(ceil_div(M, kThreads)), dim3(kThreads), 0,
                     s, M, rowptr, a_limit, c_limit, state);
  auto *const kern = perm ? &csr_dropin_kernel<TB, G, VEC, true> : &csr_dropin_kernel<TB, G, VEC, false>;
  static std::atomic<int> per_cu[2];
  // listed work on as many workgroups as the chip holds at once (fewer for
  // small M, where most of them would find nothing to take)
  const int E = std::max(1, std::min(resident_wgs(reinterpret_cast<const void *>(kern), dev,
                                                  &per_cu[perm ? 1 : 0]),
                                     ceil_div(M, 16)));
  const dim3 grid(E + ceil_div(M, SW * kDropinRounds), col_tiles);
  hipLaunchKernelGGL(kern, grid, dim3(kThreads), 0, s, M, E, a_limit, rowptr, colind, val, perm,
                     B, ldb, C, ldc, state);
  hipLaunchKernelGGL(list_counts_reset_kernel, dim3(1), dim3(kWave), 0, s, state);
}

template <typename TB, int VEC>
static void dispatch_dropin(int G, int M, int col_tiles, int dev, const int *rowptr,
                            const int *colind, const float *val, const int *perm, const TB *B,
                            int ldb, float *C, int ldc, int *state, hipStream_t s) {
#define SPMM_DL(g) \
  launch_dropin<TB, g, VEC>(M, col_tiles, dev, rowptr, colind, val, perm, B, ldb, C, ldc, state, s)
  switch (G) {
    case 1: SPMM_DL(1); break;
    case 2: SPMM_DL(2); break;
    case 4: SPMM_DL(4); break;
    case 8: SPMM_DL(8); break;
    case 16: SPMM_DL(16); break;
    case 32: SPMM_DL(32); break;
    default: SPMM_DL(64); break;
  }
#undef SPMM_DL
}

// Column slices of the plan-free form: one column tile of B (N / 16-byte
// vector lanes, 4 to 32) cut into slices of about 10 MB, when one tile of B
// takes at least 16 MB and rows hold at least 16 entries a slice on average;
// 1 (no slices) otherwise. On the 232,965-row graph this gives the
// autotuner's counts: 12 at FP32 N = 128, 6 at FP32 N = 64 and BF16 N = 128.
static int dropin_slice_count(int M, int K, int N, long long nnz, int elem_bytes) {
  const int e16 = 16 / elem_bytes, lanes = N / e16;
  if (M <= 0 || N % e16 != 0 || lanes < 4 || lanes > 32 || (lanes & (lanes - 1))) return 1;
  const double tile = static_cast<double>(K) * N * elem_bytes;
  const int S = std::min(32, static_cast<int>(tile / 10e6 + 0.5));
  if (S < 2 || nnz < 16ll * S * M) return 1;
  return S;
}

// Rows per subgroup of the sliced drop-in's subgroup class (rows sorted by
// degree bin): about 240 entries per subgroup and launch, 3 to 8 rows.
// Measured on the 232,965-row graph, ms per product at 2 / 3 / 4 / 6 rows:
// FP32 N = 128 (12 slices, 41 entries a row a slice) 4.51 / 4.04 / 3.90 /
// 3.81; FP32 N = 64 (6 slices) 2.15 / 2.11 / 2.11 / 2.23; BF16 N = 128
// (6 slices) 2.28 / 2.23 / 2.26 / 2.39; BF16 N = 64 (3 slices) 1.47 / 1.44 /
// 1.43 / 1.44. One row was slower everywhere (FP32 N = 128: 6.3 ms),
// probably because the listed rows hold every resident slot first and the
// short workgroups then queue behind them.
static int dropin_sliced_rounds(int M, long long nnz, int S) {
  const double per_slice = static_cast<double>(nnz) / (static_cast<double>(M) * S);
  return std::max(3, std::min(8, static_cast<int>(240.0 / std::max(per_slice, 1.0) + 0.5)));
}

// The planned sliced launch's class limits as per-row degree limits: a bin
// of degrees [2^(b-1), 2^b) runs a subgroup per row when 2^b <= a_limit S
// (degree below the largest power of two <= a_limit S) and a workgroup per
// row when 2^(b-1) >= c_limit S (degree at least the smallest power of two
// >= c_limit S).
static void sliced_row_limits(int G, int S, int *a_row, int *c_row) {
  int a_limit, c_limit;
  csr_class_limits(G, &a_limit, &c_limit);
  const long long a = static_cast<long long>(a_limit) * S, c = static_cast<long long>(c_limit) * S;
  long long pa = 1, pc = 1;
  while (2 * pa <= a) pa *= 2;
  while (pc < c) pc *= 2;
  *a_row = static_cast<int>(pa);
  *c_row = static_cast<int>(std::min<long long>(pc, INT_MAX));
}

template <typename TB, int G>
static void launch_dropin_sliced(int M, int S, int W, int rounds, int dev, const int *rowptr,
                                 const int *colind, const float *val, const TB *B, int ldb,
                                 float *C, int ldc, int *state, int *scratch, hipStream_t s) {
  constexpr int SW = kWavesPerWG * (kWave / G);
  int a_row, c_row;
  sliced_row_limits(G, S, &a_row, &c_row);
  // scratch: the row sort's bin counts (zero between calls), cursors and
  // class sizes at a fixed head, then the slice offsets
  int *const counts = scratch, *const cursor = counts + kBins, *const meta = cursor + kBins;
  int *const off = scratch + kDropinHead;
  int *const rows = state + kStateHead;
  const int b_sub = 31 - __builtin_clz(static_cast<unsigned>(a_row));      // last subgroup-class bin
  const int b_wg = 32 - __builtin_clz(static_cast<unsigned>(c_row));       // first workgroup-class bin
  const int grid_r = std::min(ceil_div(M, kThreads), 2048);
  hipLaunchKernelGGL(bin_count_kernel, dim3(grid_r), dim3(kThreads), 0, s, rowptr, M, 0, counts);
  hipLaunchKernelGGL(dropin_bins_kernel, dim3(1), dim3(kWave), 0, s, counts, cursor, meta, b_wg,
                     b_sub);
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid_r), dim3(kThreads), 0, s, rowptr, M, 0, cursor,
                     rows);
  hipLaunchKernelGGL(dropin_slice_offsets_kernel,
                     dim3(std::min(ceil_div(M, kWavesPerWG * kOffRows), 8192)),
                     dim3(kThreads), 0, s, M, S, W, rowptr, colind, off, state);
  auto *const k0 = &csr_dropin_sliced_kernel<TB, G, 1, false>;
  auto *const k1 = &csr_dropin_sliced_kernel<TB, G, 1, true>;
  static std::atomic<int> per_cu;
  const int E = std::max(
      1, std::min(resident_wgs(reinterpret_cast<const void *>(k1), dev, &per_cu), ceil_div(M, 16)));
  const dim3 grid(E + ceil_div(M, SW * rounds));
  const size_t Ms = static_cast<size_t>(M);
  for (int sl = 0; sl < S; ++sl) {
    const int *lo = off + sl * Ms, *hi = off + (sl + 1) * Ms;
    const int last = sl == S - 1;
    if (sl == 0)
      hipLaunchKernelGGL(k0, grid, dim3(kThreads), 0, s, M, E, rounds, rows, meta, lo, hi, rowptr,
                         colind, val, B, ldb, C, ldc, state, last);
    else
      hipLaunchKernelGGL(k1, grid, dim3(kThreads), 0, s, M, E, rounds, rows, meta, lo, hi, rowptr,
                         colind, val, B, ldb, C, ldc, state, last);
  }
}

template <typename TB>
static int run_dropin_sliced(spmm_handle_s *h, int M, int K, int N, int nnz, int S,
                             const int *rowptr,
                             const int *colind, const float *val, const TB *B, int ldb, float *C,
                             int ldc, int *state, hipStream_t s) {
  SPMM_CHECK_ARG(S >= 2 && S <= kOffMaxS, "plan-free slice count out of range");
  const size_t n = kDropinHead + (static_cast<size_t>(S) + 1) * static_cast<size_t>(M);
  const int st = grow_ints(h, &h->slice_off, &h->slice_off_cap, n, s);
  if (st) return st;
  const int W = ceil_div(K, S);
  const int rounds = dropin_sliced_rounds(M, nnz, S);
#define SPMM_DS(g)                                                                             \
  launch_dropin_sliced<TB, g>(M, S, W, rounds, h->device, rowptr, colind, val, B, ldb, C, ldc, state, \
                              h->slice_off, s)
  switch (N / Vec16<TB>::kElems) {
    case 4: SPMM_DS(4); break;
    case 8: SPMM_DS(8); break;
    case 16: SPMM_DS(16); break;
    default: SPMM_DS(32); break;
  }
#undef SPMM_DS
  SPMM_LAUNCHED();
  return SPMM_OK;
}

// The plan-free form: no host analysis, no allocation after the first call
// at a given size, no synchronisation (the narrow kernel's case aside, which
// analyses per call). perm != nullptr: entry e's value is val[perm[e]].
// Large graphs (dropin_slice_count) run column-sliced, each row summed as
// the planned sliced launch sums it at the same geometry and slice count.
template <typename TB>
static int run_dropin(spmm_handle_s *h, int M, int K, int N, int nnz,
                      const int *rowptr, const int *colind, const float *val,
                      const TB *B, int ldb, float *C, int ldc, void *stream,
                      const int *perm = nullptr) {
  SPMM_CHECK_ARG(h != nullptr, "null handle");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(rowptr && C, "null rowptr or C");
  SPMM_CHECK_ARG(nnz == 0 || (colind && val && B), "null colind, val or B");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldb >= N), "ldb/ldc smaller than N");
  hipStream_t s = as_stream(stream);
  const Geometry g = pick_geometry(nullptr, N, sizeof(TB), B, ldb, C, ldc, nullptr);
  if (g.G == 0) {
    // narrow or unaligned operands: the narrow kernel on a per-call analysis
    spmm_csr_plan_s *p = nullptr;
    int st = analyze(h, M, nnz, rowptr, 0, &p, stream);
    if (st != SPMM_OK) return st;
    st = run_planned<TB>(p, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, nullptr, stream,
                         perm);
    // the plan's row list is read by the kernel just queued
    if (hipStreamSynchronize(s) != hipSuccess && st == SPMM_OK) st = SPMM_ERR_HIP;
    spmm_csr_plan_destroy(p);
    return st;
  }
  int *state = nullptr;
  const int st = list_state(h, M, s, &state);
  if (st) return st;
  const int S = perm ? 1 : dropin_slice_count(M, K, N, nnz, sizeof(TB));
  if (S > 1)
    return run_dropin_sliced<TB>(h, M, K, N, nnz, S, rowptr, colind, val, B, ldb, C, ldc, state, s);
  if (g.VEC == 2)
    dispatch_dropin<TB, 2>(g.G, M, g.col_tiles, h->device, rowptr, colind, val, perm, B, ldb, C,
                           ldc, state, s);
  else
    dispatch_dropin<TB, 1>(g.G, M, g.col_tiles, h->device, rowptr, colind, val, perm, B, ldb, C,
                           ldc, state, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // namespace spmm

// Plan of C = A^T B: the stable transpose of A's structure (csr2csc order)
// with the entry permutation, and the row schedule of that transpose. Values
// are not copied: each call reads val[perm[e]], so A's values may change
// between calls (learnable edge weights).
struct spmm_csrt_plan_s {
  int M = 0, K = 0, nnz = 0;
  int *trowptr = nullptr, *tcol = nullptr, *perm = nullptr;
  spmm_csr_plan_s *tplan = nullptr;
};

namespace spmm {

template <typename TB>
static int run_t_planned(spmm_csrt_plan_s *p, int M, int K, int N, int nnz, const float *val,
                         const TB *B, int ldb, float *C, int ldc, const spmm_tune_t *tune,
                         void *stream) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  if (p->M != M || p->K != K || p->nnz != nnz)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "transposed plan was built for M=" + std::to_string(p->M) +
                                     " K=" + std::to_string(p->K) +
                                     " nnz=" + std::to_string(p->nnz));
  // the transposed problem: K output rows, B has M rows
  return run_planned<TB>(p->tplan, K, M, N, nnz, p->trowptr, p->tcol, val, B, ldb, C, ldc, tune,
                         stream, p->perm);
}

// Plan-free C = A^T B: A's stable transpose built into the handle's buffers
// on the stream (no synchronisation), then the plan-free SpMM over it with
// values read through the permutation.
template <typename TB>
static int run_t_dropin(spmm_handle_s *h, int M, int K, int N, int nnz, const int *rowptr,
                        const int *colind, const float *val, const TB *B, int ldb, float *C,
                        int ldc, void *stream) {
  SPMM_CHECK_ARG(h != nullptr, "null handle");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  if (K == 0 || N == 0) return SPMM_OK;
  hipStream_t s = as_stream(stream);
  const size_t need = static_cast<size_t>(K) + 1 + 2 * static_cast<size_t>(nnz > 0 ? nnz : 1);
  int st = grow_ints(h, &h->tstruct, &h->tstruct_cap, need, s);
  if (st) return st;
  int *trp = h->tstruct;
  int *tcol = trp + K + 1;
  int *perm = tcol + (nnz > 0 ? nnz : 1);
  st = transpose_async(h, M, K, nnz, rowptr, colind, nullptr, trp, tcol, nullptr, perm, s);
  if (st) return st;
  return run_dropin<TB>(h, K, M, N, nnz, trp, tcol, val, B, ldb, C, ldc, stream, perm);
}

// The fused backward's geometry: one column tile (G lanes x VEC vectors
// = N), from `tune` when it gives one, else the widest single-vector
// subgroup (two vectors past 64 lanes). G = 0: not fusable.
static Geometry fused_geometry(int N, int elem_bytes, const spmm_tune_t *tune) {
  Geometry g;
  const int e16 = 16 / elem_bytes;
  if (N % e16 != 0) return g;
  const int lanes = N / e16;
  int G = 0, vec = 1;
  if (tune && tune->lanes_per_row > 0) {
    vec = tune->vec == 2 ? 2 : 1;
    G = tune->lanes_per_row;
  } else {
    vec = lanes > 64 ? 2 : 1;
    G = lanes / vec;
  }
  if (G < 1 || G > 64 || (G & (G - 1)) != 0 || G * vec != lanes) return g;
  g.G = G;
  g.VEC = vec;
  g.col_tiles = 1;
  g.rows_scale = tune && tune->rows_scale > 0 ? tune->rows_scale : 1;
  return g;
}

template <typename TB, int G, int VEC>
static void launch_t_sddmm(const LaunchTable &t, const spmm_csrt_plan_s *p, const float *val,
                           const TB *dY, int lddy, const TB *X, int ldx, float *dX, int lddx,
                           float *dval, hipStream_t s) {
  hipLaunchKernelGGL((csr_t_sddmm_kernel<TB, G, VEC>), dim3(t.total_wg), dim3(kThreads), 0, s,
                     t, p->tplan->d_rows, p->tplan->d_span, p->tcol, val, p->perm, dY, lddy, X,
                     ldx, dX, lddx, dval);
}

template <typename TB, int VEC>
static void dispatch_t_sddmm(int G, const LaunchTable &t, const spmm_csrt_plan_s *p,
                             const float *val, const TB *dY, int lddy, const TB *X, int ldx,
                             float *dX, int lddx, float *dval, hipStream_t s) {
#define SPMM_TS(g) launch_t_sddmm<TB, g, VEC>(t, p, val, dY, lddy, X, ldx, dX, lddx, dval, s)
  switch (G) {
    case 1: SPMM_TS(1); break;
    case 2: SPMM_TS(2); break;
    case 4: SPMM_TS(4); break;
    case 8: SPMM_TS(8); break;
    case 16: SPMM_TS(16); break;
    case 32: SPMM_TS(32); break;
    default: SPMM_TS(64); break;
  }
#undef SPMM_TS
}

// dX = A^T dY and dval = SDDMM(dY, X) on A's pattern: fused when one column
// tile covers N with 16-byte aligned operands, else the transposed SpMM and
// the entry-chunked SDDMM one after the other.
template <typename TB>
static int run_t_sddmm(spmm_csrt_plan_s *p, int M, int K, int N, int nnz, const int *rowptr,
                       const int *colind, const float *val, const TB *dY, int lddy,
                       const TB *X, int ldx, float *dX, int lddx, float *dval,
                       const spmm_tune_t *tune, void *stream) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  if (p->M != M || p->K != K || p->nnz != nnz)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "transposed plan was built for M=" + std::to_string(p->M) +
                                     " K=" + std::to_string(p->K) +
                                     " nnz=" + std::to_string(p->nnz));
  SPMM_CHECK_ARG(N >= 0, "negative size");
  const int e16 = 16 / static_cast<int>(sizeof(TB));
  const Geometry g = fused_geometry(N, sizeof(TB), tune);
  const bool fusable = g.G > 0 && K > 0 && nnz > 0 && val && dY && X && dX && dval &&
                       aligned16(dY) && aligned16(X) && aligned16(dX) && lddy % e16 == 0 &&
                       ldx % e16 == 0 && lddx % 4 == 0 && lddy >= N && ldx >= N && lddx >= N;
  if (!fusable) {
    const int st = run_t_planned<TB>(p, M, K, N, nnz, val, dY, lddy, dX, lddx, nullptr, stream);
    if (st) return st;
    return run_sddmm<TB>(nullptr, M, K, N, nnz, rowptr, colind, dY, lddy, X, ldx, dval, stream);
  }
  LaunchTable t;
  build_table(p->tplan->counts, p->tplan->offs, g.G, g.rows_scale, &t);
  if (t.total_wg == 0) return SPMM_OK;
  hipStream_t s = as_stream(stream);
  if (g.VEC == 2)
    dispatch_t_sddmm<TB, 2>(g.G, t, p, val, dY, lddy, X, ldx, dX, lddx, dval, s);
  else
    dispatch_t_sddmm<TB, 1>(g.G, t, p, val, dY, lddy, X, ldx, dX, lddx, dval, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // namespace spmm

extern "C" {

int spmm_csr_t_analyze(spmm_handle_t h, int M, int K, int nnz, const int *rowptr,
                       const int *colind, spmm_csrt_plan_t *plan, void *stream) {
  SPMM_CHECK_ARG(h && plan, "null handle or plan");
  *plan = nullptr;
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && nnz >= 0, "negative size");
  auto *p = new (std::nothrow) spmm_csrt_plan_s();
  if (!p) SPMM_FAIL(SPMM_ERR_ALLOC, "host allocation failed");
  p->M = M;
  p->K = K;
  p->nnz = nnz;
  const size_t e = static_cast<size_t>(nnz > 0 ? nnz : 1);
  if (hipMalloc(&p->trowptr, sizeof(int) * (static_cast<size_t>(K) + 1)) != hipSuccess ||
      hipMalloc(&p->tcol, sizeof(int) * e) != hipSuccess ||
      hipMalloc(&p->perm, sizeof(int) * e) != hipSuccess) {
    spmm_csr_t_plan_destroy(p);
    SPMM_FAIL(SPMM_ERR_ALLOC, "hipMalloc of the transposed structure failed");
  }
  int st = csr_transpose(h, M, K, nnz, rowptr, colind, nullptr, p->trowptr, p->tcol, nullptr,
                         p->perm, stream);
  if (st == SPMM_OK) st = spmm::analyze(h, K, nnz, p->trowptr, 0, &p->tplan, stream);
  if (st != SPMM_OK) {
    const std::string msg = spmm_last_error();
    spmm_csr_t_plan_destroy(p);
    spmm::set_error(msg);
    return st;
  }
  *plan = p;
  return SPMM_OK;
}

int spmm_csr_t_plan_destroy(spmm_csrt_plan_t plan) {
  if (!plan) return SPMM_OK;
  if (plan->trowptr) (void)hipFree(plan->trowptr);
  if (plan->tcol) (void)hipFree(plan->tcol);
  if (plan->perm) (void)hipFree(plan->perm);
  spmm_csr_plan_destroy(plan->tplan);
  delete plan;
  return SPMM_OK;
}

int spmm_csr_t_autotune(spmm_csrt_plan_t plan, int M, int K, int N, int nnz, const float *val,
                        const void *B, int bf16, int ldb, float *C, int ldc, spmm_tune_t *best,
                        void *stream) {
  SPMM_CHECK_ARG(plan != nullptr, "plan is null");
  if (plan->M != M || plan->K != K || plan->nnz != nnz)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "transposed plan was built for other sizes");
  if (bf16)
    return spmm::autotune<uint16_t>(plan->tplan, K, M, N, nnz, plan->trowptr, plan->tcol, val,
                                    static_cast<const uint16_t *>(B), ldb, C, ldc, best, stream,
                                    plan->perm);
  return spmm::autotune<float>(plan->tplan, K, M, N, nnz, plan->trowptr, plan->tcol, val,
                               static_cast<const float *>(B), ldb, C, ldc, best, stream,
                               plan->perm);
}

int spmm_csr_t_f32_planned(spmm_csrt_plan_t plan, int M, int K, int N, int nnz,
                           const float *val, const float *B, int ldb, float *C, int ldc,
                           const spmm_tune_t *tune, void *stream) {
  return spmm::run_t_planned<float>(plan, M, K, N, nnz, val, B, ldb, C, ldc, tune, stream);
}

int spmm_csr_t_bf16_planned(spmm_csrt_plan_t plan, int M, int K, int N, int nnz,
                            const float *val, const uint16_t *B, int ldb, float *C, int ldc,
                            const spmm_tune_t *tune, void *stream) {
  return spmm::run_t_planned<uint16_t>(plan, M, K, N, nnz, val, B, ldb, C, ldc, tune, stream);
}

int spmm_csr_t_sddmm_f32_planned(spmm_csrt_plan_t plan, int M, int K, int N, int nnz,
                                 const int *rowptr, const int *colind, const float *val,
                                 const float *dY, int lddy, const float *X, int ldx, float *dX,
                                 int lddx, float *dval, const spmm_tune_t *tune, void *stream) {
  return spmm::run_t_sddmm<float>(plan, M, K, N, nnz, rowptr, colind, val, dY, lddy, X, ldx, dX,
                                  lddx, dval, tune, stream);
}

int spmm_csr_t_sddmm_bf16_planned(spmm_csrt_plan_t plan, int M, int K, int N, int nnz,
                                  const int *rowptr, const int *colind, const float *val,
                                  const uint16_t *dY, int lddy, const uint16_t *X, int ldx,
                                  float *dX, int lddx, float *dval, const spmm_tune_t *tune,
                                  void *stream) {
  return spmm::run_t_sddmm<uint16_t>(plan, M, K, N, nnz, rowptr, colind, val, dY, lddy, X, ldx,
                                     dX, lddx, dval, tune, stream);
}

int spmm_csr_t_f32(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                   const int *colind, const float *val, const float *B, int ldb, float *C,
                   int ldc, void *stream) {
  return spmm::run_t_dropin<float>(h, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, stream);
}

int spmm_csr_t_bf16(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                    const int *colind, const float *val, const uint16_t *B, int ldb, float *C,
                    int ldc, void *stream) {
  return spmm::run_t_dropin<uint16_t>(h, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc,
                                      stream);
}

int spmm_csr_analyze(spmm_handle_t h, int M, int nnz, const int *rowptr,
                     spmm_csr_plan_t *plan, void *stream) {
  return spmm::analyze(h, M, nnz, rowptr, 0, plan, stream);
}

int spmm_csr_plan_destroy(spmm_csr_plan_t plan) {
  if (!plan) return SPMM_OK;
  if (plan->d_rows) (void)hipFree(plan->d_rows);
  if (plan->d_span) (void)hipFree(plan->d_span);
  for (auto &ss : plan->slice_sets)
    if (ss.d_off) (void)hipFree(ss.d_off);
  if (plan->d_wide) (void)hipFree(plan->d_wide);
  delete plan;
  return SPMM_OK;
}

int spmm_csr_autotune(spmm_csr_plan_t plan, int M, int K, int N, int nnz, const int *rowptr,
                      const int *colind, const float *val, const void *B, int bf16, int ldb,
                      float *C, int ldc, spmm_tune_t *best, void *stream) {
  if (bf16)
    return spmm::autotune<uint16_t>(plan, M, K, N, nnz, rowptr, colind, val,
                                    static_cast<const uint16_t *>(B), ldb, C, ldc, best, stream);
  return spmm::autotune<float>(plan, M, K, N, nnz, rowptr, colind, val,
                               static_cast<const float *>(B), ldb, C, ldc, best, stream);
}

int spmm_csr_plan_bins(spmm_csr_plan_t plan, int *counts32) {
  SPMM_CHECK_ARG(plan && counts32, "null plan or counts");
  for (int b = 0; b < 32; ++b) counts32[b] = plan->counts[b];
  return SPMM_OK;
}

int spmm_csr_dropin_geometry(int M, int K, int N, int nnz, int elem_bytes, spmm_tune_t *out) {
  SPMM_CHECK_ARG(out != nullptr, "out is null");
  SPMM_CHECK_ARG(elem_bytes == 4 || elem_bytes == 2, "elem_bytes must be 4 (FP32) or 2 (BF16)");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  const int S = spmm::dropin_slice_count(M, K, N, nnz, elem_bytes);
  if (S > 1) {
    *out = {N / (16 / elem_bytes), 1, 1, S};
    return SPMM_OK;
  }
  alignas(16) static const float probe[4] = {0.f, 0.f, 0.f, 0.f};  // aligned operands
  const spmm::Geometry g = spmm::pick_geometry(nullptr, N, elem_bytes, probe, N, probe, N, nullptr);
  *out = {g.G, g.G ? g.VEC : 0, 1, 0};
  return SPMM_OK;
}

int spmm_csr_f32(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                 const int *colind, const float *val, const float *B, int ldb,
                 float *C, int ldc, void *stream) {
  return spmm::run_dropin<float>(h, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, stream);
}

int spmm_csr_bf16(spmm_handle_t h, int M, int K, int N, int nnz, const int *rowptr,
                  const int *colind, const float *val, const uint16_t *B, int ldb,
                  float *C, int ldc, void *stream) {
  return spmm::run_dropin<uint16_t>(h, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc,
                                    stream);
}

int spmm_csr_f32_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                         const int *rowptr, const int *colind, const float *val,
                         const float *B, int ldb, float *C, int ldc,
                         const spmm_tune_t *tune, void *stream) {
  return spmm::run_planned<float>(plan, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc,
                                  tune, stream);
}

int spmm_csr_bf16_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                          const int *rowptr, const int *colind, const float *val,
                          const uint16_t *B, int ldb, float *C, int ldc,
                          const spmm_tune_t *tune, void *stream) {
  return spmm::run_planned<uint16_t>(plan, M, K, N, nnz, rowptr, colind, val, B, ldb, C,
                                     ldc, tune, stream);
}

}  // extern "C"
