// FBF SpMM over wide B2SR tiles (T = 16, 32, 64): C = A_bin * B with FP32 or
// BF16 B and FP32 C, planned and plan-free forms (the FBF contract of
// spmm.cu:877-956 at the tile sizes of csr2bsr_batch_bsrbmv.cu:654-1070).
#include "b2sr_common.hpp"

#include <algorithm>
#include <climits>

namespace spmm {

// ------------------------------------- block-row FBF, wide tiles (T >= 16)
// A tile is T words of T bits (32, 128 or 512 bytes) and power-law graphs
// leave about one bit per tile, so the words, not the gathers, are most of
// the bytes. One workgroup per block row (the plan's block-row order).
//
// Words: the block row's word array streams through LDS a round at a time.
// Every lane loads LPL 16-byte pieces (one coalesced 8 / 16 / 32 KB round of
// 256 / 128 / 64 tiles for T = 16 / 32 / 64); the next round is requested
// before the current one is decoded; the round is stored with 16 bytes of
// padding after each lane group's tiles so the column reads below hit
// distinct banks.
//
// Decode: subgroup g reads the words of its units -- RPS whole rows, or 1/P
// of one row's tiles when there are more subgroups than rows -- out of LDS,
// lane l taking TPL consecutive tiles. Per-unit bit counts are scanned over
// the subgroup's lanes (two 16-bit counts per register) and over the units
// (row-major), which places every set bit in one round list ordered by row,
// then column.
//
// Gather: the round list is split into one equal slice per subgroup, so a
// hub row of hundreds of entries is gathered by the whole workgroup rather
// than by its owner alone. A slice sums each run of one row (8 B rows in
// flight per lane); a run wholly inside the slice goes straight into that
// row's LDS sum; the runs at either end of each slice (the only ones a row
// can share with other slices) are added after a barrier, each row's by the
// subgroup holding its first such run, in list order. Lists longer than the
// LDS window run in windows. Results do not depend on timing. Every word is
// read from HBM once.
//
// The plan-free form, and the planned one at T = 16, run decode and gather
// in one kernel, round by round. The planned form at T = 32 and 64 splits
// them: b2sr_wide_decode_kernel streams the words and writes every round's
// list to the plan's list buffer, and b2sr_wide_gather_kernel reads the
// lists back. The word stream no longer waits on gathers, a block row's
// rounds no longer follow one another, and the gather kernel, without the
// stage and the decode, holds more workgroups per CU.
template <int T, int G> struct WideDecodeShape {
  static constexpr int kTileBytes = T * T / 8;
  static constexpr int kWordBytes = T / 8;
  static constexpr int kLPL = T == 16 ? 2 : (T == 32 ? 4 : 8);  // 16-byte loads a lane a round
  static constexpr int kRoundBytes = kB2Threads * 16 * kLPL;
  static constexpr int kNT = kRoundBytes / kTileBytes;      // tiles per round
  static constexpr int kS = kWave / G, kSW = kB2Waves * kS;   // subgroups per workgroup
  static constexpr int kRPS = T >= kSW ? T / kSW : 1;       // rows per subgroup
  static constexpr int kP = kSW > T ? kSW / T : 1;          // subgroups per row
  static constexpr int kU = kSW * kRPS;                     // units (row, part), row-major
  static constexpr int kKU = (kU + kWave - 1) / kWave;      // units per lane in their scan
  static constexpr int kPairs = (kRPS + 1) / 2;             // packed 16-bit count pairs
  static constexpr int kTPL = kNT / (kP * G);               // tiles per lane per round
  static constexpr int kGroupBytes = kTPL * kTileBytes;     // one lane's tiles, padded by 16
  static constexpr int kStage = kRoundBytes + kRoundBytes / kGroupBytes * 16;
  static constexpr int kCap = 1024;                         // list window (4 KB)
  static_assert(G <= 32, "wider subgroups run as 32 lanes over more column tiles");
  static_assert(kTPL >= 1 && kNT % (kP * G) == 0, "a round splits evenly over a row's lanes");
  static_assert(kNT <= kB2Threads && T <= 64, "block columns of a round: one lane each");
  static_assert(kNT == wide_round_tiles(T), "the plan's round table uses the same rounds");
};

template <int T, typename TB, int G> struct WideShape : WideDecodeShape<T, G> {
  static constexpr int kEPL = Vec16<TB>::kElems;
  static constexpr int kCW = G * kEPL;  // columns per workgroup
};

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

constexpr int kWideRowBits = 6;  // list entry = (B row << 6) | row within the block row

// Inclusive prefix sum over aligned groups of W lanes (W a power of two up
// to 64) with DPP row shifts and row broadcasts (VALU only, no LDS path).
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t x) {  // 0 where the source is off the row
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x110 + D, 0xf, 0xf, false));
}
template <int W>
__device__ __forceinline__ uint32_t group_scan(uint32_t x, int lane) {
  const int lg = lane & (W - 1);
  if constexpr (W > 1) { const uint32_t v = row_shr<1>(x); if (lg >= 1) x += v; }
  if constexpr (W > 2) { const uint32_t v = row_shr<2>(x); if (lg >= 2) x += v; }
  if constexpr (W > 4) { const uint32_t v = row_shr<4>(x); if (lg >= 4) x += v; }
  if constexpr (W > 8) { const uint32_t v = row_shr<8>(x); if (lg >= 8) x += v; }
  if constexpr (W >= 32) {  // each odd row adds the last lane of the row before it
    x += static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x142, 0xa, 0xf, false));
  }
  if constexpr (W >= 64) {  // rows 2 and 3 add lane 31
    x += static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x143, 0xc, 0xf, false));
  }
  return x;
}

// A thread's place in the decode: subgroup, first row and unit, and lane
// within the row's lanes.
template <int T, int G> struct WideLane {
  using D = WideDecodeShape<T, G>;
  int tid, lane, lane_sub, gsub, r0, unit0, lr;
  __device__ __forceinline__ WideLane()
      : tid(static_cast<int>(threadIdx.x)), lane(tid & (kWave - 1)), lane_sub(lane & (G - 1)),
        gsub((tid / kWave) * D::kS + lane / G), r0((gsub / D::kP) * D::kRPS),
        unit0(gsub * D::kRPS), lr((gsub % D::kP) * G + lane_sub) {}
  // this lane's words in the stage: rows r0 .. r0 + RPS - 1 of tiles
  // lr TPL .. lr TPL + TPL - 1
  __device__ __forceinline__ const unsigned char *mine(const unsigned char *stage) const {
    return stage + lr * (D::kGroupBytes + 16) + r0 * D::kWordBytes;
  }
};

// A block row's words, a round at a time: HBM -> registers (fetch, issued a
// round ahead) -> LDS stage and block columns (store).
template <int T, int G> struct WideWords {
  using D = WideDecodeShape<T, G>;
  const unsigned char *wb0;
  const int *bci;
  int tile0, ntiles;
  u32x4 pre[D::kLPL];
  int pbc = 0;
  __device__ __forceinline__ WideWords(const void *words, const int *bci_, int2 sp)
      : wb0(static_cast<const unsigned char *>(words) +
            static_cast<size_t>(sp.x) * D::kTileBytes),
        bci(bci_), tile0(sp.x), ntiles(sp.y - sp.x) {}
  __device__ __forceinline__ void fetch(int t0) {  // round starting at tile t0
    const int tid = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < D::kLPL; ++k) {
      const int off = (k * kB2Threads + tid) * 16;
      pre[k] = u32x4{0u, 0u, 0u, 0u};
      if (t0 + off / D::kTileBytes < ntiles)
        pre[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(
            wb0 + static_cast<size_t>(t0) * D::kTileBytes + off));
    }
    pbc = 0;
    if (tid < D::kNT && t0 + tid < ntiles) pbc = bci[tile0 + t0 + tid];
  }
  __device__ __forceinline__ void store(unsigned char *stage, int *bcs) const {
    const int tid = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < D::kLPL; ++k) {
      const int off = (k * kB2Threads + tid) * 16;
      *reinterpret_cast<u32x4 *>(stage + off + off / D::kGroupBytes * 16) = pre[k];
    }
    if (tid < D::kNT) bcs[tid] = pbc;
  }
};

template <int T, int G>
__device__ __forceinline__ typename Word<T>::type stage_word(const unsigned char *mine, int i,
                                                             int u) {
  using D = WideDecodeShape<T, G>;
  return *reinterpret_cast<const typename Word<T>::type *>(mine + u * D::kTileBytes +
                                                           i * D::kWordBytes);
}

// Count pass of a staged round: this lane's first list position in each of
// its rows (q0) and the round list's length (returned). One barrier.
template <int T, int G>
__device__ __forceinline__ int wide_positions(const unsigned char *mine, int *utot,
                                              const WideLane<T, G> &w,
                                              int (&q0)[WideDecodeShape<T, G>::kRPS]) {
  using D = WideDecodeShape<T, G>;
  constexpr int RPS = D::kRPS, TPL = D::kTPL, U = D::kU, KU = D::kKU, NP = D::kPairs;
  uint32_t pk[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) pk[a] = 0;
#pragma unroll
  for (int i = 0; i < RPS; ++i)
#pragma unroll
    for (int u = 0; u < TPL; ++u)
      pk[i / 2] += static_cast<uint32_t>(popcount(stage_word<T, G>(mine, i, u))) << (16 * (i & 1));
  // per-unit positions inside the subgroup (two counts a register: a unit
  // holds at most 32 lanes x 512 bits a round, below 2^16)
  uint32_t inc[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) inc[a] = group_scan<G>(pk[a], w.lane);
  if (w.lane_sub == G - 1) {
#pragma unroll
    for (int i = 0; i < RPS; ++i) utot[w.unit0 + i] = (inc[i / 2] >> (16 * (i & 1))) & 0xffffu;
  }
  __syncthreads();
  // unit offsets: every wave scans the U unit totals
  int uv[KU], ue[KU], ls = 0;
#pragma unroll
  for (int k = 0; k < KU; ++k) {
    const int idx = w.lane * KU + k;
    uv[k] = idx < U ? utot[idx] : 0;
    ue[k] = ls;
    ls += uv[k];
  }
  const int linc = static_cast<int>(group_scan<kWave>(static_cast<uint32_t>(ls), w.lane));
  const int L = __shfl(linc, kWave - 1, kWave);  // the round's entries
#pragma unroll
  for (int k = 0; k < KU; ++k) ue[k] += linc - ls;
#pragma unroll
  for (int i = 0; i < RPS; ++i) {
    const int u = w.unit0 + i;
    int off = 0;
#pragma unroll
    for (int k = 0; k < KU; ++k) {
      const int v = __shfl(ue[k], u / KU, kWave);
      if (u % KU == k) off = v;
    }
    q0[i] = off + static_cast<int>(((inc[i / 2] - pk[i / 2]) >> (16 * (i & 1))) & 0xffffu);
  }
  return L;
}

// Writes this lane's entries of list positions [wlo, wlo + cap) as
// put(ok, position - wlo, entry); ok is false where the entry falls outside
// (put may then write a spare slot or nothing). A lane's entries take its
// positions q0 .. q0 + count - 1 in a fixed order: the top bit of every
// word first (branch-free), then any further bits. Words of one tile row
// seldom hold more than one bit.
template <int T, int G, typename Put>
__device__ __forceinline__ void wide_emit(const unsigned char *mine, const int *bcs,
                                          const WideLane<T, G> &w,
                                          const int (&q0)[WideDecodeShape<T, G>::kRPS], int wlo,
                                          int cap, Put put) {
  using D = WideDecodeShape<T, G>;
  constexpr int RPS = D::kRPS, TPL = D::kTPL;
#pragma unroll
  for (int i = 0; i < RPS; ++i) {
    int q = q0[i] - wlo;
    const int row = w.r0 + i;
    Bits<T> rest[TPL];
    int bcv[TPL];
    bool more = false;
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
      const Bits<T> x = stage_word<T, G>(mine, i, u);
      bcv[u] = bcs[w.lr * TPL + u];
      const int b = x ? top_bit(x) : 0;
      put(x != 0 && q >= 0 && q < cap, q, ((bcv[u] * T + T - 1 - b) << kWideRowBits) | row);
      q += x != 0;
      rest[u] = x & ~(static_cast<Bits<T>>(1) << b);
      more |= rest[u] != 0;
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < TPL; ++u) {
        Bits<T> x = rest[u];
        while (x) {
          const int b = top_bit(x);
          x ^= static_cast<Bits<T>>(1) << b;
          put(q >= 0 && q < cap, q, ((bcv[u] * T + T - 1 - b) << kWideRowBits) | row);
          ++q;
        }
      }
    }
  }
}

// Gather and boundary pass over one list window (n entries at `list`, in
// LDS); the caller has put a barrier between the list's writes and this.
// Ends after the boundary partials are added (one barrier inside).
template <int T, typename TB, int G, bool kBuf>
__device__ __forceinline__ void wide_gather(const int *list, int n, float *rowsum, float *bnd,
                                            int *bnd_row, const RowLoader<TB, kBuf> &ld,
                                            int gsub, int lane_sub) {
  using W = WideShape<T, TB, G>;
  constexpr int EPL = W::kEPL, CW = W::kCW, SW = W::kSW;
  // slice gsub of the window
  const int cs = (n + SW - 1) / SW;
  const int a0 = gsub * cs, a1 = min(n, a0 + cs);
  int cur_row = -1, hrow = -1, runs = 0;
  float cur[EPL], head[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) cur[q] = head[q] = 0.f;
  for (int k = a0; k < a1; k += 8) {
    typename Vec16<TB>::Raw raw[8];
    int ent[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      ent[u] = list[min(k + u, a1 - 1)];
      raw[u] = ld(ent[u] >> kWideRowBits);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (k + u < a1) {
        const int r = ent[u] & ((1 << kWideRowBits) - 1);
        if (r != cur_row) {  // a run ends (the same for every lane of the subgroup)
          if (cur_row >= 0) {
            if (runs == 0) {
              hrow = cur_row;
#pragma unroll
              for (int q = 0; q < EPL; ++q) head[q] = cur[q];
            } else {  // wholly inside this slice: no other slice holds this row
              float *rs = rowsum + cur_row * CW + lane_sub * EPL;
#pragma unroll
              for (int q = 0; q < EPL; ++q) rs[q] += cur[q];
            }
            ++runs;
          }
          cur_row = r;
#pragma unroll
          for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
        }
        float bv[EPL];
        Vec16<TB>::widen(raw[u], bv);
#pragma unroll
        for (int q = 0; q < EPL; ++q) cur[q] += bv[q];
      }
    }
  }
  int trow = -1;
  if (cur_row >= 0) {
    if (runs == 0) {
      hrow = cur_row;
#pragma unroll
      for (int q = 0; q < EPL; ++q) head[q] = cur[q];
    } else {
      trow = cur_row;
    }
  }
#pragma unroll
  for (int q = 0; q < EPL; ++q) {
    bnd[(2 * gsub) * CW + lane_sub * EPL + q] = head[q];
    bnd[(2 * gsub + 1) * CW + lane_sub * EPL + q] = cur[q];
  }
  if (lane_sub == 0) {
    bnd_row[2 * gsub] = hrow;
    bnd_row[2 * gsub + 1] = trow;
  }
  __syncthreads();
  // boundary partials: a row's partials sit in consecutive slots (rows
  // ascend along the list; a slice's tail slot is -1 when the slice holds
  // one run, its head slot -1 when it is empty, and empty slices come
  // last). The subgroup whose slot is the row's first adds them all, in
  // slot order, so every row has one writer.
#pragma unroll
  for (int which = 0; which < 2; ++which) {
    const int sl = 2 * gsub + which;
    const int r = bnd_row[sl];
    if (r < 0) continue;
    if (which == 0 && gsub > 0) {  // the previous valid slot: slice gsub - 1's tail, else head
      const int pt = bnd_row[sl - 1];
      if ((pt >= 0 ? pt : bnd_row[sl - 2]) == r) continue;
    }
    float acc[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = bnd[sl * CW + lane_sub * EPL + q];
    bool open = true;  // eight slots' reads in flight per step (hub rows span many)
    for (int s0 = sl + 1; open && s0 < 2 * SW; s0 += 8) {
      int rr[8];
      float vv[8][EPL];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int s2 = min(s0 + u, 2 * SW - 1);
        rr[u] = s0 + u < 2 * SW ? bnd_row[s2] : -2;
#pragma unroll
        for (int q = 0; q < EPL; ++q) vv[u][q] = bnd[s2 * CW + lane_sub * EPL + q];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (!open || (rr[u] == -1 && ((s0 + u) & 1))) continue;  // a one-run slice
        if (rr[u] != r) {  // another row, the empty slices or the end
          open = false;
          continue;
        }
#pragma unroll
        for (int q = 0; q < EPL; ++q) acc[q] += vv[u][q];
      }
    }
    float *rs = rowsum + r * CW + lane_sub * EPL;
#pragma unroll
    for (int q = 0; q < EPL; ++q) rs[q] += acc[q];
  }
}

template <int T, typename TB, int G>
__device__ __forceinline__ void wide_zero_rows(float *rowsum) {
  constexpr int CW = WideShape<T, TB, G>::kCW;
  for (int x = threadIdx.x; x < T * CW / 4; x += kB2Threads)  // the final store loop's mapping
    reinterpret_cast<f32x4v *>(rowsum)[x] = f32x4v{0.f, 0.f, 0.f, 0.f};
}

template <int T, typename TB, int G>
__device__ __forceinline__ void wide_store_rows(int br, int M, const float *rowsum, float *C,
                                                int ldc) {
  constexpr int CW = WideShape<T, TB, G>::kCW;
  for (int x = threadIdx.x; x < T * CW / 4; x += kB2Threads) {
    const int j = x / (CW / 4), c4 = x % (CW / 4);
    const int row = br * T + j;
    if (row < M)
      *reinterpret_cast<f32x4v *>(C + static_cast<size_t>(row) * ldc + blockIdx.y * CW + 4 * c4) =
          reinterpret_cast<const f32x4v *>(rowsum)[x];
  }
}

// One block row, decode and gather in one workgroup (the plan-free form).
template <int T, typename TB, int G, bool kBuf>
__device__ __forceinline__ void wide_brow(int br, int2 sp, int M, const int *__restrict__ bci,
                                          const typename Word<T>::type *__restrict__ words,
                                          const TB *__restrict__ B, int ldb, int b_bytes,
                                          float *__restrict__ C, int ldc) {
  using W = WideShape<T, TB, G>;
  constexpr int EPL = W::kEPL, CW = W::kCW, RPS = W::kRPS, NT = W::kNT, SW = W::kSW, U = W::kU;
  constexpr int CAP = W::kCap;
  __shared__ __align__(16) unsigned char stage[W::kStage];
  __shared__ __align__(16) int bcs[NT];
  __shared__ __align__(16) int list[CAP + kB2Threads];  // + one spare slot per thread
  __shared__ int utot[U];
  __shared__ __align__(16) float rowsum[T * CW];
  __shared__ __align__(16) float bnd[2 * SW * CW];  // slice g: head partial 2g, tail 2g + 1
  __shared__ int bnd_row[2 * SW];                    // their rows, -1 for none
  const WideLane<T, G> w;
  const RowLoader<TB, kBuf> ld(B, ldb, blockIdx.y * CW + w.lane_sub * EPL, b_bytes);
  wide_zero_rows<T, TB, G>(rowsum);
  WideWords<T, G> wd(words, bci, sp);
  const int ntiles = sp.y - sp.x;
  wd.fetch(0);
  for (int t0 = 0; t0 < ntiles; t0 += NT) {
    __syncthreads();  // the previous round is consumed
    wd.store(stage, bcs);
    if (t0 + NT < ntiles) wd.fetch(t0 + NT);
    __syncthreads();
    // (the words are read again from LDS by the decode: the stage holds
    // until the next round)
    const unsigned char *mine = w.mine(stage);
    int q0[RPS];
    const int L = wide_positions<T, G>(mine, utot, w, q0);
    for (int wlo = 0; wlo < L; wlo += CAP) {
      // decode: entries [wlo, wlo + CAP) of the round list (a word without
      // bits writes a spare slot)
      wide_emit<T, G>(mine, bcs, w, q0, wlo, CAP,
                      [&](bool ok, int q, int e) { list[ok ? q : CAP + w.tid] = e; });
      __syncthreads();
      wide_gather<T, TB, G, kBuf>(list, min(CAP, L - wlo), rowsum, bnd, bnd_row, ld, w.gsub,
                                  w.lane_sub);
    }
  }
  __syncthreads();
  wide_store_rows<T, TB, G>(br, M, rowsum, C, ldc);
}

// ---------------------------------------------- planned form: two kernels
// Round k of a block row is its tiles k NT .. k NT + nt - 1; the plan numbers
// all rounds (rounds[g] = their tile range, rbase[x] = the first round of
// plan unit x's block row). The decode kernel writes, for round g starting at tile t0:
//   - rc[g T + r]: the set bits of row r in the round (the list is row-major,
//     so these also place each row's run in it);
//   - lists[kWideSlots t0 + q], q < L: the round list, when its length L is
//     at most kWideSlots nt. A longer list (tiles denser than that) is not
//     written; the gather kernel then walks that round's words itself.

// Decode: one workgroup per round, so a block row's rounds run side by side.
template <int T, int G>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_decode_kernel(const int2 *__restrict__ rounds, const int *__restrict__ bci,
                        const void *__restrict__ words, int *__restrict__ lists,
                        int *__restrict__ rc) {
  using D = WideDecodeShape<T, G>;
  __shared__ __align__(16) unsigned char stage[D::kStage];
  __shared__ __align__(16) int bcs[D::kNT];
  __shared__ int utot[D::kU];
  const int2 tr = rounds[blockIdx.x];
  const WideLane<T, G> w;
  WideWords<T, G> wd(words, bci, tr);
  wd.fetch(0);
  wd.store(stage, bcs);
  __syncthreads();
  const unsigned char *mine = w.mine(stage);
  int q0[D::kRPS];
  const int L = wide_positions<T, G>(mine, utot, w, q0);  // utot is complete after it
  if (w.tid < T) {  // row r's units are r P .. r P + P - 1
    int c = 0;
#pragma unroll
    for (int p = 0; p < D::kP; ++p) c += utot[w.tid * D::kP + p];
    rc[static_cast<size_t>(blockIdx.x) * T + w.tid] = c;
  }
  if (L <= kWideSlots * (tr.y - tr.x)) {
    int *out = lists + static_cast<size_t>(kWideSlots) * tr.x;
    wide_emit<T, G>(mine, bcs, w, q0, 0, L, [&](bool ok, int q, int e) {
      if (ok) out[q] = e;
    });
  }
}

// A round whose list was not written: row i of the block row is summed by
// subgroup i mod SW, tile by tile in order, top bit first, 8 B rows in
// flight per lane.
template <int T, typename TB, int G, bool kBuf>
__device__ __forceinline__ void wide_walk_round(const typename Word<T>::type *__restrict__ words,
                                                const int *__restrict__ bci, int tile0, int nt,
                                                float *rowsum, const RowLoader<TB, kBuf> &ld,
                                                int gsub, int lane_sub) {
  using W = WideShape<T, TB, G>;
  constexpr int EPL = W::kEPL, CW = W::kCW;
  for (int i = gsub; i < T; i += W::kSW) {
    float cur[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) cur[q] = 0.f;
    for (int u = 0; u < nt; ++u) {
      Bits<T> x = words[static_cast<size_t>(tile0 + u) * T + i];
      const int bc = bci[tile0 + u];
      while (x) {
        typename Vec16<TB>::Raw raw[8];
        int m = 0;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          if (x) {
            const int b = top_bit(x);
            x ^= static_cast<Bits<T>>(1) << b;
            raw[v] = ld(bc * T + T - 1 - b);
            m = v + 1;
          }
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          if (v < m) {
            float bv[EPL];
            Vec16<TB>::widen(raw[v], bv);
#pragma unroll
            for (int q = 0; q < EPL; ++q) cur[q] += bv[q];
          }
        }
      }
    }
    float *rs = rowsum + i * CW + lane_sub * EPL;
#pragma unroll
    for (int q = 0; q < EPL; ++q) rs[q] += cur[q];
  }
}

// Gather: one workgroup per plan unit (block row) and column tile. The block
// row's rounds are taken in phases: consecutive rounds whose lists fit the
// LDS window together (wide_win(T) entries). A phase's lists are copied into
// one row-major list -- row r's entries of round j follow those of rounds
// before j -- at places computed from the row counts, then gathered in one
// pass: no barrier per round, and a hub block row's rounds are spread over
// all subgroups at once. A round whose list was not written is walked alone.
// LDS list window, entries: eight (T = 32) or four (T = 64) rounds of one
// entry per tile. At T = 64 a 1024-entry window made the gather kernel 20%
// slower; at T = 32 a 2048-entry one with 32 staged rounds (fewer
// workgroups a CU) 6% slower.
constexpr int wide_win(int T) { return T == 64 ? 2048 : 1024; }
constexpr int kWideBatch = 16;   // rounds whose row counts are staged at once

template <int T, typename TB, int G, bool kBuf>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_gather_kernel(const int *__restrict__ brows, const int2 *__restrict__ span,
                        const int *__restrict__ rbase, int M, const int *__restrict__ bci,
                        const typename Word<T>::type *__restrict__ words,
                        const int *__restrict__ lists, const int *__restrict__ rc,
                        const TB *__restrict__ B, int ldb, int b_bytes, float *__restrict__ C,
                        int ldc) {
  using W = WideShape<T, TB, G>;
  constexpr int EPL = W::kEPL, CW = W::kCW, NT = W::kNT, SW = W::kSW;
  constexpr int WIN = wide_win(T), KB = kWideBatch;
  static_assert(kWideSlots * NT <= WIN, "every written round list fits the window alone");
  __shared__ __align__(16) int list[WIN];
  __shared__ __align__(16) int cnt[KB * T];  // row counts of the staged rounds
  __shared__ int place[KB * T];              // row r of round j: list position - in-round one
  __shared__ int rlen[KB];                   // list lengths of the staged rounds
  __shared__ int cpre[KB + 1];               // 16-byte pieces of the phase's lists, prefix
  __shared__ __align__(16) float rowsum[T * CW];
  __shared__ __align__(16) float bnd[2 * SW * CW];
  __shared__ int bnd_row[2 * SW];
  const int br = brows[blockIdx.x];
  const int2 sp = span[blockIdx.x];
  const int ntiles = sp.y - sp.x;
  const int nr = (ntiles + NT - 1) / NT;
  const int g0 = rbase[blockIdx.x];  // plan unit x's first round
  const WideLane<T, G> w;
  const int wave = w.tid / kWave;
  const RowLoader<TB, kBuf> ld(B, ldb, blockIdx.y * CW + w.lane_sub * EPL, b_bytes);
  wide_zero_rows<T, TB, G>(rowsum);
  auto round_tiles = [&](int k) { return min(NT, ntiles - k * NT); };
  for (int kb = 0; kb < nr; kb += KB) {
    const int nb = min(KB, nr - kb);
    __syncthreads();  // the previous batch's tables are consumed
    for (int i = w.tid; i < nb * T / 4; i += kB2Threads)
      reinterpret_cast<int4 *>(cnt)[i] =
          reinterpret_cast<const int4 *>(rc + static_cast<size_t>(g0 + kb) * T)[i];
    __syncthreads();
    // per round: list length, and each row's first position in the round list
    for (int j = wave; j < nb; j += kB2Waves) {
      const int c = w.lane < T ? cnt[j * T + w.lane] : 0;
      const int incl = static_cast<int>(group_scan<kWave>(static_cast<uint32_t>(c), w.lane));
      if (w.lane < T) place[j * T + w.lane] = incl - c;
      if (w.lane == kWave - 1) rlen[j] = incl;
    }
    __syncthreads();
    int k = 0;
    while (k < nb) {
      if (rlen[k] > kWideSlots * round_tiles(kb + k)) {  // list not written: walk the words
        __syncthreads();  // a phase before it has added its boundary partials
        wide_walk_round<T, TB, G, kBuf>(words, bci, sp.x + (kb + k) * NT, round_tiles(kb + k),
                                        rowsum, ld, w.gsub, w.lane_sub);
        ++k;
        continue;
      }
      // the phase: rounds k .. k + nk - 1 (the same on every thread)
      int nk = 0, tot = 0;
      while (k + nk < nb) {
        const int L = rlen[k + nk];
        if (L > kWideSlots * round_tiles(kb + k + nk) || tot + L > WIN) break;
        tot += L;
        ++nk;
      }
      if (wave == 0) {  // lane r: row r's place in each round of the phase
        int run = 0;
        for (int j = k; j < k + nk; ++j) {
          if (w.lane < T) place[j * T + w.lane] = run - place[j * T + w.lane];
          run += w.lane < T ? cnt[j * T + w.lane] : 0;
        }
        const int incl = static_cast<int>(group_scan<kWave>(static_cast<uint32_t>(run), w.lane));
        for (int j = k; j < k + nk; ++j)
          if (w.lane < T) place[j * T + w.lane] += incl - run;
      } else if (wave == 1) {  // lane j: 16-byte pieces of round k + j's list
        const int pc = w.lane < nk ? (rlen[k + w.lane] + 3) / 4 : 0;
        const int incl = static_cast<int>(group_scan<kWave>(static_cast<uint32_t>(pc), w.lane));
        if (w.lane < nk) cpre[w.lane + 1] = incl;
        if (w.lane == 0) cpre[0] = 0;
      }
      __syncthreads();
      for (int c = w.tid; c < cpre[nk]; c += kB2Threads) {
        int j = 0;
        while (cpre[j + 1] <= c) ++j;
        const int q4 = 4 * (c - cpre[j]);
        const int4 v = *reinterpret_cast<const int4 *>(
            lists + static_cast<size_t>(kWideSlots) * (sp.x + (kb + k + j) * NT) + q4);
        const int L = rlen[k + j];
        const int e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (q4 + u < L)
            list[place[(k + j) * T + (e[u] & ((1 << kWideRowBits) - 1))] + q4 + u] = e[u];
      }
      __syncthreads();
      wide_gather<T, TB, G, kBuf>(list, tot, rowsum, bnd, bnd_row, ld, w.gsub, w.lane_sub);
      k += nk;
    }
  }
  __syncthreads();
  wide_store_rows<T, TB, G>(br, M, rowsum, C, ldc);
}

// Planned, decode and gather in one workgroup per plan unit (T = 16).
template <int T, typename TB, int G, bool kBuf>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_planned_kernel(const int *__restrict__ brows, const int2 *__restrict__ span, int M,
                         const int *__restrict__ bci,
                         const typename Word<T>::type *__restrict__ words,
                         const TB *__restrict__ B, int ldb, int b_bytes, float *__restrict__ C,
                         int ldc) {
  wide_brow<T, TB, G, kBuf>(brows[blockIdx.x], span[blockIdx.x], M, bci, words, B, ldb, b_bytes,
                            C, ldc);
}

// Plan-free with a handle: workgroup x takes the x-th block row of the
// device-side sort by tile count (heaviest first), as the planned kernel
// takes the plan's order.
template <int T, typename TB, int G, bool kBuf>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_sorted_kernel(const int *__restrict__ rows, const int *__restrict__ brp, int M,
                        const int *__restrict__ bci,
                        const typename Word<T>::type *__restrict__ words,
                        const TB *__restrict__ B, int ldb, int b_bytes, float *__restrict__ C,
                        int ldc) {
  const int br = rows[blockIdx.x];
  wide_brow<T, TB, G, kBuf>(br, make_int2(brp[br], brp[br + 1]), M, bci, words, B, ldb, b_bytes,
                            C, ldc);
}

// Plan-free: the block rows of at least heavy_tiles tiles, listed by
// class_list_kernel (state + kStateHead + MB, count state[2]), go first to
// the E workgroups resident from the start (x, x + E, ...); workgroup E + k
// then takes block row k unless it is listed. The last workgroup to finish
// zeroes the list count.
template <int T, typename TB, int G, bool kBuf>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_kernel(const int *__restrict__ brp, int E, int heavy_tiles, int *__restrict__ state,
                 int M, const int *__restrict__ bci,
                 const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                 int ldb, int b_bytes, float *__restrict__ C, int ldc) {
  const int x = blockIdx.x;
  const int MB = (M + T - 1) / T;
  if (x < E) {
    const int n = state[2];
    for (int j = x; j < n; j += E) {
      const int br = state[kStateHead + MB + j];
      wide_brow<T, TB, G, kBuf>(br, make_int2(brp[br], brp[br + 1]), M, bci, words, B, ldb,
                                b_bytes, C, ldc);
    }
  } else {
    const int br = x - E;
    const int2 sp = make_int2(brp[br], brp[br + 1]);
    if (sp.y - sp.x < heavy_tiles)
      wide_brow<T, TB, G, kBuf>(br, sp, M, bci, words, B, ldb, b_bytes, C, ldc);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // every workgroup has read the count and list before it counts itself
    if (atomicAdd(&state[1], 1) == static_cast<int>(gridDim.x * gridDim.y) - 1) {
      atomicExch(&state[2], 0);
      atomicExch(&state[1], 0);
    }
  }
}

// p == nullptr: the plan-free form, over the block rows sorted by tile
// count when `sorted` is given, else with the block rows of at least
// wide_heavy(T) tiles listed first on `state`; else the planned form.
template <int T, typename TB, int G>
static void launch_wide(const spmm_csr_plan_s *p, const int *brp, int *state, const int *sorted,
                        int cus, int K, const int *bci, const void *val, const TB *B, int ldb,
                        float *C, int ldc, int tiles, int M, hipStream_t s) {
  const int MB = (M + T - 1) / T;
  const auto *w = static_cast<const typename Word<T>::type *>(val);
  const int b_bytes = buffer_span(K, ldb, sizeof(TB));
  if (!p && sorted) {
    auto *const kern = b_bytes ? &b2sr_wide_sorted_kernel<T, TB, G, true>
                               : &b2sr_wide_sorted_kernel<T, TB, G, false>;
    hipLaunchKernelGGL(kern, dim3(MB, tiles), dim3(kB2Threads), 0, s, sorted, brp, M, bci, w, B,
                       ldb, b_bytes, C, ldc);
    return;
  }
  if constexpr (T == 16) {  // 16-row block rows: too short for the split to pay
    if (p) {
      auto *const kern = b_bytes ? &b2sr_wide_planned_kernel<T, TB, G, true>
                                 : &b2sr_wide_planned_kernel<T, TB, G, false>;
      hipLaunchKernelGGL(kern, dim3(MB, tiles), dim3(kB2Threads), 0, s, p->d_rows, p->d_span, M,
                         bci, w, B, ldb, b_bytes, C, ldc);
      return;
    }
  } else if (p) {
    if (p->wide_rounds > 0)
      hipLaunchKernelGGL((b2sr_wide_decode_kernel<T, G>), dim3(p->wide_rounds), dim3(kB2Threads),
                         0, s, p->d_wide_round, bci, val, p->d_wide, p->d_wide_rc);
    auto *const kern = b_bytes ? &b2sr_wide_gather_kernel<T, TB, G, true>
                               : &b2sr_wide_gather_kernel<T, TB, G, false>;
    hipLaunchKernelGGL(kern, dim3(MB, tiles), dim3(kB2Threads), 0, s, p->d_rows, p->d_span,
                       p->d_wide_rbase, M, bci, w, p->d_wide, p->d_wide_rc, B, ldb, b_bytes, C,
                       ldc);
    return;
  }
  auto *const kern =
      b_bytes ? &b2sr_wide_kernel<T, TB, G, true> : &b2sr_wide_kernel<T, TB, G, false>;
  hipLaunchKernelGGL(class_list_kernel<kB2Threads>, dim3(ceil_div(MB, kB2Threads)),
                     dim3(kB2Threads), 0, s, MB, brp, wide_heavy(T), INT_MAX, state);
  static std::atomic<int> per_cu[2];
  const int E = std::min(
      resident_wgs(reinterpret_cast<const void *>(kern), cus, &per_cu[b_bytes ? 1 : 0]), MB);
  hipLaunchKernelGGL(kern, dim3(E + MB, tiles), dim3(kB2Threads), 0, s, brp, E, wide_heavy(T),
                     state, M, bci, w, B, ldb, b_bytes, C, ldc);
}

// G = 64 runs as G = 32 over twice the column tiles (eight rows per subgroup
// at most).
template <int T, typename TB>
static void dispatch_wide(int G, int tiles, const spmm_csr_plan_s *p, const int *brp, int *state,
                          const int *sorted, int cus, int M, int K, const int *bci,
                          const void *val, const TB *B, int ldb, float *C, int ldc,
                          hipStream_t s) {
#define SPMM_WD(gg, tl) \
  launch_wide<T, TB, gg>(p, brp, state, sorted, cus, K, bci, val, B, ldb, C, ldc, tl, M, s)
  switch (G) {
    case 1: SPMM_WD(1, tiles); break;
    case 2: SPMM_WD(2, tiles); break;
    case 4: SPMM_WD(4, tiles); break;
    case 8: SPMM_WD(8, tiles); break;
    case 16: SPMM_WD(16, tiles); break;
    default: SPMM_WD(32, tiles * (G / 32)); break;
  }
#undef SPMM_WD
}

template <typename TB>
void b2sr_wide_launch(int T, int G, int tiles, const spmm_csr_plan_s *p, const int *brp,
                      int *state, const int *sorted, int cus, int M, int K, const int *bci,
                      const void *val, const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
#define SPMM_WT(t) \
  dispatch_wide<t, TB>(G, tiles, p, brp, state, sorted, cus, M, K, bci, val, B, ldb, C, ldc, s)
  switch (T) {
    case 16: SPMM_WT(16); break;
    case 32: SPMM_WT(32); break;
    default: SPMM_WT(64); break;
  }
#undef SPMM_WT
}

template void b2sr_wide_launch<float>(int, int, int, const spmm_csr_plan_s *, const int *, int *,
                                      const int *, int, int, int, const int *, const void *,
                                      const float *, int, float *, int, hipStream_t);
template void b2sr_wide_launch<uint16_t>(int, int, int, const spmm_csr_plan_s *, const int *,
                                         int *, const int *, int, int, int, const int *,
                                         const void *, const uint16_t *, int, float *, int,
                                         hipStream_t);

}  // namespace spmm
