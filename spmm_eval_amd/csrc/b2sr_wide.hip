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
template <int T, typename TB, int G> struct WideShape {
  static constexpr int kEPL = Vec16<TB>::kElems;
  static constexpr int kCW = G * kEPL;                      // columns per workgroup
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

template <int T, typename TB, int G, bool kBuf>
__device__ __forceinline__ void wide_brow(int br, int2 sp, int M, const int *__restrict__ bci,
                                          const typename Word<T>::type *__restrict__ words,
                                          const TB *__restrict__ B, int ldb, int b_bytes,
                                          float *__restrict__ C, int ldc) {
  using W = WideShape<T, TB, G>;
  using WordT = typename Word<T>::type;
  constexpr int EPL = W::kEPL, CW = W::kCW, RPS = W::kRPS, TPL = W::kTPL, LPL = W::kLPL;
  constexpr int NT = W::kNT, SW = W::kSW, U = W::kU, KU = W::kKU, NP = W::kPairs;
  constexpr int CAP = W::kCap;
  __shared__ __align__(16) unsigned char stage[W::kStage];
  __shared__ __align__(16) int bcs[NT];
  __shared__ __align__(16) int list[CAP + kB2Threads];  // + one spare slot per thread
  __shared__ int utot[U];
  __shared__ __align__(16) float rowsum[T * CW];
  __shared__ __align__(16) float bnd[2 * SW * CW];  // slice g: head partial 2g, tail 2g + 1
  __shared__ int bnd_row[2 * SW];                    // their rows, -1 for none
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int lane_sub = lane & (G - 1), sub = lane / G, gsub = wave * W::kS + sub;
  const int part = gsub % W::kP;
  const int r0 = (gsub / W::kP) * RPS;  // this subgroup's first row
  const int unit0 = gsub * RPS;         // and first unit
  const int lr = part * G + lane_sub;   // lane within the row's lanes
  const RowLoader<TB, kBuf> ld(B, ldb, blockIdx.y * CW + lane_sub * EPL, b_bytes);
  for (int x = tid; x < T * CW / 4; x += kB2Threads)  // the final store loop's mapping
    reinterpret_cast<f32x4v *>(rowsum)[x] = f32x4v{0.f, 0.f, 0.f, 0.f};

  const int ntiles = sp.y - sp.x;
  const unsigned char *wb0 =
      reinterpret_cast<const unsigned char *>(words) + static_cast<size_t>(sp.x) * W::kTileBytes;
  u32x4 pre[LPL];
  int pbc = 0;
  auto fetch = [&](int t0) {  // round starting at tile t0 -> registers
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int off = (k * kB2Threads + tid) * 16;
      pre[k] = u32x4{0u, 0u, 0u, 0u};
      if (t0 + off / W::kTileBytes < ntiles)
        pre[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(
            wb0 + static_cast<size_t>(t0) * W::kTileBytes + off));
    }
    pbc = 0;
    if (tid < NT && t0 + tid < ntiles) pbc = bci[sp.x + t0 + tid];
  };
  fetch(0);
  for (int t0 = 0; t0 < ntiles; t0 += NT) {
    __syncthreads();  // the previous round is consumed
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int off = (k * kB2Threads + tid) * 16;
      *reinterpret_cast<u32x4 *>(stage + off + off / W::kGroupBytes * 16) = pre[k];
    }
    if (tid < NT) bcs[tid] = pbc;
    if (t0 + NT < ntiles) fetch(t0 + NT);
    __syncthreads();
    // this lane's words: rows r0 .. r0 + RPS - 1 of tiles lr TPL .. lr TPL + TPL - 1
    // (read again from LDS by the decode: the stage holds until the next round)
    const unsigned char *mine = stage + lr * (W::kGroupBytes + 16) + r0 * W::kWordBytes;
    auto word = [&](int i, int u) {
      return *reinterpret_cast<const WordT *>(mine + u * W::kTileBytes + i * W::kWordBytes);
    };
    uint32_t pk[NP];
#pragma unroll
    for (int a = 0; a < NP; ++a) pk[a] = 0;
#pragma unroll
    for (int i = 0; i < RPS; ++i)
#pragma unroll
      for (int u = 0; u < TPL; ++u)
        pk[i / 2] += static_cast<uint32_t>(popcount(word(i, u))) << (16 * (i & 1));
    // per-unit positions inside the subgroup (two counts a register: a unit
    // holds at most 32 lanes x 512 bits a round, below 2^16)
    uint32_t inc[NP];
#pragma unroll
    for (int a = 0; a < NP; ++a) inc[a] = group_scan<G>(pk[a], lane);
    if (lane_sub == G - 1) {
#pragma unroll
      for (int i = 0; i < RPS; ++i) utot[unit0 + i] = (inc[i / 2] >> (16 * (i & 1))) & 0xffffu;
    }
    __syncthreads();
    // unit offsets: every wave scans the U unit totals
    int uv[KU], ue[KU], ls = 0;
#pragma unroll
    for (int k = 0; k < KU; ++k) {
      const int idx = lane * KU + k;
      uv[k] = idx < U ? utot[idx] : 0;
      ue[k] = ls;
      ls += uv[k];
    }
    const int linc = static_cast<int>(group_scan<kWave>(static_cast<uint32_t>(ls), lane));
    const int L = __shfl(linc, kWave - 1, kWave);  // the round's entries
#pragma unroll
    for (int k = 0; k < KU; ++k) ue[k] += linc - ls;
    int q0[RPS];
#pragma unroll
    for (int i = 0; i < RPS; ++i) {
      const int u = unit0 + i;
      int off = 0;
#pragma unroll
      for (int k = 0; k < KU; ++k) {
        const int v = __shfl(ue[k], u / KU, kWave);
        if (u % KU == k) off = v;
      }
      q0[i] = off + static_cast<int>(((inc[i / 2] - pk[i / 2]) >> (16 * (i & 1))) & 0xffffu);
    }
    for (int wlo = 0; wlo < L; wlo += CAP) {
      // decode: entries [wlo, wlo + CAP) of the round list. A lane's entries
      // take its positions q0 .. q0 + count - 1 in a fixed order: the top
      // bit of every word first (branch-free; a word without bits writes a
      // spare slot), then any further bits. Words of one tile row seldom
      // hold more than one bit.
#pragma unroll
      for (int i = 0; i < RPS; ++i) {
        int q = q0[i] - wlo;
        const int row = r0 + i;
        Bits<T> rest[TPL];
        int bcv[TPL];
        bool more = false;
#pragma unroll
        for (int u = 0; u < TPL; ++u) {
          const Bits<T> x = word(i, u);
          bcv[u] = bcs[lr * TPL + u];
          const int b = x ? top_bit(x) : 0;
          const bool ok = x != 0 && q >= 0 && q < CAP;
          list[ok ? q : CAP + tid] = ((bcv[u] * T + T - 1 - b) << kWideRowBits) | row;
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
              if (q >= 0 && q < CAP) list[q] = ((bcv[u] * T + T - 1 - b) << kWideRowBits) | row;
              ++q;
            }
          }
        }
      }
      __syncthreads();
      // gather: slice gsub of the window
      const int n = min(CAP, L - wlo);
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
  }
  __syncthreads();
  for (int x = tid; x < T * CW / 4; x += kB2Threads) {
    const int j = x / (CW / 4), c4 = x % (CW / 4);
    const int row = br * T + j;
    if (row < M)
      *reinterpret_cast<f32x4v *>(C + static_cast<size_t>(row) * ldc + blockIdx.y * CW + 4 * c4) =
          reinterpret_cast<const f32x4v *>(rowsum)[x];
  }
}

// Planned (kPlanFree false): workgroup x takes the plan's block row x.
// Plan-free: the block rows of at least heavy_tiles tiles, listed by
// class_list_kernel (state + kStateHead + MB, count state[2]), go first to
// the E workgroups resident from the start (x, x + E, ...); workgroup E + k
// then takes block row k unless it is listed. The last workgroup to finish
// zeroes the list count.
template <int T, typename TB, int G, bool kBuf, bool kPlanFree>
__global__ void __launch_bounds__(kB2Threads)
b2sr_wide_kernel(const int *__restrict__ brows, const int2 *__restrict__ span,
                 const int *__restrict__ brp, int E, int heavy_tiles,
                 int *__restrict__ state, int M, const int *__restrict__ bci,
                 const typename Word<T>::type *__restrict__ words, const TB *__restrict__ B,
                 int ldb, int b_bytes, float *__restrict__ C, int ldc) {
  const int x = blockIdx.x;
  if constexpr (!kPlanFree) {  // its own instantiation: no plan-free registers
    wide_brow<T, TB, G, kBuf>(brows[x], span[x], M, bci, words, B, ldb, b_bytes, C, ldc);
    return;
  }
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

// p == nullptr: the plan-free form (block rows of at least wide_heavy(T)
// tiles listed first, on `state`).
template <int T, typename TB, int G>
static void launch_wide(const spmm_csr_plan_s *p, const int *brp, int *state, int cus, int K,
                        const int *bci, const void *val, const TB *B, int ldb, float *C,
                        int ldc, int tiles, int M, hipStream_t s) {
  const int MB = (M + T - 1) / T;
  const auto *w = static_cast<const typename Word<T>::type *>(val);
  const int b_bytes = buffer_span(K, ldb, sizeof(TB));
  auto *const kern =
      p ? (b_bytes ? &b2sr_wide_kernel<T, TB, G, true, false>
                   : &b2sr_wide_kernel<T, TB, G, false, false>)
        : (b_bytes ? &b2sr_wide_kernel<T, TB, G, true, true>
                   : &b2sr_wide_kernel<T, TB, G, false, true>);
  int E = 0;
  if (!p) {
    hipLaunchKernelGGL(class_list_kernel<kB2Threads>, dim3(ceil_div(MB, kB2Threads)), dim3(kB2Threads), 0,
                       s, MB, brp, wide_heavy(T), INT_MAX, state);
    static std::atomic<int> per_cu[2];
    E = std::min(resident_wgs(reinterpret_cast<const void *>(kern), cus,
                              &per_cu[b_bytes ? 1 : 0]),
                 MB);
  }
  hipLaunchKernelGGL(kern, dim3(E + MB, tiles), dim3(kB2Threads), 0, s, p ? p->d_rows : nullptr,
                     p ? p->d_span : nullptr, brp, E, wide_heavy(T), state, M, bci, w, B, ldb,
                     b_bytes, C, ldc);
}

// G = 64 runs as G = 32 over twice the column tiles (eight rows per subgroup
// at most).
template <int T, typename TB>
static void dispatch_wide(int G, int tiles, const spmm_csr_plan_s *p, const int *brp, int *state,
                          int cus, int M, int K, const int *bci, const void *val, const TB *B,
                          int ldb, float *C, int ldc, hipStream_t s) {
#define SPMM_WD(gg, tl) \
  launch_wide<T, TB, gg>(p, brp, state, cus, K, bci, val, B, ldb, C, ldc, tl, M, s)
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
                      int *state, int cus, int M, int K, const int *bci, const void *val,
                      const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  switch (T) {
    case 16: dispatch_wide<16, TB>(G, tiles, p, brp, state, cus, M, K, bci, val, B, ldb, C, ldc, s); break;
    case 32: dispatch_wide<32, TB>(G, tiles, p, brp, state, cus, M, K, bci, val, B, ldb, C, ldc, s); break;
    default: dispatch_wide<64, TB>(G, tiles, p, brp, state, cus, M, K, bci, val, B, ldb, C, ldc, s); break;
  }
}

template void b2sr_wide_launch<float>(int, int, int, const spmm_csr_plan_s *, const int *, int *,
                                      int, int, int, const int *, const void *, const float *,
                                      int, float *, int, hipStream_t);
template void b2sr_wide_launch<uint16_t>(int, int, int, const spmm_csr_plan_s *, const int *,
                                         int *, int, int, int, const int *, const void *,
                                         const uint16_t *, int, float *, int, hipStream_t);

}  // namespace spmm
