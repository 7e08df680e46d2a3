// Device building blocks shared by the plain launches (csr_spmm.hip,
// adjoint.hip) and the fused batched launch (batched.hip): the launch table
// of row bins, the per-subgroup gather loop, the work of one 256-thread
// CSR workgroup and of one SDDMM workgroup.
#pragma once

#include "common.hpp"

namespace spmm {

constexpr int kBins = 32;
constexpr int kThreads = 256;  // 4 waves
constexpr int kWavesPerWG = kThreads / kWave;

// ------------------------------------------------------------ launch table
enum SegKind : int { kSegA = 0, kSegB = 1, kSegC = 2 };

struct Seg {
  int list_off;   // first position in the permuted row list
  int count;      // rows in this segment
  int kind;       // SegKind
  int rows_per_sub;  // R, class A only
  int wg_begin;   // first workgroup (blockIdx.x) of this segment
};

struct LaunchTable {
  int nseg;
  int total_wg;
  Seg seg[kBins];
};

// Loads of a stream read once (kNT: the non-temporal policy, so the lines do
// not push longer-lived data -- a slice of B -- out of the caches).
template <bool kNT, typename T>
__device__ __forceinline__ T ld_once(const T *p) {
  if constexpr (kNT) return __builtin_nontemporal_load(p);
  else return *p;
}

// One subgroup (G lanes) accumulates entries [e_begin, e_end) taking chunks
// of G entries at stride `chunk_stride`, starting at chunk `first`. With
// kPerm the value of entry e is val[perm[e]] (a transposed structure whose
// values stay in the source order, so learnable values need no re-gather).
// kUF: B-row loads in flight per lane (0: 8, or 4 at more than 16 elements a
// lane). Fewer, 4, suit the column-sliced launches, whose slice of B mostly
// hits L2: at 4 the 233k-row products run 1.5% (FP32) to 4% (BF16) faster
// than at 8, while the unsliced kernel on an 89k-row graph runs 4% slower.
template <typename TB, int G, int EPL, bool kPerm = false, bool kNT = false, int kUF = 0>
__device__ __forceinline__ void accumulate_range(
    int e_begin, int e_end, int first_chunk_off, int chunk_stride, int lane_sub,
    const int *__restrict__ colind, const float *__restrict__ val,
    const TB *__restrict__ Bc, int ldb, float (&acc)[EPL],
    const int *__restrict__ perm = nullptr) {
  constexpr int U0 = kUF ? kUF : (EPL > 16 ? 4 : 8);  // (8 x 32 B for BF16 at 2 vectors)
  constexpr int U = G < U0 ? G : U0;
  constexpr int V = EPL / Vec16<TB>::kElems;
  int e = e_begin + first_chunk_off;
  int cidx = 0;
  float v = 0.f;
  if (e + lane_sub < e_end) {
    cidx = ld_once<kNT>(colind + e + lane_sub);
    v = ld_once<kNT>(val + (kPerm ? perm[e + lane_sub] : e + lane_sub));
  }
  for (; e < e_end; e += chunk_stride) {
    const int n = min(G, e_end - e);
    // the next chunk's indices are requested before this chunk's gathers
    const int my_n = e + chunk_stride + lane_sub;
    int cidx_n = 0;
    float v_n = 0.f;
    if (my_n < e_end) {
      cidx_n = ld_once<kNT>(colind + my_n);
      v_n = ld_once<kNT>(val + (kPerm ? perm[my_n] : my_n));
    }
    for (int k = 0; k < n; k += U) {
      int cc[U];
      float vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cc[u] = __shfl(cidx, k + u, G);
        vv[u] = __shfl(v, k + u, G);
      }
      // All U loads are issued unconditionally (slots past n read row
      // cc = 0 of the padding lanes, always in range) and waited for once;
      // widening and accumulation are predicated.
      typename Vec16<TB>::Raw raw[U][V];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const TB *p = Bc + static_cast<size_t>(cc[u]) * ldb;
#pragma unroll
        for (int q = 0; q < V; ++q) raw[u][q] = Vec16<TB>::raw(p + q * G * Vec16<TB>::kElems);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k + u < n) {
          float b[EPL];
#pragma unroll
          for (int q = 0; q < V; ++q) Vec16<TB>::widen(raw[u][q], &b[q * Vec16<TB>::kElems]);
#pragma unroll
          for (int q = 0; q < EPL; ++q) acc[q] = fmaf(vv[u], b[q], acc[q]);
        }
      }
    }
    cidx = cidx_n;
    v = v_n;
  }
}

// Lane layout of a row segment: vector q of lane l covers columns
// q*G*E + l*E .. +E (E = elements per 16-byte vector), so each load or
// store instruction of a subgroup touches one contiguous run of G*16 bytes.
template <int G, int E, int EPL>
__device__ __forceinline__ void store_row(float *__restrict__ Crow, const float (&acc)[EPL]) {
#pragma unroll
  for (int q = 0; q < EPL / E; ++q) {
#pragma unroll
    for (int k = 0; k < E; k += 4)
      *reinterpret_cast<float4 *>(Crow + q * G * E + k) =
          make_float4(acc[q * E + k], acc[q * E + k + 1], acc[q * E + k + 2], acc[q * E + k + 3]);
  }
}

// C row segment += acc (the partial sums of one column slice added to those
// of the slices before it; same lane layout as store_row).
// kNT: C is streamed (read and written once per launch) with the
// non-temporal policy.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int G, int E, int EPL, bool kNT>
__device__ __forceinline__ void add_row(float *__restrict__ Crow, const float (&acc)[EPL]) {
#pragma unroll
  for (int q = 0; q < EPL / E; ++q) {
#pragma unroll
    for (int k = 0; k < E; k += 4) {
      f32x4 *p = reinterpret_cast<f32x4 *>(Crow + q * G * E + k);
      f32x4 c = ld_once<kNT>(p);
      c.x += acc[q * E + k];
      c.y += acc[q * E + k + 1];
      c.z += acc[q * E + k + 2];
      c.w += acc[q * E + k + 3];
      if constexpr (kNT) __builtin_nontemporal_store(c, p);
      else *p = c;
    }
  }
}

template <int G, int E, int EPL, bool kNT>
__device__ __forceinline__ void store_row_nt(float *__restrict__ Crow, const float (&acc)[EPL]) {
#pragma unroll
  for (int q = 0; q < EPL / E; ++q) {
#pragma unroll
    for (int k = 0; k < E; k += 4) {
      const f32x4 c = {acc[q * E + k], acc[q * E + k + 1], acc[q * E + k + 2], acc[q * E + k + 3]};
      f32x4 *p = reinterpret_cast<f32x4 *>(Crow + q * G * E + k);
      if constexpr (kNT) __builtin_nontemporal_store(c, p);
      else *p = c;
    }
  }
}

template <int G, int E, int EPL, bool kAcc, bool kNT = false>
__device__ __forceinline__ void put_row(float *__restrict__ Crow, const float (&acc)[EPL]) {
  if constexpr (kAcc) add_row<G, E, EPL, kNT>(Crow, acc);
  else if constexpr (kNT) store_row_nt<G, E, EPL, true>(Crow, acc);
  else store_row<G, E, EPL>(Crow, acc);
}

// Butterfly over the S subgroups of a wave: every subgroup ends with the
// same sum (IEEE addition is commutative, so partners agree bitwise).
template <int G, int EPL>
__device__ __forceinline__ void wave_reduce_subgroups(float (&acc)[EPL]) {
#pragma unroll
  for (int m = G; m < kWave; m <<= 1) {
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] += __shfl_xor(acc[q], m);
  }
}

// Segment of workgroup `wg`: binary search on wg_begin (ascending). `wg` is
// uniform, so each probe is one scalar load from the kernel-argument
// segment (~80 scalar instructions a wave, against ~300 for an unrolled
// select over all 32 entries, which made the scalar unit a bottleneck of
// launches with thousands of short workgroups).
__device__ __forceinline__ Seg find_seg(const LaunchTable &tab, int wg) {
  int lo = 0, hi = tab.nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab.seg[mid].wg_begin <= wg) lo = mid;
    else hi = mid - 1;
  }
  return tab.seg[lo];
}

// ------------------------------------------------ plan-free row classes
// The classes of one plan-free call found on the device: units (CSR rows, or
// B2SR block rows) of at least c_limit entries are appended to list 0
// (state + kStateHead, count state[0]), the others of at least a_limit
// entries to list 1 (state + kStateHead + n, count state[2]); one atomic per
// workgroup and list, positions from the waves' ballots. The kernel that
// drains the lists resets the counts.
template <int NT>
__global__ void __launch_bounds__(NT)
class_list_kernel(int n, const int *__restrict__ ptr, int a_limit, int c_limit,
                  int *__restrict__ state) {
  constexpr int NW = NT / kWave;
  __shared__ int cnt[2][NW];
  __shared__ int base[2];
  const int u = static_cast<int>(blockIdx.x) * NT + static_cast<int>(threadIdx.x);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int d = u < n ? ptr[u + 1] - ptr[u] : 0;
  const bool in[2] = {u < n && d >= c_limit, u < n && d < c_limit && d >= a_limit};
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long m[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    m[c] = __ballot(in[c]);
    if (lane == 0) cnt[c][wave] = __popcll(m[c]);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int c = threadIdx.x;
    int tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += cnt[c][w];
    base[c] = tot ? atomicAdd(&state[c == 0 ? 0 : 2], tot) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (!in[c]) continue;
    int pos = base[c] + __popcll(m[c] & below);
    for (int w = 0; w < wave; ++w) pos += cnt[c][w];
    state[kStateHead + (c == 0 ? 0 : n) + pos] = u;
  }
}

// One row of each class (the per-row code of the planned kernel and of the
// plan-free kernel, so both give the same bits at the same lane geometry).
// Class A: one subgroup of G lanes sums the row and stores it.
template <typename TB, int G, int VEC, bool kPerm, bool kAcc = false, bool kNT = false,
          int kUF = 0>
__device__ __forceinline__ void csr_row_sub(int row, int2 sp, int lane_sub, int col0,
                                            const int *__restrict__ colind,
                                            const float *__restrict__ val,
                                            const TB *__restrict__ Bc, int ldb,
                                            float *__restrict__ C, int ldc,
                                            const int *__restrict__ perm) {
  constexpr int E = Vec16<TB>::kElems;
  constexpr int EPL = VEC * E;
  float acc[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
  accumulate_range<TB, G, EPL, kPerm, kNT, kUF>(sp.x, sp.y, 0, G, lane_sub, colind, val, Bc, ldb, acc,
                                           perm);
  put_row<G, E, EPL, kAcc>(C + static_cast<size_t>(row) * ldc + col0, acc);
}

// Class B: the S subgroups of one wave split the row's chunks of G entries
// (subgroup s takes chunks s, s + S, ...), then a butterfly; subgroup 0
// stores.
template <typename TB, int G, int VEC, bool kPerm, bool kAcc = false, bool kNT = false,
          int kUF = 0>
__device__ __forceinline__ void csr_row_wave(int row, int2 sp, int sub, int lane_sub, int col0,
                                             const int *__restrict__ colind,
                                             const float *__restrict__ val,
                                             const TB *__restrict__ Bc, int ldb,
                                             float *__restrict__ C, int ldc,
                                             const int *__restrict__ perm) {
  constexpr int E = Vec16<TB>::kElems;
  constexpr int EPL = VEC * E;
  constexpr int S = kWave / G;
  float acc[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
  accumulate_range<TB, G, EPL, kPerm, kNT, kUF>(sp.x, sp.y, sub * G, S * G, lane_sub, colind, val, Bc,
                                           ldb, acc, perm);
  wave_reduce_subgroups<G, EPL>(acc);
  if (sub == 0) put_row<G, E, EPL, kAcc>(C + static_cast<size_t>(row) * ldc + col0, acc);
}

// Class C: the 4 S subgroups of one workgroup split the row, the waves meet
// in LDS in a fixed order; wave 0's subgroup 0 stores. Every thread of the
// workgroup must call it (two barriers).
template <typename TB, int G, int VEC, bool kPerm, bool kAcc = false, bool kNT = false,
          int kUF = 0>
__device__ __forceinline__ void csr_row_wg(int row, int2 sp, int wave, int sub, int lane_sub,
                                           int col0, const int *__restrict__ colind,
                                           const float *__restrict__ val,
                                           const TB *__restrict__ Bc, int ldb,
                                           float *__restrict__ C, int ldc,
                                           const int *__restrict__ perm) {
  constexpr int E = Vec16<TB>::kElems;
  constexpr int EPL = VEC * E;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  __shared__ float red[kWavesPerWG - 1][CW];
  float acc[EPL];
#pragma unroll
  for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
  const int gsub = wave * S + sub;
  accumulate_range<TB, G, EPL, kPerm, kNT, kUF>(sp.x, sp.y, gsub * G, kWavesPerWG * S * G, lane_sub,
                                           colind, val, Bc, ldb, acc, perm);
  wave_reduce_subgroups<G, EPL>(acc);
  if (wave > 0 && sub == 0) {
#pragma unroll
    for (int q = 0; q < EPL; ++q) red[wave - 1][lane_sub * EPL + q] = acc[q];
  }
  __syncthreads();
  if (wave == 0 && sub == 0) {
#pragma unroll
    for (int w = 0; w < kWavesPerWG - 1; ++w) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] += red[w][lane_sub * EPL + q];
    }
    put_row<G, E, EPL, kAcc>(C + static_cast<size_t>(row) * ldc + col0, acc);
  }
  __syncthreads();  // red[] may be reused by the next row
}

// The work of one 256-thread workgroup of the CSR launch table: virtual
// workgroup `wg`, column tile `ctile`. Used by csr_spmm_kernel and by the
// fused batched kernel.
template <typename TB, int G, int VEC, bool kPerm = false>
__device__ __forceinline__ void csr_wg(const LaunchTable &tab, int wg, int ctile,
                                       const int *__restrict__ rows,
                                       const int2 *__restrict__ span,
                                       const int *__restrict__ colind,
                                       const float *__restrict__ val,
                                       const TB *__restrict__ B, int ldb,
                                       float *__restrict__ C, int ldc,
                                       const int *__restrict__ perm = nullptr) {
  constexpr int EPL = VEC * Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;  // columns covered per column tile

  const Seg sg = find_seg(tab, wg);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;  // subgroup within the wave
  constexpr int E = Vec16<TB>::kElems;
  const int col0 = ctile * CW + lane_sub * E;
  const TB *Bc = B + col0;
  const int local_wg = wg - sg.wg_begin;

  if (sg.kind == kSegA) {
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kWavesPerWG * S;
    const int base = sg.list_off + local_wg * subs_per_wg * R;
    const int end = sg.list_off + sg.count;
    const int my_sub = wave * S + sub;
    // The plan keeps each row's entry span beside its id, so a row's id and
    // span arrive together; the next row's are requested before this row's
    // gathers, leaving colind -> gather on the critical path of each row.
    int pos = base + my_sub;
    int row = -1;
    int2 sp = make_int2(0, 0);
    if (pos < end) {
      row = rows[pos];
      sp = span[pos];
    }
    for (int k = 0; k < R && row >= 0; ++k) {
      const int pos_n = pos + subs_per_wg;
      int row_n = -1;
      int2 sp_n = make_int2(0, 0);
      if (k + 1 < R && pos_n < end) {
        row_n = rows[pos_n];
        sp_n = span[pos_n];
      }
      csr_row_sub<TB, G, VEC, kPerm>(row, sp, lane_sub, col0, colind, val, Bc, ldb, C, ldc, perm);
      row = row_n;
      sp = sp_n;
      pos = pos_n;
    }
  } else if (sg.kind == kSegB) {
    const int pos = sg.list_off + local_wg * kWavesPerWG + wave;
    if (pos >= sg.list_off + sg.count) return;  // whole wave
    csr_row_wave<TB, G, VEC, kPerm>(rows[pos], span[pos], sub, lane_sub, col0, colind, val, Bc,
                                    ldb, C, ldc, perm);
  } else {
    csr_row_wg<TB, G, VEC, kPerm>(rows[sg.list_off + local_wg], span[sg.list_off + local_wg], wave,
                                  sub, lane_sub, col0, colind, val, Bc, ldb, C, ldc, perm);
  }
}


// ------------------------------------------------ fused dX + SDDMM
// One pass over the transposed structure (row c of A^T = column c of A; its
// entries e' carry A's row i = tcol[e'] and source entry perm[e']) gives
// both halves of a training step's backward: dX[c] = sum val[perm[e']] dY[i]
// and dval[perm[e']] = dY[i] . X[c]. The dY row gathered for dX is dotted
// with X[c], which the subgroup holds in registers for the row, so the
// separate SDDMM's gather of X rows disappears. Each entry's dot product
// stays in the subgroup that gathers it (lane partials over its columns in
// column order, then a reduce-scatter and a butterfly over the G lanes), so
// no cross-subgroup reduction exists for dval; lane k + u of a chunk stores
// entry k + u.
// kTVal: the value of entry e' is tval[e'] (values already in transposed
// order) instead of val[perm[e']].
template <typename TB, int G, int EPL, bool kTVal>
__device__ __forceinline__ void accumulate_range_sddmm(
    int e_begin, int e_end, int first_chunk_off, int chunk_stride, int lane_sub,
    const int *__restrict__ tcol, const float *__restrict__ val, const int *__restrict__ perm,
    const TB *__restrict__ dYc, int lddy, const float (&x)[EPL], float (&acc)[EPL],
    float *__restrict__ dval) {
  constexpr int U0 = EPL > 16 ? 4 : 8;
  constexpr int U = G < U0 ? G : U0;
  constexpr int V = EPL / Vec16<TB>::kElems;
  int e = e_begin + first_chunk_off;
  int cidx = 0, pidx = 0;
  float v = 0.f;
  if (e + lane_sub < e_end) {
    cidx = tcol[e + lane_sub];
    pidx = perm[e + lane_sub];
    v = kTVal ? val[e + lane_sub] : val[pidx];
  }
  for (; e < e_end; e += chunk_stride) {
    const int n = min(G, e_end - e);
    const int my_n = e + chunk_stride + lane_sub;
    int cidx_n = 0, pidx_n = 0;
    float v_n = 0.f;
    if (my_n < e_end) {
      cidx_n = tcol[my_n];
      pidx_n = perm[my_n];
      v_n = kTVal ? val[my_n] : val[pidx_n];
    }
    for (int k = 0; k < n; k += U) {
      int cc[U];
      float vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cc[u] = __shfl(cidx, k + u, G);
        vv[u] = __shfl(v, k + u, G);
      }
      // all U loads unconditional (slots past n read the padding lanes' row
      // 0) and waited for once; the arithmetic is predicated
      typename Vec16<TB>::Raw raw[U][V];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const TB *p = dYc + static_cast<size_t>(cc[u]) * lddy;
#pragma unroll
        for (int q = 0; q < V; ++q) raw[u][q] = Vec16<TB>::raw(p + q * G * Vec16<TB>::kElems);
      }
      float d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        d[u] = 0.f;
        if (k + u < n) {
          float b[EPL];
#pragma unroll
          for (int q = 0; q < V; ++q) Vec16<TB>::widen(raw[u][q], &b[q * Vec16<TB>::kElems]);
#pragma unroll
          for (int q = 0; q < EPL; ++q) {
            acc[q] = fmaf(vv[u], b[q], acc[q]);
            d[u] = fmaf(b[q], x[q], d[u]);
          }
        }
      }
      // reduce-scatter over the low log2(U) lane bits (U - 1 shuffles)
      // leaves entry k + (l mod U)'s partial in lane l; a butterfly over the
      // other bits completes it, and lane l (which holds that entry's source
      // index, as l - k = l mod U) stores it
#pragma unroll
      for (int m = U / 2; m >= 1; m >>= 1) {
        const bool hi_half = (lane_sub & m) != 0;
#pragma unroll
        for (int i = 0; i < m; ++i) {
          const float send = hi_half ? d[i] : d[i + m];
          const float keep = hi_half ? d[i + m] : d[i];
          d[i] = keep + __shfl_xor(send, m, G);
        }
      }
#pragma unroll
      for (int m = U; m < G; m <<= 1) d[0] += __shfl_xor(d[0], m, G);
      if (lane_sub >= k && lane_sub < k + U && lane_sub < n) dval[pidx] = d[0];
    }
    cidx = cidx_n;
    pidx = pidx_n;
    v = v_n;
  }
}

// Virtual workgroup `wg` of the transposed launch table, with the row
// classes of csr_wg (one column tile: the subgroup's G lanes cover all N
// columns), each row's X[c] loaded beside its span. Used by
// csr_t_sddmm_kernel and by the fused batched kernel.
template <typename TB, int G, int VEC, bool kTVal>
__device__ __forceinline__ void csr_t_sddmm_wg(const LaunchTable &tab, int wg,
                                               const int *__restrict__ rows,
                                               const int2 *__restrict__ span,
                                               const int *__restrict__ tcol,
                                               const float *__restrict__ val,
                                               const int *__restrict__ perm,
                                               const TB *__restrict__ dY, int lddy,
                                               const TB *__restrict__ X, int ldx,
                                               float *__restrict__ dX, int lddx,
                                               float *__restrict__ dval) {
  constexpr int E = Vec16<TB>::kElems;
  constexpr int EPL = VEC * E;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  __shared__ float red[kWavesPerWG - 1][CW];
  const Seg sg = find_seg(tab, wg);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int col0 = lane_sub * E;
  const TB *dYc = dY + col0;
  const int local_wg = wg - sg.wg_begin;
  auto load_x = [&](int c, float (&x)[EPL]) {
    const TB *xr = X + static_cast<size_t>(c) * ldx + col0;
#pragma unroll
    for (int q = 0; q < VEC; ++q) Vec16<TB>::load(xr + q * G * E, &x[q * E]);
  };

  if (sg.kind == kSegA) {
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kWavesPerWG * S;
    const int end = sg.list_off + sg.count;
    int pos = sg.list_off + local_wg * subs_per_wg * R + wave * S + sub;
    for (int k = 0; k < R && pos < end; ++k, pos += subs_per_wg) {
      const int c = rows[pos];
      const int2 sp = span[pos];
      float x[EPL], acc[EPL];
      load_x(c, x);
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
      accumulate_range_sddmm<TB, G, EPL, kTVal>(sp.x, sp.y, 0, G, lane_sub, tcol, val, perm, dYc, lddy,
                                         x, acc, dval);
      store_row<G, E, EPL>(dX + static_cast<size_t>(c) * lddx + col0, acc);
    }
  } else if (sg.kind == kSegB) {
    const int pos = sg.list_off + local_wg * kWavesPerWG + wave;
    if (pos >= sg.list_off + sg.count) return;  // whole wave
    const int c = rows[pos];
    const int2 sp = span[pos];
    float x[EPL], acc[EPL];
    load_x(c, x);
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
    accumulate_range_sddmm<TB, G, EPL, kTVal>(sp.x, sp.y, sub * G, S * G, lane_sub, tcol, val, perm,
                                       dYc, lddy, x, acc, dval);
    wave_reduce_subgroups<G, EPL>(acc);
    if (sub == 0) store_row<G, E, EPL>(dX + static_cast<size_t>(c) * lddx + col0, acc);
  } else {
    const int c = rows[sg.list_off + local_wg];
    const int2 sp = span[sg.list_off + local_wg];
    float x[EPL], acc[EPL];
    load_x(c, x);
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
    const int gsub = wave * S + sub;
    accumulate_range_sddmm<TB, G, EPL, kTVal>(sp.x, sp.y, gsub * G, kWavesPerWG * S * G, lane_sub,
                                       tcol, val, perm, dYc, lddy, x, acc, dval);
    wave_reduce_subgroups<G, EPL>(acc);
    if (wave > 0 && sub == 0) {
#pragma unroll
      for (int q = 0; q < EPL; ++q) red[wave - 1][lane_sub * EPL + q] = acc[q];
    }
    __syncthreads();
    if (wave == 0 && sub == 0) {
#pragma unroll
      for (int w = 0; w < kWavesPerWG - 1; ++w) {
#pragma unroll
        for (int q = 0; q < EPL; ++q) acc[q] += red[w][lane_sub * EPL + q];
      }
      store_row<G, E, EPL>(dX + static_cast<size_t>(c) * lddx + col0, acc);
    }
    __syncthreads();  // red[] may be reused by the next work item
  }
}

// ---------------------------------------------------------------- SDDMM
// Entry-chunked (merge-path) SDDMM: each subgroup of G lanes (G <= 32) owns
// kSddmmChunk consecutive stored entries, whatever rows they fall in, so a
// 20k-entry hub row is spread over many subgroups and no subgroup waits on
// one. The subgroup finds the row of its first entry with a G-ary search of
// rowptr (G probes per round, ballot), then walks the rows its chunk covers.
// Per row it keeps the dY segment in registers; entries are taken L =
// min(8, G) at a time: lane l forms the partial dots of the L entries over
// its EPL columns, a reduce-scatter over the low log2(L) lane bits (L-1
// shuffles) leaves entry (l mod L)'s partial in lane l, and a butterfly over
// the remaining lane bits completes it. Lanes 0..L-1 store (coalesced).
// Column tiles beyond the first add into out. Outputs are per entry, so no
// cross-chunk fix-up exists and the result is bitwise reproducible.
constexpr int kSddmmChunk = 32;

template <typename TB, int G>
__device__ __forceinline__ void sddmm_wg(int wg, int M, int nnz, const int *__restrict__ rowptr,
                                         const int *__restrict__ colind,
                                         const TB *__restrict__ dY, int lddy,
                                         const TB *__restrict__ X, int ldx,
                                         float *__restrict__ out, int ntile) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  constexpr int L = G < 8 ? G : 8;
  const int lane = threadIdx.x & (kWave - 1);
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int gsub = (threadIdx.x / kWave) * S + sub;
  const long long e_beg_ll = (static_cast<long long>(wg) * (kThreads / G) + gsub) * kSddmmChunk;
  if (e_beg_ll >= nnz) return;
  const int e_beg = static_cast<int>(e_beg_ll);
  const int e_end = min(nnz, e_beg + kSddmmChunk);
  // Largest r with rowptr[r] <= e_beg (rowptr[0] = 0 <= e_beg < rowptr[M]).
  int lo = 0, hi = M;
  const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  while (hi - lo > 1) {
    const int step = (hi - lo + G) / (G + 1);
    const int p = lo + (lane_sub + 1) * step;
    const bool le = p < hi && rowptr[p] <= e_beg;
    const unsigned long long m = (__ballot(le) >> (sub * G)) & gmask;
    lo += __popcll(m) * step;
    hi = min(hi, lo + step);
  }
  for (int tt = 0; tt < ntile; ++tt) {
    const TB *Xc = X + tt * CW + lane_sub * EPL;
    int row = lo;
    int e = e_beg;
    while (e < e_end) {
      int rend = rowptr[row + 1];
      while (rend <= e) rend = rowptr[++row + 1];  // skip rows with no entry left
      const int stop = min(rend, e_end);
      float y[EPL];
      Vec16<TB>::load(dY + static_cast<size_t>(row) * lddy + tt * CW + lane_sub * EPL, y);
      for (; e < stop; e += L) {
        const int n = min(L, stop - e);
        const int c = lane_sub < n ? colind[e + lane_sub] : 0;
        // unconditional loads (slots past n read row 0), one wait
        typename Vec16<TB>::Raw xr[L];
#pragma unroll
        for (int u = 0; u < L; ++u) {
          const int j = __shfl(c, u, G);
          xr[u] = Vec16<TB>::raw(Xc + static_cast<size_t>(j) * ldx);
        }
        float p[L];
#pragma unroll
        for (int u = 0; u < L; ++u) {
          float d = 0.f;
          if (u < n) {
            float x[EPL];
            Vec16<TB>::widen(xr[u], x);
#pragma unroll
            for (int q = 0; q < EPL; ++q) d = fmaf(y[q], x[q], d);
          }
          p[u] = d;
        }
#pragma unroll
        for (int m = L / 2; m >= 1; m >>= 1) {
          const bool hi_half = (lane_sub & m) != 0;
#pragma unroll
          for (int i = 0; i < m; ++i) {
            const float send = hi_half ? p[i] : p[i + m];
            const float keep = hi_half ? p[i + m] : p[i];
            p[i] = keep + __shfl_xor(send, m, G);
          }
        }
#pragma unroll
        for (int m = L; m < G; m <<= 1) p[0] += __shfl_xor(p[0], m, G);
        if (lane_sub < n) {
          if (tt == 0) out[e + lane_sub] = p[0];
          else out[e + lane_sub] += p[0];
        }
      }
      e = stop;  // the last group may have been partial
      ++row;
    }
  }
}

// Row-scheduled SDDMM on a CSR plan: the rows run in the CSR classes
// (subgroup / wave / workgroup per row by degree), and a row's entries are
// dealt to its subgroups in chunks of G (lane l reads the chunk's entry l;
// the next chunk's indices are requested before this chunk's gathers). Each
// entry's dot product stays inside one subgroup (lane partials over EPL
// columns, the same reduce-scatter + butterfly as sddmm_wg, hence the same
// bits), so no cross-subgroup reduction exists. Column tiles beyond the
// first add into out.
template <typename TB, int G>
__device__ __forceinline__ void sddmm_range(int e_begin, int e_end, int first_off, int stride,
                                            int lane_sub, const int *__restrict__ colind,
                                            const float (&y)[Vec16<TB>::kElems],
                                            const TB *__restrict__ Xc, int ldx,
                                            float *__restrict__ out, bool first_tile) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int L = G < 8 ? G : 8;
  int e = e_begin + first_off;
  int c = e + lane_sub < e_end ? colind[e + lane_sub] : 0;
  for (; e < e_end; e += stride) {
    const int n = min(G, e_end - e);
    const int my_n = e + stride + lane_sub;
    const int c_n = my_n < e_end ? colind[my_n] : 0;
    for (int k = 0; k < n; k += L) {
      const int nl = min(L, n - k);
      typename Vec16<TB>::Raw xr[L];
#pragma unroll
      for (int u = 0; u < L; ++u) {
        const int j = __shfl(c, k + u < G ? k + u : 0, G);  // slots past nl read a valid row
        xr[u] = Vec16<TB>::raw(Xc + static_cast<size_t>(j) * ldx);
      }
      float p[L];
#pragma unroll
      for (int u = 0; u < L; ++u) {
        float d = 0.f;
        if (u < nl) {
          float x[EPL];
          Vec16<TB>::widen(xr[u], x);
#pragma unroll
          for (int q = 0; q < EPL; ++q) d = fmaf(y[q], x[q], d);
        }
        p[u] = d;
      }
#pragma unroll
      for (int m = L / 2; m >= 1; m >>= 1) {
        const bool hi_half = (lane_sub & m) != 0;
#pragma unroll
        for (int i = 0; i < m; ++i) {
          const float send = hi_half ? p[i] : p[i + m];
          const float keep = hi_half ? p[i + m] : p[i];
          p[i] = keep + __shfl_xor(send, m, G);
        }
      }
#pragma unroll
      for (int m = L; m < G; m <<= 1) p[0] += __shfl_xor(p[0], m, G);
      if (lane_sub < nl) {
        float *o = out + e + k + lane_sub;
        if (first_tile) *o = p[0];
        else *o += p[0];
      }
    }
    c = c_n;
  }
}

// The SDDMM work of virtual workgroup `wg` of a CSR plan's launch table.
template <typename TB, int G>
__device__ __forceinline__ void sddmm_row_wg(const LaunchTable &tab, int wg,
                                             const int *__restrict__ rows,
                                             const int2 *__restrict__ span,
                                             const int *__restrict__ colind,
                                             const TB *__restrict__ dY, int lddy,
                                             const TB *__restrict__ X, int ldx,
                                             float *__restrict__ out, int ntile) {
  constexpr int EPL = Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;
  const Seg sg = find_seg(tab, wg);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;
  const int gsub = wave * S + sub;
  const int local_wg = wg - sg.wg_begin;
  float y[EPL];
  if (sg.kind == kSegA) {
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kWavesPerWG * S;
    const int end = sg.list_off + sg.count;
    int pos = sg.list_off + local_wg * subs_per_wg * R + gsub;
    int row = -1, e0 = 0, e1 = 0;
    if (pos < end) {
      row = rows[pos];
      const int2 sp = span[pos];
      e0 = sp.x;
      e1 = sp.y;
    }
    for (int k = 0; k < R && row >= 0; ++k) {
      const int pos_n = pos + subs_per_wg;
      int row_n = -1, e0n = 0, e1n = 0;
      if (k + 1 < R && pos_n < end) {
        row_n = rows[pos_n];
        const int2 sp = span[pos_n];
        e0n = sp.x;
        e1n = sp.y;
      }
      for (int tt = 0; tt < ntile && e0 < e1; ++tt) {
        Vec16<TB>::load(dY + static_cast<size_t>(row) * lddy + tt * CW + lane_sub * EPL, y);
        sddmm_range<TB, G>(e0, e1, 0, G, lane_sub, colind, y, X + tt * CW + lane_sub * EPL, ldx,
                           out, tt == 0);
      }
      row = row_n;
      e0 = e0n;
      e1 = e1n;
      pos = pos_n;
    }
    return;
  }
  const bool wave_mode = sg.kind == kSegB;
  const int pos = wave_mode ? sg.list_off + local_wg * kWavesPerWG + wave : sg.list_off + local_wg;
  if (pos >= sg.list_off + sg.count) return;  // whole wave (no barrier here)
  const int row = rows[pos];
  const int2 sp = span[pos];
  const int e0 = sp.x, e1 = sp.y;
  const int part = wave_mode ? sub : gsub;
  const int parts = wave_mode ? S : kWavesPerWG * S;
  for (int tt = 0; tt < ntile; ++tt) {
    Vec16<TB>::load(dY + static_cast<size_t>(row) * lddy + tt * CW + lane_sub * EPL, y);
    sddmm_range<TB, G>(e0, e1, part * G, parts * G, lane_sub, colind, y,
                       X + tt * CW + lane_sub * EPL, ldx, out, tt == 0);
  }
}

// Host side.
struct Geometry {
  int G = 0;    // lanes per row (0 = narrow kernel)
  int VEC = 1;  // 16-byte vectors per lane
  int col_tiles = 1;
  int rows_scale = 1;
  int slices = 1;  // launches per column tile over ranges of B's rows
};

constexpr int kTuneSlots = 8;
struct TuneEntry {
  int N;
  int elem_bytes;
  spmm_tune_t tune;
};

// Launch table from bin counts: bins whose degrees all lie below a_limit run
// one subgroup per row (max(1, r_num / bound) rows per subgroup, r_empty for
// empty rows, times rows_scale), bins at or above c_limit one workgroup per
// row, the rest one wave per row (b_items waves / c_items workgroups per
// unit when a unit holds several rows, as a B2SR block row does).
void build_table_ex(const int *counts, const int *offs, int S, long long a_limit,
                    long long c_limit, int r_num, int r_empty, int rows_scale, LaunchTable *t,
                    int b_items = 1, int c_items = 1);
// Launch table for a CSR plan's bin counts and lane geometry G.
void build_table(const int *counts, const int *offs, int G, int rows_scale, LaunchTable *t);

}  // namespace spmm

struct spmm_csr_plan_s;
namespace spmm {
// SDDMM dval[e] = dY[i] . X[colind[e]] (plan == nullptr: entry-chunked;
// else on the rows of a CSR plan); defined for float and uint16_t.
template <typename TB>
int run_sddmm(const spmm_csr_plan_s *plan, int M, int K, int N, int nnz, const int *rowptr,
              const int *colind, const TB *dY, int lddy, const TB *X, int ldx, float *out,
              void *stream);
// Bins rows by degree rowptr[(r >> shift) + 1] - rowptr[r >> shift].
int analyze(spmm_handle_s *h, int M, int nnz, const int *rowptr, int shift,
            spmm_csr_plan_s **out, void *stream);
// The plan's column-slice offsets for S equal ranges of K columns over its
// `units` plan units (colind: their column or block-column indices,
// ascending in each unit), built on first use (synchronises) and kept in
// the plan per (S, K).
int plan_slice_offsets(spmm_csr_plan_s *p, int units, int S, int K, const int *colind,
                       hipStream_t s, const int **out);
// Plan-free: units u sorted by the bin of ptr[u + 1] - ptr[u] (bin 0 for
// none, else floor(log2) + 1), heaviest bin first, into *rows (the handle's
// drop-in scratch, after its counts, cursors and class sizes); meta[0] =
// units in bins >= b_wg, meta[1] = units in bins b_sub < b < b_wg. Three
// launches, no synchronisation; the bin counts are left at zero.
int dropin_bin_sort(spmm_handle_s *h, const int *ptr, int units, int b_wg, int b_sub,
                    hipStream_t s, const int **rows, const int **meta);
// The stable transpose queued on `s` without a synchronisation (the public
// csr_transpose adds one); capturable once the workspace has grown.
int transpose_async(spmm_handle_s *h, int M, int K, int nnz, const int *rowptr, const int *colind,
                    const float *val, int *tRowPtr, int *tColInd, float *tVal, int *perm,
                    hipStream_t s);
// Row classes of the CSR launch for G lanes per row: rows of fewer than
// *a_limit entries run one subgroup each, rows of at least *c_limit one
// workgroup each, the rest one wave each (both powers of two, so the
// classes of a degree bin never straddle a limit).
inline void csr_class_limits(int G, int *a_limit, int *c_limit) {
  *a_limit = G <= 16 ? (8 * G > 32 ? 8 * G : 32) : 2 * G;
  *c_limit = 64 * (kWave / G);
}
}  // namespace spmm

// Row schedule of one CSR structure: rows permuted into degree bins
// (highest bin first) plus the autotuned geometries.
struct spmm_csr_plan_s {
  int M = 0;
  int nnz = 0;
  int device = 0;
  int *d_rows = nullptr;
  int2 *d_span = nullptr;  // (rowptr[u], rowptr[u + 1]) of unit d_rows[i], beside it
  int counts[32] = {0};
  int offs[32] = {0};
  int ntuned = 0;
  spmm::TuneEntry tuned[spmm::kTuneSlots];
  int tile = 0;  // 0: CSR plan; T: B2SR plan (rows binned by block-row tile count)
  // column-sliced launches: per (slices S, B rows K) -- one set for each
  // slice count the autotuner tries --, the entry offsets
  // off[s * M + i] (s = 0..S) where the s-th range of B's rows starts in the
  // row at plan position i (the rows' entries are ascending in column)
  struct SliceSet {
    int S = 0, K = 0;
    int *d_off = nullptr;
  } slice_sets[6];
  int sorted = -1;  // colind ascending within every row: -1 not checked yet
  int cus = 0;  // compute units of the plan's device
  // B2SR plans for T >= 32, one allocation at d_wide: the round lists of
  // every call (kWideSlots ints per tile) and their row counts (T ints per
  // round), so calls on one such plan must not overlap; then the round
  // table (first and end tile of each round) and the first round of each
  // plan unit's block row, built at analysis
  int *d_wide = nullptr;
  int *d_wide_rc = nullptr;
  int2 *d_wide_round = nullptr;
  int *d_wide_rbase = nullptr;
  int wide_rounds = 0;
};
