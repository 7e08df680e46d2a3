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
#include "common.hpp"

#include <algorithm>
#include <new>
#include <vector>

struct spmm_csr_plan_s {
  int M = 0;
  int nnz = 0;
  int device = 0;
  int *d_rows = nullptr;  // rows permuted into bin order (highest bin first)
  int counts[32] = {0};
  int offs[32] = {0};
};

namespace spmm {

constexpr int kBins = 32;
constexpr int kThreads = 256;  // 4 waves
constexpr int kWavesPerWG = kThreads / kWave;

__device__ __forceinline__ int degree_bin(int deg) {
  return deg <= 0 ? 0 : 32 - __clz(deg);
}

__global__ void __launch_bounds__(256)
bin_count_kernel(const int *__restrict__ rowptr, int M, int *__restrict__ counts) {
  __shared__ int h[kBins];
  if (threadIdx.x < kBins) h[threadIdx.x] = 0;
  __syncthreads();
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < M;
       r += gridDim.x * blockDim.x) {
    int b = degree_bin(rowptr[r + 1] - rowptr[r]);
    atomicAdd(&h[b], 1);
  }
  __syncthreads();
  if (threadIdx.x < kBins && h[threadIdx.x])
    atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// cursor[b] starts at the bin's offset; each workgroup reserves its rows of
// each bin with one atomic and ranks them through LDS.
__global__ void __launch_bounds__(256)
bin_scatter_kernel(const int *__restrict__ rowptr, int M, int *__restrict__ cursor,
                   int *__restrict__ rows) {
  __shared__ int h[kBins];
  __shared__ int base[kBins];
  for (int r0 = blockIdx.x * blockDim.x; r0 < M; r0 += gridDim.x * blockDim.x) {
    if (threadIdx.x < kBins) h[threadIdx.x] = 0;
    __syncthreads();
    int r = r0 + threadIdx.x;
    int b = -1, rank = 0;
    if (r < M) {
      b = degree_bin(rowptr[r + 1] - rowptr[r]);
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

// One subgroup (G lanes) accumulates entries [e_begin, e_end) taking chunks
// of G entries at stride `chunk_stride`, starting at chunk `first`.
template <typename TB, int G, int EPL>
__device__ __forceinline__ void accumulate_range(
    int e_begin, int e_end, int first_chunk_off, int chunk_stride, int lane_sub,
    const int *__restrict__ colind, const float *__restrict__ val,
    const TB *__restrict__ Bc, int ldb, float (&acc)[EPL]) {
  constexpr int U0 = EPL >= 16 ? 4 : 8;  // loads in flight per lane
  constexpr int U = G < U0 ? G : U0;
  constexpr int V = EPL / Vec16<TB>::kElems;
  for (int e = e_begin + first_chunk_off; e < e_end; e += chunk_stride) {
    const int n = min(G, e_end - e);
    const int my = e + lane_sub;
    int cidx = 0;
    float v = 0.f;
    if (my < e_end) {
      cidx = colind[my];
      v = val[my];
    }
    for (int k = 0; k < n; k += U) {
      int cc[U];
      float vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cc[u] = __shfl(cidx, k + u, G);
        vv[u] = __shfl(v, k + u, G);
      }
      float b[U][EPL];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k + u < n) {
          const TB *p = Bc + static_cast<size_t>(cc[u]) * ldb;
#pragma unroll
          for (int q = 0; q < V; ++q)
            Vec16<TB>::load(p + q * Vec16<TB>::kElems, &b[u][q * Vec16<TB>::kElems]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k + u < n) {
#pragma unroll
          for (int q = 0; q < EPL; ++q) acc[q] = fmaf(vv[u], b[u][q], acc[q]);
        }
      }
    }
  }
}

template <int EPL>
__device__ __forceinline__ void store_row(float *__restrict__ Crow, const float (&acc)[EPL]) {
#pragma unroll
  for (int q = 0; q < EPL; q += 4)
    *reinterpret_cast<float4 *>(Crow + q) = make_float4(acc[q], acc[q + 1], acc[q + 2], acc[q + 3]);
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

template <typename TB, int G, int VEC>
__global__ void __launch_bounds__(kThreads)
csr_spmm_kernel(LaunchTable tab, const int *__restrict__ rows,
                const int *__restrict__ rowptr, const int *__restrict__ colind,
                const float *__restrict__ val, const TB *__restrict__ B, int ldb,
                float *__restrict__ C, int ldc) {
  constexpr int EPL = VEC * Vec16<TB>::kElems;
  constexpr int S = kWave / G;
  constexpr int CW = G * EPL;  // columns covered per column tile
  __shared__ float red[kWavesPerWG - 1][CW];

  const int wg = blockIdx.x;
  // Unrolled select keeps the table in SGPRs (a dynamic index into the
  // by-value argument would spill it to scratch).
  Seg sg = tab.seg[0];
#pragma unroll
  for (int i = 1; i < kBins; ++i) {
    if (i < tab.nseg && wg >= tab.seg[i].wg_begin) sg = tab.seg[i];
  }

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int lane_sub = lane & (G - 1);
  const int sub = lane / G;  // subgroup within the wave
  const int col0 = blockIdx.y * CW + lane_sub * EPL;
  const TB *Bc = B + col0;
  const int local_wg = wg - sg.wg_begin;

  if (sg.kind == kSegA) {
    const int R = sg.rows_per_sub;
    const int subs_per_wg = kWavesPerWG * S;
    const int base = sg.list_off + local_wg * subs_per_wg * R;
    const int end = sg.list_off + sg.count;
    const int my_sub = wave * S + sub;
    for (int k = 0; k < R; ++k) {
      const int pos = base + k * subs_per_wg + my_sub;
      if (pos >= end) break;
      const int row = rows[pos];
      float acc[EPL];
#pragma unroll
      for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
      accumulate_range<TB, G, EPL>(rowptr[row], rowptr[row + 1], 0, G, lane_sub,
                                   colind, val, Bc, ldb, acc);
      store_row<EPL>(C + static_cast<size_t>(row) * ldc + col0, acc);
    }
  } else if (sg.kind == kSegB) {
    const int pos = sg.list_off + local_wg * kWavesPerWG + wave;
    if (pos >= sg.list_off + sg.count) return;
    const int row = rows[pos];
    float acc[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
    accumulate_range<TB, G, EPL>(rowptr[row], rowptr[row + 1], sub * G, S * G,
                                 lane_sub, colind, val, Bc, ldb, acc);
    wave_reduce_subgroups<G, EPL>(acc);
    if (sub == 0) store_row<EPL>(C + static_cast<size_t>(row) * ldc + col0, acc);
  } else {
    const int row = rows[sg.list_off + local_wg];
    float acc[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) acc[q] = 0.f;
    const int gsub = wave * S + sub;
    accumulate_range<TB, G, EPL>(rowptr[row], rowptr[row + 1], gsub * G,
                                 kWavesPerWG * S * G, lane_sub, colind, val, Bc,
                                 ldb, acc);
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
      store_row<EPL>(C + static_cast<size_t>(row) * ldc + col0, acc);
    }
  }
}

// Any N, any alignment: one wave per row, lanes stride over the columns.
template <typename TB>
__global__ void __launch_bounds__(kThreads)
csr_spmm_generic_kernel(int M, int N, const int *__restrict__ rowptr,
                        const int *__restrict__ colind, const float *__restrict__ val,
                        const TB *__restrict__ B, int ldb, float *__restrict__ C,
                        int ldc) {
  const int row = blockIdx.x * kWavesPerWG + threadIdx.x / kWave;
  if (row >= M) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int e0 = rowptr[row], e1 = rowptr[row + 1];
  for (int c = lane; c < N; c += kWave) {
    float acc = 0.f;
    for (int e = e0; e < e1; ++e)
      acc = fmaf(val[e], load_scalar<TB>(B + static_cast<size_t>(colind[e]) * ldb + c), acc);
    C[static_cast<size_t>(row) * ldc + c] = acc;
  }
}

// ------------------------------------------------------------------- host
struct Geometry {
  int G = 0;    // lanes per row (0 = generic kernel)
  int VEC = 1;  // 16-byte vectors per lane
  int col_tiles = 1;
  int rows_scale = 1;
};

static Geometry pick_geometry(int N, int elem_bytes, const void *B, int ldb,
                              const void *C, int ldc, const spmm_tune_t *tune) {
  Geometry g;
  const int e16 = 16 / elem_bytes;
  const bool aligned = aligned16(B) && aligned16(C) && (ldb % e16 == 0) && (ldc % 4 == 0);
  if (tune && tune->rows_scale > 0) g.rows_scale = tune->rows_scale;
  if (!aligned || N % e16 != 0) return g;
  int vec = (tune && tune->vec == 2) ? 2 : 1;
  if (N % (e16 * vec) != 0) vec = 1;
  const int epl = e16 * vec;
  const int lanes_needed = N / epl;
  int G = 0;
  if (tune && tune->lanes_per_row > 0) {
    int t = tune->lanes_per_row;
    if ((t & (t - 1)) == 0 && t <= 64 && N % (t * epl) == 0) G = t;
  }
  if (G == 0) {
    for (int t = 64; t >= 1; t >>= 1) {
      if (t <= lanes_needed && lanes_needed % t == 0) { G = t; break; }
    }
  }
  if (G == 0) return g;
  g.G = G;
  g.VEC = vec;
  g.col_tiles = N / (G * epl);
  return g;
}

static void build_table(const spmm_csr_plan_s *p, int G, int rows_scale, LaunchTable *t) {
  const int S = kWave / G;
  const int a_limit = std::max(2 * G, 8);   // deg < a_limit -> one subgroup per row
  const int c_limit = 256 * S;              // deg >= c_limit -> one workgroup per row
  t->nseg = 0;
  int wg = 0;
  for (int b = kBins - 1; b >= 0; --b) {
    const int cnt = p->counts[b];
    if (cnt == 0) continue;
    const long long lo = b == 0 ? 0 : (1ll << (b - 1));
    const long long hi = b == 0 ? 1 : (1ll << b);  // exclusive upper bound
    Seg s;
    s.list_off = p->offs[b];
    s.count = cnt;
    s.wg_begin = wg;
    if (hi <= a_limit) {
      s.kind = kSegA;
      int R = b == 0 ? 64 : std::max(1, static_cast<int>(64 / hi));
      R = std::max(1, R * rows_scale);
      s.rows_per_sub = R;
      wg += ceil_div(cnt, kWavesPerWG * S * R);
    } else if (lo >= c_limit) {
      s.kind = kSegC;
      s.rows_per_sub = 1;
      wg += cnt;
    } else {
      s.kind = kSegB;
      s.rows_per_sub = 1;
      wg += ceil_div(cnt, kWavesPerWG);
    }
    t->seg[t->nseg++] = s;
  }
  t->total_wg = wg;
}

template <typename TB, int G, int VEC>
static void launch_vec(const LaunchTable &t, int col_tiles, const spmm_csr_plan_s *p,
                       const int *rowptr, const int *colind, const float *val,
                       const TB *B, int ldb, float *C, int ldc, hipStream_t s) {
  dim3 grid(t.total_wg, col_tiles);
  hipLaunchKernelGGL((csr_spmm_kernel<TB, G, VEC>), grid, dim3(kThreads), 0, s, t,
                     p->d_rows, rowptr, colind, val, B, ldb, C, ldc);
}

template <typename TB, int VEC>
static void dispatch_G(int G, const LaunchTable &t, int col_tiles,
                       const spmm_csr_plan_s *p, const int *rowptr,
                       const int *colind, const float *val, const TB *B, int ldb,
                       float *C, int ldc, hipStream_t s) {
  switch (G) {
    case 1: launch_vec<TB, 1, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    case 2: launch_vec<TB, 2, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    case 4: launch_vec<TB, 4, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    case 8: launch_vec<TB, 8, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    case 16: launch_vec<TB, 16, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    case 32: launch_vec<TB, 32, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
    default: launch_vec<TB, 64, VEC>(t, col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s); break;
  }
}

template <typename TB>
static int run_planned(spmm_csr_plan_s *p, int M, int K, int N, int nnz,
                       const int *rowptr, const int *colind, const float *val,
                       const TB *B, int ldb, float *C, int ldc,
                       const spmm_tune_t *tune, void *stream) {
  SPMM_CHECK_ARG(p != nullptr, "plan is null");
  SPMM_CHECK_ARG(M >= 0 && K >= 0 && N >= 0 && nnz >= 0, "negative size");
  if (p->M != M || p->nnz != nnz)
    SPMM_FAIL(SPMM_ERR_MISMATCH, "plan was built for M=" + std::to_string(p->M) +
                                     " nnz=" + std::to_string(p->nnz));
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(rowptr && C, "null rowptr or C");
  SPMM_CHECK_ARG(nnz == 0 || (colind && val && B), "null colind, val or B");
  SPMM_CHECK_ARG(ldc >= N && (K == 0 || ldb >= N), "ldb/ldc smaller than N");
  hipStream_t s = as_stream(stream);
  Geometry g = pick_geometry(N, sizeof(TB), B, ldb, C, ldc, tune);
  if (g.G == 0) {
    hipLaunchKernelGGL((csr_spmm_generic_kernel<TB>), dim3(ceil_div(M, kWavesPerWG)),
                       dim3(kThreads), 0, s, M, N, rowptr, colind, val, B, ldb, C, ldc);
    SPMM_LAUNCHED();
    return SPMM_OK;
  }
  LaunchTable t;
  build_table(p, g.G, g.rows_scale, &t);
  if (t.total_wg == 0) return SPMM_OK;
  if (g.VEC == 2)
    dispatch_G<TB, 2>(g.G, t, g.col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s);
  else
    dispatch_G<TB, 1>(g.G, t, g.col_tiles, p, rowptr, colind, val, B, ldb, C, ldc, s);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

int analyze(spmm_handle_s *h, int M, int nnz, const int *rowptr,
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
  if (hipMalloc(&p->d_rows, sizeof(int) * static_cast<size_t>(M)) != hipSuccess) {
    delete p;
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
  hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(kThreads), 0, s, rowptr, M, ws);
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
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(kThreads), 0, s, rowptr, M, ws,
                     p->d_rows);
  if (hipGetLastError() != hipSuccess) return fail("bin_scatter_kernel launch");
  // The pinned buffer and workspace are reused by the next call.
  if (hipStreamSynchronize(s) != hipSuccess) return fail("stream sync");
  *out = p;
  return SPMM_OK;
}

template <typename TB>
static int run_dropin(spmm_handle_s *h, int M, int K, int N, int nnz,
                      const int *rowptr, const int *colind, const float *val,
                      const TB *B, int ldb, float *C, int ldc, void *stream) {
  spmm_csr_plan_s *p = nullptr;
  int st = analyze(h, M, nnz, rowptr, &p, stream);
  if (st != SPMM_OK) return st;
  st = run_planned<TB>(p, M, K, N, nnz, rowptr, colind, val, B, ldb, C, ldc, nullptr, stream);
  // The plan's row list is read by the kernel just queued.
  if (hipStreamSynchronize(as_stream(stream)) != hipSuccess && st == SPMM_OK) st = SPMM_ERR_HIP;
  spmm_csr_plan_destroy(p);
  return st;
}

}  // namespace spmm

extern "C" {

int spmm_csr_analyze(spmm_handle_t h, int M, int nnz, const int *rowptr,
                     spmm_csr_plan_t *plan, void *stream) {
  return spmm::analyze(h, M, nnz, rowptr, plan, stream);
}

int spmm_csr_plan_destroy(spmm_csr_plan_t plan) {
  if (!plan) return SPMM_OK;
  if (plan->d_rows) (void)hipFree(plan->d_rows);
  delete plan;
  return SPMM_OK;
}

int spmm_csr_plan_bins(spmm_csr_plan_t plan, int *counts32) {
  SPMM_CHECK_ARG(plan && counts32, "null plan or counts");
  for (int b = 0; b < 32; ++b) counts32[b] = plan->counts[b];
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
