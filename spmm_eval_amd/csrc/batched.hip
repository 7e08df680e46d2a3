// Batched uneven graphs: G square graphs run as one block-diagonal problem.
// The plan rewrites the per-graph (local) column indices to global ones,
// analyses the union once (rows of all graphs binned together, so short and
// long rows from different graphs share waves), builds the transposed union
// for dX = A^T dY, and runs one fwd+bwd step -- Y = A X, dX = A^T dY,
// dval = SDDMM(dY, X) -- as ONE kernel launch whose work items are the
// forward CSR workgroups, the transposed CSR workgroups and the SDDMM
// workgroups (on the forward plan's rows). In persistent mode a grid sized to
// the machine (a multiple of eight workgroups) strides over the items, so a
// workgroup stays on one XCD's stretch of every class and no queue or
// counter is shared between workgroups. The plan's autotuner times every
// (lanes per row, vector width, rows per subgroup, persistent or not)
// candidate at creation and keeps the fastest; the step is captured as a
// hipGraph and replayed. Each row's arithmetic is the same as in a
// per-graph run of the same kernels, so results equal per-graph results bit
// for bit.
#include "csr_core.hpp"

#include <algorithm>
#include <atomic>
#include <new>
#include <vector>

struct spmm_batch_plan_s {
  spmm_handle_s *h = nullptr;
  int ngraphs = 0, M = 0, nnz = 0, N = 0, bf16 = 0;
  const int *rowptr = nullptr;  // caller's (global entry offsets)
  const float *val = nullptr;   // caller's
  int *gcol = nullptr;          // global column indices (owned)
  int *trowptr = nullptr, *tcol = nullptr;  // transposed union (owned)
  float *tval = nullptr;
  int *perm = nullptr;          // transposed entry -> source entry (owned)
  spmm_csr_plan_t fwd = nullptr, bwd = nullptr;
  // chosen geometry
  int G = 0, VEC = 1, rows_scale = 1, persistent = 0, sddmm_rows = 1, grid = 0;
  float tuned_ms = 0.f;
  hipStream_t cap = nullptr;  // capture stream (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *gX = nullptr, *gdY = nullptr;  // operands the graph was captured with
  float *gY = nullptr, *gdX = nullptr, *gdval = nullptr;
  // item order of the fused launch per geometry (built on the host once)
  struct Order {
    int G = 0, VEC = 0, rows_scale = 0, sddmm_rows = -1;
    int *d = nullptr;
    int n_first = 0;  // positions of the first phase
  };
  std::vector<Order> orders;
};

namespace spmm {

__global__ void globalise_cols_kernel(int ngraphs, const int *__restrict__ row_off,
                                      const int *__restrict__ nnz_off,
                                      const int *__restrict__ colind, int *__restrict__ gcol) {
  const int g = blockIdx.x;
  if (g >= ngraphs) return;
  const int off = row_off[g];
  for (int e = nnz_off[g] + threadIdx.x; e < nnz_off[g + 1]; e += blockDim.x)
    gcol[e] = colind[e] + off;
}

// SDDMM forms of the fused launch (all give dval within FP32 rounding of
// each other; each is bitwise reproducible)
constexpr int kSddChunks = 0;  // entry-chunked SDDMM items
constexpr int kSddRows = 1;    // SDDMM items on the forward plan's rows
constexpr int kSddFused = 2;   // no SDDMM items: the backward items also form dval

struct FusedArgs {
  LaunchTable fwd, bwd, sdd;  // sdd: the forward plan's rows at the SDDMM lane count
  int ctiles;                 // CSR column tiles
  int n_fwd, n_bwd, n_sddmm;  // items of each part (CSR parts include column tiles)
  int sddmm_rows;             // the SDDMM form (kSddChunks, kSddRows or kSddFused)
  int M, N, nnz, sddmm_ntile;
  const int *rows_fwd, *rows_bwd;
  const int2 *span_fwd, *span_bwd;
  const int *rowptr, *colind, *trowptr, *tcol;
  const float *val, *tval;
  const int *perm;
  const void *X, *dY;
  float *Y, *dX, *dval;
  const int *order;           // launch position -> item (see merged_item_order)
  int n_first;                // positions of the first phase (forward and SDDMM items)
};

// kSdd: the SDDMM form, one per kernel, so another form's registers do not
// lower the occupancy. kSddFused needs one column tile (ctiles == 1).
template <typename TB, int G, int VEC, int GS, int kSdd>
__device__ __forceinline__ void fused_item(const FusedArgs &a, int item) {
  const TB *X = static_cast<const TB *>(a.X);
  const TB *dY = static_cast<const TB *>(a.dY);
  if (item < a.n_fwd) {
    csr_wg<TB, G, VEC>(a.fwd, item / a.ctiles, item % a.ctiles, a.rows_fwd, a.span_fwd,
                       a.colind, a.val, X, a.N, a.Y, a.N);
    return;
  }
  item -= a.n_fwd;
  if (kSdd == kSddFused) {
    if (item < a.n_bwd)
      csr_t_sddmm_wg<TB, G, VEC, true>(a.bwd, item, a.rows_bwd, a.span_bwd, a.tcol, a.tval,
                                       a.perm, dY, a.N, X, a.N, a.dX, a.N, a.dval);
    return;
  }
  if (item < a.n_bwd) {
    csr_wg<TB, G, VEC>(a.bwd, item / a.ctiles, item % a.ctiles, a.rows_bwd, a.span_bwd,
                       a.tcol, a.tval, dY, a.N, a.dX, a.N);
    return;
  }
  item -= a.n_bwd;
  if (kSdd == kSddRows)
    sddmm_row_wg<TB, GS>(a.sdd, item, a.rows_fwd, a.span_fwd, a.colind, dY, a.N, X, a.N,
                         a.dval, a.sddmm_ntile);
  else
    sddmm_wg<TB, GS>(item, a.M, a.nnz, a.rowptr, a.colind, dY, a.N, X, a.N, a.dval,
                     a.sddmm_ntile);
}

// Launch order. Workgroup b runs on XCD b mod 8 (the dispatcher deals
// workgroups to the eight XCDs in turn). The plan merges the items of every
// run -- each degree class of the forward and transposed rows, the SDDMM
// rows or chunks -- by their fractional position in their run
// (merged_item_order); rows ascend graph by graph within every run, so
// consecutive positions of the merged order belong to the same few graphs
// in every part. XCD x then takes one contiguous stretch of the merged order
// (xcd_stretch: a bijection on [s0, s1) that gives XCD x the x-th stretch):
// at any time an XCD works on a few graphs, in all parts and classes at once,
// and their feature rows X and dY stay in its L2, read past it about once;
// every XCD gets a share of every class. Only the order changes, not any
// item's work.
__device__ __forceinline__ int xcd_count(int t, int x) {  // b in [0, t) with b mod 8 == x
  return (t + 7 - x) >> 3;
}
__device__ __forceinline__ int xcd_stretch(int b, int s0, int s1) {
  const int x = b & 7;
  int v = s0 + xcd_count(b, x) - xcd_count(s0, x);
  for (int xp = 0; xp < x; ++xp) v += xcd_count(s1, xp) - xcd_count(s0, xp);
  return v;
}

template <typename TB, int G, int VEC, int GS, bool kPersistent, int kSdd>
__global__ void __launch_bounds__(kThreads) fused_step_kernel(FusedArgs a) {
  const int total = a.n_fwd + a.n_bwd + a.n_sddmm;
  auto pos = [&](int b) {  // each phase stretched over the XCDs on its own
    return b < a.n_first ? xcd_stretch(b, 0, a.n_first) : xcd_stretch(b, a.n_first, total);
  };
  if (!kPersistent) {
    fused_item<TB, G, VEC, GS, kSdd>(a, a.order[pos(blockIdx.x)]);
    return;
  }
  // the grid is a multiple of 8, so every position b this workgroup takes
  // runs on the XCD that xcd_stretch assumes for b
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    fused_item<TB, G, VEC, GS, kSdd>(a, a.order[pos(b)]);
    __syncthreads();  // the next item may reuse the LDS of this one
  }
}

static int sddmm_G(int N, int e16) {
  const int lanes = N / e16;
  for (int t = 32; t >= 1; t >>= 1)
    if (t <= lanes && lanes % t == 0) return t;
  return 0;
}

// grid: 8 workgroups per CU; the persistent launch takes as many as the
// instantiation keeps resident (a multiple of 8, so each workgroup's items
// stay on its XCD).
template <typename TB, int G, int VEC, int GS, int kSdd>
static void launch_fused_k(const FusedArgs &a, bool persistent, int grid, hipStream_t s) {
  if (persistent) {
    auto *const kern = &fused_step_kernel<TB, G, VEC, GS, true, kSdd>;
    static std::atomic<int> per_cu{0};
    int pc = per_cu.load(std::memory_order_relaxed);
    if (pc == 0) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, reinterpret_cast<const void *>(kern),
                                                       kThreads, 0) != hipSuccess || pc <= 0)
        pc = 8;
      per_cu.store(pc, std::memory_order_relaxed);
    }
    const int g = std::max(8, (grid / 8) * std::min(pc, 8) / 8 * 8);
    hipLaunchKernelGGL(kern, dim3(g), dim3(kThreads), 0, s, a);
  } else
    hipLaunchKernelGGL((fused_step_kernel<TB, G, VEC, GS, false, kSdd>),
                       dim3(a.n_fwd + a.n_bwd + a.n_sddmm), dim3(kThreads), 0, s, a);
}

template <typename TB, int G, int VEC, int GS>
static int launch_fused_g(const FusedArgs &a, bool persistent, int grid, hipStream_t s) {
  if (a.sddmm_rows == kSddFused)
    launch_fused_k<TB, G, VEC, GS, kSddFused>(a, persistent, grid, s);
  else if (a.sddmm_rows == kSddRows)
    launch_fused_k<TB, G, VEC, GS, kSddRows>(a, persistent, grid, s);
  else
    launch_fused_k<TB, G, VEC, GS, kSddChunks>(a, persistent, grid, s);
  return hipGetLastError() == hipSuccess ? SPMM_OK : SPMM_ERR_HIP;
}

template <typename TB, int VEC, int GS>
static int launch_fused_v(int G, const FusedArgs &a, bool persistent, int grid, hipStream_t s) {
  switch (G) {
    case 4: return launch_fused_g<TB, 4, VEC, GS>(a, persistent, grid, s);
    case 8: return launch_fused_g<TB, 8, VEC, GS>(a, persistent, grid, s);
    case 16: return launch_fused_g<TB, 16, VEC, GS>(a, persistent, grid, s);
    case 32: return launch_fused_g<TB, 32, VEC, GS>(a, persistent, grid, s);
    case 64: return launch_fused_g<TB, 64, VEC, GS>(a, persistent, grid, s);
    default: return SPMM_ERR_UNSUPPORTED;
  }
}

template <typename TB, int GS>
static int launch_fused_s(int G, int VEC, const FusedArgs &a, bool persistent, int grid,
                          hipStream_t s) {
  return VEC == 2 ? launch_fused_v<TB, 2, GS>(G, a, persistent, grid, s)
                  : launch_fused_v<TB, 1, GS>(G, a, persistent, grid, s);
}

template <typename TB>
static int launch_fused(int G, int VEC, int GS, const FusedArgs &a, bool persistent, int grid,
                        hipStream_t s) {
  switch (GS) {
    case 4: return launch_fused_s<TB, 4>(G, VEC, a, persistent, grid, s);
    case 8: return launch_fused_s<TB, 8>(G, VEC, a, persistent, grid, s);
    case 16: return launch_fused_s<TB, 16>(G, VEC, a, persistent, grid, s);
    case 32: return launch_fused_s<TB, 32>(G, VEC, a, persistent, grid, s);
    default: return SPMM_ERR_UNSUPPORTED;
  }
}

// The launch tables and item counts of one geometry.
static int fill_args(const spmm_batch_plan_s *p, int G, int VEC, int rows_scale, int sddmm_rows,
                     FusedArgs *ap, int *GS_out) {
  FusedArgs &a = *ap;
  const int elem = p->bf16 ? 2 : 4;
  const int e16 = 16 / elem;
  build_table(p->fwd->counts, p->fwd->offs, G, rows_scale, &a.fwd);
  build_table(p->bwd->counts, p->bwd->offs, G, rows_scale, &a.bwd);
  a.ctiles = p->N / (G * e16 * VEC);
  a.n_fwd = a.fwd.total_wg * a.ctiles;
  a.n_bwd = a.bwd.total_wg * a.ctiles;
  const int GS = sddmm_G(p->N, e16);
  build_table(p->fwd->counts, p->fwd->offs, GS, 1, &a.sdd);
  if (sddmm_rows == kSddFused && a.ctiles != 1)
    return SPMM_ERR_UNSUPPORTED;  // the fused backward needs one column tile
  a.sddmm_rows = sddmm_rows;
  a.n_sddmm = sddmm_rows == kSddFused ? 0
              : sddmm_rows == kSddRows
                  ? a.sdd.total_wg
                  : static_cast<int>(ceil_div_ll(
                        p->nnz, static_cast<long long>(kThreads / GS) * kSddmmChunk));
  a.sddmm_ntile = p->N / (GS * e16);
  *GS_out = GS;
  return SPMM_OK;
}

// The merged order of a launch's items in two phases: first the runs that
// gather X (each degree class of the forward table; the SDDMM table's
// classes or its chunks), then the runs that gather dY (the transposed
// table's classes). Within a phase every run contributes its items at
// fractional positions (j + 1/2) / n, sorted by position (exact:
// cross-multiplied), ties in run order. Returns the first phase's length.
static int merged_item_order(const FusedArgs &a, std::vector<int> *out) {
  struct Run { int first, n; };
  auto table_runs = [](std::vector<Run> &runs, const LaunchTable &t, int ct, int off) {
    for (int i = 0; i < t.nseg; ++i) {
      const int b = t.seg[i].wg_begin, e = i + 1 < t.nseg ? t.seg[i + 1].wg_begin : t.total_wg;
      if (e > b) runs.push_back({off + b * ct, (e - b) * ct});
    }
  };
  auto merge = [&](const std::vector<Run> &runs) {
    struct It { int run, j; };
    std::vector<It> items;
    for (int r = 0; r < static_cast<int>(runs.size()); ++r)
      for (int j = 0; j < runs[r].n; ++j) items.push_back({r, j});
    std::stable_sort(items.begin(), items.end(), [&](const It &x, const It &y) {
      // (2 jx + 1) / nx < (2 jy + 1) / ny
      return (2ll * x.j + 1) * runs[y.run].n < (2ll * y.j + 1) * runs[x.run].n;
    });
    for (const It &it : items) out->push_back(runs[it.run].first + it.j);
  };
  out->clear();
  out->reserve(static_cast<size_t>(a.n_fwd) + a.n_bwd + a.n_sddmm);
  std::vector<Run> first, second;
  table_runs(first, a.fwd, a.ctiles, 0);
  if (a.sddmm_rows == kSddRows) table_runs(first, a.sdd, 1, a.n_fwd + a.n_bwd);
  else if (a.n_sddmm > 0) first.push_back({a.n_fwd + a.n_bwd, a.n_sddmm});
  table_runs(second, a.bwd, a.ctiles, a.n_fwd);
  merge(first);
  const int n_first = static_cast<int>(out->size());
  merge(second);
  return n_first;
}

// The plan's device copy of a geometry's merged order, built on first use
// (allocates and synchronises; never while s is being captured).
static int item_order(spmm_batch_plan_s *p, int G, int VEC, int rows_scale, int sddmm_rows,
                      FusedArgs *a, hipStream_t s) {
  for (const auto &o : p->orders)
    if (o.G == G && o.VEC == VEC && o.rows_scale == rows_scale && o.sddmm_rows == sddmm_rows) {
      a->order = o.d;
      a->n_first = o.n_first;
      return SPMM_OK;
    }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SPMM_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone)
    SPMM_FAIL(SPMM_ERR_UNSUPPORTED, "launch order not prepared before capture");
  std::vector<int> h;
  const int n_first = merged_item_order(*a, &h);
  spmm_batch_plan_s::Order o;
  o.n_first = n_first;
  o.G = G;
  o.VEC = VEC;
  o.rows_scale = rows_scale;
  o.sddmm_rows = sddmm_rows;
  if (hipMalloc(&o.d, sizeof(int) * std::max<size_t>(1, h.size())) != hipSuccess)
    SPMM_FAIL(SPMM_ERR_ALLOC, "hipMalloc of the launch order failed");
  hipError_t e = h.empty() ? hipSuccess
                           : hipMemcpyAsync(o.d, h.data(), sizeof(int) * h.size(),
                                            hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(o.d);
    SPMM_FAIL(SPMM_ERR_HIP, std::string("launch order upload: ") + hipGetErrorString(e));
  }
  p->orders.push_back(o);
  a->order = o.d;
  a->n_first = n_first;
  return SPMM_OK;
}

static int enqueue_step(spmm_batch_plan_s *p, int G, int VEC, int rows_scale, int persistent,
                        int sddmm_rows, const void *X, const void *dY, float *Y, float *dX,
                        float *dval, hipStream_t s) {
  FusedArgs a;
  int GS = 0;
  int st = fill_args(p, G, VEC, rows_scale, sddmm_rows, &a, &GS);
  if (st) return st;
  st = item_order(p, G, VEC, rows_scale, sddmm_rows, &a, s);
  if (st) return st;
  a.M = p->M;
  a.N = p->N;
  a.nnz = p->nnz;
  a.rows_fwd = p->fwd->d_rows;
  a.rows_bwd = p->bwd->d_rows;
  a.span_fwd = p->fwd->d_span;
  a.span_bwd = p->bwd->d_span;
  a.rowptr = p->rowptr;
  a.colind = p->gcol;
  a.trowptr = p->trowptr;
  a.tcol = p->tcol;
  a.val = p->val;
  a.tval = p->tval;
  a.perm = p->perm;
  a.X = X;
  a.dY = dY;
  a.Y = Y;
  a.dX = dX;
  a.dval = dval;
  return p->bf16 ? launch_fused<uint16_t>(G, VEC, GS, a, persistent, p->grid, s)
                 : launch_fused<float>(G, VEC, GS, a, persistent, p->grid, s);
}

static void destroy_graph(spmm_batch_plan_s *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

// The launch order of the plan's current geometry, built now if missing (so
// that neither a capture nor an eager step has to).
static int prepare_order(spmm_batch_plan_s *p) {
  FusedArgs a;
  int GS = 0;
  const int st = fill_args(p, p->G, p->VEC, p->rows_scale, p->sddmm_rows, &a, &GS);
  if (st) return st;
  return item_order(p, p->G, p->VEC, p->rows_scale, p->sddmm_rows, &a, p->cap);
}

// Frees the launch orders of every geometry but the current one.
static void prune_orders(spmm_batch_plan_s *p) {
  (void)hipStreamSynchronize(p->cap);
  std::vector<spmm_batch_plan_s::Order> keep;
  for (auto &o : p->orders) {
    if (o.G == p->G && o.VEC == p->VEC && o.rows_scale == p->rows_scale &&
        o.sddmm_rows == p->sddmm_rows)
      keep.push_back(o);
    else
      (void)hipFree(o.d);
  }
  p->orders.swap(keep);
}

// Times each geometry on scratch operands (zeros: the gather pattern, not
// the values, sets the time) and keeps the fastest.
static int tune_plan(spmm_batch_plan_s *p) {
  const int elem = p->bf16 ? 2 : 4;
  const int e16 = 16 / elem;
  const size_t feat = static_cast<size_t>(p->M) * p->N;
  void *X = nullptr, *dY = nullptr;
  float *Y = nullptr, *dX = nullptr, *dval = nullptr;
  int st = SPMM_OK;
  if (hipMalloc(&X, feat * elem) != hipSuccess || hipMalloc(&dY, feat * elem) != hipSuccess ||
      hipMalloc(&Y, feat * 4) != hipSuccess || hipMalloc(&dX, feat * 4) != hipSuccess ||
      hipMalloc(&dval, sizeof(float) * static_cast<size_t>(p->nnz > 0 ? p->nnz : 1)) != hipSuccess)
    st = SPMM_ERR_ALLOC;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st && (hipMemsetAsync(X, 0, feat * elem, p->cap) != hipSuccess ||
              hipMemsetAsync(dY, 0, feat * elem, p->cap) != hipSuccess ||
              hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = SPMM_ERR_HIP;
  float best = 1e30f;
  // one warm-up and three timed steps of a candidate (ms for the three)
  auto time3 = [&](int G, int vec, int rs, int pers, int srows, float *ms) {
    int rc = enqueue_step(p, G, vec, rs, pers, srows, X, dY, Y, dX, dval, p->cap);
    if (rc) return rc;
    if (hipEventRecord(e0, p->cap) != hipSuccess) return static_cast<int>(SPMM_ERR_HIP);
    for (int r = 0; r < 3 && !rc; ++r)
      rc = enqueue_step(p, G, vec, rs, pers, srows, X, dY, Y, dX, dval, p->cap);
    if (rc) return rc;
    if (hipEventRecord(e1, p->cap) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
      return static_cast<int>(SPMM_ERR_HIP);
    (void)hipEventElapsedTime(ms, e0, e1);
    return static_cast<int>(SPMM_OK);
  };
  // SDDMM form: fused into the backward items (2, one column tile only),
  // row-scheduled (1) or entry-chunked (0)
  for (int srows = kSddFused; srows >= kSddChunks && !st; --srows) {
    for (int vec = 1; vec <= 2 && !st; ++vec) {
      if (p->N % (e16 * vec)) continue;
      const int lanes = p->N / (e16 * vec);
      for (int G = 64; G >= 4 && !st; G >>= 1) {
        if (lanes % G || G > lanes || G * 4 < std::min(lanes, 64)) continue;
        if (srows == kSddFused && G != lanes) continue;
        for (int rs = 1; rs <= 8 && !st; rs <<= 1) {
          for (int pers = 0; pers <= 1 && !st; ++pers) {
            float ms = 0.f;
            st = time3(G, vec, rs, pers, srows, &ms);
            if (!st && ms < best) {
              best = ms;
              p->G = G;
              p->VEC = vec;
              p->rows_scale = rs;
              p->persistent = pers;
              p->sddmm_rows = srows;
            }
          }
        }
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(p->cap);
  if (X) (void)hipFree(X);
  if (dY) (void)hipFree(dY);
  if (Y) (void)hipFree(Y);
  if (dX) (void)hipFree(dX);
  if (dval) (void)hipFree(dval);
  if (st) return st;
  if (p->G == 0) return SPMM_ERR_UNSUPPORTED;
  p->tuned_ms = best / 3.0f;
  prune_orders(p);
  return SPMM_OK;
}

}  // namespace spmm

using namespace spmm;

extern "C" {

int spmm_batch_plan_destroy(spmm_batch_plan_t p) {
  if (!p) return SPMM_OK;
  destroy_graph(p);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  spmm_csr_plan_destroy(p->fwd);
  spmm_csr_plan_destroy(p->bwd);
  if (p->gcol) (void)hipFree(p->gcol);
  if (p->trowptr) (void)hipFree(p->trowptr);
  if (p->tcol) (void)hipFree(p->tcol);
  if (p->tval) (void)hipFree(p->tval);
  if (p->perm) (void)hipFree(p->perm);
  for (auto &o : p->orders) (void)hipFree(o.d);
  delete p;
  return SPMM_OK;
}

int spmm_batch_plan_create(spmm_handle_t h, int ngraphs, const int *row_off, const int *nnz_off,
                           const int *rowptr, const int *colind, const float *val, int N, int bf16,
                           spmm_batch_plan_t *out, void *stream) {
  SPMM_CHECK_ARG(h && out && row_off && nnz_off, "null handle, out or offsets");
  *out = nullptr;
  SPMM_CHECK_ARG(ngraphs > 0 && N > 0, "ngraphs and N must be positive");
  const int M = row_off[ngraphs], nnz = nnz_off[ngraphs];
  SPMM_CHECK_ARG(M >= 0 && nnz >= 0, "negative totals");
  SPMM_CHECK_ARG(rowptr && (nnz == 0 || (colind && val)), "null CSR arrays");
  for (int g = 0; g < ngraphs; ++g)
    SPMM_CHECK_ARG(row_off[g + 1] >= row_off[g] && nnz_off[g + 1] >= nnz_off[g],
                   "offsets must be non-decreasing");
  const int e16 = bf16 ? 8 : 4;
  if (N % e16 || sddmm_G(N, e16) < 4)
    SPMM_FAIL(SPMM_ERR_UNSUPPORTED, "batched plans need N a multiple of 16 (FP32) or 32 (BF16)");
  hipStream_t s = as_stream(stream);
  auto *p = new (std::nothrow) spmm_batch_plan_s();
  if (!p) SPMM_FAIL(SPMM_ERR_ALLOC, "host allocation failed");
  p->h = h;
  p->ngraphs = ngraphs;
  p->M = M;
  p->nnz = nnz;
  p->N = N;
  p->bf16 = bf16 ? 1 : 0;
  p->rowptr = rowptr;
  p->val = val;
  auto fail = [&](int code, const char *what) {
    spmm_batch_plan_destroy(p);
    set_error(std::string("spmm_batch_plan_create: ") + what);
    return code;
  };
  const size_t nn = static_cast<size_t>(nnz > 0 ? nnz : 1);
  if (hipMalloc(&p->gcol, sizeof(int) * nn) != hipSuccess ||
      hipMalloc(&p->trowptr, sizeof(int) * (static_cast<size_t>(M) + 1)) != hipSuccess ||
      hipMalloc(&p->tcol, sizeof(int) * nn) != hipSuccess ||
      hipMalloc(&p->tval, sizeof(float) * nn) != hipSuccess ||
      hipMalloc(&p->perm, sizeof(int) * nn) != hipSuccess)
    return fail(SPMM_ERR_ALLOC, "device allocation failed");
  int *d_off = static_cast<int *>(workspace(h, sizeof(int) * 2 * (static_cast<size_t>(ngraphs) + 1)));
  if (!d_off) return fail(SPMM_ERR_ALLOC, "workspace");
  if (hipMemcpyAsync(d_off, row_off, sizeof(int) * (ngraphs + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_off + ngraphs + 1, nnz_off, sizeof(int) * (ngraphs + 1), hipMemcpyHostToDevice,
                     s) != hipSuccess)
    return fail(SPMM_ERR_HIP, "offset upload");
  if (nnz > 0) {
    hipLaunchKernelGGL(globalise_cols_kernel, dim3(ngraphs), dim3(256), 0, s, ngraphs, d_off,
                       d_off + ngraphs + 1, colind, p->gcol);
    if (hipGetLastError() != hipSuccess) return fail(SPMM_ERR_HIP, "globalise launch");
  }
  if (hipStreamSynchronize(s) != hipSuccess) return fail(SPMM_ERR_HIP, "sync");
  int st = csr_transpose(h, M, M, nnz, rowptr, p->gcol, val, p->trowptr, p->tcol, p->tval,
                         p->perm, stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  st = spmm_csr_analyze(h, M, nnz, rowptr, &p->fwd, stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  st = spmm_csr_analyze(h, M, nnz, p->trowptr, &p->bwd, stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess)
    return fail(SPMM_ERR_HIP, "capture stream");
  // Persistent grid: resident workgroups of the widest instantiation.
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return fail(SPMM_ERR_HIP, "device query");
  p->grid = cus * 8;  // 8 workgroups of 4 waves per CU = 32 waves per CU (a multiple of 8)
  st = tune_plan(p);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  *out = p;
  return SPMM_OK;
}

int spmm_batch_plan_info(spmm_batch_plan_t p, spmm_tune_t *tune, int *persistent,
                         int *sddmm_rows, float *tuned_ms) {
  SPMM_CHECK_ARG(p, "null plan");
  if (tune) *tune = {p->G, p->VEC, p->rows_scale};
  if (persistent) *persistent = p->persistent;
  if (sddmm_rows) *sddmm_rows = p->sddmm_rows;
  if (tuned_ms) *tuned_ms = p->tuned_ms;
  return SPMM_OK;
}

int spmm_batch_plan_set(spmm_batch_plan_t p, const spmm_tune_t *tune, int persistent,
                        int sddmm_form) {
  SPMM_CHECK_ARG(p && tune, "null plan or tune");
  const int e16 = p->bf16 ? 8 : 4;
  const int G = tune->lanes_per_row, vec = tune->vec, rs = tune->rows_scale;
  SPMM_CHECK_ARG(vec == 1 || vec == 2, "vec must be 1 or 2");
  SPMM_CHECK_ARG(rs == 1 || rs == 2 || rs == 4 || rs == 8, "rows_scale must be 1, 2, 4 or 8");
  SPMM_CHECK_ARG(p->N % (e16 * vec) == 0, "N is not a whole number of vectors");
  const int lanes = p->N / (e16 * vec);
  SPMM_CHECK_ARG((G == 4 || G == 8 || G == 16 || G == 32 || G == 64) && lanes % G == 0,
                 "lanes_per_row must be 4..64 and divide N / (elements per vector x vec)");
  SPMM_CHECK_ARG(persistent == 0 || persistent == 1, "persistent must be 0 or 1");
  SPMM_CHECK_ARG(sddmm_form >= kSddChunks && sddmm_form <= kSddFused, "sddmm_form must be 0, 1 or 2");
  SPMM_CHECK_ARG(sddmm_form != kSddFused || G == lanes,
                 "the fused backward needs one column tile (lanes_per_row x vec x 16 B = N)");
  destroy_graph(p);  // the next graph step captures the new launch
  p->G = G;
  p->VEC = vec;
  p->rows_scale = rs;
  p->persistent = persistent;
  p->sddmm_rows = sddmm_form;
  return prepare_order(p);  // before any step of it is captured
}

int spmm_batch_fwd_bwd(spmm_batch_plan_t p, const void *X, const void *dY, float *Y, float *dX,
                       float *dval, int use_graph, void *stream) {
  SPMM_CHECK_ARG(p, "null plan");
  SPMM_CHECK_ARG(X && dY && Y && dX && (dval || p->nnz == 0), "null operand");
  hipStream_t s = as_stream(stream);
  if (!use_graph) {
    const int st = enqueue_step(p, p->G, p->VEC, p->rows_scale, p->persistent, p->sddmm_rows, X,
                                dY, Y, dX, dval, s);
    if (st) SPMM_FAIL(st, "fused launch failed");
    return SPMM_OK;
  }
  if (!p->exec || p->gX != X || p->gdY != dY || p->gY != Y || p->gdX != dX || p->gdval != dval) {
    destroy_graph(p);
    const int po = prepare_order(p);
    if (po) SPMM_FAIL(po, "launch order");
    // Capture only records the launch; the graph runs on s below.
    SPMM_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
    const int st = enqueue_step(p, p->G, p->VEC, p->rows_scale, p->persistent, p->sddmm_rows, X,
                                dY, Y, dX, dval, p->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      SPMM_FAIL(st, "capture of the fused launch failed");
    }
    SPMM_HIP(e);
    p->graph = g;
    SPMM_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    p->gX = X;
    p->gdY = dY;
    p->gY = Y;
    p->gdX = dX;
    p->gdval = dval;
  }
  SPMM_HIP(hipGraphLaunch(p->exec, s));
  return SPMM_OK;
}

}  // extern "C"
