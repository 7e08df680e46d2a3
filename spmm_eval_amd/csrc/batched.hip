// Batched uneven graphs: G square graphs run as one block-diagonal problem.
// The plan rewrites the per-graph (local) column indices to global ones,
// analyses the union once (rows of all graphs binned together, so short and
// long rows from different graphs share waves), builds the transposed union
// for dX = A^T dY and captures one fwd+bwd step -- Y = A X, dX = A^T dY,
// dval = SDDMM(dY, X) -- as a hipGraph that is replayed per step. Each row's
// arithmetic is the same as in a per-graph run of the same kernels, so the
// batched results equal per-graph results bit for bit.
#include "common.hpp"

#include <new>

struct spmm_batch_plan_s {
  spmm_handle_s *h = nullptr;
  int ngraphs = 0, M = 0, nnz = 0, N = 0, bf16 = 0;
  const int *rowptr = nullptr;  // caller's (global entry offsets)
  const float *val = nullptr;   // caller's
  int *gcol = nullptr;          // global column indices (owned)
  int *trowptr = nullptr, *tcol = nullptr;  // transposed union (owned)
  float *tval = nullptr;
  spmm_csr_plan_t fwd = nullptr, bwd = nullptr;
  hipStream_t cap = nullptr;  // capture stream (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *gX = nullptr, *gdY = nullptr;  // operands the graph was captured with
  float *gY = nullptr, *gdX = nullptr, *gdval = nullptr;
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

static void destroy_graph(spmm_batch_plan_s *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

static int enqueue_step(spmm_batch_plan_s *p, const void *X, const void *dY, float *Y, float *dX,
                        float *dval, hipStream_t s) {
  const int M = p->M, N = p->N;
  int st;
  if (p->bf16) {
    st = spmm_csr_bf16_planned(p->fwd, M, M, N, p->nnz, p->rowptr, p->gcol, p->val,
                               static_cast<const uint16_t *>(X), N, Y, N, nullptr, s);
    if (st) return st;
    st = spmm_csr_bf16_planned(p->bwd, M, M, N, p->nnz, p->trowptr, p->tcol, p->tval,
                               static_cast<const uint16_t *>(dY), N, dX, N, nullptr, s);
    if (st) return st;
    st = sddmm_csr_bf16(p->h, M, M, N, p->nnz, p->rowptr, p->gcol,
                        static_cast<const uint16_t *>(dY), N, static_cast<const uint16_t *>(X), N,
                        dval, s);
  } else {
    st = spmm_csr_f32_planned(p->fwd, M, M, N, p->nnz, p->rowptr, p->gcol, p->val,
                              static_cast<const float *>(X), N, Y, N, nullptr, s);
    if (st) return st;
    st = spmm_csr_f32_planned(p->bwd, M, M, N, p->nnz, p->trowptr, p->tcol, p->tval,
                              static_cast<const float *>(dY), N, dX, N, nullptr, s);
    if (st) return st;
    st = sddmm_csr_f32(p->h, M, M, N, p->nnz, p->rowptr, p->gcol, static_cast<const float *>(dY),
                       N, static_cast<const float *>(X), N, dval, s);
  }
  return st;
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
      hipMalloc(&p->tval, sizeof(float) * nn) != hipSuccess)
    return fail(SPMM_ERR_ALLOC, "device allocation failed");
  // Offsets to the device (the handle workspace is free until the transpose).
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
  int st = csr_transpose(h, M, M, nnz, rowptr, p->gcol, val, p->trowptr, p->tcol, p->tval, nullptr,
                         stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  st = spmm_csr_analyze(h, M, nnz, rowptr, &p->fwd, stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  st = spmm_csr_analyze(h, M, nnz, p->trowptr, &p->bwd, stream);
  if (st) { spmm_batch_plan_destroy(p); return st; }
  if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess)
    return fail(SPMM_ERR_HIP, "capture stream");
  *out = p;
  return SPMM_OK;
}

int spmm_batch_fwd_bwd(spmm_batch_plan_t p, const void *X, const void *dY, float *Y, float *dX,
                       float *dval, int use_graph, void *stream) {
  SPMM_CHECK_ARG(p, "null plan");
  SPMM_CHECK_ARG(X && dY && Y && dX && (dval || p->nnz == 0), "null operand");
  hipStream_t s = as_stream(stream);
  if (!use_graph) return enqueue_step(p, X, dY, Y, dX, dval, s);
  if (!p->exec || p->gX != X || p->gdY != dY || p->gY != Y || p->gdX != dX || p->gdval != dval) {
    destroy_graph(p);
    SPMM_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
    int st = enqueue_step(p, X, dY, Y, dX, dval, p->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->cap, &g);
    if (st) {
      if (g) (void)hipGraphDestroy(g);
      return st;
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
