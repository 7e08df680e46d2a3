// Handle lifetime, status strings and the grow-only workspace.
#include "common.hpp"

#include <new>

namespace spmm {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// True when `s` is being captured into a graph (growing a buffer then would
// allocate inside the capture and leave the graph pointing at a buffer a
// later call could outgrow).
static bool capturing(hipStream_t s) {
  if (!s) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

void *workspace(spmm_handle_s *h, size_t bytes, hipStream_t s) {
  if (bytes == 0) bytes = 256;
  if (h->ws_bytes >= bytes) return h->ws;
  if (capturing(s)) {
    set_error("workspace: the handle's workspace must grow, which cannot happen during graph "
              "capture; make one call of this size before capturing");
    return nullptr;
  }
  size_t want = bytes + bytes / 4;  // headroom against repeated growth
  void *p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    set_error("workspace: hipMalloc of " + std::to_string(want) + " bytes failed");
    return nullptr;
  }
  if (h->ws) h->retired.push_back(h->ws);  // earlier work may still use it
  h->ws = p;
  h->ws_bytes = want;
  return h->ws;
}

int grow_ints(spmm_handle_s *h, int **buf, size_t *cap, size_t ints, hipStream_t s) {
  if (*cap >= ints) return SPMM_OK;
  if (capturing(s))
    SPMM_FAIL(SPMM_ERR_UNSUPPORTED,
              "plan-free state must grow, which cannot happen during graph capture; make one "
              "call at the largest size before capturing");
  const size_t want = ints + ints / 4 + 64;
  int *p = nullptr;
  if (hipMalloc(&p, sizeof(int) * want) != hipSuccess)
    SPMM_FAIL(SPMM_ERR_ALLOC, "plan-free state allocation failed");
  if (hipMemset(p, 0, sizeof(int) * want) != hipSuccess) {
    (void)hipFree(p);
    SPMM_FAIL(SPMM_ERR_HIP, "plan-free state initialisation failed");
  }
  if (*buf) h->retired.push_back(*buf);  // earlier work may still use it
  *buf = p;
  *cap = want;
  return SPMM_OK;
}

void *pinned(spmm_handle_s *h, size_t bytes) {
  if (h->pinned_bytes >= bytes) return h->pinned;
  if (h->pinned) (void)hipHostFree(h->pinned);
  h->pinned = nullptr;
  h->pinned_bytes = 0;
  if (hipHostMalloc(&h->pinned, bytes, hipHostMallocDefault) != hipSuccess) {
    h->pinned = nullptr;
    set_error("pinned: hipHostMalloc failed");
    return nullptr;
  }
  h->pinned_bytes = bytes;
  return h->pinned;
}

}  // namespace spmm

extern "C" {

int spmm_create(spmm_handle_t *out) {
  SPMM_CHECK_ARG(out != nullptr, "out is null");
  *out = nullptr;
  auto *h = new (std::nothrow) spmm_handle_s();
  if (!h) SPMM_FAIL(SPMM_ERR_ALLOC, "host allocation failed");
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    SPMM_FAIL(SPMM_ERR_HIP, "hipGetDevice failed (no GPU visible?)");
  }
  *out = h;
  return SPMM_OK;
}

int spmm_destroy(spmm_handle_t h) {
  if (!h) return SPMM_OK;
  // work queued through the handle may still use its buffers
  (void)hipDeviceSynchronize();
  if (h->ws) (void)hipFree(h->ws);
  if (h->pinned) (void)hipHostFree(h->pinned);
  if (h->list_state) (void)hipFree(h->list_state);
  if (h->tstruct) (void)hipFree(h->tstruct);
  if (h->slice_off) (void)hipFree(h->slice_off);
  for (void *p : h->retired) (void)hipFree(p);
  delete h;
  return SPMM_OK;
}

const char *spmm_status_string(int status) {
  switch (status) {
    case SPMM_OK: return "SPMM_OK";
    case SPMM_ERR_INVALID_ARG: return "SPMM_ERR_INVALID_ARG";
    case SPMM_ERR_UNSUPPORTED: return "SPMM_ERR_UNSUPPORTED";
    case SPMM_ERR_HIP: return "SPMM_ERR_HIP";
    case SPMM_ERR_ALLOC: return "SPMM_ERR_ALLOC";
    case SPMM_ERR_MISMATCH: return "SPMM_ERR_MISMATCH";
    default: return "SPMM_ERR_UNKNOWN";
  }
}

const char *spmm_last_error(void) { return spmm::g_last_error.c_str(); }

int spmm_version(void) { return 100; }

}  // extern "C"
