// Handle lifetime, status strings and the grow-only workspace.
#include "common.hpp"

#include <new>

namespace spmm {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

void *workspace(spmm_handle_s *h, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (h->ws_bytes >= bytes) return h->ws;
  if (h->ws) {
    // Earlier asynchronous work may still read the old buffer.
    if (hipDeviceSynchronize() != hipSuccess) {
      set_error("workspace: hipDeviceSynchronize failed");
      return nullptr;
    }
    (void)hipFree(h->ws);
    h->ws = nullptr;
    h->ws_bytes = 0;
  }
  size_t want = bytes + bytes / 4;  // headroom against repeated growth
  if (hipMalloc(&h->ws, want) != hipSuccess) {
    h->ws = nullptr;
    set_error("workspace: hipMalloc of " + std::to_string(want) +
              " bytes failed");
    return nullptr;
  }
  h->ws_bytes = want;
  return h->ws;
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
  if (h->ws) (void)hipFree(h->ws);
  if (h->pinned) (void)hipHostFree(h->pinned);
  if (h->b2sr_state) (void)hipFree(h->b2sr_state);
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
