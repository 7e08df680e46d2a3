// C++ host harness over the C-ABI, in the shape of the reference's
// test_spmm_full.cu main (test_spmm_full.cu:474-514): read a Matrix-Market
// graph (readMtxCSR, :99-125) or generate the seeded power-law graph, run the
// CSR FP32 SpMM that replaces the cuSPARSE call (:212-342), convert the graph
// to B2SR-4 (CSR2B2SR, :127-169) and run the FBF SpMM (:344-369) on random
// +-1 B, then compare the two outputs exactly (verifyResultFull, :412-431)
// and print both times as the reference does (:505). Unlike the reference
// it also checks every element against a host integer product, seeds B,
// and returns a nonzero exit status on any mismatch or library error.
//
//   spmm_harness <A.mtx> <hidden>
//   spmm_harness --chung-lu <rows> <nnz> <seed> <hidden>
//   spmm_harness --host-checks
//
// --host-checks runs without a GPU: argument validation of the C-ABI (each
// rejected call returns its status code before touching the device) and the
// host-side Matrix-Market reader on a file it writes.
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "spmm_mi355x.h"

extern "C" {
// libspmm_gen.so: the seeded Chung-Lu generator of the synthetic workloads
struct gen_graph_s;
gen_graph_s *gen_chung_lu(int n, int64_t target_nnz, double gamma, double max_deg,
                          uint64_t seed);
int64_t gen_graph_nnz(const gen_graph_s *g);
int gen_graph_copy(const gen_graph_s *g, int *rowptr, int *colind);
void gen_graph_free(gen_graph_s *g);
}

namespace {

int g_checks = 0, g_failed = 0;

void expect(bool ok, const char *what) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    fprintf(stderr, "host check failed: %s (last error: %s)\n", what, spmm_last_error());
  }
}

// a call that must be refused with `want` and leave a message
void expect_status(int got, int want, const char *what) {
  expect(got == want && spmm_last_error() && spmm_last_error()[0], what);
}

int host_checks() {
  void *const bad = reinterpret_cast<void *>(16);  // never dereferenced
  int *const ibad = static_cast<int *>(bad);
  float *const fbad = static_cast<float *>(bad);
  uint32_t *const wbad = static_cast<uint32_t *>(bad);
  expect(spmm_version() >= 100, "spmm_version");
  expect(strcmp(spmm_status_string(SPMM_OK), "SPMM_OK") == 0, "status string 0");
  expect(strcmp(spmm_status_string(SPMM_ERR_MISMATCH), "SPMM_ERR_MISMATCH") == 0, "status string 5");
  // null handle / plan / output pointer
  expect_status(spmm_create(nullptr), SPMM_ERR_INVALID_ARG, "spmm_create(NULL)");
  expect_status(spmm_csr_f32(nullptr, 4, 4, 4, 4, ibad, ibad, fbad, fbad, 4, fbad, 4, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_csr_f32 without a handle");
  expect_status(spmm_csr_f32_planned(nullptr, 4, 4, 4, 4, ibad, ibad, fbad, fbad, 4, fbad, 4,
                                     nullptr, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_csr_f32_planned without a plan");
  int nb = 0;
  expect_status(csr2b2sr_nnzb(nullptr, 4, 4, 4, ibad, ibad, 4, ibad, &nb, nullptr),
                SPMM_ERR_INVALID_ARG, "csr2b2sr_nnzb without a handle");
  expect_status(spmm_batch_plan_create(nullptr, 1, nullptr, nullptr, nullptr, nullptr, nullptr, 8,
                                       0, nullptr, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_batch_plan_create without a handle");
  // unsupported tile size, negative and inconsistent sizes, short strides
  expect_status(spmm_b2sr_f32(nullptr, 5, 4, 4, 4, ibad, ibad, bad, fbad, 4, fbad, 4, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_b2sr_f32 with T = 5");
  expect_status(bmm32_xnor_f32(-1, 1, 1, wbad, wbad, fbad, 1, nullptr), SPMM_ERR_INVALID_ARG,
                "bmm32_xnor_f32 with M < 0");
  expect_status(bmm32_xnor_f32(4, 4, 1, nullptr, wbad, fbad, 4, nullptr), SPMM_ERR_INVALID_ARG,
                "bmm32_xnor_f32 with A null");
  expect_status(bmm32_xnor_bin(4, 4, -3, wbad, wbad, wbad, nullptr), SPMM_ERR_INVALID_ARG,
                "bmm32_xnor_bin with Kbits < 0");
  expect_status(pack_sign_rows_u32(4, 40, fbad, 39, wbad, nullptr), SPMM_ERR_INVALID_ARG,
                "pack_sign_rows_u32 with ldx < W");
  expect_status(sddmm_csr_f32(nullptr, 2, 2, 4, 2, ibad, ibad, fbad, 2, fbad, 4, fbad, nullptr),
                SPMM_ERR_INVALID_ARG, "sddmm_csr_f32 with lddy < N");
  expect_status(spmm_b2sr_f32_noalloc(3, 4, 4, 4, ibad, ibad, bad, fbad, 4, fbad, 4, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_b2sr_f32_noalloc with T = 3");
  expect_status(spmm_b2sr_signpacked_i32(nullptr, 4, 4, 4, 64, ibad, ibad, bad, wbad, 1,
                                         static_cast<int32_t *>(bad), 64, nullptr),
                SPMM_ERR_INVALID_ARG, "spmm_b2sr_signpacked_i32 with ldbw < ceil(N/32)");
  // zero-sized calls are no-ops
  expect(bmm32_xnor_f32(0, 5, 1, nullptr, nullptr, nullptr, 5, nullptr) == SPMM_OK, "empty bmm");
  expect(pack_sign_rows_u32(0, 0, nullptr, 0, nullptr, nullptr) == SPMM_OK, "empty pack");
  expect(spmm_b2sr_f32(nullptr, 4, 0, 4, 4, nullptr, nullptr, nullptr, nullptr, 4, nullptr, 4,
                       nullptr) == SPMM_OK, "empty B2SR SpMM");
  // the host reader: a 4 x 5 real general file with entries out of order and
  // a duplicate pair (kept, in file order, as readMtx + coo2csr keep it)
  char path[] = "/tmp/spmm_harness_XXXXXX.mtx";
  const int fd = mkstemps(path, 4);
  expect(fd >= 0, "temporary file");
  if (fd >= 0) {
    FILE *f = fdopen(fd, "w");
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%% comment\n4 5 6\n"
               "3 2 1.5\n1 5 -2\n1 1 4\n4 3 0.25\n3 2 7\n2 4 1\n");
    fclose(f);
    int M = 0, K = 0, nnz = 0, kind = -1;
    expect(spmm_mtx_read_info(path, &M, &K, &nnz, &kind) == SPMM_OK && M == 4 && K == 5 &&
               nnz == 6 && kind == 2,
           "spmm_mtx_read_info sizes");
    std::vector<int> rp(M + 1), ci(nnz);
    std::vector<float> v(nnz);
    expect(spmm_mtx_read_csr(path, M, K, nnz, rp.data(), ci.data(), v.data()) == SPMM_OK,
           "spmm_mtx_read_csr");
    const int rp_want[] = {0, 2, 3, 5, 6}, ci_want[] = {0, 4, 3, 1, 1, 2};
    const float v_want[] = {4.f, -2.f, 1.f, 1.5f, 7.f, 0.25f};
    bool same = true;
    for (int i = 0; i <= M; ++i) same &= rp[i] == rp_want[i];
    for (int e = 0; e < nnz; ++e) same &= ci[e] == ci_want[e] && v[e] == v_want[e];
    expect(same, "CSR from the file (row order, column order, duplicate pair in file order)");
    expect_status(spmm_mtx_read_csr(path, M, K, nnz + 1, rp.data(), ci.data(), v.data()),
                  SPMM_ERR_MISMATCH, "spmm_mtx_read_csr with a wrong nnz");
    remove(path);
  }
  int M = 0, K = 0, nnz = 0, kind = 0;
  expect(spmm_mtx_read_info("/nonexistent/graph.mtx", &M, &K, &nnz, &kind) != SPMM_OK,
         "missing file is an error");
  printf("host checks: %d run, %d failed\n", g_checks, g_failed);
  return g_failed ? 1 : 0;
}

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
      return 2;                                                                \
    }                                                                          \
  } while (0)
#define SPCHK(x)                                                               \
  do {                                                                         \
    int s_ = (x);                                                              \
    if (s_ != SPMM_OK) {                                                       \
      fprintf(stderr, "%s: %s (%s)\n", #x, spmm_status_string(s_), spmm_last_error()); \
      return 2;                                                                \
    }                                                                          \
  } while (0)

// Mean ms per call of `reps` back-to-back calls between two events, after
// one warm call.
template <typename F>
float time_calls(F f, int reps, int *status) {
  hipEvent_t a, b;
  float ms = -1.f;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return ms;
  *status = f();
  if (*status == SPMM_OK && hipEventRecord(a, nullptr) == hipSuccess) {
    for (int r = 0; r < reps && *status == SPMM_OK; ++r) *status = f();
    if (hipEventRecord(b, nullptr) == hipSuccess && hipEventSynchronize(b) == hipSuccess)
      (void)hipEventElapsedTime(&ms, a, b);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

int run_gpu(int M, int K, std::vector<int> &rowptr, std::vector<int> &colind, int N) {
  const int nnz = rowptr[M];
  printf("graph: %d x %d, %d nnz, hidden %d\n", M, K, nnz, N);
  // binary A (the reference forces values to 1.0, csr2bsr_batch_bsrbmv.cu:296-297)
  std::vector<float> val(nnz, 1.f), B(static_cast<size_t>(K) * N);
  uint64_t s = 0x5EED1234ull;
  for (auto &x : B) x = (splitmix(s) >> 63) ? 1.f : -1.f;
  HIPCHK(hipSetDevice(0));
  int *d_rp, *d_ci, *d_brp, *d_bci = nullptr;
  float *d_v, *d_B, *d_Ccsr, *d_Cb, *d_Cp;
  unsigned char *d_words = nullptr;
  const size_t cb = sizeof(float) * M * static_cast<size_t>(N);
  const int MB = (M + 3) / 4;
  HIPCHK(hipMalloc(&d_rp, sizeof(int) * (M + 1)));
  HIPCHK(hipMalloc(&d_ci, sizeof(int) * (nnz > 0 ? nnz : 1)));
  HIPCHK(hipMalloc(&d_v, sizeof(float) * (nnz > 0 ? nnz : 1)));
  HIPCHK(hipMalloc(&d_B, sizeof(float) * B.size()));
  HIPCHK(hipMalloc(&d_Ccsr, cb));
  HIPCHK(hipMalloc(&d_Cb, cb));
  HIPCHK(hipMalloc(&d_Cp, cb));
  HIPCHK(hipMalloc(&d_brp, sizeof(int) * (MB + 1)));
  HIPCHK(hipMemcpy(d_rp, rowptr.data(), sizeof(int) * (M + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_ci, colind.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_v, val.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_B, B.data(), sizeof(float) * B.size(), hipMemcpyHostToDevice));
  spmm_handle_t h;
  SPCHK(spmm_create(&h));
  int st = SPMM_OK;
  // CSR FP32 SpMM (the cuSPARSE call's replacement), plan-free drop-in
  const float csr_ms = time_calls(
      [&] { return spmm_csr_f32(h, M, K, N, nnz, d_rp, d_ci, d_v, d_B, N, d_Ccsr, N, nullptr); }, 10,
      &st);
  SPCHK(st);
  // CSR -> B2SR-4 on the device, then the FBF SpMM (drop-in and planned)
  int nb = 0;
  SPCHK(csr2b2sr_nnzb(h, M, K, nnz, d_rp, d_ci, 4, d_brp, &nb, nullptr));
  HIPCHK(hipMalloc(&d_bci, sizeof(int) * (nb > 0 ? nb : 1)));
  HIPCHK(hipMalloc(&d_words, 4 * static_cast<size_t>(nb > 0 ? nb : 1)));
  SPCHK(csr2b2sr(h, M, K, nnz, d_rp, d_ci, 4, d_brp, d_bci, d_words, nullptr));
  const float b2sr_ms = time_calls(
      [&] { return spmm_b2sr_f32(h, 4, M, K, N, d_brp, d_bci, d_words, d_B, N, d_Cb, N, nullptr); },
      10, &st);
  SPCHK(st);
  spmm_csr_plan_t plan;
  SPCHK(spmm_b2sr_analyze(h, 4, M, d_brp, &plan, nullptr));
  const float planned_ms = time_calls(
      [&] {
        return spmm_b2sr_f32_planned(plan, 4, M, K, N, d_brp, d_bci, d_words, d_B, N, d_Cp, N,
                                     nullptr);
      },
      10, &st);
  SPCHK(st);
  HIPCHK(hipDeviceSynchronize());
  std::vector<float> Ccsr(static_cast<size_t>(M) * N), Cb(Ccsr.size()), Cp(Ccsr.size());
  HIPCHK(hipMemcpy(Ccsr.data(), d_Ccsr, cb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Cb.data(), d_Cb, cb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(Cp.data(), d_Cp, cb, hipMemcpyDeviceToHost));
  // the reference's check (B2SR == CSR, compared with !=) plus an integer
  // product on the host for every element
  long long mismatch = 0, wrong = 0;
  std::vector<long long> acc(N);
  for (int i = 0; i < M; ++i) {
    for (int k = 0; k < N; ++k) acc[k] = 0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
      for (int k = 0; k < N; ++k) acc[k] += static_cast<long long>(B[static_cast<size_t>(colind[e]) * N + k]);
    for (int k = 0; k < N; ++k) {
      const size_t x = static_cast<size_t>(i) * N + k;
      mismatch += Cb[x] != Ccsr[x] || Cp[x] != Ccsr[x];
      wrong += Ccsr[x] != static_cast<float>(acc[k]);
    }
  }
  const bool pass = mismatch == 0 && wrong == 0;
  printf(pass ? "TEST PASSED: %d\n" : "TEST FAILED: %d\n", pass ? 1 : 0);
  printf("mismatches B2SR vs CSR: %lld, CSR vs host integer product: %lld\n", mismatch, wrong);
  printf("b2sr: %.3f, b2sr-planned: %.3f, csr-float: %.3f (ms per call)\n", b2sr_ms, planned_ms,
         csr_ms);
  spmm_csr_plan_destroy(plan);
  spmm_destroy(h);
  for (void *p : {static_cast<void *>(d_rp), static_cast<void *>(d_ci), static_cast<void *>(d_v),
                  static_cast<void *>(d_B), static_cast<void *>(d_Ccsr), static_cast<void *>(d_Cb),
                  static_cast<void *>(d_Cp), static_cast<void *>(d_brp), static_cast<void *>(d_bci),
                  static_cast<void *>(d_words)})
    (void)hipFree(p);
  return pass ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 2 && std::string(argv[1]) == "--host-checks") return host_checks();
  std::vector<int> rowptr, colind;
  int M = 0, K = 0, N = 0;
  if (argc == 3) {
    int nnz = 0, kind = 0;
    if (spmm_mtx_read_info(argv[1], &M, &K, &nnz, &kind) != SPMM_OK) {
      fprintf(stderr, "%s\n", spmm_last_error());
      return 2;
    }
    rowptr.resize(M + 1);
    colind.resize(nnz > 0 ? nnz : 1);
    if (spmm_mtx_read_csr(argv[1], M, K, nnz, rowptr.data(), colind.data(), nullptr) != SPMM_OK) {
      fprintf(stderr, "%s\n", spmm_last_error());
      return 2;
    }
    N = atoi(argv[2]);
  } else if (argc == 6 && std::string(argv[1]) == "--chung-lu") {
    M = K = atoi(argv[2]);
    const long long target = atoll(argv[3]);
    // degree cut-off sqrt(nnz), the generator's default for the workloads
    gen_graph_s *g = gen_chung_lu(M, target, 2.1, std::sqrt(static_cast<double>(target)),
                                  strtoull(argv[4], nullptr, 0));
    if (!g) {
      fprintf(stderr, "gen_chung_lu rejected its arguments\n");
      return 2;
    }
    rowptr.resize(M + 1);
    colind.resize(gen_graph_nnz(g) > 0 ? gen_graph_nnz(g) : 1);
    const int rc = gen_graph_copy(g, rowptr.data(), colind.data());
    gen_graph_free(g);
    if (rc != 0) return 2;
    N = atoi(argv[5]);
  } else {
    fprintf(stderr,
            "usage: %s <A.mtx> <hidden> | --chung-lu <rows> <nnz> <seed> <hidden> | --host-checks\n",
            argv[0]);
    return 2;
  }
  if (M != K) {
    fprintf(stderr, "the harness compares square graphs (the reference assumes nBrows = nrows)\n");
    return 2;
  }
  if (N <= 0) {
    fprintf(stderr, "hidden size must be positive\n");
    return 2;
  }
  return run_gpu(M, K, rowptr, colind, N);
}
