// TEST INFRASTRUCTURE ONLY: C entry points over the reference's own host
// CSR/CSC helpers (backend/readMtx.hpp, compiled in place from the
// reference tree by oracle/Makefile into oracle/_ref/). Used to pin the
// oracle's csr2csc restatement against the reference code itself.
#include "readMtx.hpp"

extern "C" {

// backend/readMtx.hpp:319-345
void ref_csr2csc(int M, int K, const int *rowptr, const int *colind, const float *val,
                 int *colptr, int *rowind, float *cval) {
  csr2csc<float>(colptr, rowind, cval, rowptr, colind, val, M, K);
}

// backend/readMtx.hpp:351-422 then :246-304, as test_spmm_full.cu:99-113
// runs them (directed = 0, no mtxinfo, no dat_name): sizes first...
int ref_read_mtx_info(const char *path, int *M, int *K, int *nnz) {
  std::vector<int> r, c;
  std::vector<float> v;
  return readMtx<float>(path, &r, &c, &v, M, K, nnz, 0, false, nullptr);
}

// ...then the CSR arrays (caller-allocated from the sizes above).
int ref_read_mtx_csr(const char *path, int *rowptr, int *colind, float *val) {
  std::vector<int> r, c;
  std::vector<float> v;
  int M = 0, K = 0, nnz = 0;
  const int rc = readMtx<float>(path, &r, &c, &v, &M, &K, &nnz, 0, false, nullptr);
  coo2csr<float>(rowptr, colind, val, r, c, v, M, K);
  return rc;
}

// backend/readMtx.hpp:246-304 (sorts the COO tuples by (row, col) first)
void ref_coo2csr(int M, int K, int nvals, const int *rows, const int *cols, const float *vals,
                 int *rowptr, int *colind, float *val) {
  std::vector<int> r(rows, rows + nvals), c(cols, cols + nvals);
  std::vector<float> v(vals, vals + nvals);
  coo2csr<float>(rowptr, colind, val, r, c, v, M, K);
}

}
