// Matrix Market ingest (host): the reader the reference harness runs before
// every product (test_spmm_full.cu:99-125): readMtx (backend/readMtx.hpp:
// 351-422, banner and size line as in backend/mmio.hpp:128-226) followed by
// coo2csr (readMtx.hpp:246-304). Semantics kept as the reference has them:
//   - "%%MatrixMarket matrix coordinate <type> <storage>", tokens after the
//     first compared case-insensitively; comment lines start with '%';
//   - 1-based indices become 0-based; pattern entries read as 1.0, integer
//     and real values as float (strtof, the rounding of fscanf "%f");
//   - symmetric / skew files are NOT expanded and duplicate pairs are NOT
//     merged (readMtx.hpp:413 leaves removeSelfloop disabled);
//   - entries are ordered by (row, col). The reference's std::sort leaves
//     the order of duplicate pairs unspecified; here they keep file order.
// Where the reference prints and carries on (too few entries, an index out
// of range) or reads no values (complex files), this returns an error.
// The whole file is read into memory once and parsed with strtol / strtof.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spmm_mi355x.h"

namespace spmm {
void set_error(const std::string &msg);
}

namespace {

struct MtxHeader {
  int M = 0, K = 0, nnz = 0;
  int type = 0;  // 0 pattern, 1 integer, 2 real
  int sym = 0;   // 0 general, 1 symmetric, 2 skew-symmetric
  size_t body = 0;  // offset of the first entry
};

int fail(int status, const std::string &msg) {
  spmm::set_error(msg);
  return status;
}

// Reads `path` whole; a trailing NUL ends the buffer for strtol / strtof.
int slurp(const char *path, std::vector<char> *buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(SPMM_ERR_INVALID_ARG, std::string("cannot open ") + path);
  std::vector<char> b;
  char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + n);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err) return fail(SPMM_ERR_INVALID_ARG, std::string("read error on ") + path);
  b.push_back('\0');
  buf->swap(b);
  return SPMM_OK;
}

std::string lower(std::string s) {
  for (char &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

int parse_header(const std::vector<char> &buf, MtxHeader *h) {
  const char *p = buf.data();
  const char *eol = std::strchr(p, '\n');
  const std::string line(p, eol ? eol : p + std::strlen(p));
  char tok[5][128];
  if (std::sscanf(line.c_str(), "%127s %127s %127s %127s %127s", tok[0], tok[1], tok[2], tok[3],
                  tok[4]) != 5)
    return fail(SPMM_ERR_INVALID_ARG, "Matrix Market banner: expected 5 fields");
  if (std::strncmp(tok[0], "%%MatrixMarket", 14) != 0)
    return fail(SPMM_ERR_INVALID_ARG, "not a Matrix Market file (no %%MatrixMarket banner)");
  if (lower(tok[1]) != "matrix") return fail(SPMM_ERR_UNSUPPORTED, "banner object is not 'matrix'");
  const std::string fmt = lower(tok[2]), type = lower(tok[3]), sym = lower(tok[4]);
  if (fmt != "coordinate")
    return fail(SPMM_ERR_UNSUPPORTED, "only coordinate (sparse) Matrix Market files are read");
  if (type == "pattern") h->type = 0;
  else if (type == "integer") h->type = 1;
  else if (type == "real") h->type = 2;
  else return fail(SPMM_ERR_UNSUPPORTED, "Matrix Market value type '" + type + "' not supported");
  if (sym == "general") h->sym = 0;
  else if (sym == "symmetric") h->sym = 1;
  else if (sym == "skew-symmetric") h->sym = 2;
  else return fail(SPMM_ERR_UNSUPPORTED, "Matrix Market storage '" + sym + "' not supported");
  // size line: the first line after the banner not starting with '%'
  // (a blank one is skipped, like mm_read_mtx_crd_size's fscanf retry)
  const char *q = eol ? eol + 1 : p + std::strlen(p);
  while (*q) {
    const char *e = std::strchr(q, '\n');
    const char *next = e ? e + 1 : q + std::strlen(q);
    if (*q != '%') {
      char *end = nullptr;
      errno = 0;
      const long m = std::strtol(q, &end, 10);
      if (end != q) {
        const char *a = end;
        const long k = std::strtol(a, &end, 10);
        if (end == a) return fail(SPMM_ERR_INVALID_ARG, "Matrix Market size line: expected 3 integers");
        const char *b = end;
        const long z = std::strtol(b, &end, 10);
        if (end == b) return fail(SPMM_ERR_INVALID_ARG, "Matrix Market size line: expected 3 integers");
        if (errno || m < 0 || k < 0 || z < 0 || m > INT_MAX || k > INT_MAX || z > INT_MAX)
          return fail(SPMM_ERR_INVALID_ARG, "Matrix Market size out of range");
        h->M = static_cast<int>(m);
        h->K = static_cast<int>(k);
        h->nnz = static_cast<int>(z);
        h->body = static_cast<size_t>(end - buf.data());
        return SPMM_OK;
      }
    }
    q = next;
  }
  return fail(SPMM_ERR_INVALID_ARG, "Matrix Market file has no size line");
}

int read_header(const char *path, std::vector<char> *buf, MtxHeader *h) {
  if (!path) return fail(SPMM_ERR_INVALID_ARG, "path is null");
  int st = slurp(path, buf);
  if (st) return st;
  return parse_header(*buf, h);
}

}  // namespace

extern "C" {

int spmm_mtx_read_info(const char *path, int *M, int *K, int *nnz, int *kind) {
  std::vector<char> buf;
  MtxHeader h;
  const int st = read_header(path, &buf, &h);
  if (st) return st;
  if (M) *M = h.M;
  if (K) *K = h.K;
  if (nnz) *nnz = h.nnz;
  if (kind) *kind = h.type + 4 * h.sym;
  return SPMM_OK;
}

int spmm_mtx_read_csr(const char *path, int M, int K, int nnz, int *rowptr, int *colind,
                      float *val) {
  if (!rowptr || (nnz > 0 && !colind)) return fail(SPMM_ERR_INVALID_ARG, "null rowptr or colind");
  std::vector<char> buf;
  MtxHeader h;
  int st = read_header(path, &buf, &h);
  if (st) return st;
  if (h.M != M || h.K != K || h.nnz != nnz)
    return fail(SPMM_ERR_MISMATCH, "sizes differ from the file's (" + std::to_string(h.M) + " " +
                                       std::to_string(h.K) + " " + std::to_string(h.nnz) + ")");
  // entries in file order
  std::vector<int> r(static_cast<size_t>(nnz)), c(static_cast<size_t>(nnz));
  std::vector<float> v(val ? static_cast<size_t>(nnz) : 0);
  const char *p = buf.data() + h.body;
  for (int i = 0; i < nnz; ++i) {
    char *end = nullptr;
    const long ri = std::strtol(p, &end, 10);
    if (end == p)
      return fail(SPMM_ERR_INVALID_ARG, "Matrix Market file ends after " + std::to_string(i) +
                                            " of " + std::to_string(nnz) + " entries");
    p = end;
    const long ci = std::strtol(p, &end, 10);
    if (end == p) return fail(SPMM_ERR_INVALID_ARG, "entry " + std::to_string(i) + ": no column");
    p = end;
    if (ri < 1 || ri > M || ci < 1 || ci > K)
      return fail(SPMM_ERR_INVALID_ARG, "entry " + std::to_string(i) + ": index out of range");
    r[i] = static_cast<int>(ri - 1);
    c[i] = static_cast<int>(ci - 1);
    float x = 1.0f;
    if (h.type == 1) {
      const long iv = std::strtol(p, &end, 10);
      if (end == p) return fail(SPMM_ERR_INVALID_ARG, "entry " + std::to_string(i) + ": no value");
      x = static_cast<float>(static_cast<int>(iv));
      p = end;
    } else if (h.type == 2) {
      x = std::strtof(p, &end);
      if (end == p) return fail(SPMM_ERR_INVALID_ARG, "entry " + std::to_string(i) + ": no value");
      p = end;
    }
    if (val) v[i] = x;
  }
  // stable counting sort by row, then a stable sort by column within rows
  std::fill(rowptr, rowptr + M + 1, 0);
  for (int i = 0; i < nnz; ++i) ++rowptr[r[i] + 1];
  for (int i = 0; i < M; ++i) rowptr[i + 1] += rowptr[i];
  std::vector<int> pos(rowptr, rowptr + M);
  std::vector<int> order(static_cast<size_t>(nnz));
  for (int i = 0; i < nnz; ++i) order[pos[r[i]]++] = i;
  for (int row = 0; row < M; ++row) {
    auto b = order.begin() + rowptr[row], e = order.begin() + rowptr[row + 1];
    std::stable_sort(b, e, [&](int x, int y) { return c[x] < c[y]; });
  }
  for (int i = 0; i < nnz; ++i) {
    colind[i] = c[order[i]];
    if (val) val[i] = v[order[i]];
  }
  return SPMM_OK;
}

}  // extern "C"
