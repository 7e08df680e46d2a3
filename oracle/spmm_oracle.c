/*
 * spmm_oracle.c -- CPU restatement of the reference's SpMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this library, and only as the
 * checker (or, for the *_omp functions, as the timed CPU baseline). The
 * product library never calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root). The reference itself
 * is CUDA + cuSPARSE and cannot be built or run here; see the tests for how
 * each function is pinned (closed forms, brute force, library routines,
 * hand-derived vectors).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------- BF16 --
 * Round-to-nearest-even FP32 -> BF16 bit pattern (IEEE-754 definition:
 * keep the top 16 bits, round on the dropped 16 with ties to even; NaN is
 * kept quiet). No reference counterpart (BF16 is a new form). */
uint16_t orc_f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
    return (uint16_t)((u >> 16) | 0x0040u);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

void orc_f32_to_bf16_array(int64_t n, const float *in, uint16_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = orc_f32_to_bf16(in[i]);
}

static inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* ----------------------------------------------------------------- CSR --
 * C = A * B, alpha = 1, beta = 0, A CSR zero-based int32 with float values,
 * B row-major [K][ldb], C row-major [M][ldc] in double. This is the
 * operation of the cuSPARSE SpMM in evalCSRSpmmFloatCuSPARSE
 * (test_spmm_full.cu:212-342: cusparseCreateCsr :281-284, cusparseSpMM
 * :303-307), restated with row-major B and C as the FBF kernels use
 * (backend/spmm.cu:938,951). Accumulation in double, entries in stored order. */
void orc_csr_spmm_f64(int M, int N, const int *rowptr, const int *colind,
                      const float *val, const float *B, int ldb, double *C,
                      int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    double *c = C + (int64_t)i * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const double v = val[e];
      const float *b = B + (int64_t)colind[e] * ldb;
      for (int k = 0; k < N; ++k) c[k] += v * (double)b[k];
    }
  }
}

/* Same with B given as BF16 bit patterns (widened exactly). */
void orc_csr_spmm_f64_bf16(int M, int N, const int *rowptr, const int *colind,
                           const float *val, const uint16_t *B, int ldb,
                           double *C, int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    double *c = C + (int64_t)i * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const double v = val[e];
      const uint16_t *b = B + (int64_t)colind[e] * ldb;
      for (int k = 0; k < N; ++k) c[k] += v * (double)bf16_to_f32(b[k]);
    }
  }
}

/* The timed CPU baseline: row-parallel FP32 CSR SpMM, dynamic schedule in
 * chunks of 64 rows, inner loop over N left to the compiler's vectoriser.
 * Same operation as orc_csr_spmm_f64 in FP32. */
void orc_csr_spmm_f32_omp(int M, int N, const int *rowptr, const int *colind,
                          const float *val, const float *B, int ldb, float *C,
                          int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    float *restrict c = C + (int64_t)i * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0.f;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const float v = val[e];
      const float *restrict b = B + (int64_t)colind[e] * ldb;
      for (int k = 0; k < N; ++k) c[k] += v * b[k];
    }
  }
}

/* Same with BF16 B, FP32 accumulation (baseline for the BF16 form). */
void orc_csr_spmm_bf16_omp(int M, int N, const int *rowptr, const int *colind,
                           const float *val, const uint16_t *B, int ldb,
                           float *C, int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    float *restrict c = C + (int64_t)i * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0.f;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const float v = val[e];
      const uint16_t *restrict b = B + (int64_t)colind[e] * ldb;
      for (int k = 0; k < N; ++k) c[k] += v * bf16_to_f32(b[k]);
    }
  }
}

/* Integer CSR SpMM with all values 1: C[i][k] = sum over stored (i,j) of
 * B[j][k]. This is the exact known answer the reference harness checks the
 * B2SR kernel against (cuSPARSE with +-1 B and binary A, compared with !=,
 * test_spmm_full.cu:412-431, B drawn at :487-492). */
void orc_csr_spmm_i64(int M, int N, const int *rowptr, const int *colind,
                      const int32_t *B, int ldb, int64_t *C, int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    int64_t *c = C + (int64_t)i * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const int32_t *b = B + (int64_t)colind[e] * ldb;
      for (int k = 0; k < N; ++k) c[k] += b[k];
    }
  }
}

/* ----------------------------------------------------------- csr2b2sr --
 * Block rows of T CSR rows; the last block row is padded with empty rows
 * (csr2bsr_batch_bsrbmv.cu:240,271-276). Within a block row the nonzero
 * T x T blocks are the distinct col/T of its entries, ascending, as
 * cusparseXcsr2bsrNnz / cusparseScsr2bsr emit them (:302-311). Values are
 * forced to 1.0 before packing (:296-297), so every stored entry is a set
 * bit, and the ballot packers ToBit{4,8,16,32,64}Col set bit (T-1-c) of the
 * word for block row r (backend/spmm.cu:165-194 (T=4, low nibble of a
 * uchar), :226-246 (T=8, uchar), :272-291 (T=16, ushort), :318-335 (T=32,
 * uint32), :357-383 (T=64, uint64)). Words are stored bsrVal[T*b + r]
 * (csr2bsr_batch_bsrbmv.cu:339). T=2 follows the same rule with a uchar per
 * tile row (bits 1..0): the reference's ToBit2Col (spmm.cu:133-162) stores 8
 * bytes per lane at a 4-byte stride, so neighbouring lanes overwrite each
 * other and its layout is not a usable definition. */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* Sorted distinct block columns of block row br into buf; returns count. */
static int block_cols(int M, const int *rowptr, const int *colind, int T,
                      int br, int *buf) {
  int n = 0;
  int r0 = br * T, r1 = r0 + T < M ? r0 + T : M;
  for (int r = r0; r < r1; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) buf[n++] = colind[e] / T;
  qsort(buf, n, sizeof(int), cmp_int);
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (m == 0 || buf[m - 1] != buf[i]) buf[m++] = buf[i];
  return m;
}

static int max_block_row_entries(int M, const int *rowptr, int T) {
  int MB = (M + T - 1) / T, mx = 0;
  for (int br = 0; br < MB; ++br) {
    int r1 = (br + 1) * T < M ? (br + 1) * T : M;
    int n = rowptr[r1] - rowptr[br * T];
    if (n > mx) mx = n;
  }
  return mx;
}

/* Returns nb (or -1 on bad T); writes bsrRowPtr[MB+1]. */
int orc_csr2b2sr_nnzb(int M, const int *rowptr, const int *colind, int T,
                      int *bsrRowPtr) {
  if (T != 2 && T != 4 && T != 8 && T != 16 && T != 32 && T != 64) return -1;
  int MB = (M + T - 1) / T;
  int *buf = (int *)malloc(sizeof(int) * (size_t)(max_block_row_entries(M, rowptr, T) + 1));
  bsrRowPtr[0] = 0;
  for (int br = 0; br < MB; ++br)
    bsrRowPtr[br + 1] = bsrRowPtr[br] + block_cols(M, rowptr, colind, T, br, buf);
  free(buf);
  return bsrRowPtr[MB];
}

static void set_bit(void *bsrVal, int T, int64_t word, int bit) {
  switch (T) {
    case 2:
    case 4:
    case 8: ((uint8_t *)bsrVal)[word] |= (uint8_t)(1u << bit); break;
    case 16: ((uint16_t *)bsrVal)[word] |= (uint16_t)(1u << bit); break;
    case 32: ((uint32_t *)bsrVal)[word] |= (1u << bit); break;
    default: ((uint64_t *)bsrVal)[word] |= (1ull << bit); break;
  }
}

void orc_csr2b2sr(int M, const int *rowptr, const int *colind, int T,
                  const int *bsrRowPtr, int *bsrColInd, void *bsrVal) {
  int MB = (M + T - 1) / T;
  int nb = bsrRowPtr[MB];
  size_t wbytes = T <= 8 ? 1 : (size_t)T / 8;
  memset(bsrVal, 0, wbytes * (size_t)nb * T);
  int *buf = (int *)malloc(sizeof(int) * (size_t)(max_block_row_entries(M, rowptr, T) + 1));
  for (int br = 0; br < MB; ++br) {
    int m = block_cols(M, rowptr, colind, T, br, buf);
    int b0 = bsrRowPtr[br];
    for (int i = 0; i < m; ++i) bsrColInd[b0 + i] = buf[i];
    int r1 = (br + 1) * T < M ? (br + 1) * T : M;
    for (int row = br * T; row < r1; ++row) {
      for (int e = rowptr[row]; e < rowptr[row + 1]; ++e) {
        int bc = colind[e] / T, c = colind[e] % T;
        int lo = 0, hi = m - 1;
        while (lo < hi) {
          int mid = (lo + hi) / 2;
          if (buf[mid] < bc) lo = mid + 1; else hi = mid;
        }
        set_bit(bsrVal, T, (int64_t)(b0 + lo) * T + (row - br * T), T - 1 - c);
      }
    }
  }
  free(buf);
}

static inline uint64_t get_word(const void *bsrVal, int T, int64_t w) {
  switch (T) {
    case 2:
    case 4:
    case 8: return ((const uint8_t *)bsrVal)[w];
    case 16: return ((const uint16_t *)bsrVal)[w];
    case 32: return ((const uint32_t *)bsrVal)[w];
    default: return ((const uint64_t *)bsrVal)[w];
  }
}

/* ------------------------------------------------------------ B2SR SpMM --
 * FBF SpMM: C[T*br + r][n] = sum over tiles t of block row br, over c with
 * bit (T-1-c) of bsrVal[T*t + r] set, of B[T*bsrColInd[t] + c][n]
 * (spmm4_full_full_new2, backend/spmm.cu:877-956, decode at :905-909,
 * gather-accumulate at :931-943), generalised to T in {2,4,8,16,32,64}. Rows
 * past M (padding of the last block row) are not written. Double sums. */
void orc_b2sr_spmm_f64(int T, int M, int N, const int *bsrRowPtr,
                       const int *bsrColInd, const void *bsrVal,
                       const float *B, int ldb, double *C, int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int row = 0; row < M; ++row) {
    int br = row / T, r = row % T;
    double *c = C + (int64_t)row * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0.0;
    for (int t = bsrRowPtr[br]; t < bsrRowPtr[br + 1]; ++t) {
      uint64_t w = get_word(bsrVal, T, (int64_t)t * T + r);
      for (int cc = 0; cc < T; ++cc) {
        if (!((w >> (T - 1 - cc)) & 1u)) continue;
        const float *b = B + (int64_t)(bsrColInd[t] * T + cc) * ldb;
        for (int k = 0; k < N; ++k) c[k] += (double)b[k];
      }
    }
  }
}

/* Binary A x sign-packed B: C[i][k] = sum over set bits (i, j) of
 * (bit (31 - k%32) of Bw[j][k/32] ? +1 : -1). The +-1 semantics are those of
 * the harness's B (test_spmm_full.cu:184-190 binarizeBMtxSign, x >= 0 -> +1)
 * packed as ToBit32RowDenseSign (backend/spmm.cu:75-91). */
void orc_b2sr_spmm_sign_i32(int T, int M, int N, const int *bsrRowPtr,
                            const int *bsrColInd, const void *bsrVal,
                            const uint32_t *Bw, int ldbw, int32_t *C, int ldc) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int row = 0; row < M; ++row) {
    int br = row / T, r = row % T;
    int32_t *c = C + (int64_t)row * ldc;
    for (int k = 0; k < N; ++k) c[k] = 0;
    for (int t = bsrRowPtr[br]; t < bsrRowPtr[br + 1]; ++t) {
      uint64_t w = get_word(bsrVal, T, (int64_t)t * T + r);
      for (int cc = 0; cc < T; ++cc) {
        if (!((w >> (T - 1 - cc)) & 1u)) continue;
        const uint32_t *b = Bw + (int64_t)(bsrColInd[t] * T + cc) * ldbw;
        for (int k = 0; k < N; ++k) c[k] += ((b[k / 32] >> (31 - k % 32)) & 1u) ? 1 : -1;
      }
    }
  }
}

/* -------------------------------------------------------- dense packers --
 * ToBit32RowDenseSign / ToBit32RowDenseBin (backend/spmm.cu:75-91, :57-73):
 * out[h][w] bit (31-i) = X[h][32w+i] >= 0 (sign) or > 0 (bin). The kernels
 * read a padding value of -1 past W, which gives 0 in both predicates. */
static void pack_rows(int H, int W, const float *X, int ldx, uint32_t *out, int sign) {
  int Wc = (W + 31) / 32;
  for (int h = 0; h < H; ++h)
    for (int w = 0; w < Wc; ++w) {
      uint32_t v = 0;
      for (int i = 0; i < 32; ++i) {
        int col = 32 * w + i;
        float x = col < W ? X[(int64_t)h * ldx + col] : -1.0f;
        v = (v << 1) | (uint32_t)(sign ? x >= 0 : x > 0);
      }
      out[(int64_t)h * Wc + w] = v;
    }
}
void orc_pack_sign_rows(int H, int W, const float *X, int ldx, uint32_t *out) {
  pack_rows(H, W, X, ldx, out, 1);
}
void orc_pack_bin_rows(int H, int W, const float *X, int ldx, uint32_t *out) {
  pack_rows(H, W, X, ldx, out, 0);
}

/* ToBit32ColUd (backend/spmm.cu:93-111) launched on grid
 * (ceil(H/32), ceil(W/32)): lane l of block (bx, by) writes word
 * out[by*32*ceil(H/32) + bx*32 + l] whose bit (31-i) is
 * X[32bx + l][32by + i] >= 0 (the brev'd ballot over lanes i, kept by lane
 * l). Out-of-range elements read -1, i.e. bit 0. */
void orc_pack_sign_cols(int H, int W, const float *X, int ldx, uint32_t *out) {
  int Hc = (H + 31) / 32, Wc = (W + 31) / 32;
  for (int w = 0; w < Wc; ++w)
    for (int h = 0; h < 32 * Hc; ++h) {
      uint32_t v = 0;
      for (int i = 0; i < 32; ++i) {
        int col = 32 * w + i;
        float x = (h < H && col < W) ? X[(int64_t)h * ldx + col] : -1.0f;
        v = (v << 1) | (uint32_t)(x >= 0);
      }
      out[(int64_t)w * 32 * Hc + h] = v;
    }
}

/* Bit32ColTo (backend/spmm.cu:116-130): inverse of the column-word layout
 * to 0/1 values, X[h][32w+i] = bit (31-i) of in[w*32*ceil(H/32) + h]. */
void orc_unpack_cols(int H, int W, const uint32_t *in, float *X, int ldx) {
  int Hc = (H + 31) / 32, Wc = (W + 31) / 32;
  for (int w = 0; w < Wc; ++w)
    for (int h = 0; h < H; ++h) {
      uint32_t v = in[(int64_t)w * 32 * Hc + h];
      for (int i = 0; i < 32 && 32 * w + i < W; ++i)
        X[(int64_t)h * ldx + 32 * w + i] = (float)((v >> (31 - i)) & 1u);
    }
}

/* ---------------------------------------------------------- binary mm --
 * BMM32_Arow_Brow / BMM32_MT_M_S (backend/mm.cu:17-53, :56-75):
 * C[m][n] = 32*Kw - 2 * sum_i popc(A[i*M + m] ^ B[i*N + n]). */
void orc_bmm32_xnor_f32(int M, int N, int Kw, const uint32_t *A,
                        const uint32_t *B, float *C, int ldc) {
#pragma omp parallel for schedule(static)
  for (int m = 0; m < M; ++m)
    for (int n = 0; n < N; ++n) {
      int pop = 0;
      for (int i = 0; i < Kw; ++i)
        pop += __builtin_popcount(A[(int64_t)i * M + m] ^ B[(int64_t)i * N + n]);
      C[(int64_t)m * ldc + n] = (float)(32 * Kw) - (float)pop * 2.0f;
    }
}

/* BMM32_BIN (backend/mm.cu:77-105): t = Kbits - 2*pop over ceil(Kbits/32)
 * words; bit (31 - n%32) of C[(n/32)*M + m] = (t >= 0). */
void orc_bmm32_xnor_bin(int M, int N, int Kbits, const uint32_t *A,
                        const uint32_t *B, uint32_t *C) {
  int Kw = (Kbits + 31) / 32, Nw = (N + 31) / 32;
  memset(C, 0, sizeof(uint32_t) * (size_t)Nw * M);
  for (int m = 0; m < M; ++m)
    for (int n = 0; n < N; ++n) {
      int pop = 0;
      for (int i = 0; i < Kw; ++i)
        pop += __builtin_popcount(A[(int64_t)i * M + m] ^ B[(int64_t)i * N + n]);
      float t = (float)Kbits - 2.0f * (float)pop;
      if (t >= 0) C[(int64_t)(n / 32) * M + m] |= 1u << (31 - n % 32);
    }
}

/* ------------------------------------------------------------ adjoint --
 * csr2csc (backend/readMtx.hpp:319-345): expands CSR to COO and rebuilds
 * it keyed by column with the (col, row) sort of customSort/compare
 * (:35-80) and the counting pass of coo2csr (:246-304); each output row
 * (= input column) lists source rows ascending. With no duplicate (row,
 * col) entries this is a stable counting sort by column, written that way
 * here. perm[e] = source CSR position of output entry e (may be NULL). */
void orc_csr2csc(int M, int K, const int *rowptr, const int *colind,
                 const float *val, int *colptr, int *rowind, float *cval,
                 int *perm) {
  for (int j = 0; j <= K; ++j) colptr[j] = 0;
  int nnz = rowptr[M];
  for (int e = 0; e < nnz; ++e) colptr[colind[e] + 1]++;
  for (int j = 0; j < K; ++j) colptr[j + 1] += colptr[j];
  int *cur = (int *)malloc(sizeof(int) * (size_t)(K + 1));
  memcpy(cur, colptr, sizeof(int) * (size_t)(K + 1));
  for (int i = 0; i < M; ++i)
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      int d = cur[colind[e]]++;
      rowind[d] = i;
      if (cval) cval[d] = val[e];
      if (perm) perm[d] = e;
    }
  free(cur);
}

/* SDDMM at stored positions: out[e] = sum_k dY[i][k] * X[j][k] (double).
 * The gradient of L = <A X, dY> with respect to the value of entry (i, j).
 * No reference counterpart (the adjoint is a new form). */
void orc_sddmm_f64(int M, int N, const int *rowptr, const int *colind,
                   const float *dY, int lddy, const float *X, int ldx,
                   double *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i)
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const float *y = dY + (int64_t)i * lddy;
      const float *x = X + (int64_t)colind[e] * ldx;
      double s = 0;
      for (int k = 0; k < N; ++k) s += (double)y[k] * (double)x[k];
      out[e] = s;
    }
}

/* The timed CPU baseline of the backward's SDDMM: out[e] = dY[i] . X[j] in
 * FP32 (sum in k order), rows in dynamic chunks of 64. Same operation as
 * orc_sddmm_f64 in FP32; its BF16 form reads both operands as bf16 bits. */
void orc_sddmm_f32_omp(int M, int N, const int *rowptr, const int *colind,
                       const float *dY, int lddy, const float *X, int ldx,
                       float *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    const float *restrict y = dY + (int64_t)i * lddy;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const float *restrict x = X + (int64_t)colind[e] * ldx;
      float s = 0.f;
      for (int k = 0; k < N; ++k) s += y[k] * x[k];
      out[e] = s;
    }
  }
}

void orc_sddmm_bf16_omp(int M, int N, const int *rowptr, const int *colind,
                        const uint16_t *dY, int lddy, const uint16_t *X, int ldx,
                        float *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < M; ++i) {
    const uint16_t *restrict y = dY + (int64_t)i * lddy;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const uint16_t *restrict x = X + (int64_t)colind[e] * ldx;
      float s = 0.f;
      for (int k = 0; k < N; ++k) s += bf16_to_f32(y[k]) * bf16_to_f32(x[k]);
      out[e] = s;
    }
  }
}

/* ------------------------------------------------- aligned B2SR repack --
 * B2SR-8 / 16 regrouped into 32-bit words, the reference's
 * packUchar2AlignedUnsigned / packUshort2AlignedUnsigned with their size
 * estimates countUcharUnsignedSize / countUshortUnsignedSize
 * (csr2bsr_batch_bsrbmv.cu:586-646, 837-888): each block row's tiles are
 * taken in groups of GS = 32 / T (4 for T = 8, 2 for T = 16), padded with
 * empty tiles of column 0; the group's word for tile row j is
 * tile(i)[j] << (32 - T) | tile(i+1)[j] << (32 - 2T) | ..., first tile in the
 * highest bits (:623-627, :874-876). new_rowptr counts tiles (a multiple of
 * GS per block row), new_colind[tiles] the padded columns, new_val[tiles/GS
 * * T] the words (group-major, then row). Returns the padded tile count. */
int orc_b2sr_aligned_size(int T, int MB, const int *rowptr) {
  if (T != 8 && T != 16) return -1;
  int GS = 32 / T, n = 0;
  for (int br = 0; br < MB; ++br) {
    int t = rowptr[br + 1] - rowptr[br];
    n += (t + GS - 1) / GS * GS;
  }
  return n;
}

void orc_b2sr_aligned_pack(int T, int MB, const int *rowptr, const int *colind,
                           const void *val, int *new_rowptr, int *new_colind,
                           uint32_t *new_val) {
  int GS = 32 / T;
  new_rowptr[0] = 0;
  int64_t wcnt = 0;
  int ccnt = 0;
  for (int br = 0; br < MB; ++br) {
    int t0 = rowptr[br], n = rowptr[br + 1] - t0;
    int padded = (n + GS - 1) / GS * GS;
    for (int i = 0; i < padded; i += GS) {
      for (int j = 0; j < T; ++j) {
        uint32_t w = 0;
        for (int k = 0; k < GS; ++k) {
          uint32_t part = 0;
          if (i + k < n) part = (uint32_t)get_word(val, T, (int64_t)(t0 + i + k) * T + j);
          w |= part << (32 - T * (k + 1));
        }
        new_val[wcnt++] = w;
      }
    }
    for (int i = 0; i < padded; ++i) new_colind[ccnt++] = i < n ? colind[t0 + i] : 0;
    new_rowptr[br + 1] = new_rowptr[br] + padded;
  }
}

/* Binary SpMV on the aligned layout (the consumer the reference's absent
 * bsrbmv would be; its checks compare counts and the binarised vector,
 * utility.cu:135-185): x is bit-packed per block column (x_words[bc], bit
 * T-1-c = x[T*bc + c], the tile rows' own bit order). For row r = T*br + j:
 * count[r] = sum over groups of popc(word_j & xw), xw the group's GS x words
 * in the same positions as its tiles. y_count[MB*T] (int32) and y_bin[MB]
 * (bit T-1-j set iff count[T*br + j] > 0, the layout bin2full decodes,
 * utility.cu:154-178); either may be NULL. */
void orc_spmv_b2sr_aligned(int T, int MB, const int *new_rowptr, const int *new_colind,
                           const uint32_t *new_val, const void *x_words, int32_t *y_count,
                           void *y_bin) {
  int GS = 32 / T;
  for (int br = 0; br < MB; ++br) {
    uint32_t bin = 0;
    for (int j = 0; j < T; ++j) {
      int32_t c = 0;
      for (int t = new_rowptr[br]; t < new_rowptr[br + 1]; t += GS) {
        uint32_t xw = 0;
        for (int k = 0; k < GS; ++k)
          xw |= (uint32_t)get_word(x_words, T, new_colind[t + k]) << (32 - T * (k + 1));
        c += __builtin_popcount(new_val[(int64_t)(t / GS) * T + j] & xw);
      }
      if (y_count) y_count[(int64_t)br * T + j] = c;
      if (c > 0) bin |= 1u << (T - 1 - j);
    }
    if (y_bin) {
      if (T == 8) ((uint8_t *)y_bin)[br] = (uint8_t)bin;
      else ((uint16_t *)y_bin)[br] = (uint16_t)bin;
    }
  }
}
