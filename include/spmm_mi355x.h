/*
 * spmm_mi355x.h -- C-ABI of the MI355X (gfx950) SpMM library.
 *
 * Plain C types only: device pointers are raw pointers into HIP device
 * memory, sizes are int, streams are hipStream_t (passed as void* so the
 * header does not need HIP). Every call returns an int status (SPMM_OK == 0);
 * spmm_last_error() gives the message of the last failure on the calling
 * thread. All kernels run asynchronously on the given stream unless the
 * comment says the call synchronises.
 *
 * Matrix conventions (the reference's FBF kernels use the same, see
 * backend/spmm.cu:938,951): dense B is row-major [K][ldb], dense C is
 * row-major [M][ldc]; CSR is zero-based int32 (rowptr[M+1], colind[nnz]).
 * C is fully written (beta = 0), a superset of the reference kernels'
 * "store nonzero results only" behaviour on a pre-zeroed C
 * (test_spmm_full.cu:347-348).
 *
 * B2SR ("bit-block sparse row") layout, as produced by the reference's
 * csr2bsr_batch_{4,8,16,32,64} + ToBit{4,8,16,32,64}Col
 * (backend/csr2bsr_batch_bsrbmv.cu:207-397,403-579,654-830,894-1070,
 * 1076-1253; backend/spmm.cu:165-194,226-246,272-291,318-335,357-383):
 *   bsrRowPtr[MB+1], bsrColInd[nb] ascending within a block row, and
 *   bsrVal[nb*T] words; word bsrVal[T*b + r] has bit (T-1-c) set iff
 *   A[T*br + r][T*bc + c] is a stored entry. Word type: uint8 for T=2 (bits
 *   1..0), T=4 (low nibble) and T=8, uint16 for T=16, uint32 for T=32,
 *   uint64 for T=64.  MB = ceil(M/T). (For T=2 the reference's ToBit2Col,
 *   spmm.cu:133-162, overlaps the stores of neighbouring lanes; the rule
 *   above is used.)
 *
 * BF16 operands are uint16 bit patterns (round-to-nearest-even from FP32).
 */
#ifndef SPMM_MI355X_H
#define SPMM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPMM_OK 0
#define SPMM_ERR_INVALID_ARG 1   /* null pointer, negative size, bad ld */
#define SPMM_ERR_UNSUPPORTED 2   /* tile size / N / alignment not supported */
#define SPMM_ERR_HIP 3           /* a HIP runtime call failed */
#define SPMM_ERR_ALLOC 4         /* device allocation failed */
#define SPMM_ERR_MISMATCH 5      /* plan does not match the operands given */

typedef struct spmm_handle_s *spmm_handle_t;       /* workspace + device id */
typedef struct spmm_csr_plan_s *spmm_csr_plan_t;   /* row schedule of a CSR */
typedef struct spmm_csrt_plan_s *spmm_csrt_plan_t; /* transposed CSR + schedule */
typedef struct spmm_batch_plan_s *spmm_batch_plan_t;

/* Launch geometry override (the autotuner's knobs). Zero fields = default.
 *   lanes_per_row : lanes that share one output row (power of two, 4..64)
 *   vec           : 16-byte vectors of B loaded per lane per gather (1 or 2)
 *   rows_scale    : multiplier on rows handled per subgroup for short rows
 *   slices        : CSR SpMM, and B2SR FBF variant 0 at T <= 8: 2 .. 64
 *                   runs the product as that many launches per column tile,
 *                   launch s taking the entries (tiles) whose column (block
 *                   column) lies in the s-th equal range of B's rows (block
 *                   rows), so one launch gathers from a slice of B small
 *                   enough to stay in an XCD's L2, each adding its partial
 *                   sums into C in slice order; 0 or 1 = one launch. Needs
 *                   column indices ascending within each row (bsrColInd
 *                   within each block row); the plan keeps per-row slice
 *                   offsets (built on first use, which allocates and
 *                   synchronises), so a plan used with slices is valid only
 *                   for the colind / bsrColInd it saw.
 *   variant       : B2SR FBF SpMM only (the lane mappings of the reference's
 *                   variant kernels as candidates): 0 = block-row kernels
 *                   (lanes_per_row lanes own a column segment of a block
 *                   row, S = 64 / lanes_per_row block rows per wave -- two
 *                   per wave at 32 lanes, the shape of spmm4_full_full_new3,
 *                   backend/spmm.cu:970-1050; rows_scale block rows per
 *                   subgroup for the shortest); 1 = word lanes (a lane per
 *                   tile row, four tiles a round, any N); 2 = B staged in LDS
 *                   (N <= 8, T <= 8: the mapping of
 *                   spmm4_full_full_1_1024_lessthan8, backend/spmm.cu:417-475). */
typedef struct spmm_tune_s {
  int lanes_per_row;
  int vec;
  int rows_scale;
  int slices;
  int variant;
} spmm_tune_t;

/* ---------------------------------------------------------------- handle */
int spmm_create(spmm_handle_t *out);
/* Waits for the device, then frees the handle's buffers, including those a
 * larger call replaced (kept until here: work queued earlier, or a graph
 * captured from an earlier call, may still use them). */
int spmm_destroy(spmm_handle_t h);
const char *spmm_status_string(int status);
const char *spmm_last_error(void);
int spmm_version(void);

/* ------------------------------------------------------------ CSR SpMM
 * Replaces the cuSPARSE CSR SpMM of evalCSRSpmmFloatCuSPARSE
 * (test_spmm_full.cu:212-342: cusparseCreateCsr :281-284, cusparseSpMM
 * :303-307, alpha = 1, beta = 0, op N/N, 32-bit indices, R_32F), with B and
 * C row-major instead of the column-major layout the harness converts to.
 *
 * spmm_csr_analyze builds the row schedule (rows binned by degree, longest
 * bins first) and keeps each row's entry span (rowptr[r], rowptr[r+1])
 * beside its id, so the planned calls read the spans from the plan, not
 * from rowptr: a plan is valid only for the structure it was built from
 * (same rowptr contents), as a cuSPARSE analysis is. It launches three small
 * kernels and SYNCHRONISES the stream once to read the bin counts. The
 * plan keeps no pointer to rowptr. */
int spmm_csr_analyze(spmm_handle_t h, int M, int nnz, const int *rowptr,
                     spmm_csr_plan_t *plan, void *stream);
int spmm_csr_plan_destroy(spmm_csr_plan_t plan);
/* counts[32]: number of rows in each degree bin (bin 0: empty rows,
 * bin b >= 1: degree in [2^(b-1), 2^b)). */
int spmm_csr_plan_bins(spmm_csr_plan_t plan, int *counts32);

/* Drop-in form: no host analysis and no synchronisation. A small kernel
 * classifies the rows on the device (wave- and workgroup-sized rows listed
 * in state the handle owns), then one kernel runs them with the planned
 * kernel's per-row code, so the result equals the planned form's at the same
 * lane geometry bit for bit (the default geometry when neither is tuned).
 * Large graphs (one column tile of B of about 15 MB or more, rows averaging
 * at least 16 entries a slice) run column-sliced: per-row slice offsets
 * built on the device each call, then one launch per slice of about 10 MB
 * of B; the result equals the planned form at spmm_csr_dropin_geometry's
 * geometry bit for bit, and rows whose columns do not ascend are detected
 * on the device and summed whole in the first launch instead.
 * The handle's state grows (hipMalloc) on the first call at a larger M;
 * growing while the stream is being captured returns SPMM_ERR_UNSUPPORTED,
 * so a caller that captures makes one call at the largest size first. Calls
 * through one handle must be ordered on one stream. N not a multiple of one
 * 16-byte vector, or unaligned B / C / ld, runs the narrow kernel on a
 * per-call analysis instead (synchronises). */
int spmm_csr_f32(spmm_handle_t h, int M, int K, int N, int nnz,
                 const int *rowptr, const int *colind, const float *val,
                 const float *B, int ldb, float *C, int ldc, void *stream);
int spmm_csr_bf16(spmm_handle_t h, int M, int K, int N, int nnz,
                  const int *rowptr, const int *colind, const float *val,
                  const uint16_t *B, int ldb, float *C, int ldc, void *stream);

/* The lane geometry the drop-in above uses for these sizes, with B
 * elements of elem_bytes (4: FP32, 2: BF16) and 16-byte aligned B, C, ldb and
 * ldc: lanes_per_row and vec (0 when N is not a multiple of one 16-byte
 * vector: the narrow kernel), rows_scale 1, slices (0: one launch). Host
 * only. No counterpart in the reference, whose SpMM is cuSPARSE's; it lets a
 * caller reproduce a drop-in result with the planned form. */
int spmm_csr_dropin_geometry(int M, int K, int N, int nnz, int elem_bytes, spmm_tune_t *out);

/* Planned form: fully asynchronous, no allocation. `tune` may be NULL. */
int spmm_csr_f32_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                         const int *rowptr, const int *colind,
                         const float *val, const float *B, int ldb, float *C,
                         int ldc, const spmm_tune_t *tune, void *stream);
int spmm_csr_bf16_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                          const int *rowptr, const int *colind,
                          const float *val, const uint16_t *B, int ldb,
                          float *C, int ldc, const spmm_tune_t *tune,
                          void *stream);

/* Autotuner: times every valid launch geometry (lanes per row, 16-byte
 * vectors per lane, rows per subgroup) of the planned kernel on these
 * operands (B is uint16 BF16 when bf16 != 0), records the fastest in the
 * plan for this (N, dtype) -- later planned calls with tune == NULL use it
 * -- and returns it in *best. Writes C; SYNCHRONISES. No reference
 * counterpart. */
int spmm_csr_autotune(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                      const int *rowptr, const int *colind, const float *val,
                      const void *B, int bf16, int ldb, float *C, int ldc,
                      spmm_tune_t *best, void *stream);

/* ---------------------------------------------------------------- B2SR
 * csr2b2sr_nnzb + csr2b2sr replace cusparseXcsr2bsrNnz
 * (test_spmm_full.cu:148-151) and csr2bsr_batch_{4,8,16,32}
 * (backend/csr2bsr_batch_bsrbmv.cu:207,403,654,894). CSR input is on the
 * device with column indices ascending within each row (as coo2csr emits,
 * backend/readMtx.hpp:246-304); values are ignored: every stored entry is a
 * set bit (the reference forces values to 1.0 before packing,
 * csr2bsr_batch_bsrbmv.cu:296-297). T in {2, 4, 8, 16, 32, 64}.
 * csr2b2sr_nnzb writes bsrRowPtr[MB+1] and SYNCHRONISES to return nnzb. */
int csr2b2sr_nnzb(spmm_handle_t h, int M, int K, int nnz, const int *rowptr,
                  const int *colind, int T, int *bsrRowPtr, int *nnzb,
                  void *stream);
int csr2b2sr(spmm_handle_t h, int M, int K, int nnz, const int *rowptr,
             const int *colind, int T, const int *bsrRowPtr, int *bsrColInd,
             void *bsrVal, void *stream);

/* FBF SpMM: C[M][N] = A_bin * B, replacing spmm4_full_full_new2 and its
 * siblings (backend/spmm.cu:877-956, launched at test_spmm_full.cu:351-361).
 * Any N >= 1 (the reference requires 32 | N and N | 1024). No plan: with
 * bsrVal aligned as the planned form requires, small kernels sort the block
 * rows by tile count into device state kept in the handle, and the product
 * kernel takes the heaviest first (otherwise a row-per-subgroup kernel plus
 * a hub-row kernel). Calls on one handle must therefore be ordered on one
 * stream. Asynchronous; graph-capturable after a first call
 * at the largest block-row count (which sizes the state): growing it during
 * a capture returns SPMM_ERR_UNSUPPORTED, and state replaced by a larger
 * call stays allocated until spmm_destroy, so a captured graph never points
 * at freed memory. */
int spmm_b2sr_f32(spmm_handle_t h, int T, int M, int K, int N,
                  const int *bsrRowPtr, const int *bsrColInd,
                  const void *bsrVal, const float *B, int ldb, float *C,
                  int ldc, void *stream);
int spmm_b2sr_bf16(spmm_handle_t h, int T, int M, int K, int N,
                   const int *bsrRowPtr, const int *bsrColInd,
                   const void *bsrVal, const uint16_t *B, int ldb, float *C,
                   int ldc, void *stream);

/* Allocation-free FBF SpMM (FP32 B). No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same operands and
 * results as spmm_b2sr_f32, without a handle: the block-row lists live in a
 * static device array of this library (one per GPU, about 2 MiB, held until
 * the process ends), so the call makes no hipMalloc/hipFree. At most 262144
 * block rows (ceil(M/T)); larger M returns SPMM_ERR_UNSUPPORTED. The array
 * is shared by every call on one GPU: calls that overlap on one GPU (other
 * streams, other threads) race and must be ordered by the caller. */
int spmm_b2sr_f32_noalloc(int T, int M, int K, int N, const int *bsrRowPtr,
                          const int *bsrColInd, const void *bsrVal,
                          const float *B, int ldb, float *C, int ldc,
                          void *stream);

/* Planned FBF SpMM. spmm_b2sr_analyze bins the block rows by tile count and
 * keeps each block row's tile span beside it (SYNCHRONISES once); the
 * planned call is asynchronous and reads the spans from the plan, so a plan
 * is valid only for the bsrRowPtr it was built from. For T = 32 and 64 the
 * plan also numbers the rounds of 128 / 64 tiles and owns a list buffer
 * (8 ints per tile and T per round) that each call writes (a word-decode
 * launch) and reads back (a gather launch): calls on one such plan must not
 * overlap (other streams, other threads). Other T run one launch. bsrVal
 * must be aligned to T bytes for T <= 8 and to 16 bytes for T >= 16, and
 * K < 2^27. The plan is a spmm_csr_plan_t; it keeps no pointer to
 * bsrRowPtr. */
int spmm_b2sr_analyze(spmm_handle_t h, int T, int M, const int *bsrRowPtr,
                      spmm_csr_plan_t *plan, void *stream);
int spmm_b2sr_f32_planned(spmm_csr_plan_t plan, int T, int M, int K, int N,
                          const int *bsrRowPtr, const int *bsrColInd,
                          const void *bsrVal, const float *B, int ldb,
                          float *C, int ldc, void *stream);
int spmm_b2sr_bf16_planned(spmm_csr_plan_t plan, int T, int M, int K, int N,
                           const int *bsrRowPtr, const int *bsrColInd,
                           const void *bsrVal, const uint16_t *B, int ldb,
                           float *C, int ldc, void *stream);

/* B2SR autotuner: times the candidate lane mappings of the planned FBF SpMM
 * (variant 0 at every lanes_per_row dividing N's 16-byte vectors x
 * rows_scale 1 / 2 / 4, variant 1, and variant 2 where N <= 8 and T <= 8;
 * then, for T <= 8 when the best is variant 0 and B is 8 MB or more, that
 * mapping at 2 / 4 / 6 / 8 / 12 / 16 column slices) on these operands (B
 * is uint16 BF16 when bf16 != 0), records the fastest in
 * the plan for this (N, dtype) -- later planned calls use it -- and returns
 * it in *best. Writes C; SYNCHRONISES. spmm_b2sr_plan_set pins a candidate
 * for (N, dtype) instead (SPMM_ERR_INVALID_ARG for one that cannot run at
 * that N); a pinned variant 0 needs the planned form's alignment at the
 * call, else the call runs variant 1. No reference counterpart: the
 * reference launches one kernel, and keeps its variants in the source. */
int spmm_b2sr_autotune(spmm_csr_plan_t plan, int T, int M, int K, int N,
                       const int *bsrRowPtr, const int *bsrColInd,
                       const void *bsrVal, const void *B, int bf16, int ldb,
                       float *C, int ldc, spmm_tune_t *best, void *stream);
int spmm_b2sr_plan_set(spmm_csr_plan_t plan, int N, int bf16,
                       const spmm_tune_t *tune);

/* Binary A x sign-packed B -> int32: C[i][k] = sum over stored (i,j) of
 * (+1 if B[j][k] >= 0 else -1), with B packed by pack_sign_rows_u32 into
 * Bw[K][ldbw] (ldbw >= ceil(N/32) words). With a handle and 0 < K < 2^24 it
 * expands B to +-1 BF16 in the handle's workspace (K*N*2 bytes; the first
 * call at a larger size allocates, so capture after one such call) and runs
 * the FBF drop-in into C's memory, converting C to int32 in place; calls on
 * one handle must be ordered on one stream. No reference counterpart (its
 * bsrbmv.cu is absent, csr2bsr_batch_bsrbmv.cu:1). */
int spmm_b2sr_signpacked_i32(spmm_handle_t h, int T, int M, int K, int N,
                             const int *bsrRowPtr, const int *bsrColInd,
                             const void *bsrVal, const uint32_t *Bw, int ldbw,
                             int32_t *C, int ldc, void *stream);

/* ------------------------------------------------------ aligned B2SR-8/16
 * Regroups a B2SR-8 or B2SR-16 matrix into uint32 words: each block row's
 * tiles in groups of 32/T (padded with empty tiles of column 0), word
 * newVal[T*g + j] = row j of the group's tiles, first tile in the highest
 * bits. b2sr_aligned_nnzb replaces countUcharUnsignedSize /
 * countUshortUnsignedSize (backend/csr2bsr_batch_bsrbmv.cu:586-600,
 * 837-851) and also writes newRowPtr[MB+1] (padded tile offsets); it
 * SYNCHRONISES to return the padded tile count (words = nnzb_aligned /
 * (32/T) * T). b2sr_aligned_pack replaces packUchar2AlignedUnsigned /
 * packUshort2AlignedUnsigned (:607-647, :858-888). */
int b2sr_aligned_nnzb(spmm_handle_t h, int T, int MB, const int *bsrRowPtr,
                      int *newRowPtr, int *nnzb_aligned, void *stream);
int b2sr_aligned_pack(spmm_handle_t h, int T, int MB, int nnzb_aligned,
                      const int *bsrRowPtr, const int *bsrColInd,
                      const void *bsrVal, const int *newRowPtr,
                      int *newColInd, uint32_t *newVal, void *stream);

/* Binary SpMV on the aligned layout, y = A_bin * x for the M rows of A.
 * The reference's bsrbmv kernels are absent from its tree
 * (csr2bsr_batch_bsrbmv.cu:1); its checks of their output are
 * backend/utility.cu:135-185 (full-value vectors against cuSPARSE csrmv,
 * and the binarised vector through bin2full).
 * _bin: x_bits holds one T-bit word per block column (bit T-1-c =
 * x[T*bc + c]); y_count[M] (int32, may be NULL) gets exact counts and
 * y_bin[ceil(M/T)] (T-bit words, may be NULL) gets bit T-1-j of word br set
 * iff row T*br + j has a non-zero count.
 * _f32: x[K] FP32, y[M] FP32 (sum over the row's set bits). */
int spmv_b2sr_aligned_bin(spmm_handle_t h, int T, int M, int K,
                          int nnzb_aligned, const int *newRowPtr,
                          const int *newColInd, const uint32_t *newVal,
                          const void *x_bits, int32_t *y_count, void *y_bin,
                          void *stream);
int spmv_b2sr_aligned_f32(spmm_handle_t h, int T, int M, int K,
                          int nnzb_aligned, const int *newRowPtr,
                          const int *newColInd, const uint32_t *newVal,
                          const float *x, float *y, void *stream);

/* Dense bit packers (backend/spmm.cu:57-130).
 * pack_sign_rows_u32 / pack_bin_rows_u32: X[H][W] (row stride ldx) ->
 *   out[H][ceil(W/32)], bit (31-i) of word w = X[h][32w+i] >= 0 (sign) or
 *   > 0 (bin); bits past W are 0 (ToBit32RowDenseSign/Bin :75-91, :57-73).
 * pack_sign_cols_u32: X[H][W] -> out[ceil(W/32)][32*ceil(H/32)], word
 *   out[w][h] has bit (31-i) = X[h][32w+i] >= 0; rows past H are 0
 *   (ToBit32ColUd :93-111 with grid (ceil(H/32), ceil(W/32))).
 * unpack_cols_u32_f32: inverse of pack_sign_cols_u32 to 0/1 floats
 *   (Bit32ColTo :116-130). */
int pack_sign_rows_u32(int H, int W, const float *X, int ldx, uint32_t *out,
                       void *stream);
int pack_bin_rows_u32(int H, int W, const float *X, int ldx, uint32_t *out,
                      void *stream);
int pack_sign_cols_u32(int H, int W, const float *X, int ldx, uint32_t *out,
                       void *stream);
int unpack_cols_u32_f32(int H, int W, const uint32_t *in, float *X, int ldx,
                        void *stream);

/* ------------------------------------------------------- binary matmul
 * XNOR/popcount matmul of backend/mm.cu. A is word-major [Kw][M]
 * (A[i*M + m]), B is word-major [Kw][N].
 * bmm32_xnor_f32: C[m][n] = 32*Kw - 2*sum_i popc(A[i][m] ^ B[i][n]), row-major
 *   with stride ldc (BMM32_Arow_Brow :17-53 and BMM32_MT_M_S :56-75; bits
 *   past K in the last word count as matches exactly as there).
 * bmm32_xnor_bin: bit (31-j%32) of C[j/32][m] = (Kbits - 2*popcount >= 0),
 *   word-major [ceil(N/32)][M] (BMM32_BIN :77-105). */
int bmm32_xnor_f32(int M, int N, int Kw, const uint32_t *A, const uint32_t *B,
                   float *C, int ldc, void *stream);
int bmm32_xnor_bin(int M, int N, int Kbits, const uint32_t *A,
                   const uint32_t *B, uint32_t *C, void *stream);

/* ------------------------------------------------------------- adjoint
 * csr_transpose: stable device transpose (entries of each output row in
 * ascending source-row order), the device form of csr2csc
 * (backend/readMtx.hpp:319-345). Outputs tRowPtr[K+1], tColInd[nnz],
 * tVal[nnz] (may be NULL to skip values) and, if perm != NULL, perm[e] =
 * the CSR position of transposed entry e. Synchronises the stream. */
int csr_transpose(spmm_handle_t h, int M, int K, int nnz, const int *rowptr,
                  const int *colind, const float *val, int *tRowPtr,
                  int *tColInd, float *tVal, int *perm, void *stream);

/* Transposed SpMM C[K][N] = A^T B, B [M][ldb], from A's own CSR (no
 * counterpart in the reference; the transposed structure is csr2csc's,
 * backend/readMtx.hpp:319-345). spmm_csr_t_analyze builds the stable
 * transpose, its entry permutation and its row schedule once (SYNCHRONISES);
 * the planned calls read A's values through the permutation, so val may
 * change between calls (the backward pass dX = A^T dY of learnable edge
 * weights). The drop-in forms build the transpose into buffers the handle
 * owns (kernels only: no synchronisation; they grow on the first call at a
 * larger size, refused during graph capture as for spmm_csr_f32) and run the
 * plan-free SpMM over it. The planned form takes every geometry of the CSR
 * SpMM, column slices included (the transpose's rows ascend in A's row
 * index); spmm_csr_t_autotune tries them as spmm_csr_autotune does. */
int spmm_csr_t_analyze(spmm_handle_t h, int M, int K, int nnz,
                       const int *rowptr, const int *colind,
                       spmm_csrt_plan_t *plan, void *stream);
int spmm_csr_t_plan_destroy(spmm_csrt_plan_t plan);
int spmm_csr_t_autotune(spmm_csrt_plan_t plan, int M, int K, int N, int nnz,
                        const float *val, const void *B, int bf16, int ldb,
                        float *C, int ldc, spmm_tune_t *best, void *stream);
int spmm_csr_t_f32_planned(spmm_csrt_plan_t plan, int M, int K, int N,
                           int nnz, const float *val, const float *B,
                           int ldb, float *C, int ldc,
                           const spmm_tune_t *tune, void *stream);
int spmm_csr_t_bf16_planned(spmm_csrt_plan_t plan, int M, int K, int N,
                            int nnz, const float *val, const uint16_t *B,
                            int ldb, float *C, int ldc,
                            const spmm_tune_t *tune, void *stream);
int spmm_csr_t_f32(spmm_handle_t h, int M, int K, int N, int nnz,
                   const int *rowptr, const int *colind, const float *val,
                   const float *B, int ldb, float *C, int ldc, void *stream);
int spmm_csr_t_bf16(spmm_handle_t h, int M, int K, int N, int nnz,
                    const int *rowptr, const int *colind, const float *val,
                    const uint16_t *B, int ldb, float *C, int ldc,
                    void *stream);

/* A training step's backward in one pass over A's transposed plan (no
 * counterpart in the reference, which has no backward): dX[K][N] = A^T dY
 * and dval[e] = sum_k dY[i][k] * X[j][k] for stored (i, j) = entry e. The
 * dY row gathered for dX is dotted with X[j], held in registers, so the
 * separate SDDMM's gather of X rows is not needed. Fused when one subgroup
 * covers N (tune: lanes x vec x 16 bytes = N, or null for the default) with
 * 16-byte aligned dY, X, dX and leading dimensions; otherwise the transposed
 * SpMM and the plan-free SDDMM run in turn (rowptr / colind: A's CSR, read
 * only then). dY [M][lddy], X [K][ldx], dX [K][lddx]. Asynchronous. */
int spmm_csr_t_sddmm_f32_planned(spmm_csrt_plan_t plan, int M, int K, int N,
                                 int nnz, const int *rowptr, const int *colind,
                                 const float *val, const float *dY, int lddy,
                                 const float *X, int ldx, float *dX, int lddx,
                                 float *dval, const spmm_tune_t *tune,
                                 void *stream);
int spmm_csr_t_sddmm_bf16_planned(spmm_csrt_plan_t plan, int M, int K, int N,
                                  int nnz, const int *rowptr, const int *colind,
                                  const float *val, const uint16_t *dY, int lddy,
                                  const uint16_t *X, int ldx, float *dX, int lddx,
                                  float *dval, const spmm_tune_t *tune,
                                  void *stream);

/* SDDMM at the stored positions: dval[e] = sum_k dY[i][k] * X[j][k] for
 * stored (i, j) = entry e, FP32 accumulation. dY [M][lddy], X [K][ldx]. */
int sddmm_csr_f32(spmm_handle_t h, int M, int K, int N, int nnz,
                  const int *rowptr, const int *colind, const float *dY,
                  int lddy, const float *X, int ldx, float *dval,
                  void *stream);
int sddmm_csr_bf16(spmm_handle_t h, int M, int K, int N, int nnz,
                   const int *rowptr, const int *colind, const uint16_t *dY,
                   int lddy, const uint16_t *X, int ldx, float *dval,
                   void *stream);
/* The same on the row schedule of a CSR plan of (rowptr, colind) (the
 * forward's plan in a training step): rows by degree class, each entry's
 * dot product in one subgroup. Bitwise equal to the plan-free form. */
int sddmm_csr_f32_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                          const int *rowptr, const int *colind,
                          const float *dY, int lddy, const float *X, int ldx,
                          float *dval, void *stream);
int sddmm_csr_bf16_planned(spmm_csr_plan_t plan, int M, int K, int N, int nnz,
                           const int *rowptr, const int *colind,
                           const uint16_t *dY, int lddy, const uint16_t *X,
                           int ldx, float *dval, void *stream);

/* ---------------------------------------------------- Matrix Market
 * The reference harness's ingest (test_spmm_full.cu:99-125): readMtx
 * (backend/readMtx.hpp:351-422, mmio.hpp:128-226) then coo2csr
 * (readMtx.hpp:246-304), on the host. As there: 1-based indices become
 * 0-based, pattern entries read as 1.0, integer/real values as float (the
 * rounding of fscanf "%f"), symmetric files are NOT expanded and duplicate
 * pairs NOT merged, entries ordered by (row, col) (duplicate pairs keep file
 * order). Coordinate files with pattern / integer / real values only (the
 * reference reads no values for complex ones): SPMM_ERR_UNSUPPORTED
 * otherwise; a truncated file or an index out of range is an error.
 * spmm_mtx_read_info gives the sizes (kind = type + 4 * symmetry, type 0
 * pattern / 1 integer / 2 real, symmetry 0 general / 1 symmetric / 2 skew)
 * so the caller can allocate rowptr[M+1], colind[nnz] and val[nnz] (val
 * may be NULL) for spmm_mtx_read_csr, which checks M, K, nnz again. */
int spmm_mtx_read_info(const char *path, int *M, int *K, int *nnz, int *kind);
int spmm_mtx_read_csr(const char *path, int M, int K, int nnz, int *rowptr,
                      int *colind, float *val);

/* ------------------------------------------------------------- batched
 * Many CSR graphs run as one block-diagonal problem. Graph g has rows
 * [row_off[g], row_off[g+1]) and entries [nnz_off[g], nnz_off[g+1]) of the
 * concatenated arrays; its column indices are local (0-based in g), and its
 * feature rows follow the same row offsets (square graphs). The plan holds
 * the analysed forward and transposed schedules and runs one fwd+bwd step
 * (Y = A X, dX = A^T dY, dval = SDDMM(dY, X)) as ONE fused kernel launch
 * (optionally persistent: a grid of 8 workgroups per CU striding over the
 * work items), captured as a hipGraph.
 * Creation autotunes the launch geometry and SYNCHRONISES. N must be a
 * multiple of 16 (FP32 features) or 32 (BF16). */
int spmm_batch_plan_create(spmm_handle_t h, int ngraphs, const int *row_off,
                           const int *nnz_off, const int *rowptr,
                           const int *colind, const float *val, int N,
                           int bf16, spmm_batch_plan_t *plan, void *stream);
int spmm_batch_plan_destroy(spmm_batch_plan_t plan);
/* The geometry the plan's autotuner chose (lanes per row, vector width,
 * rows per subgroup), whether the fused launch is persistent, its SDDMM form
 * (0: fixed entry chunks; 1: the rows of the forward plan, the same bits as
 * 0; 2: no SDDMM items, the backward items form dval beside dX in one pass
 * over the transposed rows, which needs lanes x vec x 16 bytes = N; dval
 * then equals spmm_csr_t_sddmm_*_planned per graph bit for bit and the
 * other forms to FP32 rounding), and the tuned time of one step in ms. Any
 * output pointer may be NULL. */
int spmm_batch_plan_info(spmm_batch_plan_t plan, spmm_tune_t *tune,
                         int *persistent, int *sddmm_rows, float *tuned_ms);
/* Overrides the autotuner's choice (same meanings and limits as above;
 * SPMM_ERR_INVALID_ARG for a combination the launch cannot run). */
int spmm_batch_plan_set(spmm_batch_plan_t plan, const spmm_tune_t *tune,
                        int persistent, int sddmm_form);
/* One fwd+bwd step; features are uint16 BF16 when the plan is bf16, else
 * float. Y, dX are FP32 [total_rows][N]; dval FP32 [total_nnz]. */
int spmm_batch_fwd_bwd(spmm_batch_plan_t plan, const void *X, const void *dY,
                       float *Y, float *dX, float *dval, int use_graph,
                       void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPMM_MI355X_H */
