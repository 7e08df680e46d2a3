// XNOR/popcount binary matmul (the kernels of backend/mm.cu), on the VALU
// popcount path: C[m][n] = 32*Kw - 2 * sum_i popc(A[i][m] ^ B[i][n]) with A
// and B word-major ([Kw][M], [Kw][N]).
//
// A workgroup of 256 threads covers 64 rows x 4 column words (128 columns):
// the 64 lanes of a wave are 64 consecutive rows (coalesced A reads), the 4
// waves are 4 column words. Each thread keeps 32 popcount accumulators for
// its row against the 32 columns of its word; the B words of a 32-deep K
// slice are staged in LDS and read as wave-wide broadcasts. Rows and columns
// past M / N are masked, so any shape is accepted (the reference kernels
// require multiples of 32, mm.cu:13-15).
#include "common.hpp"

namespace spmm {

constexpr int kBmmThreads = 256;
constexpr int kBmmKC = 32;  // K words per LDS stage

// FP32 output: each wave's 64 x 32 result tile is staged in LDS (row stride
// 36 floats) and written as 16-byte stores, eight lanes per 128-byte row
// segment (a lane writing its own row's 32 values would touch 64 rows per
// store instruction). Workgroups at the right edge or with an unaligned C /
// ldc store element by element.
constexpr int kBmmCS = 36;
constexpr int kBmmLdsWords = 4 * 64 * kBmmCS;  // >= kBmmKC * 128 (the B stage)

template <bool kBin>
__global__ void __launch_bounds__(kBmmThreads)
bmm32_kernel(int M, int N, int Kw, int Kbits, const uint32_t *__restrict__ A,
             const uint32_t *__restrict__ B, float *__restrict__ Cf, int ldc,
             uint32_t *__restrict__ Cb, int vec_c) {
  static_assert(kBmmLdsWords >= kBmmKC * 128, "the C stage reuses the B stage");
  __shared__ uint32_t lds[kBmmLdsWords];
  uint32_t(*Bs)[4 * 32] = reinterpret_cast<uint32_t(*)[4 * 32]>(lds);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = blockIdx.x * 64 + lane;
  const int n0 = (blockIdx.y * 4 + wave) * 32;  // first column of this thread's word
  uint32_t pop[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) pop[j] = 0;

  for (int k0 = 0; k0 < Kw; k0 += kBmmKC) {
    __syncthreads();
    // Stage B[k0 .. k0+KC)[blockIdx.y*128 .. +128): 4096 words, 16 per thread.
    for (int x = threadIdx.x; x < kBmmKC * 128; x += kBmmThreads) {
      const int i = x / 128, c = x % 128;
      const int n = blockIdx.y * 128 + c;
      Bs[i][c] = (k0 + i < Kw && n < N) ? B[static_cast<size_t>(k0 + i) * N + n] : 0u;
    }
    __syncthreads();
    const int kc = min(kBmmKC, Kw - k0);
    for (int i = 0; i < kc; ++i) {
      const uint32_t a = m < M ? A[static_cast<size_t>(k0 + i) * M + m] : 0u;
#pragma unroll
      for (int j = 0; j < 32; ++j) pop[j] += __popc(a ^ Bs[i][wave * 32 + j]);
    }
  }
  if (kBin) {
    if (m >= M || n0 >= N) return;
    uint32_t word = 0;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const float t = static_cast<float>(Kbits) - 2.0f * static_cast<float>(pop[j]);
      word = (word << 1) | ((n0 + j < N && t >= 0.f) ? 1u : 0u);
    }
    Cb[static_cast<size_t>(n0 / 32) * M + m] = word;
    return;
  }
  const float full = static_cast<float>(32 * Kw);
  // one decision per workgroup (its barrier below needs every wave): the
  // staged path when all 128 of its columns lie inside N
  if (!vec_c || static_cast<int>(blockIdx.y) * 128 + 128 > N) {  // edge tile or unaligned C
    if (m >= M || n0 >= N) return;
    float *row = Cf + static_cast<size_t>(m) * ldc + n0;
#pragma unroll
    for (int j = 0; j < 32; ++j)
      if (n0 + j < N) row[j] = full - static_cast<float>(pop[j]) * 2.0f;
    return;
  }
  __syncthreads();  // the B stage is reused for the C tiles
  float *ct = reinterpret_cast<float *>(lds) + wave * 64 * kBmmCS;
#pragma unroll
  for (int j = 0; j < 32; j += 4)
    *reinterpret_cast<float4 *>(ct + lane * kBmmCS + j) =
        make_float4(full - 2.0f * static_cast<float>(pop[j]),
                    full - 2.0f * static_cast<float>(pop[j + 1]),
                    full - 2.0f * static_cast<float>(pop[j + 2]),
                    full - 2.0f * static_cast<float>(pop[j + 3]));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int m0 = blockIdx.x * 64;
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int idx = it * 64 + lane, r = idx >> 3, c4 = (idx & 7) * 4;
    if (m0 + r < M)
      *reinterpret_cast<float4 *>(Cf + static_cast<size_t>(m0 + r) * ldc + n0 + c4) =
          *reinterpret_cast<const float4 *>(ct + r * kBmmCS + c4);
  }
}

}  // namespace spmm

using namespace spmm;

extern "C" {

int bmm32_xnor_f32(int M, int N, int Kw, const uint32_t *A, const uint32_t *B, float *C, int ldc,
                   void *stream) {
  SPMM_CHECK_ARG(M >= 0 && N >= 0 && Kw >= 0, "negative size");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(C && ldc >= N, "null C or ldc < N");
  SPMM_CHECK_ARG(Kw == 0 || (A && B), "null A or B");
  dim3 grid(ceil_div(M, 64), ceil_div(N, 128));
  const int vec_c = aligned16(C) && ldc % 4 == 0;
  hipLaunchKernelGGL((bmm32_kernel<false>), grid, dim3(kBmmThreads), 0, as_stream(stream), M, N,
                     Kw, 32 * Kw, A, B, C, ldc, nullptr, vec_c);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

int bmm32_xnor_bin(int M, int N, int Kbits, const uint32_t *A, const uint32_t *B, uint32_t *C,
                   void *stream) {
  SPMM_CHECK_ARG(M >= 0 && N >= 0 && Kbits >= 0, "negative size");
  if (M == 0 || N == 0) return SPMM_OK;
  SPMM_CHECK_ARG(C, "null C");
  const int Kw = ceil_div(Kbits, 32);
  SPMM_CHECK_ARG(Kw == 0 || (A && B), "null A or B");
  dim3 grid(ceil_div(M, 64), ceil_div(N, 128));
  hipLaunchKernelGGL((bmm32_kernel<true>), grid, dim3(kBmmThreads), 0, as_stream(stream), M, N, Kw,
                     Kbits, A, B, nullptr, 0, C, 0);
  SPMM_LAUNCHED();
  return SPMM_OK;
}

}  // extern "C"
