// Float32 GEMM on the matrix cores (v_mfma_f32_32x32x2_f32) for the binned weight gradient
// and pre-spike errors of the dense backward:
//   C[m][n] = alpha * sum_{k in split} A(k, m) * B[k][n]
// with A(k, m) = A[k * lda + m] (kAK: A stored k-major, both operands "column panels") or
// A[m * lda + k] (A stored m-major), B[k * ldb + n]. blockIdx.z is the split of the K range
// (k_per_split each); split s writes its own C slice at C + s * c_split_stride, so partial
// sums are reduced afterwards in a fixed order (deterministic).
//
// A workgroup of 4 waves computes a 128 x 128 tile; wave (wm, wn) owns a 64 x 64 block as
// 2 x 2 MFMA 32 x 32 accumulators. K advances 32 at a time through two LDS buffers: the next
// chunk's global loads are in flight while the current chunk is multiplied, one barrier per
// chunk. The MFMA step s of a chunk takes k = 2 s + (lane / 32) from each lane, so LDS holds
// each operand as blocks (k / 8, k % 2) of rows m, a row holding the four k of that block
// and parity that one lane consumes in four consecutive steps: one 16-byte read per operand
// per four MFMA steps. A row of 4 floats after every 8 rows keeps the transposing stores off
// shared banks. Products are rounded to float32 and accumulated in float32 (the MFMA's
// arithmetic).
#pragma once

#include "common.hpp"

namespace bats {

typedef float v16f32 __attribute__((ext_vector_type(16)));

constexpr int kGemmBM = 128;
constexpr int kGemmBN = 128;
constexpr int kGemmBK = 32;
constexpr int kGemmThreads = 256;
constexpr int kGemmSlots = 512;  // resident workgroups on the 256 CUs (two per CU by LDS)
constexpr int kGemmRows = kGemmBM + kGemmBM / 8;  // padded rows of one (k / 8, k % 2) block
constexpr int kGemmBlk = kGemmRows * 4;           // floats of one block
constexpr int kGemmOperand = kGemmBlk * (kGemmBK / 8) * 2;  // floats of one operand chunk

// LDS float index of element (k, m) of a chunk (0 <= k < kGemmBK, 0 <= m < 128).
__device__ __forceinline__ int gemm_lds_at(int k, int m) {
  return (((k >> 3) * 2 + (k & 1)) * kGemmRows + m + (m >> 3)) * 4 + ((k >> 1) & 3);
}

// Four consecutive floats of row `r` starting at column `c0` of a row-major matrix with `ld`
// columns, 0 outside [0, rows) x [0, cols); one 16-byte load when the row is aligned and the
// four are in range.
__device__ __forceinline__ float4 load4(const float *__restrict__ p, int ld, int r, int c0,
                                        int rows, int cols, bool vec) {
  if (r >= rows) return make_float4(0.f, 0.f, 0.f, 0.f);
  const float *q = p + (size_t)r * ld + c0;
  if (vec && c0 + 3 < cols) return *reinterpret_cast<const float4 *>(q);
  float4 v;
  v.x = c0 < cols ? q[0] : 0.f;
  v.y = c0 + 1 < cols ? q[1] : 0.f;
  v.z = c0 + 2 < cols ? q[2] : 0.f;
  v.w = c0 + 3 < cols ? q[3] : 0.f;
  return v;
}

template <bool kAK>
__global__ __launch_bounds__(kGemmThreads) void gemm_f32_kernel(
    int M, int N, int K, const float *__restrict__ A, int lda, const float *__restrict__ B,
    int ldb, float *__restrict__ C, int ldc, int64_t c_split_stride, int k_per_split,
    float alpha, int a_vec, int b_vec) {
  __shared__ __align__(16) float As[2][kGemmOperand];
  __shared__ __align__(16) float Bs[2][kGemmOperand];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kGemmBM, n0 = blockIdx.x * kGemmBN;
  const int kb = blockIdx.z * k_per_split, ke = min(K, kb + k_per_split);
  const int nchunks = ke > kb ? (ke - kb + kGemmBK - 1) / kGemmBK : 0;
  const bool av = a_vec != 0, bv = b_vec != 0;
  // chunk loads, 32 x 128 = 1024 float4 of each operand, four per thread. A k-major operand:
  // a wave's 64 float4 cover k in [8 (w % 4), +8) and m in [32 (w / 4), +32), lane l at
  // k = 2 ((l >> 3) & 3) + (l >> 5), m = 4 (l & 7) (rows of 8 lanes are 128 contiguous bytes).
  // An m-major A: lane f at m = f / 8, k = 4 (f % 8).
  float4 ra[4], rb[4];
  auto kmaj_k = [&](int f) { return ((f >> 6) & 3) * 8 + 2 * ((f >> 3) & 3) + ((f >> 5) & 1); };
  auto kmaj_m = [&](int f) { return (f >> 8) * 32 + 4 * (f & 7); };
  auto load = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int f = tid + h * kGemmThreads;
      if constexpr (kAK) {
        ra[h] = load4(A, lda, k0 + kmaj_k(f), m0 + kmaj_m(f), ke, M, av);
      } else {
        ra[h] = load4(A, lda, m0 + f / 8, k0 + 4 * (f % 8), M, ke, av);
      }
      rb[h] = load4(B, ldb, k0 + kmaj_k(f), n0 + kmaj_m(f), ke, N, bv);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int f = tid + h * kGemmThreads;
      const float *va = reinterpret_cast<const float *>(&ra[h]);
      const float *vb = reinterpret_cast<const float *>(&rb[h]);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if constexpr (kAK) {
          As[buf][gemm_lds_at(kmaj_k(f), kmaj_m(f) + e)] = va[e];
        } else {
          As[buf][gemm_lds_at(4 * (f % 8) + e, f / 8)] = va[e];
        }
        Bs[buf][gemm_lds_at(kmaj_k(f), kmaj_m(f) + e)] = vb[e];
      }
    }
  };
  v16f32 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v16f32{};
  if (nchunks > 0) {
    load(kb);
    store(0);
  }
  __syncthreads();
  const int kr = lane >> 5, col = lane & 31;
  const int ma = wm * 64 + col, nb = wn * 64 + col;
  for (int c = 0; c < nchunks; ++c) {
    const int buf = c & 1;
    if (c + 1 < nchunks) load(kb + (c + 1) * kGemmBK);
#pragma unroll
    for (int p = 0; p < kGemmBK / 8; ++p) {
      const int blk = (p * 2 + kr) * kGemmRows;
      const float4 a0 = *reinterpret_cast<const float4 *>(&As[buf][(blk + ma + (ma >> 3)) * 4]);
      const float4 a1 =
          *reinterpret_cast<const float4 *>(&As[buf][(blk + ma + 32 + ((ma + 32) >> 3)) * 4]);
      const float4 b0 = *reinterpret_cast<const float4 *>(&Bs[buf][(blk + nb + (nb >> 3)) * 4]);
      const float4 b1 =
          *reinterpret_cast<const float4 *>(&Bs[buf][(blk + nb + 32 + ((nb + 32) >> 3)) * 4]);
      const float *pa0 = reinterpret_cast<const float *>(&a0);
      const float *pa1 = reinterpret_cast<const float *>(&a1);
      const float *pb0 = reinterpret_cast<const float *>(&b0);
      const float *pb1 = reinterpret_cast<const float *>(&b1);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(pa0[q], pb0[q], acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(pa0[q], pb1[q], acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(pa1[q], pb0[q], acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(pa1[q], pb1[q], acc[1][1], 0, 0, 0);
      }
    }
    if (c + 1 < nchunks) store(buf ^ 1);
    __syncthreads();
  }
  // accumulator element r of lane l is row 8*(r/4) + 4*(l/32) + r%4, column l%32 of the
  // 32 x 32 block
  float *Cs = C + (int64_t)blockIdx.z * c_split_stride;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 64 + j * 32 + col;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * kr + (r & 3);
        if (m < M && n < N) Cs[(size_t)m * ldc + n] = alpha * acc[i][j][r];
      }
    }
  }
}

// Host launcher. splits >= 1; C must hold `splits` slices of c_split_stride floats.
template <bool kAK>
inline int launch_gemm_f32(hipStream_t st, int M, int N, int K, const float *A, int lda,
                           const float *B, int ldb, float *C, int ldc, int64_t c_split_stride,
                           int splits, float alpha) {
  if (M <= 0 || N <= 0) return 0;
  splits = std::max(1, splits);
  int kps = (K + splits - 1) / splits;
  kps = (kps + kGemmBK - 1) / kGemmBK * kGemmBK;
  const int a_vec = (lda % 4 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  const int b_vec = (ldb % 4 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
  dim3 grid((N + kGemmBN - 1) / kGemmBN, (M + kGemmBM - 1) / kGemmBM, splits);
  hipLaunchKernelGGL(gemm_f32_kernel<kAK>, grid, dim3(kGemmThreads), 0, st, M, N, K, A, lda, B,
                     ldb, C, ldc, c_split_stride, kps, alpha, a_vec, b_vec);
  return check_launch("gemm_f32_kernel");
}

}  // namespace bats
