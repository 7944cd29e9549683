// SPDX-License-Identifier: GPL-3.0-or-later
//
// 16x16 matrix-core tiles for the SCF megakernel's dense products (commutator, density, the
// eigensolver's warm-start transform and back-transform).
//
// One wave computes one 16x16 output tile D = sum_k A(i, k) B(k, j) with
// v_mfma_f64_16x16x4f64 (or v_mfma_f32_16x16x4f32 in the reduced-precision phase). Per
// instruction, lane l (r = l % 16, q = l / 16) supplies A(i0 + r, k0 + q) and B(k0 + q, j0 + r);
// the accumulator element t of lane l then holds
//   FP64: D(q + 4t, r)        FP32: D(4q + t, r)
// (checked on gfx950 with integer-valued inputs). Operands are fetched through caller lambdas
// that return 0 outside the matrix, so ragged edges need no padding.
#pragma once

#include <hip/hip_runtime.h>

namespace pm7 {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Row of accumulator element t of this lane within the tile.
__device__ inline int tile_row_f64(int q, int t) { return q + 4 * t; }
__device__ inline int tile_row_f32(int q, int t) { return 4 * q + t; }

template <class FA, class FB>
__device__ inline dbl4 tile16_f64(int K, int i0, int j0, FA fa, FB fb) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  int k0 = 0;
  // Two steps per trip: both pairs of loads are in flight before the first product.
  for (; k0 + 4 < K; k0 += 8) {
    const double a0 = fa(i0 + r, k0 + q), b0 = fb(k0 + q, j0 + r);
    const double a1 = fa(i0 + r, k0 + 4 + q), b1 = fb(k0 + 4 + q, j0 + r);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc, 0, 0, 0);
  }
  for (; k0 < K; k0 += 4) {
    const double a0 = fa(i0 + r, k0 + q), b0 = fb(k0 + q, j0 + r);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc, 0, 0, 0);
  }
  return acc;
}

template <class FA, class FB>
__device__ inline flt4 tile16_f32(int K, int i0, int j0, FA fa, FB fb) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  flt4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  int k0 = 0;
  for (; k0 + 4 < K; k0 += 8) {
    const float a0 = fa(i0 + r, k0 + q), b0 = fb(k0 + q, j0 + r);
    const float a1 = fa(i0 + r, k0 + 4 + q), b1 = fb(k0 + 4 + q, j0 + r);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc, 0, 0, 0);
  }
  for (; k0 < K; k0 += 4) {
    const float a0 = fa(i0 + r, k0 + q), b0 = fb(k0 + q, j0 + r);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc, 0, 0, 0);
  }
  return acc;
}

// One wave computes a 32x32 block D = sum_k A(i, k) B(k, j), i in [i0, i0 + 32), j in [j0, j0 + 32),
// as four 16x16 tiles acc[2 a + b] (rows i0 + 16 a, columns j0 + 16 b). Each A fragment feeds two
// tiles and each B fragment two, and 4 k-steps (16 loads per lane) are issued before their 16
// products, so the operand latency is paid once per 16 products instead of once per 2. Each
// tile accumulates its k-steps in increasing k, one product at a time, exactly as tile16_f64
// does, so the results are bit-identical to four tile16_f64 calls.
template <class FA, class FB>
__device__ inline void block32_f64(int K, int i0, int j0, FA fa, FB fb, dbl4 (&acc)[4]) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  int k0 = 0;
  for (; k0 + 12 < K; k0 += 16) {
    double a[4][2], b[4][2];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 4 * u + q;
      a[u][0] = fa(i0 + r, k);
      a[u][1] = fa(i0 + 16 + r, k);
      b[u][0] = fb(k, j0 + r);
      b[u][1] = fb(k, j0 + 16 + r);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][0], b[u][0], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][0], b[u][1], acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][1], b[u][0], acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][1], b[u][1], acc[3], 0, 0, 0);
    }
  }
  for (; k0 < K; k0 += 4) {
    const int k = k0 + q;
    const double a0 = fa(i0 + r, k), a1 = fa(i0 + 16 + r, k), b0 = fb(k, j0 + r), b1 = fb(k, j0 + 16 + r);
    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[3], 0, 0, 0);
  }
}

// The FP32 block (v_mfma_f32_16x16x4f32, FP32 accumulation), same layout and order as tile16_f32.
template <class FA, class FB>
__device__ inline void block32_f32(int K, int i0, int j0, FA fa, FB fb, flt4 (&acc)[4]) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
  int k0 = 0;
  for (; k0 + 12 < K; k0 += 16) {
    float a[4][2], b[4][2];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 4 * u + q;
      a[u][0] = fa(i0 + r, k);
      a[u][1] = fa(i0 + 16 + r, k);
      b[u][0] = fb(k, j0 + r);
      b[u][1] = fb(k, j0 + 16 + r);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][0], b[u][0], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][0], b[u][1], acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][1], b[u][0], acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][1], b[u][1], acc[3], 0, 0, 0);
    }
  }
  for (; k0 < K; k0 += 4) {
    const int k = k0 + q;
    const float a0 = fa(i0 + r, k), a1 = fa(i0 + 16 + r, k), b0 = fb(k, j0 + r), b1 = fb(k, j0 + 16 + r);
    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0], 0, 0, 0);
    acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[1], 0, 0, 0);
    acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[2], 0, 0, 0);
    acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[3], 0, 0, 0);
  }
}

// D = A B over an m x m output (k-extent K), every 16x16 tile on the matrix cores, tiles dealt
// to the NT/64 waves in turn; store(i, j, value) is called for i, j < m. The caller
// synchronises before D is read.
template <int NT, class FA, class FB, class ST>
__device__ inline void gemm_f64(int m, int K, FA fa, FB fb, ST store) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
  const int nb = (m + 31) >> 5;  // 32 x 32 blocks
  for (int t = wave; t < nb * nb; t += NW) {
    const int I = t / nb, J = t - I * nb, i0 = 32 * I, j0 = 32 * J;
    dbl4 x[4];
    block32_f64(K, i0, j0, fa, fb, x);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + 16 * (s >> 1) + tile_row_f64(q, u), j = j0 + 16 * (s & 1) + r;
        if (i < m && j < m) store(i, j, x[s][u]);
      }
  }
}

}  // namespace pm7
