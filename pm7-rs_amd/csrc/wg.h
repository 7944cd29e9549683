// SPDX-License-Identifier: GPL-3.0-or-later
//
// Workgroup reductions with a fixed order (shuffle tree inside a wave, then wave order), so a
// reduction gives the same bits whichever other workgroups share the device.
#pragma once

#include <hip/hip_runtime.h>

namespace pm7 {

template <int NT>
struct Wg {
  static constexpr int NW = NT / 64;
};

// Sum NV values over the workgroup (deterministic: fixed shuffle tree, then wave order).
template <int NT, int NV>
__device__ inline void wg_sum_n(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[wave * NV + k] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < Wg<NT>::NW; ++w) s += red[w * NV + threadIdx.x];
    red[Wg<NT>::NW * NV + threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = red[Wg<NT>::NW * NV + k];
  __syncthreads();
}

template <int NT>
__device__ inline double wg_sum(double x, double* red) {
  double v[1] = {x};
  wg_sum_n<NT, 1>(v, red);
  return v[0];
}

// Maximum over the workgroup (max is order-independent).
template <int NT>
__device__ inline double wg_max(double x, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_down(x, off, 64));
  if (lane == 0) red[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < Wg<NT>::NW; ++w) m = fmax(m, red[w]);
    red[Wg<NT>::NW] = m;
  }
  __syncthreads();
  const double m = red[Wg<NT>::NW];
  __syncthreads();
  return m;
}

}  // namespace pm7
