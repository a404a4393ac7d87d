// Pieces shared by the operator kernels (spectral.hip) and the panel layout
// (panel.hip).
#pragma once

#include "common.h"

namespace snapgpu {
namespace spectral {

constexpr int kB = 32;  // dense block width (columns of V, Z, Y)

// Stored values are uint32 counts (tile matrices) or FP32 bit patterns.
template <bool FV>
__device__ __forceinline__ double val64(const uint32_t* v, int64_t e) {
  return FV ? (double)__uint_as_float(v[e]) : (double)v[e];
}
template <bool FV>
__device__ __forceinline__ float val32(const uint32_t* v, int64_t e) {
  return FV ? __uint_as_float(v[e]) : (float)v[e];
}

// Vs[i][c] = r_i * V[i][c] as FP32 (row stride BW), columns >= b zero-filled.
template <int BW>
__global__ void scale_in_kernel(const double* V, int64_t ldv, int64_t n, int b, const float* r,
                                float* Vs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / BW;
  const int c = (int)(t % BW);
  if (i >= n) return;
  Vs[i * BW + c] = c < b ? (float)((double)r[i] * V[i * ldv + c]) : 0.f;
}

}  // namespace spectral
}  // namespace snapgpu
