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

// One lane's share of a wave-per-row sparse dot product: for the entries
// e, e + 64, e + 128, ... below e1, in that order, step(acc, value, gather(idx)).
// Entries go kLaneBatch at a time: the batch's index and value loads are issued
// together, then its gathers, then the steps, so each lane has kLaneBatch gathers
// in flight rather than one behind each dependent index load. The order of the
// steps, and so every rounding, is the plain loop's.
constexpr int kLaneBatch = 8;

template <bool FV, typename Acc, typename Gather, typename Step>
__device__ __forceinline__ void lane_entries(const int32_t* __restrict__ idx, const uint32_t* __restrict__ val,
                                             int64_t e, int64_t e1, Acc& acc, Gather gather, Step step) {
  using G = decltype(gather(0));
  for (; e + (kLaneBatch - 1) * 64 < e1; e += kLaneBatch * 64) {
    int32_t j[kLaneBatch];
    double v[kLaneBatch];
#pragma unroll
    for (int u = 0; u < kLaneBatch; ++u) {
      j[u] = idx[e + 64 * u];
      v[u] = val64<FV>(val, e + 64 * u);
    }
    G g[kLaneBatch];
#pragma unroll
    for (int u = 0; u < kLaneBatch; ++u) g[u] = gather(j[u]);
#pragma unroll
    for (int u = 0; u < kLaneBatch; ++u) step(acc, v[u], g[u]);
  }
  for (; e < e1; e += 64) {
    const double v = val64<FV>(val, e);
    step(acc, v, gather(idx[e]));
  }
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

namespace panel {
// out[row][c] = scale[row] * (sum over the layout's entries of value * dense row),
// 16-column FP32 rows (scale null: 1), through the panel product and its per-row
// reduction (panel.hip): the CSC product of the mixed layout (window.hip, and the
// adjoint over that layout in spectral.hip).
void product16_scaled(const snapgpu_panel_layout& L, const float* dense, float* partial, const float* scale,
                      float* out, hipStream_t st);
}  // namespace panel

namespace window {
// The same over the window layout (window.hip): the CSR product of the adjoint over
// the mixed layout.
void product16_scaled(const snapgpu_window_layout& L, const float* dense, const float* scale, float* out,
                      hipStream_t st);
}  // namespace window
}  // namespace snapgpu
