// Shared runtime pieces of libsnapgpu: status handling, wave-level helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "snapgpu.h"

namespace snapgpu {

// Thread-local last error message, exposed through snapgpu_last_error().
void set_error(const std::string& msg);
void clear_error();

#define SG_HIP(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::snapgpu::set_error(std::string("HIP error in " #expr ": ") +           \
                           hipGetErrorString(e_));                             \
      return SNAPGPU_ERR_HIP;                                                  \
    }                                                                          \
  } while (0)

#define SG_ARG(cond, msg)                                                      \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::snapgpu::set_error(std::string("invalid argument: ") + (msg));         \
      return SNAPGPU_ERR_ARG;                                                  \
    }                                                                          \
  } while (0)

#define SG_LAUNCH_CHECK()                                                      \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::snapgpu::set_error(std::string("kernel launch failed: ") +             \
                           hipGetErrorString(e_));                             \
      return SNAPGPU_ERR_HIP;                                                  \
    }                                                                          \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Column counts (entries per column) of a CSR matrix into d_out (zeroed by the
// caller) when every row's column indices ascend; sets *sorted = false, with
// d_out then incomplete, otherwise. d_status: 4 bytes of device scratch.
// Synchronises `st`.
int col_counts_sorted(const int64_t* d_indptr, int64_t n, const int32_t* d_idx, int64_t nnz, int64_t m,
                      unsigned long long* d_out, int32_t* d_status, hipStream_t st, bool* sorted);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Device helpers (CDNA: 64-lane wavefronts)
// ---------------------------------------------------------------------------

// Orders LDS traffic between lanes of one wavefront. A wave executes its LDS
// instructions in order; this only stops the compiler from moving memory
// operations across the point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint64_t lanemask_gt(int lane) {
  return lane == 63 ? 0ull : (~0ull << (lane + 1));
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Index of the lowest set bit (x != 0).
__device__ __forceinline__ int lowbit64(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) {
  return __shfl(v, src, 64);
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = __shfl((uint32_t)v, src, 64);
  uint32_t hi = __shfl((uint32_t)(v >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = shfl_u64(v, lane >= d ? lane - d : lane);
    if (lane >= d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    uint32_t o = __shfl_xor(v, d, 64);
    v = v < o ? v : o;
  }
  return v;
}

}  // namespace snapgpu
