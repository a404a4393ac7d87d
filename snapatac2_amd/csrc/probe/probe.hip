// Measurement aids for bench.py (libsnapgpu_probe.so): a copy kernel of known
// traffic and a gather-rate probe with the operator's access pattern.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "snapgpu_probe.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int launch_status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(2, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

// Plain streaming copy (16-byte vectors).
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Each group of 16 lanes streams 64 row indices per batch (one coalesced dword
// per lane and quarter), broadcasts them, and reads one 16-float row (64 B) of
// `dense` per index, eight rows in flight per lane and four rows per wave
// instruction: the gathers of the panel products without their bookkeeping.
constexpr int kProbeBlocks = 4096;
__global__ void __launch_bounds__(256) gather_kernel(const uint32_t* __restrict__ idx, int64_t count,
                                                     const float* __restrict__ dense, float* __restrict__ out) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  float acc = 0.f;
  for (int64_t base = wave * 256; base < count; base += nwaves * 256) {
    const int64_t seg = base + g * 64;  // this group's 64 indices
    uint32_t e[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t q = seg + s * 16 + c;
      e[s] = q < count ? __builtin_nontemporal_load(idx + q) : 0u;
    }
#pragma unroll
    for (int j0 = 0; j0 < 64; j0 += 8) {
      float z[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = j0 + u;
        const uint32_t r = (uint32_t)__shfl((int)e[j / 16], g * 16 + j % 16, 64);
        z[u] = seg + j < count ? dense[(int64_t)r * 16 + c] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += z[u];
    }
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

extern "C" {

const char* snapgpu_probe_last_error(void) { return g_err.c_str(); }

int snapgpu_probe_copy_bytes(const void* d_src, void* d_dst, int64_t bytes, void* stream) {
  g_err.clear();
  if (bytes < 0 || bytes % 16 != 0) return fail(1, "bytes must be a multiple of 16");
  if (((uintptr_t)d_src % 16) != 0 || ((uintptr_t)d_dst % 16) != 0)
    return fail(1, "pointers must be 16-byte aligned");
  if (bytes == 0) return 0;
  const int64_t n16 = bytes / 16;
  const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 8192);
  copy_kernel<<<(unsigned)blocks, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
      static_cast<const uint4*>(d_src), static_cast<uint4*>(d_dst), n16);
  return launch_status();
}

int snapgpu_probe_gather(const uint32_t* d_idx, int64_t count, const float* d_dense, int64_t rows,
                         float* d_out, int64_t out_len, void* stream) {
  g_err.clear();
  if (rows < 1 || count < 0) return fail(1, "bad probe size");
  if (out_len < (int64_t)kProbeBlocks * 256) return fail(1, "out must hold 4096 * 256 floats");
  if (!d_dense || !d_out || (count > 0 && !d_idx)) return fail(1, "null pointer");
  gather_kernel<<<kProbeBlocks, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(d_idx, count, d_dense, d_out);
  return launch_status();
}

}  // extern "C"
