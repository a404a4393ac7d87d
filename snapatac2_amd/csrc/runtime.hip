// Status reporting and device-wide prefix sums.
#include "common.h"

namespace snapgpu {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// ---------------------------------------------------------------------------
// Inclusive scan of int64: three-phase (per-block totals, recursive scan of the
// totals, per-block rescan with carried offset). Fixed-order integer sums, so
// the result is deterministic.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = kScanThreads * kScanItems;  // 4096

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* lds_warp, int64_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t incl = wave_incl_scan_u64((uint64_t)v, lane);
  if (lane == 63) lds_warp[wid] = (int64_t)incl;
  __syncthreads();
  int64_t wave_off = 0, tot = 0;
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wid) wave_off += lds_warp[w];
    tot += lds_warp[w];
  }
  *total = tot;
  return wave_off + (int64_t)incl - v;
}

__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(const int64_t* data, int64_t n,
                                                                   int64_t* block_sums) {
  __shared__ int64_t lds[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
    if (idx < n) s += data[idx];
  }
  int64_t total;
  block_excl_scan(s, lds, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Each thread owns kScanItems consecutive elements.
__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(int64_t* data, int64_t n,
                                                                  const int64_t* block_offsets) {
  __shared__ int64_t lds[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + i;
    v[i] = idx < n ? data[idx] : 0;
    s += v[i];
  }
  int64_t total;
  int64_t off = block_excl_scan(s, lds, &total);
  if (block_offsets != nullptr && blockIdx.x > 0) off += block_offsets[blockIdx.x - 1];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    off += v[i];
    int64_t idx = base + i;
    if (idx < n) data[idx] = off;
  }
}

// Plain streaming copy (16-byte vectors); used to calibrate hardware byte counters.
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Gather-rate probe with the access pattern of the operator's sparse products,
// minus their pair bookkeeping: each group of 16 lanes streams 64 row indices
// per batch (one coalesced dword per lane and quarter), broadcasts them, and
// reads one 16-float row (64 B) of `dense` per index, eight rows in flight per
// lane and four rows per wave instruction. With an L2-sized `dense` the rate is
// the gather ceiling the products can approach.
constexpr int kProbeBlocks = 4096;
__global__ void __launch_bounds__(256) gather_probe_kernel(const uint32_t* __restrict__ idx, int64_t count,
                                                           const float* __restrict__ dense,
                                                           float* __restrict__ out) {
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

static int64_t scan_ws_elems(int64_t n) {
  int64_t total = 0;
  while (n > kScanTile) {
    n = ceil_div(n, kScanTile);
    total += n;
  }
  return total;
}

static int scan_rec(int64_t* data, int64_t n, int64_t* ws, hipStream_t st) {
  if (n <= 0) return SNAPGPU_OK;
  if (n <= kScanTile) {
    scan_apply_kernel<<<1, kScanThreads, 0, st>>>(data, n, nullptr);
    SG_LAUNCH_CHECK();
    return SNAPGPU_OK;
  }
  const int64_t nb = ceil_div(n, kScanTile);
  scan_reduce_kernel<<<(unsigned)nb, kScanThreads, 0, st>>>(data, n, ws);
  SG_LAUNCH_CHECK();
  int rc = scan_rec(ws, nb, ws + nb, st);
  if (rc) return rc;
  scan_apply_kernel<<<(unsigned)nb, kScanThreads, 0, st>>>(data, n, ws);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // namespace snapgpu

using namespace snapgpu;

extern "C" {

const char* snapgpu_last_error(void) { return g_last_error.c_str(); }

int snapgpu_abi_version(void) { return 1; }

int snapgpu_copy_bytes(const void* d_src, void* d_dst, int64_t bytes, void* stream) {
  clear_error();
  SG_ARG(bytes >= 0 && bytes % 16 == 0, "bytes must be a multiple of 16");
  SG_ARG(((uintptr_t)d_src % 16) == 0 && ((uintptr_t)d_dst % 16) == 0, "pointers must be 16-byte aligned");
  if (bytes == 0) return SNAPGPU_OK;
  const int64_t n16 = bytes / 16;
  copy_kernel<<<(unsigned)std::min<int64_t>(ceil_div(n16, 256), 8192), 256, 0, as_stream(stream)>>>(
      static_cast<const uint4*>(d_src), static_cast<uint4*>(d_dst), n16);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_gather_probe(const uint32_t* d_idx, int64_t count, const float* d_dense, int64_t rows,
                         float* d_out, int64_t out_len, void* stream) {
  clear_error();
  SG_ARG(rows >= 1 && count >= 0, "bad probe size");
  SG_ARG(out_len >= (int64_t)kProbeBlocks * 256, "out must hold 4096 * 256 floats");
  SG_ARG(d_dense && d_out && (count == 0 || d_idx), "null pointer");
  (void)rows;  // indices are the caller's, each < rows
  gather_probe_kernel<<<kProbeBlocks, 256, 0, as_stream(stream)>>>(d_idx, count, d_dense, d_out);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_scan_workspace_bytes(int64_t n) { return scan_ws_elems(n) * 8 + 8; }

int snapgpu_scan_inclusive_i64(int64_t* d_data, int64_t n, void* d_workspace,
                               int64_t workspace_bytes, void* stream) {
  clear_error();
  SG_ARG(n >= 0, "n < 0");
  if (n == 0) return SNAPGPU_OK;
  SG_ARG(d_data != nullptr, "null data");
  SG_ARG(workspace_bytes >= snapgpu_scan_workspace_bytes(n), "workspace too small");
  SG_ARG(n <= kScanTile || d_workspace != nullptr, "null workspace");
  return scan_rec(d_data, n, static_cast<int64_t*>(d_workspace), as_stream(stream));
}

}  // extern "C"
