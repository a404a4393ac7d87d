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
