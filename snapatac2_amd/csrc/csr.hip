// Column reductions over a CSR count matrix.
#include "common.h"

namespace snapgpu {

template <typename IdxT>
__global__ void __launch_bounds__(256) col_sum_kernel(const IdxT* __restrict__ idx,
                                                       const uint32_t* __restrict__ val, int64_t nnz,
                                                       unsigned long long* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride)
    atomicAdd(&out[idx[e]], (unsigned long long)val[e]);
}

// ---------------------------------------------------------------------------
// Column sums of a CSR matrix with ascending column indices in every row.
// A block owns kWinRows consecutive rows and sweeps the columns in windows of
// kWinCols: per window, teams of 16 lanes walk each row from its cursor while
// the column is inside the window (32 entries per step, kWinR rows per team at
// once), add into a 32-bit LDS histogram, and the window's non-zero bins go to
// global memory with one coalesced 64-bit atomic each. Global atomics drop from
// one per entry to one per (block, column). A step's entries past the window
// are read again in the next window, so the window is as wide as LDS allows:
// 32-bit bins give twice the columns of 64-bit ones. A bin can only wrap when
// its block's values add up to 2^32 or more, which the block checks. An entry
// whose column lies below the current window (a row out of order, a negative
// index) and a row not exhausted after the last window (an index >= m) set
// status bit 0; a block with 2^32 or more entries sets bit 1, and one whose
// values add up to 2^32 or more bit 2. The caller then discards the output.
// ---------------------------------------------------------------------------
constexpr int kWinCols = 32768;   // 128 KB of u32 bins
constexpr int kWinRows = 3968;    // rows per block: u32 cursors + row starts fit the rest of LDS
constexpr int kWinThreads = 1024;
constexpr int kWinTeams = kWinThreads / 16;
constexpr int kWinR = 4;

template <typename IdxT, bool COUNT>
__global__ void __launch_bounds__(kWinThreads) col_window_kernel(
    const int64_t* __restrict__ indptr, int64_t n, const IdxT* __restrict__ idx,
    const uint32_t* __restrict__ val, int64_t m, unsigned long long* __restrict__ out,
    int32_t* __restrict__ status) {
  __shared__ uint32_t hist[kWinCols];
  __shared__ uint32_t cur[kWinRows];
  __shared__ uint32_t start[kWinRows + 1];
  __shared__ unsigned long long block_sum;
  const int tid = threadIdx.x;
  const int64_t g0 = (int64_t)blockIdx.x * kWinRows;
  const int nr = (int)min<int64_t>(kWinRows, n - g0);
  const int64_t base = indptr[g0];
  if (indptr[g0 + nr] - base >= (1ll << 32)) {  // block-uniform
    if (tid == 0) atomicOr(status, 2);
    return;
  }
  for (int i = tid; i <= nr; i += kWinThreads) {
    const uint32_t o = (uint32_t)(indptr[g0 + i] - base);
    start[i] = o;
    if (i < nr) cur[i] = o;
  }
  if (tid == 0) block_sum = 0ull;
  const int team = tid >> 4, tl = tid & 15;
  const int tshift = tid & 48;  // this team's 16 bits in the wave's ballot
  bool bad = false;
  unsigned long long mine = 0ull;  // values this thread added (the block's total bounds every bin)
  for (int64_t c0 = 0; c0 < m; c0 += kWinCols) {
    for (int j = tid; j < kWinCols; j += kWinThreads) hist[j] = 0u;
    __syncthreads();
    const int64_t c1 = c0 + kWinCols;
    for (int i0 = team; i0 < nr; i0 += kWinTeams * kWinR) {
      int64_t e[kWinR], ee[kWinR];
#pragma unroll
      for (int r = 0; r < kWinR; ++r) {
        const int i = i0 + kWinTeams * r;
        const bool ok = i < nr;
        e[r] = ok ? base + cur[i] : 0;
        ee[r] = ok ? base + start[i + 1] : 0;
      }
      bool more = true;
      while (more) {
        // unconditional loads at clamped positions (nnz >= 1)
        IdxT ca[kWinR], cb[kWinR];
        uint32_t va[kWinR], vb[kWinR];
#pragma unroll
        for (int r = 0; r < kWinR; ++r) {
          const int64_t hi = max<int64_t>(ee[r] - 1, 0);
          const int64_t p0 = min(e[r] + tl, hi), p1 = min(e[r] + 16 + tl, hi);
          ca[r] = idx[p0];
          cb[r] = idx[p1];
          if (!COUNT) {
            va[r] = val[p0];
            vb[r] = val[p1];
          }
        }
        more = false;
#pragma unroll
        for (int r = 0; r < kWinR; ++r) {
          const bool t0 = (e[r] + tl < ee[r]) & ((int64_t)ca[r] < c1);
          const bool t1 = (e[r] + 16 + tl < ee[r]) & ((int64_t)cb[r] < c1);
          const uint32_t q = (uint32_t)((ballot64(t0) >> tshift) & 0xFFFFu) |
                             ((uint32_t)((ballot64(t1) >> tshift) & 0xFFFFu) << 16);
          const int take = q == 0xFFFFFFFFu ? 32 : __builtin_ctz(~q);  // leading run of the row
          if (tl < take) {
            const int64_t c = (int64_t)ca[r];
            if (c < c0) bad = true;
            else {
              atomicAdd(&hist[c - c0], COUNT ? 1u : va[r]);
              mine += COUNT ? 0u : va[r];
            }
          }
          if (16 + tl < take) {
            const int64_t c = (int64_t)cb[r];
            if (c < c0) bad = true;
            else {
              atomicAdd(&hist[c - c0], COUNT ? 1u : vb[r]);
              mine += COUNT ? 0u : vb[r];
            }
          }
          e[r] += take;
          more |= take == 32;
        }
      }
#pragma unroll
      for (int r = 0; r < kWinR; ++r) {
        const int i = i0 + kWinTeams * r;
        if (i < nr && tl == 0) cur[i] = (uint32_t)(e[r] - base);
      }
    }
    __syncthreads();
    for (int j = tid; j < kWinCols; j += kWinThreads) {
      const uint32_t h = hist[j];
      if (h != 0u && c0 + j < m) atomicAdd(&out[c0 + j], (unsigned long long)h);
    }
    __syncthreads();
  }
  for (int i = tid; i < nr; i += kWinThreads) bad |= cur[i] != start[i + 1];
  if (bad) atomicOr(status, 1);
  if (!COUNT) {  // counts cannot wrap: a block holds fewer than 2^32 entries
    atomicAdd(&block_sum, mine);
    __syncthreads();
    if (tid == 0 && block_sum >= (1ull << 32)) atomicOr(status, 4);
  }
}

template <typename IdxT, bool COUNT>
static int col_window_sums(const int64_t* d_indptr, int64_t n, const IdxT* d_idx, const uint32_t* d_val,
                           int64_t nnz, int64_t m, unsigned long long* d_out, int32_t* d_status,
                           hipStream_t st, bool* sorted) {
  // d_out must be zero; on return *sorted says whether d_out holds the sums
  *sorted = true;
  if (nnz == 0 || n == 0) return SNAPGPU_OK;
  SG_HIP(hipMemsetAsync(d_status, 0, 4, st));
  col_window_kernel<IdxT, COUNT><<<(unsigned)ceil_div(n, kWinRows), kWinThreads, 0, st>>>(
      d_indptr, n, d_idx, d_val, m, d_out, d_status);
  SG_LAUNCH_CHECK();
  int32_t h = 0;
  SG_HIP(hipMemcpyAsync(&h, d_status, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  *sorted = h == 0;
  return SNAPGPU_OK;
}

int col_counts_sorted(const int64_t* d_indptr, int64_t n, const int32_t* d_idx, int64_t nnz, int64_t m,
                      unsigned long long* d_out, int32_t* d_status, hipStream_t st, bool* sorted) {
  return col_window_sums<int32_t, true>(d_indptr, n, d_idx, nullptr, nnz, m, d_out, d_status, st, sorted);
}

}  // namespace snapgpu

using namespace snapgpu;

extern "C" {

int snapgpu_csr_col_sum_u32(const void* d_indices, int32_t index_bits, const uint32_t* d_values,
                            int64_t nnz, uint64_t* d_out, void* stream) {
  clear_error();
  SG_ARG(nnz >= 0, "nnz < 0");
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  if (nnz == 0) return SNAPGPU_OK;
  SG_ARG(d_indices && d_values && d_out, "null pointer");
  const int64_t blocks = std::min<int64_t>(ceil_div(nnz, 256), 256 * 64);
  auto* out = reinterpret_cast<unsigned long long*>(d_out);
  if (index_bits == 32)
    col_sum_kernel<int32_t><<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
        static_cast<const int32_t*>(d_indices), d_values, nnz, out);
  else
    col_sum_kernel<int64_t><<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
        static_cast<const int64_t*>(d_indices), d_values, nnz, out);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_csr_col_sum_rows_u32(const int64_t* d_indptr, int64_t n_rows, const void* d_indices,
                                 int32_t index_bits, const uint32_t* d_values, int64_t n_cols,
                                 uint64_t* d_out, void* d_ws16, void* stream) {
  clear_error();
  SG_ARG(n_rows >= 0 && n_cols >= 1, "bad shape");
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  SG_ARG(d_indptr && d_out && d_ws16, "null pointer");
  hipStream_t st = as_stream(stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_out);
  SG_HIP(hipMemsetAsync(out, 0, n_cols * 8, st));
  int64_t nnz = 0;
  if (n_rows > 0) {
    SG_HIP(hipMemcpyAsync(&nnz, d_indptr + n_rows, 8, hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
  }
  if (nnz == 0) return SNAPGPU_OK;
  SG_ARG(d_indices && d_values, "null pointer");
  bool sorted = true;
  int rc;
  if (index_bits == 32)
    rc = col_window_sums<int32_t, false>(d_indptr, n_rows, static_cast<const int32_t*>(d_indices),
                                         d_values, nnz, n_cols, out, (int32_t*)d_ws16, st, &sorted);
  else
    rc = col_window_sums<int64_t, false>(d_indptr, n_rows, static_cast<const int64_t*>(d_indices),
                                         d_values, nnz, n_cols, out, (int32_t*)d_ws16, st, &sorted);
  if (rc) return rc;
  if (sorted) return SNAPGPU_OK;
  // rows out of order: one atomic per entry (indices must still lie in [0, n_cols))
  SG_HIP(hipMemsetAsync(out, 0, n_cols * 8, st));
  return snapgpu_csr_col_sum_u32(d_indices, index_bits, d_values, nnz, d_out, stream);
}

}  // extern "C"
