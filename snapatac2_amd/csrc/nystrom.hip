// Nyström spectral embedding for large cell counts: a landmark subset is
// embedded with the full operator, then every cell is placed from its
// similarities to the landmarks.
//
// Reference: spectral_embedding_nystrom, nystrom, idf_from_chunks
// (snapatac2-python/src/embedding.rs:60-131, 193-268, 288-312).
// Pieces here: row gather (the landmark matrix), column entry counts (the IDF
// of all cells), row norms of the weighted matrix (normalize, embedding.rs:
// 314-326), a row/column-scaled CSR x dense FP64 product (both of nystrom's
// sparse products), and nystrom's per-chunk degree normalisation.
#include <algorithm>

#include "common.h"
#include "spectral_common.h"

namespace snapgpu {
namespace nys {

constexpr int kThreads = 256;

__global__ void take_len_kernel(const int64_t* __restrict__ indptr, int64_t n_rows, const int64_t* __restrict__ rows,
                                int64_t n_take, int64_t* __restrict__ out_len, int32_t* __restrict__ status) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_take) return;
  const int64_t row = rows[r];
  if (row < 0 || row >= n_rows) {
    out_len[r] = 0;
    atomicOr(status, 1);
    return;
  }
  out_len[r] = indptr[row + 1] - indptr[row];
}

// One wave per taken row copies its entries.
__global__ void __launch_bounds__(kThreads) take_fill_kernel(const int64_t* __restrict__ indptr,
                                                             const int32_t* __restrict__ idx,
                                                             const uint32_t* __restrict__ val,
                                                             const int64_t* __restrict__ rows, int64_t n_take,
                                                             const int64_t* __restrict__ out_indptr,
                                                             int32_t* __restrict__ out_idx,
                                                             uint32_t* __restrict__ out_val) {
  const int lane = threadIdx.x & 63;
  const int64_t r = ((int64_t)blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (r >= n_take) return;
  const int64_t row = rows[r];
  const int64_t a = indptr[row], len = indptr[row + 1] - a, o = out_indptr[r];
  for (int64_t e = lane; e < len; e += 64) {
    out_idx[o + e] = idx[a + e];
    out_val[o + e] = val[a + e];
  }
}

__global__ void col_count_atomic_kernel(const int32_t* __restrict__ idx, int64_t nnz,
                                        unsigned long long* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&out[idx[e]], 1ull);
}

// s_i = 1 / ||x_i o w||_2 (0 for an empty row); one wave per row.
__global__ void __launch_bounds__(kThreads) row_inv_norm_kernel(const int64_t* __restrict__ indptr,
                                                                const int32_t* __restrict__ idx,
                                                                const uint32_t* __restrict__ cnt, int64_t n,
                                                                const double* __restrict__ w,
                                                                double* __restrict__ s) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (row >= n) return;
  double acc = 0.0;
  spectral::lane_entries<false>(
      idx, cnt, indptr[row] + lane, indptr[row + 1], acc, [&](int32_t j) { return w[j]; },
      [](double& a, double v, double wj) {
        const double x = v * wj;
        a = fma(x, x, a);
      });
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) s[row] = acc > 0.0 ? 1.0 / sqrt(acc) : 0.0;
}

// C[i, :] = a_i * sum_e cnt_e * b[idx_e] * B[idx_e, :] over the entries of row i.
// One wave per row: lane % KP is the column inside a KP-column block and the
// 64 / KP lane groups take every (64 / KP)-th entry; the groups' sums are added
// by a butterfly over the group bits. KP is the smallest power of two >= k (at
// most 32), so a narrow block keeps every lane busy; the summation order is
// fixed for a given k. A lane's entries are taken kBatch at a time: their
// indices and counts are loaded first, then every gather, then the multiply-adds
// in entry order, so kBatch gathers per lane are in flight instead of one after
// each dependent index load. Lanes past column k gather column k - 1 (in
// bounds, never stored), which keeps every load unconditional.
constexpr int kBatch = 8;

template <int KP, bool CS>
__device__ __forceinline__ double scaled_row_sum(const int32_t* __restrict__ idx, const uint32_t* __restrict__ cnt,
                                                 const double* __restrict__ b, const double* __restrict__ B,
                                                 int64_t ldb, int c, int64_t e, int64_t e1) {
  constexpr int G = 64 / KP;
  double acc = 0.0;
  for (; e + (kBatch - 1) * G < e1; e += kBatch * G) {
    int32_t j[kBatch];
    uint32_t q[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      j[u] = idx[e + u * G];
      q[u] = cnt[e + u * G];
    }
    double s[kBatch], z[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      s[u] = CS ? b[j[u]] : 1.0;
      z[u] = B[(int64_t)j[u] * ldb + c];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) acc = fma((double)q[u] * s[u], z[u], acc);
  }
  for (; e < e1; e += G) {
    const int32_t j = idx[e];
    acc = fma((double)cnt[e] * (CS ? b[j] : 1.0), B[(int64_t)j * ldb + c], acc);
  }
  return acc;
}

template <int KP, bool CS>
__global__ void __launch_bounds__(kThreads) spmm_scaled_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ idx, const uint32_t* __restrict__ cnt,
    int64_t n, const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ B,
    int64_t ldb, int k, double* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int grp = lane / KP, col = lane % KP;
  const int64_t row = ((int64_t)blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (row >= n) return;
  const int64_t e0 = indptr[row], e1 = indptr[row + 1];
  const double ai = a ? a[row] : 1.0;
  for (int c0 = 0; c0 < k; c0 += KP) {
    const int c = c0 + col;
    const bool on = c < k;
    double acc = scaled_row_sum<KP, CS>(idx, cnt, b, B, ldb, on ? c : k - 1, e0 + grp, e1);
#pragma unroll
    for (int off = KP; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
    if (grp == 0 && on) C[row * ldc + c] = ai * acc;
  }
}

// Per chunk of `chunk` rows: T[ch][c] = evals[c] * sum_i Q[i][c] (fixed-order
// sums: 8 row groups x 32 columns per block, then the groups in order).
__global__ void __launch_bounds__(kThreads) chunk_colsum_kernel(const double* __restrict__ Q, int64_t ldq, int64_t n,
                                                                int k, int64_t chunk,
                                                                const double* __restrict__ evals,
                                                                double* __restrict__ T) {
  __shared__ double part[8][33];
  const int64_t ch = blockIdx.x;
  const int rg = threadIdx.x >> 5, cl = threadIdx.x & 31;
  const int64_t r0 = ch * chunk, r1 = r0 + chunk < n ? r0 + chunk : n;
  for (int c0 = 0; c0 < k; c0 += 32) {
    const int c = c0 + cl;
    double acc = 0.0;
    if (c < k)
      for (int64_t i = r0 + rg; i < r1; i += 8) acc += Q[i * ldq + c];
    part[rg][cl] = acc;
    __syncthreads();
    if (rg == 0 && c < k) {
      double s = 0.0;
      for (int g = 0; g < 8; ++g) s += part[g][cl];
      T[ch * k + c] = s * evals[c];
    }
    __syncthreads();
  }
}

// d_i = Q_i . T[chunk(i)]; per chunk the smallest positive d (positive
// doubles order like their bit patterns) and whether any d <= 0.
__global__ void degree_kernel(const double* __restrict__ Q, int64_t ldq, int64_t n, int k, int64_t chunk,
                              const double* __restrict__ T, double* __restrict__ d,
                              unsigned long long* __restrict__ dmin, int32_t* __restrict__ nonpos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t ch = i / chunk;
  double acc = 0.0;
  for (int c = 0; c < k; ++c) acc = fma(Q[i * ldq + c], T[ch * k + c], acc);
  d[i] = acc;
  if (acc > 0.0) atomicMin(&dmin[ch], (unsigned long long)__double_as_longlong(acc));
  else atomicOr(&nonpos[ch], 1);
}

__global__ void scale_rows_kernel(double* __restrict__ Q, int64_t ldq, int64_t n, int k, int64_t chunk,
                                  const double* __restrict__ d, const unsigned long long* __restrict__ dmin,
                                  int32_t* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double di = d[i];
  if (!(di > 0.0)) {
    const unsigned long long m = dmin[i / chunk];
    if (m == ~0ull) {
      atomicOr(status, 2);  // no positive degree in the chunk
      return;
    }
    di = __longlong_as_double((long long)m);
  }
  const double r = 1.0 / sqrt(di);
  for (int c = 0; c < k; ++c) Q[i * ldq + c] *= r;
}

// E[i][c] = U[i][c] * sqrt(dinv[i]) / evals[c] (nystrom, embedding.rs:203-214:
// spectral_mf returns the degrees d = 1 / dinv (embedding.rs:189-190) and
// nystrom multiplies row i by 1 / sqrt(d_i) = sqrt(dinv_i), then divides
// column c by its eigenvalue).
__global__ void landmark_scale_kernel(const double* __restrict__ U, int64_t ldu, int64_t n, int k,
                                      const double* __restrict__ dinv, const double* __restrict__ evals,
                                      double* __restrict__ E, int64_t lde) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  const int64_t i = t / k;
  const int c = (int)(t % k);
  E[i * lde + c] = U[i * ldu + c] * sqrt(dinv[i]) / evals[c];
}

inline int waves_grid(int64_t n) { return (int)ceil_div(n > 0 ? n : 1, kThreads / 64); }
inline int grid(int64_t n) { return (int)ceil_div(n > 0 ? n : 1, kThreads); }

}  // namespace nys
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::nys;

extern "C" {

int64_t snapgpu_csr_take_rows_workspace_bytes(int64_t n_take) {
  return snapgpu_scan_workspace_bytes(n_take > 0 ? n_take : 1) + 256;
}

int snapgpu_csr_take_rows_count(const int64_t* d_indptr, int64_t n_rows, const int64_t* d_rows, int64_t n_take,
                                int64_t* d_out_indptr, void* d_ws, int64_t ws_bytes, int64_t* h_nnz, void* stream) {
  clear_error();
  SG_ARG(n_rows >= 0 && n_take >= 0 && d_out_indptr != nullptr && h_nnz != nullptr, "bad arguments");
  SG_ARG(ws_bytes >= snapgpu_csr_take_rows_workspace_bytes(n_take), "workspace too small");
  hipStream_t st = as_stream(stream);
  const int64_t zero = 0;
  SG_HIP(hipMemcpyAsync(d_out_indptr, &zero, 8, hipMemcpyHostToDevice, st));
  *h_nnz = 0;
  if (n_take == 0) return SNAPGPU_OK;
  int32_t* status = reinterpret_cast<int32_t*>(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  void* sws = status + 64;
  SG_HIP(hipMemsetAsync(status, 0, 4, st));
  take_len_kernel<<<grid(n_take), kThreads, 0, st>>>(d_indptr, n_rows, d_rows, n_take, d_out_indptr + 1, status);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(d_out_indptr + 1, n_take, sws, snapgpu_scan_workspace_bytes(n_take), stream);
  if (rc) return rc;
  int32_t hs = 0;
  SG_HIP(hipMemcpyAsync(&hs, status, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(h_nnz, d_out_indptr + n_take, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  SG_ARG(hs == 0, "row index out of range");
  return SNAPGPU_OK;
}

int snapgpu_csr_take_rows_fill(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_val,
                               const int64_t* d_rows, int64_t n_take, const int64_t* d_out_indptr, int32_t* d_out_idx,
                               uint32_t* d_out_val, void* stream) {
  clear_error();
  SG_ARG(n_take >= 0, "n_take < 0");
  if (n_take == 0) return SNAPGPU_OK;
  take_fill_kernel<<<waves_grid(n_take), kThreads, 0, as_stream(stream)>>>(d_indptr, d_idx, d_val, d_rows, n_take,
                                                                           d_out_indptr, d_out_idx, d_out_val);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_csr_col_counts(const int64_t* d_indptr, int64_t n, const int32_t* d_idx, int64_t nnz, int64_t m,
                           uint64_t* d_out, int32_t* d_ws16, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && nnz >= 0 && m >= 0 && d_out != nullptr, "bad arguments");
  hipStream_t st = as_stream(stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_out);
  SG_HIP(hipMemsetAsync(out, 0, 8 * m, st));
  if (nnz == 0 || m == 0) return SNAPGPU_OK;
  bool sorted = false;
  int rc = col_counts_sorted(d_indptr, n, d_idx, nnz, m, out, d_ws16, st, &sorted);
  if (rc) return rc;
  if (!sorted) {  // rows out of order: one atomic per entry
    SG_HIP(hipMemsetAsync(out, 0, 8 * m, st));
    const int64_t blocks = std::min<int64_t>(ceil_div(nnz, kThreads), 256 * 64);
    col_count_atomic_kernel<<<(unsigned)blocks, kThreads, 0, st>>>(d_idx, nnz, out);
    SG_LAUNCH_CHECK();
  }
  return SNAPGPU_OK;
}

int snapgpu_csr_row_inv_norm(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                             const double* d_w, double* d_s, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && d_w != nullptr && d_s != nullptr, "bad arguments");
  if (n == 0) return SNAPGPU_OK;
  row_inv_norm_kernel<<<waves_grid(n), kThreads, 0, as_stream(stream)>>>(d_indptr, d_idx, d_cnt, n, d_w, d_s);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_csr_spmm_scaled_f64(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                                const double* d_row_scale, const double* d_col_scale, const double* d_B, int64_t ldb,
                                int32_t k, double* d_C, int64_t ldc, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && k >= 1 && ldb >= k && ldc >= k && d_B != nullptr && d_C != nullptr, "bad arguments");
  if (n == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
#define SG_SPMM(KP)                                                                                          \
  do {                                                                                                       \
    if (d_col_scale)                                                                                         \
      spmm_scaled_kernel<KP, true><<<waves_grid(n), kThreads, 0, st>>>(d_indptr, d_idx, d_cnt, n, d_row_scale, \
                                                                       d_col_scale, d_B, ldb, k, d_C, ldc);  \
    else                                                                                                     \
      spmm_scaled_kernel<KP, false><<<waves_grid(n), kThreads, 0, st>>>(d_indptr, d_idx, d_cnt, n,            \
                                                                        d_row_scale, nullptr, d_B, ldb, k,   \
                                                                        d_C, ldc);                           \
  } while (0)
  if (k <= 1) SG_SPMM(1);
  else if (k <= 2) SG_SPMM(2);
  else if (k <= 4) SG_SPMM(4);
  else if (k <= 8) SG_SPMM(8);
  else if (k <= 16) SG_SPMM(16);
  else SG_SPMM(32);
#undef SG_SPMM
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_nystrom_landmark_scale(const double* d_U, int64_t ldu, int64_t n, int32_t k, const double* d_dinv,
                                   const double* d_evals, double* d_E, int64_t lde, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && k >= 1 && ldu >= k && lde >= k && d_U && d_dinv && d_evals && d_E, "bad arguments");
  if (n == 0) return SNAPGPU_OK;
  landmark_scale_kernel<<<grid(n * k), kThreads, 0, as_stream(stream)>>>(d_U, ldu, n, k, d_dinv, d_evals, d_E, lde);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_nystrom_normalize_workspace_bytes(int64_t n, int32_t k, int64_t chunk_rows) {
  const int64_t nc = chunk_rows > 0 ? ceil_div(n > 0 ? n : 1, chunk_rows) : 1;
  return 8 * nc * (k > 0 ? k : 1) + 8 * nc + 4 * nc + 8 * (n > 0 ? n : 1) + 5 * 256;
}

int snapgpu_nystrom_normalize(double* d_Q, int64_t ldq, int64_t n, int32_t k, const double* d_evals,
                              int64_t chunk_rows, void* d_ws, int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && k >= 1 && ldq >= k && chunk_rows >= 1 && d_Q != nullptr && d_evals != nullptr, "bad arguments");
  SG_ARG(ws_bytes >= snapgpu_nystrom_normalize_workspace_bytes(n, k, chunk_rows), "workspace too small");
  if (n == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  const int64_t nc = ceil_div(n, chunk_rows);
  char* p = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  auto take = [&](int64_t bytes) {
    char* at = p;
    p += (bytes + 255) & ~(int64_t)255;
    return at;
  };
  double* T = reinterpret_cast<double*>(take(8 * nc * k));
  auto* dmin = reinterpret_cast<unsigned long long*>(take(8 * nc));
  int32_t* nonpos = reinterpret_cast<int32_t*>(take(4 * nc));
  double* d = reinterpret_cast<double*>(take(8 * n));
  int32_t* status = reinterpret_cast<int32_t*>(take(4));
  SG_HIP(hipMemsetAsync(dmin, 0xFF, 8 * nc, st));
  SG_HIP(hipMemsetAsync(nonpos, 0, 4 * nc, st));
  SG_HIP(hipMemsetAsync(status, 0, 4, st));
  chunk_colsum_kernel<<<(unsigned)nc, kThreads, 0, st>>>(d_Q, ldq, n, k, chunk_rows, d_evals, T);
  SG_LAUNCH_CHECK();
  degree_kernel<<<grid(n), kThreads, 0, st>>>(d_Q, ldq, n, k, chunk_rows, T, d, dmin, nonpos);
  SG_LAUNCH_CHECK();
  scale_rows_kernel<<<grid(n), kThreads, 0, st>>>(d_Q, ldq, n, k, chunk_rows, d, dmin, status);
  SG_LAUNCH_CHECK();
  int32_t hs = 0;
  SG_HIP(hipMemcpyAsync(&hs, status, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (hs) {
    set_error("a chunk of cells has no positive Nystrom degree (min of an empty array)");
    return SNAPGPU_ERR_DATA;
  }
  return SNAPGPU_OK;
}

}  // extern "C"
