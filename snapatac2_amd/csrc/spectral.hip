// Path B: the normalized-cosine operator of spectral() and its preparation.
//
// Reference (snapatac2-python/src/embedding.rs):
//   spectral_embedding :24-59  column slice by the feature mask, u32 -> f64,
//                              weights = feature_weights or idf(X), normalize
//   idf                :269-286 df_j = stored entries of column j; all equal ->
//                              w = 1; else df 0 -> 1, df n -> n-1, w = ln(n / df)
//   normalize          :314-326 x_ij *= w_j, then each row / its L2 norm
//   spectral_mf        :131-153 c = column sums of X~, d = X~ c - 1,
//                              dinv = 1/d, X^ = diag(sqrt(dinv)) X~
//                      :155-171 matvec f(v) = X^ (X^T v) - dinv * v
//
// Storage: the selected matrix stays as integer counts, in CSR (rows = cells)
// and in a CSC copy built by a deterministic stable transpose. Values are never
// materialised: with r_i = s_i sqrt(dinv_i) (s_i = 1 / ||x_i o w||),
//     A v = r o (X (w^2 o (X^T (r o v)))) - dinv o v,
// so each operator application streams the column/row indices and counts once
// in each direction and gathers 32-wide FP32 rows of the dense operand.
// FP32 arithmetic inside the two sparse products, FP64 at the interfaces, fixed
// summation order everywhere (no floating-point atomics): results are
// bitwise reproducible.
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <string>

#include "spectral_common.h"

namespace snapgpu {
namespace spectral {

// ---------------------------------------------------------------------------
// Column selection of a CSR count matrix (feature mask -> new column ids)
// ---------------------------------------------------------------------------
// The selection is a table of 2 words per 32 columns, sel[2w] = bit mask of the
// selected columns among 32w .. 32w + 31 and sel[2w + 1] = number of selected
// columns below 32w: 8 bytes per 32 columns (1.5 MB for 6.2M tiles) stays in
// L2, where a 4-byte-per-column map would not. New id of a selected column c =
// sel[2w + 1] + popcount(sel[2w] & bits below c % 32), i.e. its rank.
__global__ void col_select_table_kernel(const int64_t* __restrict__ cols, int64_t k, int64_t words,
                                        uint32_t* __restrict__ sel) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  auto lower = [&](int64_t v) {  // first i with cols[i] >= v
    int64_t lo = 0, hi = k;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (cols[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const int64_t a = lower(32 * w), b = lower(32 * w + 32);
  uint32_t mask = 0;
  for (int64_t i = a; i < b; ++i) {
    const int64_t d = cols[i] - 32 * w;
    if (d >= 0 && d < 32) mask |= 1u << d;
  }
  sel[2 * w] = mask;
  sel[2 * w + 1] = (uint32_t)a;
}

template <typename IdxT>
__device__ __forceinline__ uint2 sel_word(const uint32_t* __restrict__ sel, int64_t words, IdxT c) {
  const int64_t w = (int64_t)c >> 5;
  if (c < 0 || w >= words) return make_uint2(0u, 0u);
  return *reinterpret_cast<const uint2*>(sel + 2 * w);
}

template <typename IdxT>
__global__ void __launch_bounds__(256) select_count_kernel(const int64_t* indptr, const IdxT* idx,
                                                           int64_t n, const uint32_t* sel, int64_t words,
                                                           int64_t* row_nnz) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  int64_t cnt = 0;
  for (int64_t e = indptr[row] + lane; e < indptr[row + 1]; e += 64) {
    const IdxT c = idx[e];
    cnt += (sel_word(sel, words, c).x >> ((uint32_t)c & 31u)) & 1u;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (lane == 0) row_nnz[row] = cnt;
}

template <typename IdxT>
__global__ void __launch_bounds__(256) select_fill_kernel(const int64_t* indptr, const IdxT* idx,
                                                          const uint32_t* val, int64_t n,
                                                          const uint32_t* sel, int64_t words,
                                                          const int64_t* new_indptr,
                                                          int32_t* new_idx, uint32_t* new_val) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  int64_t out = new_indptr[row];
  const int64_t e1 = indptr[row + 1];
  for (int64_t e0 = indptr[row]; e0 < e1; e0 += 64) {
    const int64_t e = e0 + lane;
    int32_t nc = -1;
    if (e < e1) {
      const IdxT c = idx[e];
      const uint2 s = sel_word(sel, words, c);
      const uint32_t bit = (uint32_t)c & 31u;
      if ((s.x >> bit) & 1u) nc = (int32_t)(s.y + __popc(s.x & ((1u << bit) - 1u)));
    }
    const uint64_t m = ballot64(nc >= 0);
    if (nc >= 0) {
      const int64_t o = out + popc64(m & lanemask_lt(lane));
      new_idx[o] = nc;
      new_val[o] = val[e];
    }
    out += popc64(m);
  }
}

// Both selection passes for 32-bit column indices, with the table in LDS. With
// one table lookup per entry straight from L2, every lane's lookup is its own
// L2 request (a row's neighbouring entries lie hundreds of columns apart), and
// the request rate, not HBM, sets the pace. Here a block owns kSelRows rows and
// sweeps the columns in windows of kSelWinWords table words (262,144 columns):
// it stages the window's slice of the table in LDS, then each wave walks its
// rows from their cursors while the columns stay below the window's end, 64
// entries per step, and looks the columns up in LDS. Entries are visited once,
// in row order, so counts and output equal the per-entry kernels' for any input:
// a column below the window (a row out of order, or negative) is looked up in
// the global table, and the last window takes every remaining entry (columns
// past the table select nothing).
constexpr int kSelWinWords = 8192;  // 64 KB of table words
constexpr int kSelRows = 512;
constexpr int kSelThreads = 1024;

template <bool FILL>
__global__ void __launch_bounds__(kSelThreads) select_window_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ idx, const uint32_t* __restrict__ val,
    int64_t n, const uint32_t* __restrict__ sel, int64_t words, int64_t* __restrict__ row_nnz,
    const int64_t* __restrict__ new_indptr, int32_t* __restrict__ new_idx, uint32_t* __restrict__ new_val) {
  __shared__ uint2 tab[kSelWinWords];
  __shared__ int64_t cur[kSelRows];  // next entry of the row
  __shared__ int64_t end[kSelRows];
  __shared__ int64_t out[kSelRows];  // count: selected so far; fill: next output slot
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * kSelRows;
  const int nr = (int)min<int64_t>(kSelRows, n - r0);
  for (int i = tid; i < nr; i += kSelThreads) {
    cur[i] = indptr[r0 + i];
    end[i] = indptr[r0 + i + 1];
    out[i] = FILL ? new_indptr[r0 + i] : 0;
  }
  const int64_t nwin = (words + kSelWinWords - 1) / kSelWinWords;
  for (int64_t w = 0; w < nwin; ++w) {
    const int64_t w0 = w * kSelWinWords;
    const int nw = (int)min<int64_t>(kSelWinWords, words - w0);
    __syncthreads();  // the previous window's walks are done with tab (and the cursors are set)
    for (int j = tid; j < nw; j += kSelThreads) tab[j] = *reinterpret_cast<const uint2*>(sel + 2 * (w0 + j));
    __syncthreads();
    const int64_t c_lo = w0 * 32;
    const int64_t c_hi = w + 1 == nwin ? INT64_MAX : (w0 + nw) * 32;
    for (int i = wid; i < nr; i += kSelThreads / 64) {
      int64_t e = cur[i], o = out[i];
      const int64_t e1 = end[i];
      int32_t c = 0;
      uint32_t v = 0;
      if (e + lane < e1) {
        c = idx[e + lane];
        if (FILL) v = val[e + lane];
      }
      while (e < e1) {
        const int64_t ee = e + lane;
        // the next step's entries are loaded before this step is used (values
        // too: at a quarter of the columns selected nearly every 64-byte
        // segment of values holds a selected entry, so they cost no traffic)
        int32_t cn = 0;
        uint32_t vn = 0;
        if (ee + 64 < e1) {
          cn = idx[ee + 64];
          if (FILL) vn = val[ee + 64];
        }
        const uint64_t tk = ballot64(ee < e1 && (int64_t)c < c_hi);
        const int take = ~tk == 0 ? 64 : __builtin_ctzll(~tk);  // leading run of entries in the window
        int32_t nc = -1;
        if (lane < take) {
          uint2 s;
          const int64_t wd = (int64_t)(c >> 5);
          if ((int64_t)c < c_lo) s = sel_word(sel, words, c);
          else if (wd < w0 + nw) s = tab[wd - w0];
          else s = make_uint2(0u, 0u);  // past the table (last window only)
          const uint32_t bit = (uint32_t)c & 31u;
          if ((s.x >> bit) & 1u) nc = (int32_t)(s.y + __popc(s.x & ((1u << bit) - 1u)));
        }
        const uint64_t m = ballot64(nc >= 0);
        if (FILL && nc >= 0) {
          const int64_t p = o + popc64(m & lanemask_lt(lane));
          new_idx[p] = nc;
          new_val[p] = v;
        }
        o += popc64(m);
        e += take;
        if (take < 64) break;  // the row's end, or an entry at or past the window's end
        c = cn;
        v = vn;
      }
      if (lane == 0) {
        cur[i] = e;
        out[i] = o;
      }
    }
  }
  if (!FILL) {
    __syncthreads();
    for (int i = tid; i < nr; i += kSelThreads) row_nnz[r0 + i] = out[i];
  }
}

// ---------------------------------------------------------------------------
// Stable CSR -> CSC transpose (rows stay ascending inside each column)
// ---------------------------------------------------------------------------
__global__ void col_count_kernel(const int32_t* idx, int64_t nnz, unsigned long long* counts) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride)
    atomicAdd(&counts[idx[e]], 1ull);
}

// (row << 32 | value) of every entry of rows [r0, r1) (entries [indptr[r0], indptr[r1])):
// the sort by column carries row and value along, so placing an entry reads them
// in sorted order instead of gathering them by original position
__global__ void __launch_bounds__(256) entry_pack_kernel(const int64_t* indptr, int64_t r0, int64_t r1,
                                                         const uint32_t* val, uint64_t* packed) {
  const int lane = threadIdx.x & 63;
  const int64_t row = r0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (row >= r1) return;
  const int64_t base = indptr[r0];
  for (int64_t e = indptr[row] + lane; e < indptr[row + 1]; e += 64)
    packed[e - base] = ((uint64_t)row << 32) | (uint64_t)val[e];
}

// Runs of equal keys in a sorted chunk: cursor[col] is the next free CSC slot of
// column col. rebase: cursor[col] -= t0 at the run's first position t0, so the
// slot of sorted position t is cursor[key[t]] + t; advance: cursor[col] += t + 1
// at the run's last position (the column's next free slot again). One thread
// touches each column present, so no atomics.
__global__ void cursor_rebase_kernel(const uint32_t* keys, int64_t cn, int64_t* cursor) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cn; t += stride) {
    const uint32_t col = keys[t];
    if (t == 0 || keys[t - 1] != col) cursor[col] -= t;
  }
}

__global__ void scatter_chunk_kernel(const uint32_t* skeys, const uint64_t* spacked, int64_t cn,
                                     const int64_t* cursor, int32_t* rowidx, uint32_t* out_val) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cn; t += stride) {
    const int64_t o = cursor[skeys[t]] + t;
    const uint64_t v = spacked[t];
    rowidx[o] = (int32_t)(v >> 32);
    out_val[o] = (uint32_t)v;
  }
}

__global__ void cursor_advance_kernel(const uint32_t* keys, int64_t cn, int64_t* cursor) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cn; t += stride) {
    const uint32_t col = keys[t];
    if (t + 1 == cn || keys[t + 1] != col) cursor[col] += t + 1;
  }
}

// ---------------------------------------------------------------------------
// IDF, row scales, degrees
// ---------------------------------------------------------------------------
__global__ void df_minmax_kernel(const int64_t* colptr, int64_t m, unsigned long long* mm) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long lo = ~0ull, hi = 0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    const unsigned long long df = (unsigned long long)(colptr[j + 1] - colptr[j]);
    lo = df < lo ? df : lo;
    hi = df > hi ? df : hi;
  }
  atomicMin(&mm[0], lo);
  atomicMax(&mm[1], hi);
}

__global__ void idf_kernel(const int64_t* colptr, int64_t m, int64_t n, const unsigned long long* mm,
                           double* w) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  if (mm[0] == mm[1]) {  // all document frequencies equal (embedding.rs:273-274)
    w[j] = 1.0;
    return;
  }
  double df = (double)(colptr[j + 1] - colptr[j]);
  const double nn = (double)n;
  if (df == 0.0) df = 1.0;
  else if (df == nn) df = nn - 1.0;
  w[j] = log(nn / df);
}

// s_i = 1 / ||(count_ij * w_j)_j||_2 (0 for an empty row); one wave per row.
template <bool FV>
__global__ void __launch_bounds__(256) row_scale_kernel(const int64_t* indptr, const int32_t* idx,
                                                        const uint32_t* cnt, int64_t n,
                                                        const double* w, double* s) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  double acc = 0.0;
  lane_entries<FV>(
      idx, cnt, indptr[row] + lane, indptr[row + 1], acc, [&](int32_t j) { return w[j]; },
      [](double& a, double v, double wj) {
        const double x = v * wj;
        a = fma(x, x, a);
      });
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) s[row] = acc > 0.0 ? 1.0 / sqrt(acc) : 0.0;
}

// c_j = sum_i x~_ij = w_j * sum_i count_ij s_i; one wave per column (CSC).
template <bool FV>
__global__ void __launch_bounds__(256) colsum_kernel(const int64_t* colptr, const int32_t* rowidx,
                                                     const uint32_t* cnt, int64_t m, const double* w,
                                                     const double* s, double* c) {
  const int lane = threadIdx.x & 63;
  const int64_t col = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (col >= m) return;
  double acc = 0.0;
  lane_entries<FV>(
      rowidx, cnt, colptr[col] + lane, colptr[col + 1], acc, [&](int32_t i) { return s[i]; },
      [](double& a, double v, double si) { a = fma(v, si, a); });
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) c[col] = acc * w[col];
}

// d_i = x~_i . c - 1; dinv = 1/d; r = s sqrt(dinv). wc[j] = (w_j, c_j): one
// 16-byte gather per entry. Flags rows with d <= 0 that have entries (a cell
// sharing no feature with any other cell).
template <bool FV>
__global__ void __launch_bounds__(256) degree_kernel(const int64_t* indptr, const int32_t* idx,
                                                     const uint32_t* cnt, int64_t n, const double2* wc,
                                                     const double* s, double* dinv, double* r64, float* r32,
                                                     int32_t* status) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  const int64_t e0 = indptr[row], e1 = indptr[row + 1];
  double acc = 0.0;
  lane_entries<FV>(
      idx, cnt, e0 + lane, e1, acc, [&](int32_t j) { return wc[j]; },
      [](double& a, double v, double2 q) { a = fma(v * q.x, q.y, a); });
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) {
    const double d = acc * s[row] - 1.0;
    const double di = 1.0 / d;
    dinv[row] = di;
    double rr = 0.0;
    if (e1 > e0) {
      if (!(d > 0.0)) {
        if (atomicCAS(&status[0], 0, 1) == 0) {
          status[1] = (int32_t)(row & 0xFFFFFFFF);
          status[2] = (int32_t)(row >> 32);
        }
      } else {
        rr = s[row] * sqrt(di);
      }
    }
    r64[row] = rr;
    r32[row] = (float)rr;
  }
}

__global__ void square_f32_kernel(const double* w, int64_t m, float* w2) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) w2[j] = (float)(w[j] * w[j]);
}

// ---------------------------------------------------------------------------
// Operator application (FP32 sparse products)
// ---------------------------------------------------------------------------
// One wave per row. Lanes form G = 64/BW groups of BW lanes; group g handles
// entries g, g+G, g+2G, ... of the row and lane c of the group column c of the
// dense operand (rows of BW floats: a 4*BW-byte gather per entry). Batches of 64
// (column, count) pairs are loaded coalesced and broadcast to the groups by
// ds_bpermute; gathers are issued 8 steps at a time so several are in flight.
// The G partial sums are combined by a fixed xor tree: the summation order is
// fixed, so results are bitwise reproducible.
//   EPI 0: out32[row][c] = rowscale[row] * sum                    (FP32, stride BW)
//   EPI 1: Y[row][c] = rowscale[row] * sum - dinv[row] * V[row][c] (FP64)
template <int BW, int EPI, bool FV>
__global__ void __launch_bounds__(256) spmm_kernel(const int64_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ col,
                                                   const uint32_t* __restrict__ cnt, int64_t nrows,
                                                   const float* __restrict__ dense,
                                                   const float* __restrict__ rowscale,
                                                   float* __restrict__ out32,
                                                   const double* __restrict__ V, int64_t ldv,
                                                   const double* __restrict__ dinv,
                                                   double* __restrict__ Y, int64_t ldy, int b) {
  constexpr int G = 64 / BW;
  const int lane = threadIdx.x & 63;
  const int g = lane / BW;
  const int c = lane % BW;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= nrows) return;
  const int64_t e0 = ptr[row], e1 = ptr[row + 1];
  float acc = 0.f;
  int32_t j = 0;
  float v = 0.f;
  if (e0 + lane < e1) {
    j = col[e0 + lane];
    v = val32<FV>(cnt, e0 + lane);
  }
  for (int64_t base = e0; base < e1; base += 64) {
    int32_t jn = 0;
    float vn = 0.f;
    const int64_t en = base + 64 + lane;
    if (en < e1) {  // prefetch the next batch
      jn = col[en];
      vn = val32<FV>(cnt, en);
    }
    const int nb = (int)min((int64_t)64, e1 - base);
    const int steps = (nb + G - 1) / G;  // entries s*G + g, s < steps
    int s = 0;
    for (; s + 8 <= steps; s += 8) {
      float z[8], vv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = (s + u) * G + g;
        const int32_t jt = __shfl(j, t, 64);
        const float vt = __shfl(v, t, 64);
        vv[u] = t < nb ? vt : 0.f;
        z[u] = dense[(int64_t)(t < nb ? jt : 0) * BW + c];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = fmaf(vv[u], z[u], acc);
    }
    for (; s < steps; ++s) {
      const int t = s * G + g;
      const int32_t jt = __shfl(j, t, 64);
      const float vt = __shfl(v, t, 64);
      if (t < nb) acc = fmaf(vt, dense[(int64_t)jt * BW + c], acc);
    }
    j = jn;
    v = vn;
  }
#pragma unroll
  for (int o = BW; o < 64; o <<= 1) acc += __shfl_xor(acc, o, 64);
  if (g != 0) return;
  if (EPI == 0) {
    out32[row * BW + c] = rowscale ? acc * rowscale[row] : acc;
  } else if (c < b) {
    Y[row * ldy + c] = (double)(acc * rowscale[row]) - dinv[row] * V[row * ldv + c];
  }
}

template <int BW, bool FV>
static void launch_apply(const int64_t* indptr, const int32_t* idx, const uint32_t* cnt, int64_t n,
                         const int64_t* colptr, const int32_t* rowidx, const uint32_t* ccnt,
                         int64_t m, const float* r32, const float* w2, const double* dinv,
                         const double* V, int64_t ldv, int b, double* Y, int64_t ldy, float* Vs,
                         float* Z, hipStream_t st) {
  scale_in_kernel<BW><<<(unsigned)ceil_div(n * BW, 256), 256, 0, st>>>(V, ldv, n, b, r32, Vs);
  // Z = w^2 o (X^T Vs): CSC rows are features
  spmm_kernel<BW, 0, FV><<<(unsigned)ceil_div(m, 4), 256, 0, st>>>(colptr, rowidx, ccnt, m, Vs, w2, Z,
                                                                   nullptr, 0, nullptr, nullptr, 0, b);
  // Y = r o (X Z) - dinv o V
  spmm_kernel<BW, 1, FV><<<(unsigned)ceil_div(n, 4), 256, 0, st>>>(indptr, idx, cnt, n, Z, r32, nullptr,
                                                                   V, ldv, dinv, Y, ldy, b);
}

// ---------------------------------------------------------------------------
// FP64 tall-skinny products for the block Lanczos basis
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

// The products below run on v_mfma_f64_16x16x4_f64 (D[16x16] += A[16x4] B[4x16],
// FP64 in and out). Lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16], r < 4. Which matrix index a lane's operand stands for is
// free as long as both operands agree on the K index, so every lane can load a
// contiguous run of one row of the tall matrix.

// --- C = A^T B --------------------------------------------------------------
// MFMA M = columns of A, N = columns of B, K = rows. A block (4 waves) owns a
// chunk of rows and 16*TA columns of A; lane l loads TA consecutive columns
// a0 + TA*(l%16) .. +TA-1 of row (l/16) of each 4-row step (MFMA t then stands
// for column a0 + TA*i + t of M index i) and one value of B (column l%16 of
// each 16-wide tile). Rows are clamped to the chunk and their B values zeroed,
// so every load is unconditional and in bounds.
constexpr int kTnWaves = 4;

template <int TA, int TC, int U, bool VEC>
__device__ __forceinline__ void tn_load(const double* __restrict__ A, int64_t lda,
                                        const double* __restrict__ B, int64_t ldb, int ka, int kb,
                                        int a_lane, int c_lane, int64_t row0, int64_t r1,
                                        double (&va)[U][TA], double (&vb)[U][TC]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t row = row0 + 4 * u;
    const bool rv = row < r1;
    const int64_t rc = rv ? row : r1 - 1;
    const double* ap = A + rc * lda + a_lane;
    if constexpr (VEC) {
      static_assert(TA % 2 == 0, "vector loads take column pairs");
#pragma unroll
      for (int t = 0; t < TA; t += 2) {
        const double2 v = *reinterpret_cast<const double2*>(ap + t);
        va[u][t] = v.x;
        va[u][t + 1] = v.y;
      }
    } else {
      // columns past ka read column ka - 1; they only reach outputs never stored
#pragma unroll
      for (int t = 0; t < TA; ++t) va[u][t] = A[rc * lda + min(a_lane + t, ka - 1)];
    }
    // rows past the chunk contribute zero; a multiply rather than a select keeps
    // the load unconditional (a select lets the compiler branch around the load
    // and wait for it in the branch)
    const double keep = rv ? 1.0 : 0.0;
#pragma unroll
    for (int tc = 0; tc < TC; ++tc) vb[u][tc] = B[rc * ldb + min(16 * tc + c_lane, kb - 1)] * keep;
  }
}

template <int TA, int TC, int U>
__device__ __forceinline__ void tn_mfma(f64x4 (&acc)[TA][TC], const double (&va)[U][TA],
                                        const double (&vb)[U][TC]) {
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
      for (int tc = 0; tc < TC; ++tc)
        acc[t][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[u][t], vb[u][tc], acc[t][tc], 0, 0, 0);
}

// Rows of the chunk in steps of 2 * 4 waves * U * 4; two register sets, each
// loaded one step ahead of its use.
// lrow = this lane's row offset within a step: wave * U * 4 + l / 16.
template <int TA, int TC, int U, bool VEC>
__device__ __forceinline__ void tn_rows(const double* __restrict__ A, int64_t lda,
                                        const double* __restrict__ B, int64_t ldb, int ka, int kb,
                                        int a_lane, int c_lane, int64_t r0, int64_t r1, int lrow,
                                        f64x4 (&acc)[TA][TC]) {
  constexpr int kStep = kTnWaves * U * 4;
  double xa[U][TA], xb[U][TC], ya[U][TA], yb[U][TC];
  tn_load<TA, TC, U, VEC>(A, lda, B, ldb, ka, kb, a_lane, c_lane, r0 + lrow, r1, xa, xb);
  for (int64_t rb = r0; rb < r1; rb += 2 * kStep) {  // wave-uniform trip count
    tn_load<TA, TC, U, VEC>(A, lda, B, ldb, ka, kb, a_lane, c_lane, rb + kStep + lrow, r1, ya, yb);
    tn_mfma<TA, TC, U>(acc, xa, xb);
    tn_load<TA, TC, U, VEC>(A, lda, B, ldb, ka, kb, a_lane, c_lane, rb + 2 * kStep + lrow, r1, xa, xb);
    tn_mfma<TA, TC, U>(acc, ya, yb);
  }
}

// partial[chunk][a][c] = sum over the chunk's rows of A[i][a] B[i][c]; the four
// waves' sums are added in wave order.
template <int TA, int TC>
__global__ void __launch_bounds__(256) gemm_tn_kernel(const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb, int64_t n,
                                                      int ka, int kb, int64_t rows_per_chunk, int vec_ok,
                                                      double* __restrict__ partial) {
  constexpr int U = TA * TC >= 4 ? 2 : 4;  // 4-row steps per register set
  __shared__ double red[TA * TC * 4 * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t chunk = blockIdx.x;
  const int a0 = blockIdx.y * 16 * TA;
  const int64_t r0 = chunk * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
  f64x4 acc[TA][TC];
#pragma unroll
  for (int t = 0; t < TA; ++t)
#pragma unroll
    for (int tc = 0; tc < TC; ++tc) acc[t][tc] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int a_lane = a0 + TA * (lane & 15);
  const int c_lane = lane & 15;
  const int lrow = wave * U * 4 + (lane >> 4);
  if (r0 < r1) {
    bool vec = false;
    if constexpr (TA % 2 == 0) vec = vec_ok && a0 + 16 * TA <= ka;
    if (vec) {
      if constexpr (TA % 2 == 0)
        tn_rows<TA, TC, U, true>(A, lda, B, ldb, ka, kb, a_lane, c_lane, r0, r1, lrow, acc);
    } else {
      tn_rows<TA, TC, U, false>(A, lda, B, ldb, ka, kb, a_lane, c_lane, r0, r1, lrow, acc);
    }
  }
  for (int s = 1; s < kTnWaves; ++s) {
    __syncthreads();
    if (wave == s) {
#pragma unroll
      for (int t = 0; t < TA; ++t)
#pragma unroll
        for (int tc = 0; tc < TC; ++tc)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[((t * TC + tc) * 4 + r) * 64 + lane] = acc[t][tc][r];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int t = 0; t < TA; ++t)
#pragma unroll
        for (int tc = 0; tc < TC; ++tc)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[t][tc][r] += red[((t * TC + tc) * 4 + r) * 64 + lane];
    }
  }
  if (wave != 0) return;
  double* P = partial + chunk * (int64_t)ka * kb;
#pragma unroll
  for (int t = 0; t < TA; ++t)
#pragma unroll
    for (int tc = 0; tc < TC; ++tc)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = a0 + TA * ((lane >> 4) + 4 * r) + t;
        const int c = 16 * tc + c_lane;
        if (a < ka && c < kb) P[(int64_t)a * kb + c] = acc[t][tc][r];
      }
}

// C[i] = sum over chunks of partial[chunk][i], chunks in ascending order within
// 16 segments, the segments then added in order.
constexpr int kRedSegs = 16;
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double* __restrict__ partial,
                                                              int64_t nchunks, int64_t count,
                                                              double* __restrict__ C) {
  __shared__ double red[kRedSegs][16];
  const int o = threadIdx.x & 15, seg = threadIdx.x >> 4;
  const int64_t i = (int64_t)blockIdx.x * 16 + o;
  double acc = 0.0;
  if (i < count) {
    const int64_t q0 = nchunks * seg / kRedSegs, q1 = nchunks * (seg + 1) / kRedSegs;
#pragma unroll 8
    for (int64_t q = q0; q < q1; ++q) acc += partial[q * count + i];
  }
  red[seg][o] = acc;
  __syncthreads();
  if (seg == 0 && i < count) {
    double s = red[0][o];
    for (int k = 1; k < kRedSegs; ++k) s += red[k][o];
    C[i] = s;
  }
}

// --- C = beta C + alpha A B -------------------------------------------------
// MFMA M = columns of C (16 per tile, TC tiles), N = rows, K = columns of A.
// B^T for up to kNnSlab columns of A is staged in LDS once per block. A wave
// owns 2 tiles of 16 rows; for each 32-column slice of A, lane l loads columns
// a0 + 8*(l/16) .. +7 of row l%16 of each tile (64 contiguous bytes), and MFMA
// e (e < 8) stands for column a0 + 8*(l/16) + e in K slot l/16. The sum over A's
// columns runs slice by slice in a fixed order.
constexpr int kNnWaves = 4;
constexpr int kNnTiles = 2;                          // 16-row tiles per wave
constexpr int kNnRows = kNnWaves * kNnTiles * 16;    // rows per block step
template <int TC>
struct NnSlab {
  static constexpr int cols = TC == 1 ? 384 : 192;   // columns of A per launch
  static constexpr int zero_slice = cols / 32;       // an all-zero slice after them
  static constexpr int stride = cols + 32 + 2;       // LDS row stride (doubles)
};

template <bool VEC>
__device__ __forceinline__ void nn_load(const double* __restrict__ A, int64_t lda, int ka,
                                        const int64_t (&rows)[kNnTiles], int col0,
                                        double (&v)[kNnTiles][8]) {
#pragma unroll
  for (int tt = 0; tt < kNnTiles; ++tt) {
    const double* ap = A + rows[tt] * lda;
    if constexpr (VEC) {
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const double2 x = *reinterpret_cast<const double2*>(ap + col0 + e);
        v[tt][e] = x.x;
        v[tt][e + 1] = x.y;
      }
    } else {
      // columns past ka read column ka - 1 and meet zero rows of the staged B^T
#pragma unroll
      for (int e = 0; e < 8; ++e) v[tt][e] = ap[min(col0 + e, ka - 1)];
    }
  }
}

template <int TC>
__device__ __forceinline__ void nn_mfma(f64x4 (&acc)[kNnTiles][TC], const double* bt, int lane, int s,
                                        const double (&v)[kNnTiles][8]) {
  // B^T values of this lane: row 16*tc + l%16, columns 32*s + 8*(l/16) + e
  const double* bp = bt + (lane & 15) * NnSlab<TC>::stride + 32 * s + 8 * (lane >> 4);
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
#pragma unroll
    for (int tc = 0; tc < TC; ++tc) {
      const double2 b = *reinterpret_cast<const double2*>(bp + 16 * tc * NnSlab<TC>::stride + e);
#pragma unroll
      for (int tt = 0; tt < kNnTiles; ++tt) {
        acc[tt][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(b.x, v[tt][e], acc[tt][tc], 0, 0, 0);
        acc[tt][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(b.y, v[tt][e + 1], acc[tt][tc], 0, 0, 0);
      }
    }
  }
}

// Slices [s0, s1) of 32 columns; two register sets, each loaded one slice ahead.
// An odd count ends with a pair whose second half meets the zero slice of B^T,
// so the loop body has no branch between a load and its use.
template <int TC, bool VEC>
__device__ __forceinline__ void nn_slices(const double* __restrict__ A, int64_t lda, int ka,
                                          const int64_t (&rows)[kNnTiles], const double* bt, int lane,
                                          int s0, int s1, f64x4 (&acc)[kNnTiles][TC]) {
  if (s0 >= s1) return;
  const int kofs = 8 * (lane >> 4);
  double x[kNnTiles][8], y[kNnTiles][8];
  nn_load<VEC>(A, lda, ka, rows, 32 * s0 + kofs, x);
  for (int s = s0; s < s1; s += 2) {
    nn_load<VEC>(A, lda, ka, rows, 32 * min(s + 1, s1 - 1) + kofs, y);
    nn_mfma<TC>(acc, bt, lane, s, x);
    nn_load<VEC>(A, lda, ka, rows, 32 * min(s + 2, s1 - 1) + kofs, x);
    nn_mfma<TC>(acc, bt, lane, s + 1 < s1 ? s + 1 : NnSlab<TC>::zero_slice, y);
  }
}

template <int TC>
__global__ void __launch_bounds__(256) gemm_nn_kernel(const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t n, int ka,
                                                      int kb, double alpha, double beta, int vec_ok,
                                                      double* __restrict__ C, int64_t ldc) {
  constexpr int S = NnSlab<TC>::stride;
  __shared__ __align__(16) double bt[16 * TC * S];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // B^T, zero-padded to 16*TC rows and whole 32-column slices, then the zero slice
  const int nsl = (ka + 31) / 32;
  for (int e = threadIdx.x; e < 16 * TC * nsl * 32; e += 256) {
    const int a = e / (16 * TC), c = e % (16 * TC);
    bt[c * S + a] = (a < ka && c < kb) ? B[(int64_t)a * kb + c] : 0.0;
  }
  for (int e = threadIdx.x; e < 16 * TC * 32; e += 256)
    bt[(e >> 5) * S + 32 * NnSlab<TC>::zero_slice + (e & 31)] = 0.0;
  __syncthreads();
  const int nfull = vec_ok ? ka / 32 : 0;
  for (int64_t rb = (int64_t)blockIdx.x * kNnRows; rb < n; rb += (int64_t)gridDim.x * kNnRows) {
    int64_t rows[kNnTiles];
#pragma unroll
    for (int tt = 0; tt < kNnTiles; ++tt)
      rows[tt] = min(n - 1, rb + (wave * kNnTiles + tt) * 16 + (lane & 15));
    f64x4 acc[kNnTiles][TC];
#pragma unroll
    for (int tt = 0; tt < kNnTiles; ++tt)
#pragma unroll
      for (int tc = 0; tc < TC; ++tc) acc[tt][tc] = f64x4{0.0, 0.0, 0.0, 0.0};
    nn_slices<TC, true>(A, lda, ka, rows, bt, lane, 0, nfull, acc);
    nn_slices<TC, false>(A, lda, ka, rows, bt, lane, nfull, nsl, acc);
#pragma unroll
    for (int tt = 0; tt < kNnTiles; ++tt) {
      const int64_t i = rb + (wave * kNnTiles + tt) * 16 + (lane & 15);
      if (i >= n) continue;
#pragma unroll
      for (int tc = 0; tc < TC; ++tc)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int c = 16 * tc + (lane >> 4) + 4 * r;
          if (c < kb) {
            const double prev = beta == 0.0 ? 0.0 : beta * C[i * ldc + c];
            C[i * ldc + c] = prev + alpha * acc[tt][tc][r];
          }
        }
    }
  }
}

__global__ void random_kernel(uint64_t seed, int64_t n, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)i * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  out[i] = ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0) - 0.5;
}

// ---------------------------------------------------------------------------
// Adjoint of the operator with respect to the feature weights. With u = row
// scales, d = degrees, a = d^-1/2, r = u a, e = X^T u, P = X^T (r o V) and
// Q = X^T (r o G), the derivative of f = <G, A(w) V> is
//   h_j   = <P_j, Q_j>
//   alpha = rowdot(G, X (w^2 o P)) + rowdot(V, X (w^2 o Q)),  beta = rowdot(G, V)
//   gamma = -alpha u d^-3/2 / 2 + beta d^-2
//   phi   = X^T (gamma o u),  psi = X (w^2 o phi)
//   kappa = alpha a u^3 + gamma u^2 (d + 1) + u^3 psi,  sigma = (X o X)^T kappa
//   df/dw = w o (2 h + 2 e o phi - sigma)
// covering the explicit w^2, the row norms u(w) and the degrees d(w). X W is
// never formed: every term is one pass over the stored counts.
// ---------------------------------------------------------------------------
__global__ void rowdot_f32_kernel(const float* A, const float* B, int64_t rows, int BW, int b,
                                  double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (int c = 0; c < b; ++c) s += (double)A[i * BW + c] * (double)B[i * BW + c];
  out[i] = s;
}

__global__ void scale_rows_f32_kernel(float* A, int64_t rows, int BW, const float* sc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * BW) return;
  A[t] *= sc[t / BW];
}

// [r o V | r o G | 0] in BW-wide FP32 rows (2b <= BW): both operands of the
// adjoint's sparse products in one pass each
template <int BW>
__global__ void scale_in_pair_kernel(const double* V, int64_t ldv, const double* G, int64_t ldg, int64_t n,
                                     int b, const float* r, float* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / BW;
  const int c = (int)(t % BW);
  if (i >= n) return;
  float v = 0.f;
  if (c < b) v = (float)((double)r[i] * V[i * ldv + c]);
  else if (c < 2 * b) v = (float)((double)r[i] * G[i * ldg + c - b]);
  out[i * BW + c] = v;
}

// alpha, gamma and t = gamma u per cell (all zero for an empty row)
__global__ void adj_rows_kernel(const double* G, int64_t ldg, const double* V, int64_t ldv,
                                const float* XP, const float* XQ, int BW, int b, int64_t n,
                                const double* u, const double* dinv, double* alpha, double* gamma,
                                double* t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double al = 0.0, be = 0.0;
  for (int c = 0; c < b; ++c) {
    const double g = G[i * ldg + c], v = V[i * ldv + c];
    al += g * (double)XP[i * BW + c] + v * (double)XQ[i * BW + c];
    be += g * v;
  }
  double ga = 0.0;
  if (u[i] > 0.0) ga = -0.5 * al * u[i] * dinv[i] * sqrt(dinv[i]) + be * dinv[i] * dinv[i];
  alpha[i] = al;
  gamma[i] = ga;
  t[i] = ga * u[i];
}

// pair[i] = (a[i], b[i]): two vectors gathered with one 16-byte load
__global__ void pack_pair_kernel(const double* a, const double* b, int64_t n, double2* pair) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pair[i] = make_double2(a[i], b[i]);
}

// pair[j] = (w_j^2, x_j)
__global__ void pack_w2x_kernel(const double* w, const double* x, int64_t m, double2* pair) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) pair[j] = make_double2(w[j] * w[j], x[j]);
}

// y1_j = sum_i cnt_ij^POW x1_i (and y2_j from the second component when x12 is
// given instead: x12[i] = (x1_i, x2_i)); one wave per column.
template <bool FV, int POW>
__global__ void __launch_bounds__(256) csc_spmv2_kernel(const int64_t* colptr, const int32_t* rowidx,
                                                        const uint32_t* cnt, int64_t m, const double* x1,
                                                        const double2* x12, double* y1, double* y2) {
  const int lane = threadIdx.x & 63;
  const int64_t col = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (col >= m) return;
  double2 a = make_double2(0.0, 0.0);
  const int64_t e0 = colptr[col] + lane, e1 = colptr[col + 1];
  if (x12 != nullptr) {
    lane_entries<FV>(rowidx, cnt, e0, e1, a, [&](int32_t i) { return x12[i]; },
                     [](double2& s, double v, double2 x) {
                       if (POW == 2) v *= v;
                       s.x = fma(v, x.x, s.x);
                       s.y = fma(v, x.y, s.y);
                     });
  } else {
    lane_entries<FV>(rowidx, cnt, e0, e1, a, [&](int32_t i) { return x1[i]; },
                     [](double2& s, double v, double x) {
                       if (POW == 2) v *= v;
                       s.x = fma(v, x, s.x);
                     });
  }
  double a1 = a.x, a2 = a.y;
  for (int o = 32; o > 0; o >>= 1) {
    a1 += __shfl_xor(a1, o, 64);
    a2 += __shfl_xor(a2, o, 64);
  }
  if (lane == 0) {
    y1[col] = a1;
    if (y2 != nullptr) y2[col] = a2;
  }
}

// y_i = sum_j cnt_ij w_j^2 x_j with w2x[j] = (w_j^2, x_j); one wave per row.
template <bool FV>
__global__ void __launch_bounds__(256) csr_spmv_w2_kernel(const int64_t* indptr, const int32_t* idx,
                                                          const uint32_t* cnt, int64_t n, const double2* w2x,
                                                          double* y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  double a = 0.0;
  lane_entries<FV>(idx, cnt, indptr[row] + lane, indptr[row + 1], a, [&](int32_t j) { return w2x[j]; },
                   [](double& s, double v, double2 q) { s = fma(v * q.x, q.y, s); });
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if (lane == 0) y[row] = a;
}

__global__ void kappa_kernel(int64_t n, const double* u, const double* dinv, const double* alpha,
                             const double* gamma, const double* psi, double* kappa) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ui = u[i];
  double k = 0.0;
  if (ui > 0.0) {
    const double u2 = ui * ui, u3 = u2 * ui;
    k = alpha[i] * sqrt(dinv[i]) * u3 + gamma[i] * u2 * (1.0 / dinv[i] + 1.0) + u3 * psi[i];
  }
  kappa[i] = k;
}

__global__ void dw_kernel(int64_t m, const double* w, const double* h, const double* e,
                          const double* phi, const double* sigma, double* dw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  dw[j] = w[j] * (2.0 * h[j] + 2.0 * e[j] * phi[j] - sigma[j]);
}

// Workspace of the adjoint: two n x 32 and two m x 32 FP32 blocks, then
// 5 n-vectors and 4 m-vectors in FP64, each 256-byte aligned.
static int64_t vjp_ws_bytes(int64_t n, int64_t m) {
  auto al = [](int64_t b) { return (b + 255) / 256 * 256; };
  return 2 * al(n * kB * 4) + 2 * al(m * kB * 4) + 5 * al(n * 8) + 4 * al(m * 8) + 256;
}

// The two FP32 products of the adjoint over the mixed layout (CSC panels, CSR
// windows) instead of the row-parallel spmm_kernel; 16-column rows only.
struct MixedProducts {
  const snapgpu_window_layout* csr;
  const snapgpu_panel_layout* csc;
  float* partial;  // the CSC panel product's pair sums
};

template <int BW, bool FV>
static void launch_vjp(const int64_t* indptr, const int32_t* idx, const uint32_t* cnt, int64_t n,
                       const int64_t* colptr, const int32_t* rowidx, const uint32_t* ccnt, int64_t m,
                       const double* u, const double* dinv, const double* w, const float* r32,
                       const float* w2, const double* V, int64_t ldv, const double* G, int64_t ldg,
                       int b, double* dw, void* ws, hipStream_t st, const MixedProducts* mixed = nullptr) {
  char* q = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  auto take = [&q](int64_t bytes) {
    char* p = q;
    q += (bytes + 255) / 256 * 256;
    return p;
  };
  float* Vs = (float*)take(n * kB * 4);
  float* Gs = (float*)take(n * kB * 4);
  float* Pm = (float*)take(m * kB * 4);
  float* Qm = (float*)take(m * kB * 4);
  double* alpha = (double*)take(n * 8);
  double* gamma = (double*)take(n * 8);
  double* t = (double*)take(n * 8);
  double* psi = (double*)take(n * 8);
  double* kappa = (double*)take(n * 8);
  double* h = (double*)take(m * 8);
  double* e = (double*)take(m * 8);
  double* phi = (double*)take(m * 8);
  double* sigma = (double*)take(m * 8);
  const unsigned gn = (unsigned)ceil_div(n, 4), gm = (unsigned)ceil_div(m, 4);
  if (2 * b <= BW) {
    // V and G side by side: one CSC and one CSR pass
    scale_in_pair_kernel<BW><<<(unsigned)ceil_div(n * BW, 256), 256, 0, st>>>(V, ldv, G, ldg, n, b, r32, Vs);
    // [P | Q] = X^T (r o [V | G])
    if (BW == 16 && mixed)
      panel::product16_scaled(*mixed->csc, Vs, mixed->partial, nullptr, Pm, st);
    else
      spmm_kernel<BW, 0, FV><<<gm, 256, 0, st>>>(colptr, rowidx, ccnt, m, Vs, nullptr, Pm, nullptr, 0,
                                                 nullptr, nullptr, 0, 2 * b);
    rowdot_f32_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(Pm, Pm + b, m, BW, b, h);
    scale_rows_f32_kernel<<<(unsigned)ceil_div(m * BW, 256), 256, 0, st>>>(Pm, m, BW, w2);
    // X (w^2 o [P | Q]) into Vs (free after the CSC pass)
    if (BW == 16 && mixed)
      window::product16_scaled(*mixed->csr, Pm, nullptr, Vs, st);
    else
      spmm_kernel<BW, 0, FV><<<gn, 256, 0, st>>>(indptr, idx, cnt, n, Pm, nullptr, Vs, nullptr, 0,
                                                 nullptr, nullptr, 0, 2 * b);
    adj_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(G, ldg, V, ldv, Vs, Vs + b, BW, b, n, u,
                                                                dinv, alpha, gamma, t);
  } else {
    scale_in_kernel<BW><<<(unsigned)ceil_div(n * BW, 256), 256, 0, st>>>(V, ldv, n, b, r32, Vs);
    scale_in_kernel<BW><<<(unsigned)ceil_div(n * BW, 256), 256, 0, st>>>(G, ldg, n, b, r32, Gs);
    // P = X^T (r o V), Q = X^T (r o G)
    spmm_kernel<BW, 0, FV><<<gm, 256, 0, st>>>(colptr, rowidx, ccnt, m, Vs, nullptr, Pm, nullptr, 0,
                                               nullptr, nullptr, 0, b);
    spmm_kernel<BW, 0, FV><<<gm, 256, 0, st>>>(colptr, rowidx, ccnt, m, Gs, nullptr, Qm, nullptr, 0,
                                               nullptr, nullptr, 0, b);
    rowdot_f32_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(Pm, Qm, m, BW, b, h);
    scale_rows_f32_kernel<<<(unsigned)ceil_div(m * BW, 256), 256, 0, st>>>(Pm, m, BW, w2);
    scale_rows_f32_kernel<<<(unsigned)ceil_div(m * BW, 256), 256, 0, st>>>(Qm, m, BW, w2);
    // X (w^2 o P) into Vs and X (w^2 o Q) into Gs (both free after the CSC pass)
    spmm_kernel<BW, 0, FV><<<gn, 256, 0, st>>>(indptr, idx, cnt, n, Pm, nullptr, Vs, nullptr, 0,
                                               nullptr, nullptr, 0, b);
    spmm_kernel<BW, 0, FV><<<gn, 256, 0, st>>>(indptr, idx, cnt, n, Qm, nullptr, Gs, nullptr, 0,
                                               nullptr, nullptr, 0, b);
    adj_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(G, ldg, V, ldv, Vs, Gs, BW, b, n, u, dinv,
                                                                alpha, gamma, t);
  }
  // phi = X^T (gamma o u) and e = X^T u in one pass over the CSC
  // (t, u) side by side in Gs and (w^2, phi) in Qm, both free by now
  double2* tu = reinterpret_cast<double2*>(Gs);
  double2* w2phi = reinterpret_cast<double2*>(Qm);
  pack_pair_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(t, u, n, tu);
  csc_spmv2_kernel<FV, 1><<<gm, 256, 0, st>>>(colptr, rowidx, ccnt, m, nullptr, tu, phi, e);
  pack_w2x_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(w, phi, m, w2phi);
  csr_spmv_w2_kernel<FV><<<gn, 256, 0, st>>>(indptr, idx, cnt, n, w2phi, psi);
  kappa_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, u, dinv, alpha, gamma, psi, kappa);
  csc_spmv2_kernel<FV, 2><<<gm, 256, 0, st>>>(colptr, rowidx, ccnt, m, kappa, nullptr, sigma, nullptr);
  dw_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, w, h, e, phi, sigma, dw);
}

}  // namespace spectral
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::spectral;

static inline unsigned grid_for(int64_t work, int per_block, int64_t cap = 1 << 20) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, per_block), cap));
}

// --- C = A^T B over row chunks ------------------------------------------------
// Up to 512 chunks of rows; each writes its ka x kb partial product, then the
// partials are added in chunk order.
static int64_t gemm_tn_chunks(int64_t n) {
  return std::max<int64_t>(1, std::min<int64_t>(512, ceil_div(n, 4096)));
}

static int64_t gemm_tn_partial_bytes(int64_t n, int32_t ka, int32_t kb) {
  return gemm_tn_chunks(n) * (int64_t)ka * kb * 8;
}

static int gemm_tn_launch(const double* d_A, int64_t lda, const double* d_B, int64_t ldb, int64_t n,
                          int32_t ka, int32_t kb, double* d_C, double* part, hipStream_t st) {
  const int64_t chunks = gemm_tn_chunks(n);
  const int64_t rpc = ceil_div(std::max<int64_t>(n, 1), chunks);
  // 16*TA columns of A per block: as few column groups as 128-wide ones need,
  // each as narrow as covers ka
  const int groups = (int)ceil_div(ka, 128);
  int ta = (int)ceil_div(ka, 16 * groups);
  if (ta > 1 && (ta & 1)) ++ta;
  const int tc = kb <= 16 ? 1 : 2;
  const int vec_ok = (lda % 2 == 0) && (((uintptr_t)d_A & 15) == 0);
  const dim3 grid((unsigned)chunks, (unsigned)ceil_div(ka, 16 * ta));
#define SG_TN(TA, TC)                                                                              \
  gemm_tn_kernel<TA, TC><<<grid, 256, 0, st>>>(d_A, lda, d_B, ldb, n, ka, kb, rpc, vec_ok, part)
#define SG_TN_TA(TC)      \
  switch (ta) {           \
    case 1: SG_TN(1, TC); break; \
    case 2: SG_TN(2, TC); break; \
    case 4: SG_TN(4, TC); break; \
    case 6: SG_TN(6, TC); break; \
    default: SG_TN(8, TC); break; \
  }
  if (tc == 1) {
    SG_TN_TA(1)
  } else {
    SG_TN_TA(2)
  }
#undef SG_TN_TA
#undef SG_TN
  SG_LAUNCH_CHECK();
  const int64_t count = (int64_t)ka * kb;
  reduce_partials_kernel<<<(unsigned)ceil_div(count, 16), 256, 0, st>>>(part, chunks, count, d_C);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

// Static scratch of snapgpu_gemm_tn_f64_noalloc: one copy per GPU, held until the
// process ends. In host code `scratch` only names the symbol; its device address
// on the current GPU comes from hipGetSymbolAddress, looked up once per device
// and kept in g_scratch_addr.
static __device__ __align__(256) char scratch[SNAPGPU_GEMM_TN_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

static int scratch_address(double** out) {
  int dev = -1;
  SG_HIP(hipGetDevice(&dev));
  SG_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SG_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[dev] = p;
  }
  *out = (double*)g_scratch_addr[dev];
  return SNAPGPU_OK;
}

extern "C" {

int snapgpu_col_select_table(const int64_t* d_cols, int64_t k, int64_t m, uint32_t* d_sel, void* stream) {
  clear_error();
  SG_ARG(k >= 0 && m >= 1 && k <= m && k < (1ll << 31), "bad selection size");
  SG_ARG(d_sel && (k == 0 || d_cols), "null pointer");
  const int64_t words = ceil_div(m, 32);
  col_select_table_kernel<<<(unsigned)ceil_div(words, 256), 256, 0, as_stream(stream)>>>(d_cols, k, words,
                                                                                         d_sel);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_csr_select_count(const int64_t* d_indptr, const void* d_idx, int32_t index_bits,
                             int64_t n, const uint32_t* d_sel, int64_t m, int64_t* d_row_nnz,
                             void* stream) {
  clear_error();
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  if (n == 0) return SNAPGPU_OK;
  const unsigned g = (unsigned)ceil_div(n, 4);
  const int64_t words = ceil_div(m, 32);
  if (index_bits == 32)
    select_window_kernel<false><<<(unsigned)ceil_div(n, kSelRows), kSelThreads, 0, as_stream(stream)>>>(
        d_indptr, (const int32_t*)d_idx, nullptr, n, d_sel, words, d_row_nnz, nullptr, nullptr, nullptr);
  else
    select_count_kernel<int64_t><<<g, 256, 0, as_stream(stream)>>>(d_indptr, (const int64_t*)d_idx, n, d_sel,
                                                                   words, d_row_nnz);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_csr_select_fill(const int64_t* d_indptr, const void* d_idx, int32_t index_bits,
                            const uint32_t* d_val, int64_t n, const uint32_t* d_sel, int64_t m,
                            const int64_t* d_new_indptr, int32_t* d_new_idx, uint32_t* d_new_val,
                            void* stream) {
  clear_error();
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  if (n == 0) return SNAPGPU_OK;
  const unsigned g = (unsigned)ceil_div(n, 4);
  const int64_t words = ceil_div(m, 32);
  if (index_bits == 32)
    select_window_kernel<true><<<(unsigned)ceil_div(n, kSelRows), kSelThreads, 0, as_stream(stream)>>>(
        d_indptr, (const int32_t*)d_idx, d_val, n, d_sel, words, nullptr, d_new_indptr, d_new_idx, d_new_val);
  else
    select_fill_kernel<int64_t><<<g, 256, 0, as_stream(stream)>>>(d_indptr, (const int64_t*)d_idx, d_val, n,
                                                                  d_sel, words, d_new_indptr, d_new_idx,
                                                                  d_new_val);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_csr_transpose_workspace_bytes(int64_t m, int64_t chunk_nnz) {
  size_t tmp = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (int)std::min<int64_t>(chunk_nnz, INT32_MAX));
  return (int64_t)m * 8 + 256 + chunk_nnz * 24 + 5 * 256 + (int64_t)tmp + 256;
}

// Stable transpose. colptr[m+1] out; rows ascending inside each column.
// chunk_nnz bounds the entries sorted at once (whole rows per chunk).
int snapgpu_csr_transpose(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_val,
                          int64_t n, int64_t m, int64_t nnz, const int64_t* h_indptr,
                          int64_t chunk_nnz, int64_t* d_colptr, int32_t* d_rowidx,
                          uint32_t* d_cval, void* d_ws, int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(n >= 0 && m >= 1 && nnz >= 0, "bad shape");
  SG_ARG(m < (1ll << 32) && n < (1ll << 32), "shape too large for 32-bit ids");
  SG_ARG(chunk_nnz >= 1 && chunk_nnz <= INT32_MAX, "bad chunk size");
  SG_ARG(ws_bytes >= snapgpu_csr_transpose_workspace_bytes(m, chunk_nnz), "workspace too small");
  SG_ARG(h_indptr != nullptr, "host copy of indptr required");
  hipStream_t st = as_stream(stream);
  char* base = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  int64_t* cursor = (int64_t*)base;
  char* p = base + (m * 8 + 255) / 256 * 256;
  auto take = [&](int64_t bytes) { char* q = p; p += (bytes + 255) / 256 * 256; return q; };
  uint32_t* keys_in = (uint32_t*)take(chunk_nnz * 4);
  uint32_t* keys_out = (uint32_t*)take(chunk_nnz * 4);
  uint64_t* vals_in = (uint64_t*)take(chunk_nnz * 8);
  uint64_t* vals_out = (uint64_t*)take(chunk_nnz * 8);
  int32_t* wstatus = (int32_t*)take(4);
  void* tmp = p;
  size_t tmp_bytes = (size_t)(ws_bytes - (p - (char*)d_ws));
  // column counts -> colptr
  SG_HIP(hipMemsetAsync(d_colptr, 0, (m + 1) * 8, st));
  if (nnz > 0) {
    bool sorted = true;
    int rc = col_counts_sorted(d_indptr, n, d_idx, nnz, m, (unsigned long long*)(d_colptr + 1), wstatus,
                               st, &sorted);
    if (rc) return rc;
    if (!sorted) {  // rows out of order: one atomic per entry
      SG_HIP(hipMemsetAsync(d_colptr, 0, (m + 1) * 8, st));
      col_count_kernel<<<grid_for(nnz, 256, 16384), 256, 0, st>>>(d_idx, nnz,
                                                                  (unsigned long long*)(d_colptr + 1));
      SG_LAUNCH_CHECK();
    }
  }
  const int64_t sws = snapgpu_scan_workspace_bytes(m);
  SG_ARG((int64_t)tmp_bytes >= sws, "workspace too small for the scan");
  int rc = snapgpu_scan_inclusive_i64(d_colptr + 1, m, tmp, sws, stream);
  if (rc) return rc;
  SG_HIP(hipMemcpyAsync(cursor, d_colptr, m * 8, hipMemcpyDeviceToDevice, st));
  int end_bit = 1;
  while (end_bit < 32 && (1ll << end_bit) < m) ++end_bit;
  int64_t r0 = 0;
  while (r0 < n) {
    // whole rows, at most chunk_nnz entries (a single longer row is an error)
    int64_t lo = r0 + 1, hi = n;
    if (h_indptr[r0 + 1] - h_indptr[r0] > chunk_nnz) {
      set_error("a row has more entries than the transpose chunk");
      return SNAPGPU_ERR_CAPACITY;
    }
    while (lo < hi) {  // last r1 with indptr[r1] - indptr[r0] <= chunk_nnz
      const int64_t mid = (lo + hi + 1) >> 1;
      if (h_indptr[mid] - h_indptr[r0] <= chunk_nnz) lo = mid; else hi = mid - 1;
    }
    const int64_t r1 = lo;
    const int64_t e0 = h_indptr[r0], cn = h_indptr[r1] - e0;
    if (cn > 0) {
      SG_HIP(hipMemcpyAsync(keys_in, d_idx + e0, cn * 4, hipMemcpyDeviceToDevice, st));
      entry_pack_kernel<<<(unsigned)ceil_div(r1 - r0, 4), 256, 0, st>>>(d_indptr, r0, r1, d_val, vals_in);
      SG_LAUNCH_CHECK();
      SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                (int)cn, 0, end_bit, st));
      cursor_rebase_kernel<<<grid_for(cn, 256, 16384), 256, 0, st>>>(keys_out, cn, cursor);
      SG_LAUNCH_CHECK();
      scatter_chunk_kernel<<<grid_for(cn, 256, 16384), 256, 0, st>>>(keys_out, vals_out, cn, cursor, d_rowidx,
                                                                     d_cval);
      SG_LAUNCH_CHECK();
      cursor_advance_kernel<<<grid_for(cn, 256, 16384), 256, 0, st>>>(keys_out, cn, cursor);
      SG_LAUNCH_CHECK();
    }
    r0 = r1;
  }
  return SNAPGPU_OK;
}

int snapgpu_spectral_idf(const int64_t* d_colptr, int64_t m, int64_t n, double* d_w, void* d_ws16,
                         void* stream) {
  clear_error();
  SG_ARG(m >= 1 && n >= 1, "bad shape");
  hipStream_t st = as_stream(stream);
  unsigned long long init[2] = {~0ull, 0ull};
  SG_HIP(hipMemcpyAsync(d_ws16, init, 16, hipMemcpyHostToDevice, st));
  df_minmax_kernel<<<grid_for(m, 256, 1024), 256, 0, st>>>(d_colptr, m, (unsigned long long*)d_ws16);
  SG_LAUNCH_CHECK();
  idf_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(d_colptr, m, n, (const unsigned long long*)d_ws16, d_w);
  SG_LAUNCH_CHECK();
  SG_HIP(hipStreamSynchronize(st));  // `init` lives on this stack frame
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_prepare_workspace_bytes(int64_t m) { return m * (int64_t)sizeof(double2) + 256; }

int snapgpu_spectral_prepare(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt,
                             int64_t n, const int64_t* d_colptr, const int32_t* d_rowidx,
                             const uint32_t* d_ccnt, int64_t m, const double* d_w, double* d_s,
                             double* d_c, double* d_dinv, double* d_r64, float* d_r32, float* d_w2,
                             int32_t value_type, int32_t* d_status, void* d_ws, int64_t ws_bytes,
                             void* stream) {
  clear_error();
  SG_ARG(n >= 1 && m >= 1, "bad shape");
  SG_ARG(value_type == 0 || value_type == 1, "value_type must be 0 (uint32) or 1 (float32)");
  SG_ARG(d_ws != nullptr && ws_bytes >= snapgpu_spectral_prepare_workspace_bytes(m), "workspace too small");
  hipStream_t st = as_stream(stream);
  double2* wc = (double2*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  SG_HIP(hipMemsetAsync(d_status, 0, 8 * sizeof(int32_t), st));
  const unsigned gn = (unsigned)ceil_div(n, 4), gm = (unsigned)ceil_div(m, 4);
  if (value_type == 0) {
    row_scale_kernel<false><<<gn, 256, 0, st>>>(d_indptr, d_idx, d_cnt, n, d_w, d_s);
    colsum_kernel<false><<<gm, 256, 0, st>>>(d_colptr, d_rowidx, d_ccnt, m, d_w, d_s, d_c);
  } else {
    row_scale_kernel<true><<<gn, 256, 0, st>>>(d_indptr, d_idx, d_cnt, n, d_w, d_s);
    colsum_kernel<true><<<gm, 256, 0, st>>>(d_colptr, d_rowidx, d_ccnt, m, d_w, d_s, d_c);
  }
  SG_LAUNCH_CHECK();
  pack_pair_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(d_w, d_c, m, wc);
  SG_LAUNCH_CHECK();
  if (value_type == 0)
    degree_kernel<false><<<gn, 256, 0, st>>>(d_indptr, d_idx, d_cnt, n, wc, d_s, d_dinv, d_r64, d_r32, d_status);
  else
    degree_kernel<true><<<gn, 256, 0, st>>>(d_indptr, d_idx, d_cnt, n, wc, d_s, d_dinv, d_r64, d_r32, d_status);
  SG_LAUNCH_CHECK();
  square_f32_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(d_w, m, d_w2);
  SG_LAUNCH_CHECK();
  int32_t h[8];
  SG_HIP(hipMemcpyAsync(h, d_status, sizeof(h), hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (h[0] != 0) {
    const int64_t row = (int64_t)(uint32_t)h[1] | ((int64_t)h[2] << 32);
    set_error("cell " + std::to_string(row) +
              " shares no selected feature with any other cell (degree <= 0)");
    return SNAPGPU_ERR_DATA;
  }
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_apply_workspace_bytes(int64_t n, int64_t m) {
  return (n * kB + m * kB) * 4 + 512;
}

int snapgpu_spectral_apply(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt,
                           int64_t n, const int64_t* d_colptr, const int32_t* d_rowidx,
                           const uint32_t* d_ccnt, int64_t m, const float* d_r32, const float* d_w2,
                           const double* d_dinv, const double* d_V, int64_t ldv, int32_t b,
                           double* d_Y, int64_t ldy, int32_t value_type, void* d_ws, int64_t ws_bytes,
                           void* stream) {
  clear_error();
  SG_ARG(b >= 1 && b <= kB, "block width must be in [1, 32]");
  SG_ARG(value_type == 0 || value_type == 1, "value_type must be 0 (uint32) or 1 (float32)");
  SG_ARG(ldv >= b && ldy >= b, "leading dimension < b");
  SG_ARG(ws_bytes >= snapgpu_spectral_apply_workspace_bytes(n, m), "workspace too small");
  hipStream_t st = as_stream(stream);
  float* Vs = (float*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  float* Z = Vs + n * kB;
  const bool fv = value_type == 1;
#define SG_APPLY(BW)                                                                       \
  if (fv) launch_apply<BW, true>(d_indptr, d_idx, d_cnt, n, d_colptr, d_rowidx, d_ccnt, m, d_r32, \
                                 d_w2, d_dinv, d_V, ldv, b, d_Y, ldy, Vs, Z, st);              \
  else launch_apply<BW, false>(d_indptr, d_idx, d_cnt, n, d_colptr, d_rowidx, d_ccnt, m, d_r32, \
                               d_w2, d_dinv, d_V, ldv, b, d_Y, ldy, Vs, Z, st);
  if (b <= 8) {
    SG_APPLY(8)
  } else if (b <= 16) {
    SG_APPLY(16)
  } else {
    SG_APPLY(32)
  }
#undef SG_APPLY
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_vjp_workspace_bytes(int64_t n, int64_t m) { return vjp_ws_bytes(n, m); }

int snapgpu_spectral_vjp_w(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                           const int64_t* d_colptr, const int32_t* d_rowidx, const uint32_t* d_ccnt,
                           int64_t m, const double* d_w, const double* d_s, const double* d_dinv,
                           const float* d_r32, const float* d_w2, const double* d_V, int64_t ldv,
                           const double* d_G, int64_t ldg, int32_t b, double* d_dw, int32_t value_type,
                           void* d_ws, int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(n >= 1 && m >= 1, "bad shape");
  SG_ARG(b >= 1 && b <= kB, "block width must be in [1, 32]");
  SG_ARG(value_type == 0 || value_type == 1, "value_type must be 0 (uint32) or 1 (float32)");
  SG_ARG(ldv >= b && ldg >= b, "leading dimension < b");
  SG_ARG(d_ws != nullptr && ws_bytes >= vjp_ws_bytes(n, m), "workspace too small");
  hipStream_t st = as_stream(stream);
  const bool fv = value_type == 1;
#define SG_VJP(BW)                                                                                   \
  if (fv) launch_vjp<BW, true>(d_indptr, d_idx, d_cnt, n, d_colptr, d_rowidx, d_ccnt, m, d_s, d_dinv, \
                               d_w, d_r32, d_w2, d_V, ldv, d_G, ldg, b, d_dw, d_ws, st);            \
  else launch_vjp<BW, false>(d_indptr, d_idx, d_cnt, n, d_colptr, d_rowidx, d_ccnt, m, d_s, d_dinv,   \
                             d_w, d_r32, d_w2, d_V, ldv, d_G, ldg, b, d_dw, d_ws, st);
  // width of the FP32 operand rows: V and G side by side while 2b fits 32
  const int need = b <= 16 ? 2 * b : b;
  if (need <= 8) {
    SG_VJP(8)
  } else if (need <= 16) {
    SG_VJP(16)
  } else {
    SG_VJP(32)
  }
#undef SG_VJP
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_vjp_mixed_workspace_bytes(int64_t pairs_csc) {
  return (std::max<int64_t>(pairs_csc, 1) * 16 * 4 + 255) / 256 * 256 + 256;
}

int snapgpu_spectral_vjp_w_mixed(const snapgpu_window_layout* csr_windows, const snapgpu_panel_layout* csc_panels,
                                 const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                                 const int64_t* d_colptr, const int32_t* d_rowidx, const uint32_t* d_ccnt,
                                 int64_t m, const double* d_w, const double* d_s, const double* d_dinv,
                                 const float* d_r32, const float* d_w2, const double* d_V, int64_t ldv,
                                 const double* d_G, int64_t ldg, int32_t b, double* d_dw, void* d_ws,
                                 int64_t ws_bytes, void* d_pws, int64_t pws_bytes, void* stream) {
  clear_error();
  SG_ARG(csr_windows != nullptr && csc_panels != nullptr, "null layout");
  SG_ARG(n >= 1 && m >= 1, "bad shape");
  SG_ARG(csr_windows->n_rows == n && csr_windows->dense_rows == m && csc_panels->n_rows == m &&
             csc_panels->dense_rows == n,
         "layouts do not describe this matrix and its transpose");
  SG_ARG(csc_panels->packed >= 0 && csc_panels->packed <= 2 && csc_panels->value_type == 0,
         "the mixed layout holds integer counts");
  SG_ARG(b >= 1 && b <= 8, "block width must be in [1, 8] (V and G side by side in 16 columns)");
  SG_ARG(ldv >= b && ldg >= b, "leading dimension < b");
  SG_ARG(d_ws != nullptr && ws_bytes >= vjp_ws_bytes(n, m), "workspace too small");
  SG_ARG(d_pws != nullptr && pws_bytes >= snapgpu_spectral_vjp_mixed_workspace_bytes(csc_panels->n_pairs),
         "panel workspace too small");
  hipStream_t st = as_stream(stream);
  MixedProducts mp{csr_windows, csc_panels,
                   reinterpret_cast<float*>(((uintptr_t)d_pws + 255) & ~(uintptr_t)255)};
  launch_vjp<16, false>(d_indptr, d_idx, d_cnt, n, d_colptr, d_rowidx, d_ccnt, m, d_s, d_dinv, d_w, d_r32, d_w2,
                        d_V, ldv, d_G, ldg, b, d_dw, d_ws, st, &mp);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_gemm_tn_workspace_bytes(int64_t n, int32_t ka, int32_t kb) {
  return gemm_tn_partial_bytes(n, ka, kb) + 256;
}

// C[ka x kb] = A[n x ka]^T B[n x kb] (row-major, leading dims lda/ldb).
int snapgpu_gemm_tn_f64(const double* d_A, int64_t lda, const double* d_B, int64_t ldb, int64_t n,
                        int32_t ka, int32_t kb, double* d_C, void* d_ws, int64_t ws_bytes,
                        void* stream) {
  clear_error();
  SG_ARG(ka >= 1 && kb >= 1 && kb <= 32, "need 1 <= kb <= 32");
  SG_ARG(n >= 0 && lda >= ka && ldb >= kb, "need n >= 0, lda >= ka, ldb >= kb");
  SG_ARG(ws_bytes >= snapgpu_gemm_tn_workspace_bytes(n, ka, kb), "workspace too small");
  return gemm_tn_launch(d_A, lda, d_B, ldb, n, ka, kb, d_C, (double*)d_ws, as_stream(stream));
}

// The same product with the partial products in the library's static scratch
// array of this GPU (no workspace argument, no allocation).
int snapgpu_gemm_tn_f64_noalloc(const double* d_A, int64_t lda, const double* d_B, int64_t ldb,
                                int64_t n, int32_t ka, int32_t kb, double* d_C, void* stream) {
  clear_error();
  SG_ARG(ka >= 1 && kb >= 1 && kb <= 32, "need 1 <= kb <= 32");
  SG_ARG(n >= 0 && lda >= ka && ldb >= kb, "need n >= 0, lda >= ka, ldb >= kb");
  const int64_t need = gemm_tn_partial_bytes(n, ka, kb);
  if (need > SNAPGPU_GEMM_TN_SCRATCH_BYTES) {
    set_error("capacity: snapgpu_gemm_tn_f64_noalloc needs " + std::to_string(need) +
              " bytes of partial products (512 row chunks x ka x kb doubles at most), more than its " +
              std::to_string((long long)SNAPGPU_GEMM_TN_SCRATCH_BYTES) +
              "-byte static scratch; call snapgpu_gemm_tn_f64 with a workspace instead");
    return SNAPGPU_ERR_CAPACITY;
  }
  double* part = nullptr;
  const int rc = scratch_address(&part);
  if (rc) return rc;
  return gemm_tn_launch(d_A, lda, d_B, ldb, n, ka, kb, d_C, part, as_stream(stream));
}

// C[n x kb] = beta*C + alpha * A[n x ka] B[ka x kb] (B row-major, contiguous).
// Columns of A beyond one LDS slab of B^T go through further launches that add
// to C.
int snapgpu_gemm_nn_f64(const double* d_A, int64_t lda, const double* d_B, int64_t n, int32_t ka,
                        int32_t kb, double alpha, double beta, double* d_C, int64_t ldc,
                        void* stream) {
  clear_error();
  SG_ARG(ka >= 1 && kb >= 1 && kb <= 32, "need 1 <= kb <= 32");
  SG_ARG(lda >= ka && ldc >= kb, "need lda >= ka, ldc >= kb");
  if (n == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  const int tc = kb <= 16 ? 1 : 2;
  const int slab = tc == 1 ? NnSlab<1>::cols : NnSlab<2>::cols;
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, kNnRows), 4096);
  for (int a0 = 0; a0 < ka; a0 += slab) {
    const int w = std::min(slab, ka - a0);
    const double* A = d_A + a0;
    const int vec_ok = (lda % 2 == 0) && (((uintptr_t)A & 15) == 0);
    const double bt = a0 == 0 ? beta : 1.0;
    if (tc == 1)
      gemm_nn_kernel<1><<<grid, 256, 0, st>>>(A, lda, d_B + (int64_t)a0 * kb, n, w, kb, alpha, bt,
                                               vec_ok, d_C, ldc);
    else
      gemm_nn_kernel<2><<<grid, 256, 0, st>>>(A, lda, d_B + (int64_t)a0 * kb, n, w, kb, alpha, bt,
                                               vec_ok, d_C, ldc);
    SG_LAUNCH_CHECK();
  }
  return SNAPGPU_OK;
}

int snapgpu_random_f64(uint64_t seed, int64_t count, double* d_out, void* stream) {
  clear_error();
  if (count <= 0) return SNAPGPU_OK;
  random_kernel<<<(unsigned)ceil_div(count, 256), 256, 0, as_stream(stream)>>>(seed, count, d_out);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
