// Panel layout of a sparse count matrix for the operator's two sparse products.
//
// Each product gathers one FP32 row of a dense operand (b <= 32 columns) per
// stored entry. The dense operand (up to 64 MB at 1M cells) does not fit one
// XCD's 4 MB L2, so the matrix is re-laid out panel-major: panel p holds the
// entries whose column lies in [p * 32768, (p + 1) * 32768), rows ascending,
// columns ascending within a row. One launch walks all panels in order, so the
// waves resident at any time gather from one 2 MB (b = 16) slice of the dense
// operand, which stays in L2.
//
// Work is split by entries, not rows: the entries of each panel are cut into
// group-chunks of about kChunk entries, aligned to (row, panel) segment starts,
// and each group of b lanes streams one group-chunk. Every (row, panel) segment
// with at least one entry is a "pair"; its FP32 sum is stored once in
// partial[pair] by the group that owns it. A second kernel adds a row's pairs in
// panel order and applies the epilogue. No floating-point atomics, fixed
// summation order: results are bitwise reproducible.
//
// Entry encodings (`packed`):
//   2 (counts < 128): bits 6-20 column within the panel times 64 (the byte
//     offset of its 16-float row), bits 24-30 the count, bit 31 first entry of
//     a pair. The 16-column product adds the lane's column and the panel's
//     base to get a 32-bit gather offset with one instruction, and converts the
//     count with one more.
//   1 (counts < 65536): bits 0-14 column within the panel, bit 15 first entry
//     of a pair, bits 16-31 the count.
//   0 (float values or larger counts): column | first << 31 in `ent`, the raw
//     32-bit value in `val`.
#include <algorithm>
#include <utility>
#include <vector>

#include "spectral_common.h"

namespace snapgpu {
namespace panel {

using spectral::kB;
using spectral::scale_in_kernel;

#ifndef SNAPGPU_PANEL_ROWS
#define SNAPGPU_PANEL_ROWS 32768
#endif
#ifndef SNAPGPU_PANEL_CHUNK
#define SNAPGPU_PANEL_CHUNK 4096
#endif
constexpr int64_t kPanel = SNAPGPU_PANEL_ROWS;  // dense rows per panel (<= 32768: 15-bit column)
constexpr int64_t kChunk = SNAPGPU_PANEL_CHUNK;  // entries per group-chunk (aligned to pair starts)
#ifndef SNAPGPU_PANEL_INFLIGHT
#define SNAPGPU_PANEL_INFLIGHT 8
#endif
constexpr int kInflight = SNAPGPU_PANEL_INFLIGHT;  // gathers in flight per lane
#ifndef SNAPGPU_PANEL_INFLIGHT16
#define SNAPGPU_PANEL_INFLIGHT16 8
#endif
constexpr int kInflight16 = SNAPGPU_PANEL_INFLIGHT16;  // the same, 16-column packed-8 product
static_assert(kPanel <= 32768, "the packed entry holds a 15-bit column");

// seg[row][p] = offset within the row of its first entry with idx >= p * kPanel
__global__ void __launch_bounds__(256) segment_kernel(const int64_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx, int64_t nrows,
                                                      int P, uint32_t* __restrict__ seg) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < nrows; row += nwaves) {
    const int64_t e0 = ptr[row], len = ptr[row + 1] - e0;
    for (int pb = lane; pb <= P; pb += 64) {
      const int64_t bound = (int64_t)pb * kPanel;
      int64_t lo = 0, hi = len;  // first entry with idx >= bound
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)idx[e0 + mid] < bound) lo = mid + 1; else hi = mid;
      }
      seg[row * (P + 1) + pb] = (uint32_t)lo;
    }
  }
}

// q = p * nrows + row: F[q + 1] = segment non-empty, L[q + 1] = segment length;
// F[0] = L[0] = 0, so inclusive scans give exclusive prefixes at q.
__global__ void pair_count_kernel(const uint32_t* __restrict__ seg, int64_t nrows, int P,
                                  int64_t* __restrict__ F, int64_t* __restrict__ L) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t Q = nrows * P;
  if (q == 0) {
    F[0] = 0;
    L[0] = 0;
  }
  if (q >= Q) return;
  const int64_t p = q / nrows, row = q - p * nrows;
  const int64_t len = (int64_t)seg[row * (P + 1) + p + 1] - (int64_t)seg[row * (P + 1) + p];
  F[q + 1] = len > 0 ? 1 : 0;
  L[q + 1] = len;
}

__global__ void max_u32_kernel(const uint32_t* __restrict__ v, int64_t n, unsigned int* __restrict__ out) {
  unsigned int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, v[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// info[p] = {S(p, 0), E(p, 0)} for p = 0..P (row 0 of panel p; p = P is the total)
__global__ void panel_info_kernel(const int64_t* __restrict__ S, const int64_t* __restrict__ E, int64_t nrows,
                                  int P, int64_t* __restrict__ info) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  info[2 * p] = S[(int64_t)p * nrows];
  info[2 * p + 1] = E[(int64_t)p * nrows];
}

// rp_ptr[row + 1] = number of panels in which the row has entries
__global__ void row_pair_count_kernel(const uint32_t* __restrict__ seg, int64_t nrows, int P,
                                      int64_t* __restrict__ rp_ptr) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row == 0) rp_ptr[0] = 0;
  if (row >= nrows) return;
  int64_t c = 0;
  for (int p = 0; p < P; ++p) c += seg[row * (P + 1) + p + 1] > seg[row * (P + 1) + p];
  rp_ptr[row + 1] = c;
}

// rp_ids: the row's pair ids in panel order
__global__ void row_pair_fill_kernel(const uint32_t* __restrict__ seg, const int64_t* __restrict__ S,
                                     int64_t nrows, int P, const int64_t* __restrict__ rp_ptr,
                                     int32_t* __restrict__ rp_ids) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  int64_t o = rp_ptr[row];
  for (int p = 0; p < P; ++p)
    if (seg[row * (P + 1) + p + 1] > seg[row * (P + 1) + p]) rp_ids[o++] = (int32_t)S[(int64_t)p * nrows + row];
}

// pstart[pair] = first entry of the pair in the panel-major order; pstart[K] = nnz
__global__ void pair_start_kernel(const int64_t* __restrict__ S, const int64_t* __restrict__ E, int64_t Q,
                                  int64_t* __restrict__ pstart) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > Q) return;
  if (q == Q || S[q + 1] > S[q]) pstart[S[q]] = E[q];
}

// One wave per row: every entry moves to its panel-major position. A count
// that does not fit the packed encoding sets *bad.
template <int MODE>
__global__ void __launch_bounds__(256) scatter_kernel(const int64_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ val, int64_t nrows, int P,
                                                      const uint32_t* __restrict__ seg,
                                                      const int64_t* __restrict__ E, uint32_t* __restrict__ ent,
                                                      uint32_t* __restrict__ val_out, unsigned int* __restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < nrows; row += nwaves) {
    const int64_t e0 = ptr[row], e1 = ptr[row + 1];
    for (int64_t e = e0 + lane; e < e1; e += 64) {
      const int64_t col = idx[e];
      const int64_t p = col / kPanel;
      const int64_t i = (e - e0) - (int64_t)seg[row * (P + 1) + p];
      const int64_t dst = E[p * nrows + row] + i;
      const uint32_t local = (uint32_t)(col - p * kPanel);
      if (MODE == 2) {
        const uint32_t v = val[e];
        if (v >= 128u) atomicOr(bad, 1u);
        ent[dst] = (local << 6) | ((v & 0x7Fu) << 24) | (i == 0 ? 0x80000000u : 0u);
      } else if (MODE == 1) {
        const uint32_t v = val[e];
        if (v >= 65536u) atomicOr(bad, 1u);
        ent[dst] = local | (i == 0 ? 0x8000u : 0u) | (v << 16);
      } else {
        ent[dst] = local | (i == 0 ? 0x80000000u : 0u);
        val_out[dst] = val[e];
      }
    }
  }
}

// Group-chunk c of panel p starts at the first pair whose first entry is at or
// after E_p + j * kChunk. cinfo[p] = {chunk base, S_p, S_p+1, E_p, E_p+1}.
__global__ void chunk_kernel(const int64_t* __restrict__ pstart, const int64_t* __restrict__ cinfo, int P,
                             int64_t nchunks, int64_t K, int64_t nnz, int64_t* __restrict__ cstart,
                             int32_t* __restrict__ cpair, int32_t* __restrict__ cpanel) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nchunks) return;
  if (c == nchunks) {
    cstart[c] = nnz;
    cpair[c] = (int32_t)K;
    return;
  }
  int p = 0;
  while (p + 1 < P && cinfo[5 * (p + 1)] <= c) ++p;
  const int64_t j = c - cinfo[5 * p];
  const int64_t s0 = cinfo[5 * p + 1], s1 = cinfo[5 * p + 2];
  const int64_t t = cinfo[5 * p + 3] + j * kChunk;
  int64_t lo = s0, hi = s1;  // first pair in [s0, s1) with pstart >= t
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (pstart[mid] < t) lo = mid + 1; else hi = mid;
  }
  cpair[c] = (int32_t)lo;
  cstart[c] = lo < s1 ? pstart[lo] : cinfo[5 * p + 4];
  cpanel[c] = p;
}

// ---------------------------------------------------------------------------
// The product. Group g of the wave (b lanes, lane c = dense column c) streams
// group-chunk wave * G + g: 4 * BW entries per batch, one coalesced dword per
// lane per quarter; entry j of the batch is broadcast to the group from lane
// j % BW, register j / BW. Gathers are issued kInflight at a time. A pair's first entry
// carries a flag: the running sum of the previous pair is stored and reset.
// ---------------------------------------------------------------------------
template <int MODE, bool FV>
__device__ __forceinline__ float entry_value(uint32_t e, uint32_t v) {
  if (MODE == 2) return (float)((e >> 24) & 0x7Fu);
  if (MODE == 1) return (float)(e >> 16);
  return FV ? __uint_as_float(v) : (float)v;
}

template <int MODE>
__device__ __forceinline__ uint32_t entry_column(uint32_t e) {
  if (MODE == 2) return (e >> 6) & 0x7FFFu;
  if (MODE == 1) return e & 0x7FFFu;
  return e & 0x7FFFFFFFu;
}

template <int BW, int MODE, bool FV>
__global__ void __launch_bounds__(256) product_kernel(const uint32_t* __restrict__ ent,
                                                      const uint32_t* __restrict__ val,
                                                      const int64_t* __restrict__ cstart,
                                                      const int32_t* __restrict__ cpair,
                                                      const int32_t* __restrict__ cpanel, int64_t nchunks,
                                                      const float* __restrict__ dense,
                                                      float* __restrict__ partial) {
  constexpr int G = 64 / BW;
  constexpr int NB = 4 * BW;  // entries per group batch
  constexpr bool PACK = MODE != 0;
  constexpr uint32_t kFirst = MODE == 1 ? 0x8000u : 0x80000000u;
  const int lane = threadIdx.x & 63;
  const int g = lane / BW;
  const int c = lane % BW;
  // Blocks go to the 8 XCDs round-robin; give each XCD a contiguous run of
  // group-chunks, so the chunks resident on one XCD share a panel and its L2.
  const int64_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const int64_t xcd = blockIdx.x % 8;
  const int64_t blk = xcd * q8 + min(xcd, r8) + blockIdx.x / 8;
  const int64_t chunk = ((blk * blockDim.x + threadIdx.x) >> 6) * G + g;
  int64_t pos = 0, end = 0;
  int64_t k = 0;
  const float* db = dense;
  if (chunk < nchunks) {
    pos = cstart[chunk];
    end = cstart[chunk + 1];
    k = cpair[chunk];
    db = dense + (int64_t)cpanel[chunk] * kPanel * BW;
  }
  bool have = false;  // a pair is open (its first entry was seen)
  float acc = 0.f;
  while (__ballot(pos < end) != 0) {
    uint32_t e[4], v[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t q = pos + s * BW + c;
      e[s] = q < end ? __builtin_nontemporal_load(ent + q) : 0u;
      v[s] = (!PACK && q < end) ? __builtin_nontemporal_load(val + q) : 0u;
    }
    const int nvalid = (int)min((int64_t)NB, max((int64_t)0, end - pos));
#pragma unroll
    for (int j0 = 0; j0 < NB; j0 += kInflight) {
      uint32_t ej[kInflight], vj[kInflight];
      float z[kInflight];
#pragma unroll
      for (int u = 0; u < kInflight; ++u) {
        const int j = j0 + u;
        ej[u] = (uint32_t)__shfl((int)e[j / BW], g * BW + j % BW, 64);
        vj[u] = PACK ? 0u : (uint32_t)__shfl((int)v[j / BW], g * BW + j % BW, 64);
        z[u] = j < nvalid ? db[(int64_t)entry_column<MODE>(ej[u]) * BW + c] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kInflight; ++u) {
        if (ej[u] & kFirst) {
          if (have) partial[k++ * BW + c] = acc;
          have = true;
          acc = 0.f;
        }
        acc = fmaf(entry_value<MODE, FV>(ej[u], vj[u]), z[u], acc);
      }
    }
    pos += NB;
  }
  if (have) partial[k * BW + c] = acc;
}

// ---------------------------------------------------------------------------
// The 16-column product over packed-8 entries, with the same work split and
// summation order as product_kernel<16, 2, false> (so the same bits), and
// fewer instructions per gather:
//   - entry j of a group's 64-entry batch reaches the group's 16 lanes by a DPP
//     row broadcast (the group is one 16-lane DPP row), not a cross-lane
//     permute through LDS;
//   - its gather offset is (entry & column field) | (panel base + 4 * lane
//     column), a 32-bit offset from the dense operand's base (which therefore
//     must be below 4 GiB: at most 2^26 rows of 16 floats);
//   - its count is byte 3 of the entry (one conversion), the first-of-pair flag
//     the sign bit (one compare);
//   - batches that lie entirely inside the group-chunk skip the per-entry
//     bounds test; the next batch's entries load while this one is gathered.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v) {
  // lane N of each 16-lane row, to every lane of that row (DPP row_newbcast);
  // every source lane is active, so no old value is needed
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x150 + N, 0xF, 0xF, true);
}

template <int J0, int... U>
__device__ __forceinline__ void bcast_entries(uint32_t (&v)[sizeof...(U)], const uint32_t (&e)[4],
                                              std::integer_sequence<int, U...>) {
  ((v[U] = row_bcast<(J0 + U) % 16>(e[(J0 + U) / 16])), ...);
}

// INF entries of the batch, starting at J0. A first-of-pair entry carries
// count + 128 in byte 3: its pair's sum starts at -128 * z, so that
// fma(count + 128, z, -128 * z) rounds (count + 128) * z - 128 * z = count * z
// exactly as the generic kernel's fma(count, z, 0) does (-128 * z is exact).
// The per-entry flag tests run only when some lane of the wave meets a pair
// boundary within these INF entries.
template <int J0, int INF, bool TAIL>
__device__ __forceinline__ void gather16_step(const uint32_t (&e)[4], const char* __restrict__ dense, uint32_t pbase,
                                              int nvalid, float* __restrict__ partial, int c, int32_t& k,
                                              bool& have, float& acc) {
  uint32_t v[INF];
  float z[INF];
  bcast_entries<J0>(v, e, std::make_integer_sequence<int, INF>{});
#pragma unroll
  for (int u = 0; u < INF; ++u) {
    const uint32_t off = (v[u] & 0x1FFFC0u) | pbase;
    if (TAIL)
      z[u] = J0 + u < nvalid ? *reinterpret_cast<const float*>(dense + off) : 0.f;
    else
      z[u] = *reinterpret_cast<const float*>(dense + off);
  }
  uint32_t any = 0;
#pragma unroll
  for (int u = 0; u < INF; ++u) any |= v[u];
  if (__builtin_amdgcn_ballot_w64((int32_t)any < 0) == 0) {
#pragma unroll
    for (int u = 0; u < INF; ++u) acc = fmaf((float)(v[u] >> 24), z[u], acc);
  } else {
#pragma unroll
    for (int u = 0; u < INF; ++u) {
      if ((int32_t)v[u] < 0) {
        if (have) partial[(int64_t)k++ * 16 + c] = acc;
        have = true;
        acc = -128.f * z[u];
      }
      acc = fmaf((float)(v[u] >> 24), z[u], acc);
    }
  }
}

template <int INF, bool TAIL, int... S>
__device__ __forceinline__ void gather16_batch(const uint32_t (&e)[4], const char* __restrict__ dense,
                                               uint32_t pbase, int nvalid, float* __restrict__ partial, int c,
                                               int32_t& k, bool& have, float& acc, std::integer_sequence<int, S...>) {
  (gather16_step<S * INF, INF, TAIL>(e, dense, pbase, nvalid, partial, c, k, have, acc), ...);
}

template <int INF>
__global__ void __launch_bounds__(256) product16_kernel(const uint32_t* __restrict__ ent,
                                                        const int64_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ cpair,
                                                        const int32_t* __restrict__ cpanel, int64_t nchunks,
                                                        const float* __restrict__ dense,
                                                        float* __restrict__ partial) {
  constexpr int NB = 64;  // entries per group batch
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int c = lane & 15;
  const int64_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const int64_t xcd = blockIdx.x % 8;
  const int64_t blk = xcd * q8 + min(xcd, r8) + blockIdx.x / 8;
  const int64_t chunk = ((blk * blockDim.x + threadIdx.x) >> 6) * 4 + g;
  int64_t pos = 0, end = 0;
  int32_t k = 0;  // pair index (pairs < 2^31)
  uint32_t pbase = (uint32_t)c << 2;
  if (chunk < nchunks) {
    pos = cstart[chunk];
    end = cstart[chunk + 1];
    k = cpair[chunk];
    pbase |= (uint32_t)cpanel[chunk] << 21;
  }
  const char* db = reinterpret_cast<const char*>(dense);
  bool have = false;
  float acc = 0.f;
  uint32_t e[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int64_t q = pos + s * 16 + c;
    e[s] = q < end ? __builtin_nontemporal_load(ent + q) : 0u;
  }
  while (__ballot(pos < end) != 0) {
    if (pos < end) {
      uint32_t en[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int64_t q = pos + NB + s * 16 + c;
        en[s] = q < end ? __builtin_nontemporal_load(ent + q) : 0u;
      }
      if (pos + NB <= end) {
        gather16_batch<INF, false>(e, db, pbase, NB, partial, c, k, have, acc,
                                   std::make_integer_sequence<int, NB / INF>{});
      } else {
        gather16_batch<INF, true>(e, db, pbase, (int)(end - pos), partial, c, k, have, acc,
                                  std::make_integer_sequence<int, NB / INF>{});
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) e[s] = en[s];
      pos += NB;
    }
  }
  if (have) partial[(int64_t)k * 16 + c] = acc;
}

// out row = sum of the row's pairs in panel order, then the epilogue:
//   EPI 0: out32[row][c] = scale[row] * sum (sum when scale is null) (FP32, stride BW)
//   EPI 1: Y[row][c] = scale[row] * sum - dinv[row] * V[row][c]  (FP64, c < b)
template <int BW, int EPI>
__global__ void __launch_bounds__(256) reduce_kernel(const int64_t* __restrict__ rp_ptr,
                                                     const int32_t* __restrict__ rp_ids, int64_t nrows,
                                                     const float* __restrict__ partial,
                                                     const float* __restrict__ scale, float* __restrict__ out32,
                                                     const double* __restrict__ V, int64_t ldv,
                                                     const double* __restrict__ dinv, double* __restrict__ Y,
                                                     int64_t ldy, int b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = t / BW;
  const int c = (int)(t % BW);
  if (row >= nrows) return;
  // pairs in panel order; four loads in flight, added in the same order
  const int64_t i0 = rp_ptr[row], i1 = rp_ptr[row + 1];
  float acc = 0.f;
  int64_t i = i0;
  for (; i + 4 <= i1; i += 4) {
    const int32_t p0 = rp_ids[i], p1 = rp_ids[i + 1], p2 = rp_ids[i + 2], p3 = rp_ids[i + 3];
    const float a0 = partial[(int64_t)p0 * BW + c], a1 = partial[(int64_t)p1 * BW + c];
    const float a2 = partial[(int64_t)p2 * BW + c], a3 = partial[(int64_t)p3 * BW + c];
    acc += a0;
    acc += a1;
    acc += a2;
    acc += a3;
  }
  for (; i < i1; ++i) acc += partial[(int64_t)rp_ids[i] * BW + c];
  if (EPI == 0) {
    out32[row * BW + c] = scale ? acc * scale[row] : acc;
  } else if (c < b) {
    Y[row * ldy + c] = (double)(acc * scale[row]) - dinv[row] * V[row * ldv + c];
  }
}

template <int BW, int EPI>
static void launch_product(const snapgpu_panel_layout& L, const float* dense, float* partial,
                           const float* scale, float* out32, const double* V, int64_t ldv, const double* dinv,
                           double* Y, int64_t ldy, int b, hipStream_t st) {
  constexpr int G = 64 / BW;
  const int64_t waves = ceil_div(L.n_chunks, G);
  if (waves > 0) {
    const unsigned grid = (unsigned)ceil_div(waves, 4);
    const bool fv = L.value_type == 1;
    // 32-bit gather offsets: the dense operand (BW floats per row) below 4 GiB
    const bool off32 = L.dense_rows <= ((int64_t)1 << 26);
    if (BW == 16 && L.packed == 2 && off32)
      product16_kernel<kInflight16><<<grid, 256, 0, st>>>(L.ent, L.chunk_start, L.chunk_pair, L.chunk_panel,
                                                          L.n_chunks, dense, partial);
    else if (L.packed == 2)
      product_kernel<BW, 2, false><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                         L.chunk_panel, L.n_chunks, dense, partial);
    else if (L.packed == 1)
      product_kernel<BW, 1, false><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                         L.chunk_panel, L.n_chunks, dense, partial);
    else if (fv)
      product_kernel<BW, 0, true><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                        L.chunk_panel, L.n_chunks, dense, partial);
    else
      product_kernel<BW, 0, false><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                         L.chunk_panel, L.n_chunks, dense, partial);
  }
  reduce_kernel<BW, EPI><<<(unsigned)ceil_div(L.n_rows * BW, 256), 256, 0, st>>>(
      L.row_pair_ptr, L.row_pair_ids, L.n_rows, partial, scale, out32, V, ldv, dinv, Y, ldy, b);
}

void product16_scaled(const snapgpu_panel_layout& L, const float* dense, float* partial, const float* scale,
                      float* out, hipStream_t st) {
  launch_product<16, 0>(L, dense, partial, scale, out, nullptr, 0, nullptr, nullptr, 0, 16, st);
}

// ---------------------------------------------------------------------------
// FP64 products over the same layout: a dense operand of two FP64 columns (a
// 16-byte row per stored entry, so one L2 request like the FP32 products') and
// FP64 pair sums. Groups of two lanes, one column each, stream the group-chunks
// exactly as product_kernel does (the same pairs, each summed in entry order);
// SQ multiplies by the squared value. reduce_f64_kernel adds a row's pairs in
// panel order and applies an optional row scale.
// ---------------------------------------------------------------------------
template <int MODE, bool FV, bool SQ>
__global__ void __launch_bounds__(256) product_f64_kernel(const uint32_t* __restrict__ ent,
                                                          const uint32_t* __restrict__ val,
                                                          const int64_t* __restrict__ cstart,
                                                          const int32_t* __restrict__ cpair,
                                                          const int32_t* __restrict__ cpanel, int64_t nchunks,
                                                          const double* __restrict__ dense,
                                                          double* __restrict__ partial) {
  constexpr int BW = 2;
  constexpr int G = 64 / BW;
  constexpr int S = 8;         // entry loads per lane per batch
  constexpr int NB = S * BW;   // entries per group batch (gathers in flight per lane)
  constexpr bool PACK = MODE != 0;
  constexpr uint32_t kFirst = MODE == 1 ? 0x8000u : 0x80000000u;
  const int lane = threadIdx.x & 63;
  const int g = lane / BW;
  const int c = lane % BW;
  const int64_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const int64_t xcd = blockIdx.x % 8;
  const int64_t blk = xcd * q8 + min(xcd, r8) + blockIdx.x / 8;
  const int64_t chunk = ((blk * blockDim.x + threadIdx.x) >> 6) * G + g;
  int64_t pos = 0, end = 0;
  int64_t k = 0;
  const double* db = dense;
  if (chunk < nchunks) {
    pos = cstart[chunk];
    end = cstart[chunk + 1];
    k = cpair[chunk];
    db = dense + (int64_t)cpanel[chunk] * kPanel * BW;
  }
  bool have = false;
  double acc = 0.0;
  while (__ballot(pos < end) != 0) {
    uint32_t e[S], v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int64_t q = pos + s * BW + c;
      e[s] = q < end ? __builtin_nontemporal_load(ent + q) : 0u;
      v[s] = (!PACK && q < end) ? __builtin_nontemporal_load(val + q) : 0u;
    }
    const int nvalid = (int)min((int64_t)NB, max((int64_t)0, end - pos));
    uint32_t ej[NB], vj[NB];
    double z[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      ej[j] = (uint32_t)__shfl((int)e[j / BW], g * BW + j % BW, 64);
      vj[j] = PACK ? 0u : (uint32_t)__shfl((int)v[j / BW], g * BW + j % BW, 64);
      z[j] = j < nvalid ? db[(int64_t)entry_column<MODE>(ej[j]) * BW + c] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (ej[j] & kFirst) {
        if (have) partial[k++ * BW + c] = acc;
        have = true;
        acc = 0.0;
      }
      // exact for every stored value (a count of 2^24 or more is not a float)
      double x = MODE == 2   ? (double)((ej[j] >> 24) & 0x7Fu)
                 : MODE == 1 ? (double)(ej[j] >> 16)
                 : FV        ? (double)__uint_as_float(vj[j])
                             : (double)vj[j];
      if (SQ) x *= x;
      acc = fma(x, z[j], acc);
    }
    pos += NB;
  }
  if (have) partial[k * BW + c] = acc;
}

__global__ void __launch_bounds__(256) reduce_f64_kernel(const int64_t* __restrict__ rp_ptr,
                                                         const int32_t* __restrict__ rp_ids, int64_t nrows,
                                                         const double* __restrict__ partial,
                                                         const double* __restrict__ scale,
                                                         double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = t >> 1;
  const int c = (int)(t & 1);
  if (row >= nrows) return;
  const int64_t i0 = rp_ptr[row], i1 = rp_ptr[row + 1];
  double acc = 0.0;
  int64_t i = i0;
  for (; i + 4 <= i1; i += 4) {
    const int32_t p0 = rp_ids[i], p1 = rp_ids[i + 1], p2 = rp_ids[i + 2], p3 = rp_ids[i + 3];
    const double a0 = partial[(int64_t)p0 * 2 + c], a1 = partial[(int64_t)p1 * 2 + c];
    const double a2 = partial[(int64_t)p2 * 2 + c], a3 = partial[(int64_t)p3 * 2 + c];
    acc += a0;
    acc += a1;
    acc += a2;
    acc += a3;
  }
  for (; i < i1; ++i) acc += partial[(int64_t)rp_ids[i] * 2 + c];
  out[row * 2 + c] = scale ? scale[row] * acc : acc;
}

template <bool SQ>
static void launch_product_f64(const snapgpu_panel_layout& L, const double* dense, double* partial,
                               const double* scale, double* out, hipStream_t st) {
  const int64_t waves = ceil_div(L.n_chunks, 32);
  if (waves > 0) {
    const unsigned grid = (unsigned)ceil_div(waves, 4);
    if (L.packed == 2)
      product_f64_kernel<2, false, SQ><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                             L.chunk_panel, L.n_chunks, dense, partial);
    else if (L.packed == 1)
      product_f64_kernel<1, false, SQ><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                             L.chunk_panel, L.n_chunks, dense, partial);
    else if (L.value_type == 1)
      product_f64_kernel<0, true, SQ><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                            L.chunk_panel, L.n_chunks, dense, partial);
    else
      product_f64_kernel<0, false, SQ><<<grid, 256, 0, st>>>(L.ent, L.val, L.chunk_start, L.chunk_pair,
                                                             L.chunk_panel, L.n_chunks, dense, partial);
  }
  reduce_f64_kernel<<<(unsigned)ceil_div(L.n_rows * 2, 256), 256, 0, st>>>(L.row_pair_ptr, L.row_pair_ids,
                                                                          L.n_rows, partial, scale, out);
}

// Build workspace: seg, F/S, L/E, scan workspace, pstart, small tables.
struct BuildWs {
  uint32_t* seg;
  int64_t* S;
  int64_t* E;
  int64_t* pstart;
  int64_t* info;
  unsigned int* maxc;
  void* scan_ws;
  int64_t scan_ws_bytes;
};

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

static int64_t build_ws_bytes(int64_t nrows, int P, int64_t nnz) {
  const int64_t Q = nrows * P;
  return align256(nrows * (P + 1) * 4) + 2 * align256((Q + 1) * 8) + align256((std::min(Q, nnz) + 1) * 8) +
         align256((int64_t)(P + 1) * 5 * 8) + 256 + align256(snapgpu_scan_workspace_bytes(std::max(Q + 1, nrows + 1))) +
         256;
}

static BuildWs carve(void* ws, int64_t nrows, int P, int64_t nnz) {
  const int64_t Q = nrows * P;
  char* q = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  BuildWs w;
  w.seg = (uint32_t*)q;
  q += align256(nrows * (P + 1) * 4);
  w.S = (int64_t*)q;
  q += align256((Q + 1) * 8);
  w.E = (int64_t*)q;
  q += align256((Q + 1) * 8);
  w.pstart = (int64_t*)q;
  q += align256((std::min(Q, nnz) + 1) * 8);
  w.info = (int64_t*)q;
  q += align256((int64_t)(P + 1) * 5 * 8);
  w.maxc = (unsigned int*)q;
  q += 256;
  w.scan_ws = q;
  w.scan_ws_bytes = snapgpu_scan_workspace_bytes(std::max(Q + 1, nrows + 1));
  return w;
}

}  // namespace panel
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::panel;

extern "C" {

int32_t snapgpu_panel_count(int64_t dense_rows) {
  return (int32_t)std::max<int64_t>(1, ceil_div(dense_rows, kPanel));
}

int64_t snapgpu_panel_build_workspace_bytes(int64_t nrows, int64_t dense_rows, int64_t nnz) {
  return build_ws_bytes(nrows, snapgpu_panel_count(dense_rows), nnz);
}

int snapgpu_panel_sizes(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                        int64_t dense_rows, int64_t nnz, int32_t value_type, void* d_ws, int64_t ws_bytes,
                        int64_t* n_pairs, int64_t* n_chunks, int32_t* packed, void* stream) {
  clear_error();
  SG_ARG(nrows >= 1 && dense_rows >= 1 && nnz >= 0, "bad shape");
  SG_ARG(value_type == 0 || value_type == 1, "value_type must be 0 (uint32) or 1 (float32)");
  const int P = snapgpu_panel_count(dense_rows);
  SG_ARG(ws_bytes >= build_ws_bytes(nrows, P, nnz), "workspace too small");
  SG_ARG((int64_t)nrows * P < ((int64_t)1 << 31) && nnz < ((int64_t)1 << 40), "matrix too large for the panel layout");
  hipStream_t st = as_stream(stream);
  BuildWs w = carve(d_ws, nrows, P, nnz);
  const int64_t Q = nrows * P;
  const unsigned gw = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nrows, 4), 16384));
  segment_kernel<<<gw, 256, 0, st>>>(d_ptr, d_idx, nrows, P, w.seg);
  pair_count_kernel<<<(unsigned)ceil_div(Q, 256), 256, 0, st>>>(w.seg, nrows, P, w.S, w.E);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(w.S, Q + 1, w.scan_ws, w.scan_ws_bytes, stream);
  if (rc) return rc;
  rc = snapgpu_scan_inclusive_i64(w.E, Q + 1, w.scan_ws, w.scan_ws_bytes, stream);
  if (rc) return rc;
  SG_HIP(hipMemsetAsync(w.maxc, 0, 4, st));
  if (value_type == 0 && nnz > 0)
    max_u32_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nnz, 256), 4096), 256, 0, st>>>(d_val, nnz, w.maxc);
  panel_info_kernel<<<(unsigned)ceil_div(P + 1, 64), 64, 0, st>>>(w.S, w.E, nrows, P, w.info);
  SG_LAUNCH_CHECK();
  std::vector<int64_t> info(2 * (P + 1));
  unsigned int maxc = 0;
  SG_HIP(hipMemcpyAsync(info.data(), w.info, info.size() * 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&maxc, w.maxc, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (info[2 * P + 1] != nnz) {
    set_error("panel layout: entry count mismatch (indices not sorted within rows or out of range)");
    return SNAPGPU_ERR_DATA;
  }
  int64_t nc = 0;
  for (int p = 0; p < P; ++p) nc += ceil_div(info[2 * (p + 1) + 1] - info[2 * p + 1], kChunk);
  *n_pairs = info[2 * P];
  *n_chunks = nc;
  *packed = value_type != 0 ? 0 : maxc < 128u ? 2 : maxc < 65536u ? 1 : 0;
  return SNAPGPU_OK;
}

int snapgpu_panel_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                        int64_t dense_rows, int64_t nnz, int32_t packed, void* d_ws, int64_t ws_bytes,
                        uint32_t* d_ent, uint32_t* d_val_out, int64_t* d_chunk_start, int32_t* d_chunk_pair,
                        int32_t* d_chunk_panel, int64_t n_chunks, int64_t* d_row_pair_ptr,
                        int32_t* d_row_pair_ids, void* stream) {
  clear_error();
  const int P = snapgpu_panel_count(dense_rows);
  SG_ARG(ws_bytes >= build_ws_bytes(nrows, P, nnz), "workspace too small");
  SG_ARG(packed >= 0 && packed <= 2, "packed must be 0, 1 or 2");
  SG_ARG(packed || d_val_out != nullptr, "unpacked layout needs a value array");
  hipStream_t st = as_stream(stream);
  BuildWs w = carve(d_ws, nrows, P, nnz);
  const int64_t Q = nrows * P;
  // panel table (S, E at row 0 of each panel), chunk bases
  std::vector<int64_t> info(2 * (P + 1));
  SG_HIP(hipMemcpyAsync(info.data(), w.info, info.size() * 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  const int64_t K = info[2 * P];
  std::vector<int64_t> cinfo(5 * (P + 1), 0);
  int64_t cb = 0;
  for (int p = 0; p < P; ++p) {
    cinfo[5 * p] = cb;
    cinfo[5 * p + 1] = info[2 * p];
    cinfo[5 * p + 2] = info[2 * (p + 1)];
    cinfo[5 * p + 3] = info[2 * p + 1];
    cinfo[5 * p + 4] = info[2 * (p + 1) + 1];
    cb += ceil_div(info[2 * (p + 1) + 1] - info[2 * p + 1], kChunk);
  }
  cinfo[5 * P] = cb;
  SG_ARG(cb == n_chunks, "n_chunks does not match snapgpu_panel_sizes");
  SG_HIP(hipMemcpyAsync(w.info, cinfo.data(), cinfo.size() * 8, hipMemcpyHostToDevice, st));
  // row -> pairs (panel order)
  row_pair_count_kernel<<<(unsigned)ceil_div(nrows, 256), 256, 0, st>>>(w.seg, nrows, P, d_row_pair_ptr);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(d_row_pair_ptr + 1, nrows, w.scan_ws, w.scan_ws_bytes, stream);
  if (rc) return rc;
  row_pair_fill_kernel<<<(unsigned)ceil_div(nrows, 256), 256, 0, st>>>(w.seg, w.S, nrows, P, d_row_pair_ptr,
                                                                      d_row_pair_ids);
  pair_start_kernel<<<(unsigned)ceil_div(Q + 1, 256), 256, 0, st>>>(w.S, w.E, Q, w.pstart);
  const unsigned gw = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nrows, 4), 16384));
  unsigned int* bad = w.maxc + 1;
  SG_HIP(hipMemsetAsync(bad, 0, 4, st));
  if (packed == 2)
    scatter_kernel<2><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, P, w.seg, w.E, d_ent, nullptr, bad);
  else if (packed == 1)
    scatter_kernel<1><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, P, w.seg, w.E, d_ent, nullptr, bad);
  else
    scatter_kernel<0><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, P, w.seg, w.E, d_ent, d_val_out, bad);
  chunk_kernel<<<(unsigned)ceil_div(n_chunks + 1, 256), 256, 0, st>>>(w.pstart, w.info, P, n_chunks, K, nnz,
                                                                     d_chunk_start, d_chunk_pair, d_chunk_panel);
  SG_LAUNCH_CHECK();
  unsigned int hbad = 0;
  SG_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));  // cinfo lives on this stack frame until the copy has run
  if (hbad) {
    set_error(packed == 2 ? "packed = 2 needs counts below 128" : "packed = 1 needs counts below 65536");
    return SNAPGPU_ERR_ARG;
  }
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_apply_panel_workspace_bytes(int64_t n, int64_t m, int64_t pairs_csr,
                                                     int64_t pairs_csc) {
  return align256(n * kB * 4) + align256(m * kB * 4) + align256(std::max(pairs_csr, pairs_csc) * kB * 4) + 256;
}

int snapgpu_spectral_apply_panel(const snapgpu_panel_layout* csr, const snapgpu_panel_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv, const double* d_V,
                                 int64_t ldv, int32_t b, double* d_Y, int64_t ldy, void* d_ws, int64_t ws_bytes,
                                 void* stream) {
  clear_error();
  SG_ARG(csr != nullptr && csc != nullptr, "null layout");
  const int64_t n = csr->n_rows, m = csc->n_rows;
  SG_ARG(csr->dense_rows == m && csc->dense_rows == n, "layouts do not describe one matrix and its transpose");
  SG_ARG(b >= 1 && b <= kB, "block width must be in [1, 32]");
  SG_ARG(ldv >= b && ldy >= b, "leading dimension < b");
  SG_ARG(ws_bytes >= snapgpu_spectral_apply_panel_workspace_bytes(n, m, csr->n_pairs, csc->n_pairs),
         "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  float* Vs = (float*)q;
  q += align256(n * kB * 4);
  float* Z = (float*)q;
  q += align256(m * kB * 4);
  float* part = (float*)q;
#define SG_APPLY_PANEL(BW)                                                                                 \
  scale_in_kernel<BW><<<(unsigned)ceil_div(n * BW, 256), 256, 0, st>>>(d_V, ldv, n, b, d_r32, Vs);         \
  launch_product<BW, 0>(*csc, Vs, part, d_w2, Z, nullptr, 0, nullptr, nullptr, 0, b, st);                 \
  launch_product<BW, 1>(*csr, Z, part, d_r32, nullptr, d_V, ldv, d_dinv, d_Y, ldy, b, st);
  if (b <= 8) {
    SG_APPLY_PANEL(8)
  } else if (b <= 16) {
    SG_APPLY_PANEL(16)
  } else {
    SG_APPLY_PANEL(32)
  }
#undef SG_APPLY_PANEL
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_panel_spmm_f64_workspace_bytes(int64_t n_pairs) { return align256(std::max<int64_t>(n_pairs, 1) * 16) + 256; }

int snapgpu_panel_spmm_f64(const snapgpu_panel_layout* L, const double* d_dense, int32_t square_values,
                           const double* d_row_scale, double* d_out, void* d_ws, int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(L != nullptr && d_dense != nullptr && d_out != nullptr && d_ws != nullptr, "null pointer");
  SG_ARG(L->packed >= 0 && L->packed <= 2, "bad packing");
  SG_ARG(ws_bytes >= snapgpu_panel_spmm_f64_workspace_bytes(L->n_pairs), "workspace too small");
  if (L->n_rows == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  if (square_values)
    launch_product_f64<true>(*L, d_dense, part, d_row_scale, d_out, st);
  else
    launch_product_f64<false>(*L, d_dense, part, d_row_scale, d_out, st);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
