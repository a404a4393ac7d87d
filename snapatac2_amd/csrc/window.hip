// Window layout of a sparse count matrix for the operator's two sparse products.
//
// The panel layout gathers each stored entry's 64-byte dense row from L2, and
// those gathers bound it (about 160G rows/s at 1M cells). Here a workgroup owns
// the accumulators of up to 2048 output rows in LDS and walks the dense operand
// in windows of 1024 rows (64 KB at 16 FP32 columns): within a window each
// dense row is read by about (rows per workgroup) x (density) entries, so most
// gathers hit L1 instead of L2.
//
// Slots and pieces. Within a window an output row's entries (a "run") are split
// into pieces of at most 16 entries, each summed by one quad (4 lanes, a float4
// of the 16 columns each) and added to an accumulator slot of its own; a row has
// as many slots as its most split window needs. Per window, every slot receives
// at most one piece, so quads never update the same slot concurrently, and the
// row's result is the sum of its slots in slot order: the sums are bitwise
// reproducible whatever order the build places pieces in.
//
// Runs are cut where the row's 64-entry build chunks end (the build walks rows
// 64 entries at a time), so one (row, window) run may be two segments; pieces are
// numbered across them.
//
// Entry (32 bits): bits 0-9 column within the window, bits 10-20 slot, bit 21
// first entry of a piece, bits 24-31 the count (counts < 256; other matrices use
// the panel layout).
//
// Tables: bucket_start[b * n_windows + w] (int64, n_blocks * n_windows + 1) =
// first entry of (workgroup b, window w); quad_off[(b * n_windows + w) * 257 + k]
// (uint32) = start of quad k's piece-aligned chunk relative to the bucket (k =
// 256: the bucket length); block_row[b] (int64, n_blocks + 1) = first row of
// workgroup b; row_slot[r] = first slot | (number of slots) << 16.
#include <algorithm>
#include <vector>

#include "spectral_common.h"

namespace snapgpu {
namespace window {

using spectral::scale_in_kernel;

#ifndef SNAPGPU_WINDOW_BITS
#define SNAPGPU_WINDOW_BITS 10
#endif
constexpr int kWinBits = SNAPGPU_WINDOW_BITS;
constexpr int64_t kWin = int64_t(1) << kWinBits;  // dense rows per window
constexpr int kSlots = 2048;                       // accumulator slots per workgroup (128 KB)
constexpr int kPiece = 16;                         // entries per piece at most
constexpr int kChunk = 64;                         // build chunk of a row
constexpr int kThreads = 1024;
constexpr int kQuads = kThreads / 4;
constexpr int kMaxWindows = 16384;                 // LDS counters of the build (64 KB)
constexpr uint32_t kFlag = 1u << (kWinBits + 11);  // after the column and the 11-bit slot
static_assert(kWinBits + 12 <= 24, "column, slot and flag must stay below the count byte");

// ---------------------------------------------------------------------------
// Build
// ---------------------------------------------------------------------------

// Per row (one wave): the number of slots = the most pieces any window's run
// needs, with runs cut at the row's 64-entry chunk ends. status[0] |= 1 for a
// count >= 256, |= 2 for a column out of range or descending.
__global__ void __launch_bounds__(256) row_slots_kernel(const int64_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        const uint32_t* __restrict__ val, int64_t nrows,
                                                        int64_t dense_rows, int32_t* __restrict__ s,
                                                        unsigned int* __restrict__ status) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= nrows) return;
  const int64_t e0 = ptr[row], e1 = ptr[row + 1];
  int64_t prev_w = -1;   // window of the previous chunk's last entry
  int pieces = 0;        // pieces of the open (row, window) run so far
  int best = 0;
  int32_t prev_col = -1;
  unsigned int bad = 0;
  for (int64_t base = e0; base < e1; base += kChunk) {
    const int64_t e = base + lane;
    const bool valid = e < e1;
    const int32_t col = valid ? idx[e] : 0;
    if (valid && val[e] >= 256u) bad |= 1u;
    int32_t pc = __shfl_up(col, 1, 64);
    if (lane == 0) pc = prev_col;
    if (valid && (col < 0 || (int64_t)col >= dense_rows || col < pc)) bad |= 2u;
    const int64_t w = valid ? ((int64_t)col >> kWinBits) : -2;
    int64_t wp = __shfl_up(w, 1, 64);
    if (lane == 0) wp = -1;  // a chunk start always starts a segment
    const bool st = valid && w != wp;
    const uint64_t S = __ballot(st);
    const uint64_t V = __ballot(valid);
    // segment length: to the next segment start in the chunk, or to the chunk's end
    const uint64_t above = S & (lane == 63 ? 0ull : (~0ull << (lane + 1)));
    const int nvalid = __popcll(V);
    const int nxt = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
    const int seg = st ? nxt - lane : 0;
    const int segp = (seg + kPiece - 1) / kPiece;
    // columns ascend, so a window is one segment per chunk; only the chunk's first
    // segment (lane 0) can continue the previous chunk's run
    const int64_t w0 = __shfl(w, 0, 64);
    const int carry = (w0 == prev_w) ? pieces : 0;
    const int cnt = segp + ((lane == 0) ? carry : 0);  // pieces of the run up to this segment
    best = max(best, cnt);
    const int last = 63 - __clzll((long long)S);  // last segment start of the chunk
    const int last_cnt = __shfl(cnt, last, 64);
    prev_w = __shfl(w, nvalid - 1, 64);
    pieces = last_cnt;
    prev_col = __shfl(col, nvalid - 1, 64);
  }
  for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
  for (int o = 32; o > 0; o >>= 1) bad |= (unsigned int)__shfl_xor((int)bad, o, 64);
  if (lane == 0) {
    s[row] = best;
    if (bad) atomicOr(status, bad);
  }
}

// Per workgroup b (one block of rows): entries per window, counted in LDS.
__global__ void __launch_bounds__(kThreads) bucket_count_kernel(const int64_t* __restrict__ ptr,
                                                                const int32_t* __restrict__ idx,
                                                                const int64_t* __restrict__ block_row,
                                                                int64_t nw, int64_t* __restrict__ counts) {
  __shared__ unsigned int cnt[kMaxWindows];
  for (int64_t i = threadIdx.x; i < nw; i += kThreads) cnt[i] = 0;
  __syncthreads();
  const int64_t b = blockIdx.x;
  const int64_t e0 = ptr[block_row[b]], e1 = ptr[block_row[b + 1]];
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) {
    const int64_t w = (int64_t)idx[e] >> kWinBits;
    if (w >= 0 && w < nw) atomicAdd(&cnt[w], 1u);  // columns were checked by the plan
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < nw; i += kThreads) counts[1 + b * nw + i] = cnt[i];
  if (b == 0 && threadIdx.x == 0) counts[0] = 0;
}

// Per workgroup b: every row's segments (maximal same-window stretches of a
// 64-entry chunk) get contiguous space in their bucket (LDS cursors), split into
// pieces of <= 16 entries with slots numbered across the (row, window) run.
__global__ void __launch_bounds__(kThreads) scatter_kernel(const int64_t* __restrict__ ptr,
                                                           const int32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ val,
                                                           const int64_t* __restrict__ block_row,
                                                           const uint32_t* __restrict__ row_slot, int64_t nw,
                                                           const int64_t* __restrict__ bstart,
                                                           uint32_t* __restrict__ ent,
                                                           unsigned int* __restrict__ status) {
  __shared__ unsigned int cursor[kMaxWindows];
  for (int64_t i = threadIdx.x; i < nw; i += kThreads) cursor[i] = 0;
  __syncthreads();
  const int64_t b = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r0 = block_row[b], r1 = block_row[b + 1];
  for (int64_t row = r0 + wave; row < r1; row += kThreads / 64) {
    const int64_t e0 = ptr[row], e1 = ptr[row + 1];
    const uint32_t rs = row_slot[row];
    const uint32_t slot0 = rs & 0xFFFFu, nslot = rs >> 16;
    int64_t prev_w = -1;
    int pieces = 0;
    for (int64_t base = e0; base < e1; base += kChunk) {
      const int64_t e = base + lane;
      const bool valid = e < e1;
      const int32_t col = valid ? idx[e] : 0;
      const int64_t w = valid ? ((int64_t)col >> kWinBits) : -2;
      int64_t wp = __shfl_up(w, 1, 64);
      if (lane == 0) wp = -1;
      const bool st = valid && w != wp;
      const uint64_t S = __ballot(st);
      const int nvalid = __popcll(__ballot(valid));
      const uint64_t above = S & (lane == 63 ? 0ull : (~0ull << (lane + 1)));
      const int nxt = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
      const int seg = st ? nxt - lane : 0;
      const int64_t w0 = __shfl(w, 0, 64);
      const int carry = (w0 == prev_w) ? pieces : 0;
      // the segment's base in its bucket, and its first piece number
      unsigned int off = 0;
      const bool inw = w >= 0 && w < nw;  // columns were checked by the plan
      if (valid && !inw) atomicOr(status, 2u);
      if (st && inw) off = atomicAdd(&cursor[w], (unsigned int)seg);
      const int first_piece = (lane == 0) ? carry : 0;
      // every lane: its segment's start lane (the highest start at or below it)
      const uint64_t upto = S & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
      const int sl = valid ? 63 - __clzll((long long)upto) : 0;
      const unsigned int soff = (unsigned int)__shfl((int)off, sl, 64);
      const int sfp = __shfl(first_piece, sl, 64);
      if (valid && inw) {
        const int i = lane - sl;  // position within the segment
        const int piece = sfp + i / kPiece;
        if ((uint32_t)piece >= nslot) atomicOr(status, 4u);  // the plan does not match
        const uint32_t word = ((uint32_t)col & (uint32_t)(kWin - 1)) |
                              ((slot0 + (uint32_t)piece) << kWinBits) | ((i % kPiece) == 0 ? kFlag : 0u) |
                              (val[e] << 24);
        ent[bstart[b * nw + w] + soff + i] = word;
      }
      const int last = 63 - __clzll((long long)S);
      const int lseg = __shfl(seg, last, 64);
      const int lfp = __shfl(first_piece, last, 64);
      pieces = lfp + (lseg + kPiece - 1) / kPiece;
      prev_w = __shfl(w, nvalid - 1, 64);
    }
  }
}

// quad_off[q * 257 + k]: quad k's chunk of bucket q starts at the piece start
// nearest to k * len / 256 (the earlier one on a tie; the bucket's end counts as
// a piece start). Nearest points of a fixed set move monotonically with the
// target, so the chunks are in order.
__global__ void quad_split_kernel(const int64_t* __restrict__ bstart, const uint32_t* __restrict__ ent,
                                  int64_t nbuckets, uint32_t* __restrict__ qoff) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbuckets * (kQuads + 1)) return;
  const int64_t q = t / (kQuads + 1);
  const int k = (int)(t % (kQuads + 1));
  const int64_t bs = bstart[q], len = bstart[q + 1] - bs;
  const int64_t target = len * k / kQuads;  // floor(k * len / 256)
  if (target >= len) {  // k = 256, or an empty bucket
    qoff[t] = (uint32_t)len;
    return;
  }
  int64_t up = target, down = target;
  while (up < len && (ent[bs + up] & kFlag) == 0) ++up;
  while (down > 0 && (ent[bs + down] & kFlag) == 0) --down;  // ent[bs] is a piece start
  qoff[t] = (uint32_t)((target - down <= up - target) ? down : up);
}

// ---------------------------------------------------------------------------
// The product. Workgroup b walks the windows; quad k sums its chunk's pieces
// (4 lanes, lane j holding columns 4j..4j+3) and adds each piece to its slot;
// a barrier closes each window. The epilogue adds each row's slots in order:
//   EPI 0: out32[row][c] = scale[row] * sum (sum when scale is null) (FP32, stride 16)
//   EPI 1: Y[row][c] = scale[row] * sum - dinv[row] * V[row][c]  (FP64, c < b)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void add_piece(float4* acc, uint32_t slot, int q, const float4& a) {
  float4& d = acc[slot * 4 + q];
  float4 t = d;
  t.x += a.x;
  t.y += a.y;
  t.z += a.z;
  t.w += a.w;
  d = t;
}

// One entry of a quad's chunk: a piece start adds the previous piece to its slot.
__device__ __forceinline__ void gather_entry(float4* acc, int q, uint32_t v, const float4& z, bool& open,
                                             uint32_t& slot, float4& a) {
  if (v & kFlag) {
    if (open) add_piece(acc, slot, q, a);
    open = true;
    slot = (v >> kWinBits) & (uint32_t)(kSlots - 1);
    a = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const float x = (float)(v >> 24);
  a.x = fmaf(x, z.x, a.x);
  a.y = fmaf(x, z.y, a.y);
  a.z = fmaf(x, z.z, a.z);
  a.w = fmaf(x, z.w, a.w);
}

template <int EPI>
__global__ void __launch_bounds__(kThreads) product_kernel(
    const uint32_t* __restrict__ ent, const int64_t* __restrict__ bstart, const uint32_t* __restrict__ qoff,
    int64_t nw, const int64_t* __restrict__ block_row, const uint32_t* __restrict__ row_slot,
    const float* __restrict__ dense, const float* __restrict__ scale, float* __restrict__ out32,
    const double* __restrict__ V, int64_t ldv, const double* __restrict__ dinv, double* __restrict__ Y,
    int64_t ldy, int b) {
  __shared__ float4 acc[kSlots * 4];
  const int tid = threadIdx.x, quad = tid >> 2, q = tid & 3;
  for (int i = tid; i < kSlots * 4; i += kThreads) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  const int64_t blk = blockIdx.x;
  const char* db = reinterpret_cast<const char*>(dense);
  const uint32_t q16 = (uint32_t)q << 4;
  // Window w + 2's bounds and window w + 1's first 24 entries are loaded while
  // window w is summed, so neither dependent load (bounds, then entries) sits
  // on a window's critical path. Entries past a chunk's end read as 0 (no piece
  // start, count 0): they add 0 * z to the open piece, z being a row of the
  // same window.
  constexpr int kPre = 24;
  const int64_t bq0 = blk * nw;
  int64_t bs = bstart[bq0];
  uint32_t p0 = qoff[bq0 * (kQuads + 1) + quad], p1 = qoff[bq0 * (kQuads + 1) + quad + 1];
  int64_t bs1 = 0;
  uint32_t p01 = 0, p11 = 0;
  if (nw > 1) {
    bs1 = bstart[bq0 + 1];
    p01 = qoff[(bq0 + 1) * (kQuads + 1) + quad];
    p11 = qoff[(bq0 + 1) * (kQuads + 1) + quad + 1];
  }
  uint32_t vpre[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) vpre[u] = p0 + u < p1 ? ent[bs + p0 + u] : 0u;
  for (int64_t w = 0; w < nw; ++w) {
    const bool more1 = w + 1 < nw, more2 = w + 2 < nw;
    // window w + 1's first entries, window w + 2's bounds
    uint32_t vn[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) vn[u] = (more1 && p01 + u < p11) ? ent[bs1 + p01 + u] : 0u;
    int64_t bs2 = 0;
    uint32_t p02 = 0, p12 = 0;
    if (more2) {
      const int64_t bq2 = bq0 + w + 2;
      bs2 = bstart[bq2];
      p02 = qoff[bq2 * (kQuads + 1) + quad];
      p12 = qoff[bq2 * (kQuads + 1) + quad + 1];
    }
    const uint32_t* eb = ent + bs;
    const uint32_t wbase = ((uint32_t)w << (kWinBits + 6)) | q16;  // window w, this lane's 16 bytes
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t slot = 0;
    bool open = false;  // the chunk starts at a piece start, so its first entry opens one
#pragma unroll
    for (int h = 0; h < kPre; h += 8) {
      if (p0 + h < p1) {
        float4 z[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          z[u] = *reinterpret_cast<const float4*>(db + (((vpre[h + u] & (uint32_t)(kWin - 1)) << 6) | wbase));
#pragma unroll
        for (int u = 0; u < 8; ++u) gather_entry(acc, q, vpre[h + u], z[u], open, slot, a);
      }
    }
    for (uint32_t p = p0 + kPre; p < p1; p += 8) {
      uint32_t v[8];
      float4 z[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p + u < p1 ? eb[p + u] : 0u;
#pragma unroll
      for (int u = 0; u < 8; ++u) z[u] = *reinterpret_cast<const float4*>(db + (((v[u] & (uint32_t)(kWin - 1)) << 6) | wbase));
#pragma unroll
      for (int u = 0; u < 8; ++u) gather_entry(acc, q, v[u], z[u], open, slot, a);
    }
    if (open) add_piece(acc, slot, q, a);
    __syncthreads();
    bs = bs1;
    p0 = p01;
    p1 = p11;
    bs1 = bs2;
    p01 = p02;
    p11 = p12;
#pragma unroll
    for (int u = 0; u < kPre; ++u) vpre[u] = vn[u];
  }
  // epilogue: 4 lanes per row
  const int64_t r0 = block_row[blk], r1 = block_row[blk + 1];
  for (int64_t row = r0 + quad; row < r1; row += kQuads) {
    const uint32_t rs = row_slot[row];
    const uint32_t s0 = rs & 0xFFFFu, ns = rs >> 16;
    float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t j = 0; j < ns; ++j) {
      const float4 a = acc[(s0 + j) * 4 + q];
      t.x += a.x;
      t.y += a.y;
      t.z += a.z;
      t.w += a.w;
    }
    const float sc = scale ? scale[row] : 1.f;
    const float o[4] = {t.x * sc, t.y * sc, t.z * sc, t.w * sc};
    if (EPI == 0) {
      *reinterpret_cast<float4*>(out32 + row * 16 + q * 4) = make_float4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = q * 4 + j;
        if (c < b) Y[row * ldy + c] = (double)o[j] - dinv[row] * V[row * ldv + c];
      }
    }
  }
}

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

void product16_scaled(const snapgpu_window_layout& L, const float* dense, const float* scale, float* out,
                      hipStream_t st) {
  product_kernel<0><<<(unsigned)L.n_blocks, kThreads, 0, st>>>(L.ent, L.bucket_start, L.quad_off, L.n_windows,
                                                               L.block_row, L.row_slot, dense, scale, out, nullptr,
                                                               0, nullptr, nullptr, 0, 16);
}

struct PlanWs {
  int32_t* s;          // nrows
  unsigned int* status;
  int64_t* counts;     // n_blocks * nw + 1 (build)
  void* scan_ws;
  int64_t scan_ws_bytes;
};

static int64_t windows_of(int64_t dense_rows) { return std::max<int64_t>(1, ceil_div(dense_rows, kWin)); }

}  // namespace window
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::window;

extern "C" {

int64_t snapgpu_window_count(int64_t dense_rows) { return windows_of(dense_rows); }

int64_t snapgpu_window_plan_workspace_bytes(int64_t nrows) {
  return align256(std::max<int64_t>(nrows, 1) * 4) + 256 + 256;
}

int snapgpu_window_plan(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                        int64_t dense_rows, int64_t nnz, void* d_ws, int64_t ws_bytes, int64_t* h_block_row,
                        int64_t max_blocks, int64_t* n_blocks, uint32_t* h_row_slot, void* stream) {
  clear_error();
  SG_ARG(nrows >= 1 && dense_rows >= 1 && nnz >= 0, "bad shape");
  SG_ARG(ws_bytes >= snapgpu_window_plan_workspace_bytes(nrows), "workspace too small");
  SG_ARG(h_block_row != nullptr && n_blocks != nullptr && h_row_slot != nullptr, "null output");
  if (windows_of(dense_rows) > kMaxWindows || dense_rows > (int64_t(1) << 26)) {
    set_error("window layout: the dense operand has more rows than the layout's windows and 32-bit offsets hold");
    return SNAPGPU_ERR_CAPACITY;
  }
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  int32_t* s = (int32_t*)q;
  q += align256(nrows * 4);
  unsigned int* status = (unsigned int*)q;
  SG_HIP(hipMemsetAsync(status, 0, 4, st));
  const unsigned grid = (unsigned)ceil_div(nrows, 4);
  row_slots_kernel<<<grid, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, dense_rows, s, status);
  SG_LAUNCH_CHECK();
  std::vector<int32_t> hs(nrows);
  unsigned int hst = 0;
  SG_HIP(hipMemcpyAsync(hs.data(), s, nrows * 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&hst, status, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (hst & 2u) {
    set_error("window layout: column out of range or not sorted within a row");
    return SNAPGPU_ERR_DATA;
  }
  if (hst & 1u) {
    set_error("window layout: counts must be below 256 (use the panel layout)");
    return SNAPGPU_ERR_CAPACITY;
  }
  // rows packed greedily, in order, into workgroups of at most kSlots slots;
  // h_block_row holds max_blocks + 1 entries
  int64_t nb = 0, used = 0;
  h_block_row[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) {
    const int64_t sr = hs[r];
    if (sr > kSlots) {
      set_error("window layout: a row needs more than 2048 slots");
      return SNAPGPU_ERR_CAPACITY;
    }
    if (used + sr > kSlots) {
      if (nb + 1 >= max_blocks) {
        set_error("window layout: more workgroups than max_blocks");
        return SNAPGPU_ERR_CAPACITY;
      }
      h_block_row[++nb] = r;
      used = 0;
    }
    h_row_slot[r] = (uint32_t)used | ((uint32_t)sr << 16);
    used += sr;
  }
  h_block_row[++nb] = nrows;
  *n_blocks = nb;
  return SNAPGPU_OK;
}

int64_t snapgpu_window_build_workspace_bytes(int64_t n_blocks, int64_t dense_rows) {
  const int64_t nbk = n_blocks * windows_of(dense_rows);
  return align256((nbk + 1) * 8) + 256 + align256(snapgpu_scan_workspace_bytes(nbk + 1)) + 256;
}

int snapgpu_window_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                         int64_t dense_rows, int64_t nnz, const int64_t* d_block_row, int64_t n_blocks,
                         const uint32_t* d_row_slot, void* d_ws, int64_t ws_bytes, uint32_t* d_ent,
                         int64_t* d_bucket_start, uint32_t* d_quad_off, void* stream) {
  clear_error();
  SG_ARG(nrows >= 1 && n_blocks >= 1 && nnz >= 0, "bad shape");
  SG_ARG(ws_bytes >= snapgpu_window_build_workspace_bytes(n_blocks, dense_rows), "workspace too small");
  if (windows_of(dense_rows) > kMaxWindows) {
    set_error("window layout: the dense operand has more rows than the layout's windows hold");
    return SNAPGPU_ERR_CAPACITY;
  }
  hipStream_t st = as_stream(stream);
  const int64_t nw = windows_of(dense_rows);
  const int64_t nbk = n_blocks * nw;
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  unsigned int* status = (unsigned int*)q;
  q += 256;
  void* scan_ws = q;
  SG_HIP(hipMemsetAsync(status, 0, 4, st));
  bucket_count_kernel<<<(unsigned)n_blocks, kThreads, 0, st>>>(d_ptr, d_idx, d_block_row, nw, d_bucket_start);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(d_bucket_start, nbk + 1, scan_ws, snapgpu_scan_workspace_bytes(nbk + 1),
                                      stream);
  if (rc) return rc;
  scatter_kernel<<<(unsigned)n_blocks, kThreads, 0, st>>>(d_ptr, d_idx, d_val, d_block_row, d_row_slot, nw,
                                                           d_bucket_start, d_ent, status);
  SG_LAUNCH_CHECK();
  const int64_t nq = nbk * (kQuads + 1);
  quad_split_kernel<<<(unsigned)ceil_div(nq, 256), 256, 0, st>>>(d_bucket_start, d_ent, nbk, d_quad_off);
  SG_LAUNCH_CHECK();
  unsigned int hst = 0;
  int64_t total = 0;
  SG_HIP(hipMemcpyAsync(&hst, status, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&total, d_bucket_start + nbk, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (hst) {
    set_error((hst & 2u) ? "window layout: column out of range"
                         : "window layout: a piece went past its row's slots (the plan does not match the matrix)");
    return SNAPGPU_ERR_DATA;
  }
  if (total != nnz) {
    set_error("window layout: entry count mismatch");
    return SNAPGPU_ERR_DATA;
  }
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_apply_window_workspace_bytes(int64_t n, int64_t m) {
  return align256(n * 16 * 4) + align256(m * 16 * 4) + 256;
}

int snapgpu_spectral_apply_window(const snapgpu_window_layout* csr, const snapgpu_window_layout* csc,
                                  const float* d_r32, const float* d_w2, const double* d_dinv, const double* d_V,
                                  int64_t ldv, int32_t b, double* d_Y, int64_t ldy, void* d_ws, int64_t ws_bytes,
                                  void* stream) {
  clear_error();
  SG_ARG(csr != nullptr && csc != nullptr, "null layout");
  const int64_t n = csr->n_rows, m = csc->n_rows;
  SG_ARG(csr->dense_rows == m && csc->dense_rows == n, "layouts do not describe one matrix and its transpose");
  SG_ARG(b >= 1 && b <= 16, "block width must be in [1, 16]");
  SG_ARG(ldv >= b && ldy >= b, "leading dimension < b");
  SG_ARG(ws_bytes >= snapgpu_spectral_apply_window_workspace_bytes(n, m), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  float* Vs = (float*)q;
  q += align256(n * 16 * 4);
  float* Z = (float*)q;
  scale_in_kernel<16><<<(unsigned)ceil_div(n * 16, 256), 256, 0, st>>>(d_V, ldv, n, b, d_r32, Vs);
  product_kernel<0><<<(unsigned)csc->n_blocks, kThreads, 0, st>>>(
      csc->ent, csc->bucket_start, csc->quad_off, csc->n_windows, csc->block_row, csc->row_slot, Vs, d_w2, Z,
      nullptr, 0, nullptr, nullptr, 0, b);
  product_kernel<1><<<(unsigned)csr->n_blocks, kThreads, 0, st>>>(
      csr->ent, csr->bucket_start, csr->quad_off, csr->n_windows, csr->block_row, csr->row_slot, Z, d_r32,
      nullptr, d_V, ldv, d_dinv, d_Y, ldy, b);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_spectral_apply_mixed_workspace_bytes(int64_t n, int64_t m, int64_t pairs_csc) {
  return align256(n * 16 * 4) + align256(m * 16 * 4) + align256(std::max<int64_t>(pairs_csc, 1) * 16 * 4) + 256;
}

int snapgpu_spectral_apply_mixed(const snapgpu_window_layout* csr, const snapgpu_panel_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv, const double* d_V,
                                 int64_t ldv, int32_t b, double* d_Y, int64_t ldy, void* d_ws, int64_t ws_bytes,
                                 void* stream) {
  clear_error();
  SG_ARG(csr != nullptr && csc != nullptr, "null layout");
  const int64_t n = csr->n_rows, m = csc->n_rows;
  SG_ARG(csr->dense_rows == m && csc->dense_rows == n, "layouts do not describe one matrix and its transpose");
  SG_ARG(csc->packed >= 0 && csc->packed <= 2, "bad packing");
  SG_ARG(b >= 1 && b <= 16, "block width must be in [1, 16]");
  SG_ARG(ldv >= b && ldy >= b, "leading dimension < b");
  SG_ARG(ws_bytes >= snapgpu_spectral_apply_mixed_workspace_bytes(n, m, csc->n_pairs), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  float* Vs = (float*)q;
  q += align256(n * 16 * 4);
  float* Z = (float*)q;
  q += align256(m * 16 * 4);
  float* part = (float*)q;
  scale_in_kernel<16><<<(unsigned)ceil_div(n * 16, 256), 256, 0, st>>>(d_V, ldv, n, b, d_r32, Vs);
  // Z = w^2 o (X^T Vs) over the CSC panels, then Y = r o (X Z) - dinv o V over the CSR windows
  panel::product16_scaled(*csc, Vs, part, d_w2, Z, st);
  product_kernel<1><<<(unsigned)csr->n_blocks, kThreads, 0, st>>>(
      csr->ent, csr->bucket_start, csr->quad_off, csr->n_windows, csr->block_row, csr->row_slot, Z, d_r32,
      nullptr, d_V, ldv, d_dinv, d_Y, ldy, b);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
