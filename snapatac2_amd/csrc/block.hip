// Block-tiled layout of a sparse count matrix for the operator's two sparse
// products, with the dense operand staged through LDS.
//
// Each product is out[o] = sum over the entries (o, g, count) of count * D[g]
// for a dense operand D of 16-float rows (g = the gather side: cells for
// X^T V, features for X Z). Gathering D[g] from L2 for every entry is bound by
// the rate of 64-byte L2 requests; here one workgroup owns a block of output
// rows, keeps their sums in LDS, and sweeps the gather side in panels of kBR
// rows: each panel is read from memory once per block (contiguous 64 KB,
// prefetched into registers while the previous tile runs), stored in LDS, and
// every entry of the tile (block, panel) gathers its 64-byte row from LDS.
//
// Output rows are ordered by decreasing entry count and cut into blocks of
// about equal work (at most kBC accumulator slots each), so a block's rows have
// similar densities. A row's entries inside one panel (its "segment") are cut
// into pieces of at most kPiece entries, and every piece of a row gets its own
// accumulator slot (a row has as many slots as its longest segment has
// pieces). Inside a tile the pieces are ordered by decreasing length, and
// entry i of every piece long enough is stored in row i of the tile: thread k
// of the workgroup owns the k-th piece, reads its entries with loads that are
// coalesced across the wave, gathers all 16 columns of each entry's dense row
// from LDS and adds the piece's sum to its slot. No piece is shared and every
// slot has one writer per tile; the slot sums are added per row in slot order
// at the end: fixed summation order, no floating-point atomics, bitwise
// reproducible results.
//
// Entry encoding: bits 0-9 gather row within the panel, bits 10-19 the count
// (packed layouts, counts < 1024), bits 20-30 the slot. Unpacked (float values
// or large counts): the value in `val` instead of bits 10-19.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "spectral_common.h"

namespace snapgpu {
namespace block {

constexpr int kBR = 1024;          // gather rows per panel (10-bit local index)
constexpr int kBC = 1024;          // accumulator slots per block (= threads: one piece per thread)
constexpr int kPiece = 16;         // entries per piece at most (= rows of a tile)
constexpr int kThreads = 1024;     // 16 waves
constexpr int kRowStride = 20;     // floats per panel row in LDS (80 B: rows start on varied banks)
constexpr size_t kLds = (size_t)kBR * kRowStride * sizeof(float) + (size_t)kBC * 16 * sizeof(float);  // 144 KB
static_assert(kBC <= kThreads && kBC < 2048 && kBR == 1024, "slot and gather fields");

typedef float f4 __attribute__((ext_vector_type(4)));

struct Layout {
  const uint32_t* ent;
  const uint32_t* val;
  const int64_t* tile_start;  // [n_blocks * n_panels + 1]
  const uint16_t* tile_off;   // [n_blocks * n_panels][16] start of row i + 1 of each tile
  const int32_t* row0;        // [n_blocks + 1] first out position of each block
  const int32_t* out_row;     // [n_out] original row of each out position
  const int32_t* vslot;       // [n_out] first slot of each out position, block-relative
  const int32_t* nslots;      // [n_blocks] slots of each block
  int64_t n_blocks, n_panels, n_out, n_gather, n_entries;
  int32_t packed, value_type;
};

template <bool PACK, bool FV>
__device__ __forceinline__ float entry_value(uint32_t e, uint32_t v) {
  if (PACK) return (float)((e >> 10) & 0x3FF);
  return FV ? __uint_as_float(v) : (float)v;
}

// Epilogues (per output row o, its 16 sums s[c]):
//   EPI 0: out32[o][c] = scale[o] * s[c]                           (FP32)
//   EPI 1: Y[o][c] = (double)(scale[o] * s[c]) - dinv[o] * V[o][c]  (FP64, c < b)
struct Epi {
  float* out32;
  const float* scale;
  const double* V;
  int64_t ldv;
  const double* dinv;
  double* Y;
  int64_t ldy;
  int b;
};

template <bool PACK, bool FV, int EPI>
__global__ void __launch_bounds__(kThreads) product_kernel(Layout L, const f4* __restrict__ dense, Epi ep) {
  extern __shared__ f4 smem[];
  float* panel = reinterpret_cast<float*>(smem);                 // kBR rows x kRowStride floats
  f4* acc_l = smem + (size_t)kBR * kRowStride / 4;                // kBC slots x 4 float4
  const int64_t blk = blockIdx.x;
  const int tid = threadIdx.x;
  const int r0 = L.row0[blk], r1 = L.row0[blk + 1];
  const int nslots = L.nslots[blk];
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int i = tid; i < nslots * 4; i += kThreads) acc_l[i] = zero;
  const int64_t P = L.n_panels;
  // panel prefetch: float4 number tid + k * kThreads of the panel, k = 0..3
  // (row (tid >> 2) + 256 k, quarter tid & 3)
  const int64_t prow = tid >> 2;
  const int pq = tid & 3;
#define SG_PANEL_LOAD(P0)                                                                  \
  {                                                                                        \
    const int64_t rb = (P0) * kBR + prow;                                                  \
    pre0 = rb < L.n_gather ? dense[rb * 4 + pq] : zero;                                    \
    pre1 = rb + 256 < L.n_gather ? dense[(rb + 256) * 4 + pq] : zero;                      \
    pre2 = rb + 512 < L.n_gather ? dense[(rb + 512) * 4 + pq] : zero;                      \
    pre3 = rb + 768 < L.n_gather ? dense[(rb + 768) * 4 + pq] : zero;                      \
  }
  // Thread tid's piece of tile P0: its length LEN (rows i whose count of pieces
  // exceeds tid) and its entries EB[i] = ent[start + off_i + tid], i < LEN.
#define SG_PIECE_LOAD(P0, LEN, EB, VB)                                                     \
  {                                                                                        \
    const int64_t t_ = blk * P + (P0);                                                     \
    const int64_t st_ = L.tile_start[t_];                                                  \
    const uint16_t* of_ = L.tile_off + t_ * 16;                                            \
    int prev_ = 0;                                                                         \
    LEN = 0;                                                                               \
    _Pragma("unroll") for (int i = 0; i < kPiece; ++i) {                                   \
      const int o_ = of_[i];                                                               \
      const bool in_ = tid < o_ - prev_;                                                   \
      LEN += in_ ? 1 : 0;                                                                  \
      const int64_t at_ = st_ + prev_ + tid;                                               \
      EB[i] = (in_ && at_ < L.n_entries) ? __builtin_nontemporal_load(L.ent + at_) : 0u;   \
      if (!PACK) VB[i] = (in_ && at_ < L.n_entries) ? __builtin_nontemporal_load(L.val + at_) : 0u; \
      prev_ = o_;                                                                          \
    }                                                                                      \
  }
  f4 pre0, pre1, pre2, pre3;
  uint32_t ecur[kPiece], vcur[kPiece], enxt[kPiece], vnxt[kPiece];
  int len_cur = 0, len_nxt = 0;
  SG_PIECE_LOAD(0, len_cur, ecur, vcur)
  SG_PANEL_LOAD(0)
  for (int64_t p = 0; p < P; ++p) {
    __syncthreads();
    {
      float* rowp = panel + prow * kRowStride + pq * 4;
      *reinterpret_cast<f4*>(rowp) = pre0;
      *reinterpret_cast<f4*>(rowp + 256 * kRowStride) = pre1;
      *reinterpret_cast<f4*>(rowp + 512 * kRowStride) = pre2;
      *reinterpret_cast<f4*>(rowp + 768 * kRowStride) = pre3;
    }
    __syncthreads();
    // the next tile's piece and panel are in flight while this tile runs
    if (p + 1 < P) {
      SG_PIECE_LOAD(p + 1, len_nxt, enxt, vnxt)
      SG_PANEL_LOAD(p + 1)
    }
    if (len_cur > 0) {
      f4 a0 = zero, a1 = zero, a2 = zero, a3 = zero;
#pragma unroll
      for (int i = 0; i < kPiece; ++i) {
        if (i < len_cur) {
          const uint32_t e = ecur[i];
          const float c = entry_value<PACK, FV>(e, vcur[i]);
          const float* rp = panel + (e & 0x3FF) * kRowStride;
          a0 += c * *reinterpret_cast<const f4*>(rp);
          a1 += c * *reinterpret_cast<const f4*>(rp + 4);
          a2 += c * *reinterpret_cast<const f4*>(rp + 8);
          a3 += c * *reinterpret_cast<const f4*>(rp + 12);
        }
      }
      const int slot = (int)((ecur[0] >> 20) & 0x7FF);
      if (slot < nslots) {  // always, for a layout from snapgpu_block_build
        acc_l[slot * 4] += a0;
        acc_l[slot * 4 + 1] += a1;
        acc_l[slot * 4 + 2] += a2;
        acc_l[slot * 4 + 3] += a3;
      }
    }
    len_cur = len_nxt;
#pragma unroll
    for (int i = 0; i < kPiece; ++i) {
      ecur[i] = enxt[i];
      vcur[i] = vnxt[i];
    }
  }
#undef SG_PIECE_LOAD
#undef SG_PANEL_LOAD
  __syncthreads();
  // epilogue: row r (out position r0 + i) adds its slots in order
  for (int t = tid; t < (r1 - r0) * 4; t += kThreads) {
    const int i = t >> 2, qq = t & 3;
    const int r = r0 + i;
    const int s0 = L.vslot[r];
    const int s1 = (r + 1 < r1) ? L.vslot[r + 1] : nslots;
    f4 s = zero;
    for (int k = s0; k < s1; ++k) s += acc_l[k * 4 + qq];
    const int64_t o = L.out_row[r];
    if (EPI == 0) {
      const float sc = ep.scale[o];
      reinterpret_cast<f4*>(ep.out32)[o * 4 + qq] = sc * s;
    } else {
      const float sc = ep.scale[o];
      const double di = ep.dinv[o];
      const int c = qq * 4;
      double* y = ep.Y + o * ep.ldy + c;
      const double* v = ep.V + o * ep.ldv + c;
      if (c < ep.b) y[0] = (double)(sc * s.x) - di * v[0];
      if (c + 1 < ep.b) y[1] = (double)(sc * s.y) - di * v[1];
      if (c + 2 < ep.b) y[2] = (double)(sc * s.z) - di * v[2];
      if (c + 3 < ep.b) y[3] = (double)(sc * s.w) - di * v[3];
    }
  }
}

// ---------------------------------------------------------------------------
// Build
// ---------------------------------------------------------------------------

// key = ~nnz (descending entry counts), ties by row id (stable radix sort)
__global__ void order_keys_kernel(const int64_t* ptr, int64_t n, uint32_t* keys, uint32_t* ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t len = ptr[i + 1] - ptr[i];
  keys[i] = ~(uint32_t)min(len, (int64_t)0xFFFFFFFF);
  ids[i] = (uint32_t)i;
}

// Runs of one row's entries by panel (the row's columns ascend, so each panel's
// entries are one run), 64 entries at a time. For each entry: its run's first
// entry (s) and whether it is the run's last. One wave per row.
struct RunScan {
  int64_t run_start;
  __device__ void step(const int32_t* __restrict__ idx, int64_t e0, int64_t e1, int64_t b, int lane,
                       int64_t& e, bool& has, int32_t& col, int& pnl, int64_t& s, bool& last,
                       uint64_t& starts) {
    e = b + lane;
    has = e < e1;
    col = has ? idx[e] : 0;
    pnl = has ? col / kBR : -1;
    int prev = __shfl_up(pnl, 1, 64);
    if (lane == 0) prev = (b == e0) ? -1 : (int)(idx[b - 1] / kBR);
    int next = __shfl_down(pnl, 1, 64);
    if (lane == 63) next = (e + 1 < e1) ? (int)(idx[e + 1] / kBR) : -2;
    if (e + 1 >= e1) next = -2;
    starts = ballot64(has && pnl != prev);
    const uint64_t upto = starts & (lanemask_lt(lane) | (1ull << lane));
    s = upto ? b + (63 - __clzll((unsigned long long)upto)) : run_start;
    last = has && next != pnl;
  }
  __device__ void advance(int64_t b, uint64_t starts) {
    if (starts) run_start = b + (63 - __clzll((unsigned long long)starts));
  }
};

// Per out position q (row order[q]): its entry count and pieces[q] = the largest
// number of pieces any of its panel segments needs.
__global__ void __launch_bounds__(256) row_pieces_kernel(const int64_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ idx,
                                                         const uint32_t* __restrict__ order, int64_t n,
                                                         int64_t* __restrict__ nnz_q, int32_t* __restrict__ pieces) {
  const int lane = threadIdx.x & 63;
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (q >= n) return;
  const int64_t row = order[q];
  const int64_t e0 = ptr[row], e1 = ptr[row + 1];
  RunScan rs{e0};
  int best = 0;
  for (int64_t b = e0; b < e1; b += 64) {
    int64_t e, s;
    bool has, last;
    int32_t col;
    int pnl;
    uint64_t starts;
    rs.step(idx, e0, e1, b, lane, e, has, col, pnl, s, last, starts);
    if (last) best = max(best, (int)((e + 1 - s + kPiece - 1) / kPiece));
    rs.advance(b, starts);
  }
  for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
  if (lane == 0) {
    nnz_q[q] = e1 - e0;
    pieces[q] = best;
  }
}

// Piece lengths Lq at (block b, panel p, slot s) = base[b] + p * nslots[b] + s.
// One wave per out position; qblock[q] = its block, vslot[q] = its first slot.
struct PosTables {
  const uint32_t* order;    // out position -> row
  const int32_t* qblock;    // out position -> block
  const int32_t* vslot;     // out position -> first slot (block-relative)
  const int64_t* base;      // block -> first index of its (panel, slot) table
  const int32_t* nslots;    // block -> slots
};

__device__ __forceinline__ int64_t table_pos(const PosTables& T, int blk, int64_t p, int s) {
  return T.base[blk] + p * (int64_t)T.nslots[blk] + s;
}

// FILL = false: the last entry of each (row, panel) run writes the lengths of
// the run's pieces into Lq (zeroed beforehand). FILL = true: every entry goes
// to E[its piece] + its offset in the piece.
// FILL: entry i of a piece goes to row i of its tile, at the piece's rank:
// tile start + start of row i + rank.
struct FillTables {
  const int64_t* E;          // exclusive scan of the piece lengths (tile starts at slot 0)
  const int32_t* rank;       // rank of each piece within its tile (same indexing as Lq)
  const uint16_t* tile_off;  // [tile][16]: start of row i + 1
  int64_t P;
};

template <bool FILL, bool PACK>
__global__ void __launch_bounds__(256) pieces_kernel(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                     const uint32_t* __restrict__ val, int64_t n, PosTables T,
                                                     int64_t table_size, int64_t nnz, int64_t* __restrict__ Lq,
                                                     FillTables F, uint32_t* __restrict__ ent,
                                                     uint32_t* __restrict__ val_out, int32_t* __restrict__ err) {
  const int lane = threadIdx.x & 63;
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (q >= n) return;
  const int64_t row = T.order[q];
  const int blk = T.qblock[q];
  const int vs = T.vslot[q];
  const int64_t e0 = ptr[row], e1 = ptr[row + 1];
  RunScan rs{e0};
  for (int64_t b = e0; b < e1; b += 64) {
    int64_t e, s;
    bool has, last;
    int32_t col;
    int pnl;
    uint64_t starts;
    rs.step(idx, e0, e1, b, lane, e, has, col, pnl, s, last, starts);
    if (has) {
      const int64_t i = e - s;  // position within the (row, panel) run
      if (!FILL) {
        if (last) {
          const int64_t len = e + 1 - s;
          for (int64_t k = 0; k * kPiece < len; ++k) {
            const int64_t tp = table_pos(T, blk, pnl, vs + (int)k);
            if (tp >= 0 && tp < table_size) Lq[tp] = min((int64_t)kPiece, len - k * kPiece);
            else atomicOr(err, 1);
          }
        }
      } else {
        const int piece = (int)(i / kPiece);
        const int row = (int)(i % kPiece);
        const int64_t tp = table_pos(T, blk, pnl, vs + piece);
        const int64_t tile = (int64_t)blk * F.P + pnl;
        int64_t dst = -1;
        if (tp >= 0 && tp < table_size) {
          const int64_t tstart = F.E[table_pos(T, blk, pnl, 0)];
          const int64_t roff = row == 0 ? 0 : (int64_t)F.tile_off[tile * 16 + row - 1];
          dst = tstart + roff + F.rank[tp];
        }
        if (dst < 0 || dst >= nnz) atomicOr(err, 2);
        const uint32_t lo = (uint32_t)(col - pnl * kBR) | ((uint32_t)(vs + piece) << 20);
        if (dst >= 0 && dst < nnz) {
          if (PACK) {
            ent[dst] = lo | (val[e] << 10);
          } else {
            ent[dst] = lo;
            val_out[dst] = val[e];
          }
        }
      }
    }
    rs.advance(b, starts);
  }
}

// One wave per tile (block b, panel p): the pieces (slots with a length > 0)
// ordered by decreasing length, then slot; rank[slot] = position in that order,
// tile_off[tile][i] = start of row i + 1 (row i holds entry i of every piece
// longer than i, so it has c_i = #pieces longer than i entries).
__global__ void __launch_bounds__(256) tile_order_kernel(const int64_t* __restrict__ Lq, const int64_t* __restrict__ base,
                                                         const int32_t* __restrict__ nslots, int64_t n_blocks, int64_t P,
                                                         int32_t* __restrict__ rank, uint16_t* __restrict__ tile_off) {
  const int lane = threadIdx.x & 63;
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= n_blocks * P) return;
  const int64_t b = t / P, p = t - b * P;
  const int ns = nslots[b];
  const int64_t t0 = base[b] + p * ns;
  // counts of pieces per length 1..kPiece (wave-uniform)
  int cnt[kPiece + 1];
#pragma unroll
  for (int v = 0; v <= kPiece; ++v) cnt[v] = 0;
  for (int s0 = 0; s0 < ns; s0 += 64) {
    const int s = s0 + lane;
    const int len = s < ns ? (int)Lq[t0 + s] : 0;
#pragma unroll
    for (int v = 1; v <= kPiece; ++v) cnt[v] += popc64(ballot64(len == v));
  }
  // c_i = #pieces longer than i; the pieces of length v start at rank c_v
  int c[kPiece + 1];
  c[kPiece] = 0;
#pragma unroll
  for (int v = kPiece - 1; v >= 0; --v) c[v] = c[v + 1] + cnt[v + 1];
  if (lane == 0) {
    int off = 0;
#pragma unroll
    for (int i = 0; i < kPiece; ++i) {
      off += c[i];
      tile_off[t * 16 + i] = (uint16_t)off;
    }
  }
  int run[kPiece + 1];
#pragma unroll
  for (int v = 0; v <= kPiece; ++v) run[v] = 0;
  const uint64_t lt = lanemask_lt(lane);
  for (int s0 = 0; s0 < ns; s0 += 64) {
    const int s = s0 + lane;
    const int len = s < ns ? (int)Lq[t0 + s] : 0;
    int r = 0;
#pragma unroll
    for (int v = 1; v <= kPiece; ++v) {
      const uint64_t m = ballot64(len == v);
      if (len == v) r = c[v] + run[v] + popc64(m & lt);
      run[v] += popc64(m);
    }
    if (s < ns) rank[t0 + s] = len > 0 ? r : -1;
  }
}

// Lq (inclusive scan) -> exclusive starts in place: E[t] = Lq_incl[t] - len, computed as
// the inclusive scan shifted by one (E[0] = 0)
__global__ void shift_kernel(const int64_t* incl, int64_t count, int64_t* excl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > count) return;
  excl[t] = t == 0 ? 0 : incl[t - 1];
}

__global__ void tile_start_kernel(const int64_t* E, const int64_t* base, const int32_t* nslots, int64_t n_blocks,
                                  int64_t P, int64_t total, int64_t* tile_start) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_blocks * P) return;
  if (t == n_blocks * P) {
    tile_start[t] = total;
    return;
  }
  const int64_t b = t / P, p = t - b * P;
  tile_start[t] = E[base[b] + p * nslots[b]];
}

__global__ void max_u32_kernel(const uint32_t* __restrict__ v, int64_t n, unsigned int* __restrict__ out) {
  unsigned int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, v[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

}  // namespace block
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::block;

namespace {

// Plan workspace (carved from the caller's buffer)
struct PlanWs {
  uint32_t *keys, *keys2, *ids, *order;
  int64_t* nnz_q;
  int32_t *pieces, *qblock, *vslot, *row0, *nslots;
  int64_t* base;
  unsigned int* maxc;
  void* sort_tmp;
  size_t sort_bytes;
};

int64_t plan_layout(int64_t nrows, PlanWs* w, char* p0) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> char* {
    char* p = p0 ? p0 + off : nullptr;
    off += align256(bytes);
    return p;
  };
  PlanWs t;
  const int64_t r = std::max<int64_t>(nrows, 1);
  t.keys = (uint32_t*)take(r * 4);
  t.keys2 = (uint32_t*)take(r * 4);
  t.ids = (uint32_t*)take(r * 4);
  t.order = (uint32_t*)take(r * 4);
  t.nnz_q = (int64_t*)take(r * 8);
  t.pieces = (int32_t*)take(r * 4);
  t.qblock = (int32_t*)take(r * 4);
  t.vslot = (int32_t*)take(r * 4);
  t.row0 = (int32_t*)take((r + 1) * 4);
  t.nslots = (int32_t*)take(r * 4);
  t.base = (int64_t*)take((r + 1) * 8);
  t.maxc = (unsigned int*)take(4);
  size_t sb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sb, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)std::min<int64_t>(r, INT32_MAX));
  t.sort_bytes = sb;
  t.sort_tmp = take((int64_t)sb + 256);
  if (w) *w = t;
  return off + 256;
}

}  // namespace

extern "C" {

int64_t snapgpu_block_plan_workspace_bytes(int64_t nrows) { return plan_layout(nrows, nullptr, nullptr); }

int snapgpu_block_plan(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                       int64_t dense_rows, int64_t nnz, int32_t value_type, int64_t target_blocks, void* d_ws,
                       int64_t ws_bytes, int64_t* n_blocks, int64_t* table_size, int32_t* packed, void* stream) {
  clear_error();
  SG_ARG(nrows >= 1 && nrows < INT32_MAX && dense_rows >= 1 && nnz >= 0, "bad shape");
  SG_ARG(value_type == 0 || value_type == 1, "value_type must be 0 (uint32) or 1 (float32)");
  SG_ARG(target_blocks >= 1, "target_blocks must be >= 1");
  SG_ARG(ws_bytes >= snapgpu_block_plan_workspace_bytes(nrows), "workspace too small");
  hipStream_t st = as_stream(stream);
  PlanWs w;
  plan_layout(nrows, &w, (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255));
  const unsigned g = (unsigned)ceil_div(nrows, 256);
  order_keys_kernel<<<g, 256, 0, st>>>(d_ptr, nrows, w.keys, w.ids);
  SG_LAUNCH_CHECK();
  size_t sb = w.sort_bytes;
  SG_HIP(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, sb, w.keys, w.keys2, w.ids, w.order, (int)nrows, 0, 32, st));
  row_pieces_kernel<<<(unsigned)ceil_div(nrows * 64, 256), 256, 0, st>>>(d_ptr, d_idx, w.order, nrows, w.nnz_q,
                                                                       w.pieces);
  SG_HIP(hipMemsetAsync(w.maxc, 0, 4, st));
  if (value_type == 0 && nnz > 0)
    max_u32_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nnz, 256), 4096), 256, 0, st>>>(d_val, nnz, w.maxc);
  SG_LAUNCH_CHECK();
  std::vector<int64_t> nz(nrows);
  std::vector<int32_t> pc(nrows);
  unsigned int maxc = 0;
  SG_HIP(hipMemcpyAsync(nz.data(), w.nnz_q, nrows * 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(pc.data(), w.pieces, nrows * 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&maxc, w.maxc, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  int64_t total = 0;
  for (int64_t q = 0; q < nrows; ++q) total += nz[q];
  if (total != nnz) {
    set_error("block layout: entry count mismatch");
    return SNAPGPU_ERR_DATA;
  }
  // blocks: consecutive out positions (densest rows first), at most kBC slots
  // and about nnz / target_blocks entries each
  const int64_t P = ceil_div(dense_rows, kBR);
  const int64_t target = std::max<int64_t>(1, ceil_div(std::max<int64_t>(nnz, 1), target_blocks));
  std::vector<int32_t> qblock(nrows), vslot(nrows), row0, nslots;
  std::vector<int64_t> base;
  int64_t cur_entries = 0;
  int cur_slots = 0;
  int64_t tsize = 0;
  row0.push_back(0);
  for (int64_t q = 0; q < nrows; ++q) {
    if (pc[q] > kBC) {
      set_error("block layout: a row needs more accumulator slots than a block holds");
      return SNAPGPU_ERR_CAPACITY;
    }
    if (q > row0.back() && (cur_slots + pc[q] > kBC || cur_entries + nz[q] > target)) {
      nslots.push_back(cur_slots);
      base.push_back(tsize);
      tsize += (int64_t)cur_slots * P;
      row0.push_back((int32_t)q);
      cur_slots = 0;
      cur_entries = 0;
    }
    qblock[q] = (int32_t)(row0.size() - 1);
    vslot[q] = cur_slots;
    cur_slots += pc[q];
    cur_entries += nz[q];
  }
  nslots.push_back(cur_slots);
  base.push_back(tsize);
  tsize += (int64_t)cur_slots * P;
  row0.push_back((int32_t)nrows);
  const int64_t nb = (int64_t)nslots.size();
  base.push_back(tsize);
  SG_HIP(hipMemcpyAsync(w.qblock, qblock.data(), nrows * 4, hipMemcpyHostToDevice, st));
  SG_HIP(hipMemcpyAsync(w.vslot, vslot.data(), nrows * 4, hipMemcpyHostToDevice, st));
  SG_HIP(hipMemcpyAsync(w.row0, row0.data(), (nb + 1) * 4, hipMemcpyHostToDevice, st));
  SG_HIP(hipMemcpyAsync(w.nslots, nslots.data(), nb * 4, hipMemcpyHostToDevice, st));
  SG_HIP(hipMemcpyAsync(w.base, base.data(), (nb + 1) * 8, hipMemcpyHostToDevice, st));
  SG_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  *n_blocks = nb;
  *table_size = tsize;
  *packed = (value_type == 0 && maxc < 1024u) ? 1 : 0;
  return SNAPGPU_OK;
}

int64_t snapgpu_block_build_workspace_bytes(int64_t table_size) {
  const int64_t t = std::max<int64_t>(table_size, 0) + 1;
  return 2 * align256(t * 8) + align256(t * 4) + align256(snapgpu_scan_workspace_bytes(t)) + 768;
}

int snapgpu_block_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                        int64_t dense_rows, int64_t nnz, int32_t packed, const void* d_plan_ws, int64_t plan_ws_bytes,
                        int64_t n_blocks, int64_t table_size, void* d_ws, int64_t ws_bytes, uint32_t* d_ent,
                        uint32_t* d_val_out, int64_t* d_tile_start, uint16_t* d_tile_off, int32_t* d_row0,
                        int32_t* d_out_row, int32_t* d_vslot, int32_t* d_nslots, void* stream) {
  clear_error();
  SG_ARG(plan_ws_bytes >= snapgpu_block_plan_workspace_bytes(nrows), "plan workspace too small");
  SG_ARG(ws_bytes >= snapgpu_block_build_workspace_bytes(table_size), "workspace too small");
  SG_ARG(packed || d_val_out != nullptr, "unpacked layout needs a value array");
  hipStream_t st = as_stream(stream);
  PlanWs w;
  plan_layout(nrows, &w, (char*)(((uintptr_t)d_plan_ws + 255) & ~(uintptr_t)255));
  const int64_t P = ceil_div(dense_rows, kBR);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  const int64_t t = table_size + 1;
  int64_t* Lq = (int64_t*)q;
  q += align256(t * 8);
  int64_t* E = (int64_t*)q;
  q += align256(t * 8);
  int32_t* R = (int32_t*)q;
  q += align256(t * 4);
  int32_t* err = (int32_t*)q;
  q += 256;
  void* scan_ws = q;
  const int64_t scan_b = snapgpu_scan_workspace_bytes(t);
  PosTables T{w.order, w.qblock, w.vslot, w.base, w.nslots};
  SG_HIP(hipMemsetAsync(Lq, 0, t * 8, st));
  SG_HIP(hipMemsetAsync(err, 0, 4, st));
  const unsigned gw = (unsigned)ceil_div(nrows * 64, 256);
  const FillTables F{E, R, d_tile_off, P};
  pieces_kernel<false, true><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, T, table_size, nnz, Lq, F, nullptr,
                                                 nullptr, err);
  SG_LAUNCH_CHECK();
  if (n_blocks * P > 0)
    tile_order_kernel<<<(unsigned)ceil_div(n_blocks * P * 64, 256), 256, 0, st>>>(Lq, w.base, w.nslots, n_blocks, P,
                                                                                R, d_tile_off);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(Lq, t, scan_ws, scan_b, stream);
  if (rc) return rc;
  shift_kernel<<<(unsigned)ceil_div(t + 1, 256), 256, 0, st>>>(Lq, table_size, E);
  SG_LAUNCH_CHECK();
  if (packed)
    pieces_kernel<true, true><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, T, table_size, nnz, nullptr, F, d_ent,
                                                  nullptr, err);
  else
    pieces_kernel<true, false><<<gw, 256, 0, st>>>(d_ptr, d_idx, d_val, nrows, T, table_size, nnz, nullptr, F,
                                                   d_ent, d_val_out, err);
  tile_start_kernel<<<(unsigned)ceil_div(n_blocks * P + 1, 256), 256, 0, st>>>(E, w.base, w.nslots, n_blocks, P, nnz,
                                                                            d_tile_start);
  SG_LAUNCH_CHECK();
  SG_HIP(hipMemcpyAsync(d_row0, w.row0, (n_blocks + 1) * 4, hipMemcpyDeviceToDevice, st));
  SG_HIP(hipMemcpyAsync(d_out_row, w.order, nrows * 4, hipMemcpyDeviceToDevice, st));
  SG_HIP(hipMemcpyAsync(d_vslot, w.vslot, nrows * 4, hipMemcpyDeviceToDevice, st));
  SG_HIP(hipMemcpyAsync(d_nslots, w.nslots, n_blocks * 4, hipMemcpyDeviceToDevice, st));
  SG_HIP(hipStreamSynchronize(st));
  // the last entry of E must be nnz: every entry was placed exactly once
  int64_t last = 0;
  int32_t herr = 0;
  SG_HIP(hipMemcpy(&last, E + table_size, 8, hipMemcpyDeviceToHost));
  SG_HIP(hipMemcpy(&herr, err, 4, hipMemcpyDeviceToHost));
  if (last != nnz || herr != 0) {
    set_error("block layout: piece lengths do not add up to the entry count (columns not sorted within rows?)");
    return SNAPGPU_ERR_DATA;
  }
  return SNAPGPU_OK;
}

}  // extern "C"

namespace {

Layout to_layout(const snapgpu_block_layout& B) {
  Layout L;
  L.ent = B.ent;
  L.val = B.val;
  L.tile_start = B.tile_start;
  L.tile_off = B.tile_off;
  L.row0 = B.row0;
  L.out_row = B.out_row;
  L.vslot = B.vslot;
  L.nslots = B.nslots;
  L.n_blocks = B.n_blocks;
  L.n_panels = B.n_panels;
  L.n_out = B.n_rows;
  L.n_gather = B.dense_rows;
  L.n_entries = B.n_entries;
  L.packed = B.packed;
  L.value_type = B.value_type;
  return L;
}

template <int EPI>
int launch_block_product(const snapgpu_block_layout& B, const float* dense, const Epi& ep, hipStream_t st) {
  if (B.n_blocks <= 0) return SNAPGPU_OK;
  const Layout L = to_layout(B);
  const f4* d4 = reinterpret_cast<const f4*>(dense);
  const dim3 grid((unsigned)B.n_blocks), blk(kThreads);
  if (B.packed) {
    (void)hipFuncSetAttribute((const void*)product_kernel<true, false, EPI>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds);
    product_kernel<true, false, EPI><<<grid, blk, kLds, st>>>(L, d4, ep);
  } else if (B.value_type == 1) {
    (void)hipFuncSetAttribute((const void*)product_kernel<false, true, EPI>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds);
    product_kernel<false, true, EPI><<<grid, blk, kLds, st>>>(L, d4, ep);
  } else {
    (void)hipFuncSetAttribute((const void*)product_kernel<false, false, EPI>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds);
    product_kernel<false, false, EPI><<<grid, blk, kLds, st>>>(L, d4, ep);
  }
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // namespace

extern "C" {

int32_t snapgpu_block_panel_rows(void) { return kBR; }

int64_t snapgpu_spectral_apply_block_workspace_bytes(int64_t n, int64_t m) {
  return align256(n * 16 * 4) + align256(m * 16 * 4) + 256;
}

int snapgpu_spectral_apply_block(const snapgpu_block_layout* csr, const snapgpu_block_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv, const double* d_V,
                                 int64_t ldv, int32_t b, double* d_Y, int64_t ldy, void* d_ws, int64_t ws_bytes,
                                 void* stream) {
  clear_error();
  SG_ARG(csr != nullptr && csc != nullptr, "null layout");
  const int64_t n = csr->n_rows, m = csc->n_rows;
  SG_ARG(csr->dense_rows == m && csc->dense_rows == n, "layouts do not describe one matrix and its transpose");
  SG_ARG(csr->n_panels == ceil_div(m, kBR) && csc->n_panels == ceil_div(n, kBR), "panel count mismatch");
  SG_ARG(b >= 1 && b <= 16, "block width must be in [1, 16]");
  SG_ARG(csr->n_entries == csc->n_entries, "layouts hold different entry counts");
  SG_ARG(ldv >= b && ldy >= b, "leading dimension < b");
  SG_ARG(ws_bytes >= snapgpu_spectral_apply_block_workspace_bytes(n, m), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  float* Vs = (float*)q;
  q += align256(n * 16 * 4);
  float* Z = (float*)q;
  spectral::scale_in_kernel<16><<<(unsigned)ceil_div(n * 16, 256), 256, 0, st>>>(d_V, ldv, n, b, d_r32, Vs);
  SG_LAUNCH_CHECK();
  Epi e0{Z, d_w2, nullptr, 0, nullptr, nullptr, 0, b};
  int rc = launch_block_product<0>(*csc, Vs, e0, st);
  if (rc) return rc;
  Epi e1{nullptr, d_r32, d_V, ldv, d_dinv, d_Y, ldy, b};
  return launch_block_product<1>(*csr, Z, e1, st);
}

}  // extern "C"
