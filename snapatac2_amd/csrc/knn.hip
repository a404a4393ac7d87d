// Path C: exact k-nearest-neighbour graph (Euclidean, self excluded by index).
//
// Reference: snapatac2-core/src/utils/knn.rs:9-32 (kd-tree; neighbours in
// increasing squared Euclidean distance, index == i skipped, k taken, distance =
// sqrt) and :58-78 (each row sorted by neighbour index); binding
// snapatac2-python/src/knn.rs:7-16.
//
// Pipeline (all on the device):
//  1. prepare: FP64 rows -> augmented 32-wide fp32 rows x~ = (x, |x|^2, 0) for
//     the database and q~ = (x, -1/2, 0) for the queries, so s = x~.q~ =
//     x.q - |x|^2/2 and d^2 = |q|^2 - 2 s; |x|^2 and the largest norm.
//  2. candidates. BF16 mode: the points are put in kd-tree leaf order (leaves of
//     128), each leaf is shifted by its own centre, and a workgroup of 8 waves
//     (128 queries = one leaf) streams database leaves through LDS, each as
//     three bf16 terms (x = hi + mid + lo + O(2^-24 x)) multiplied with six
//     v_mfma_f32_16x16x32_bf16 per 16x16 tile (every product of the split terms
//     down to order 2^-24, fp32 accumulation), so the approximation error is at
//     fp32 level even where d^2 is much smaller than |q|^2. A leaf whose
//     bounding sphere lies beyond every query's current threshold is skipped
//     (the threshold is kept in the certificate, so skipping stays exact).
//     FP32 mode: SIMT FMAs over every point in the original order. In both,
//     lane l holds query l%16 and 4 points of a tile, so the 4 lanes of a query
//     cover 4 disjoint stripes of the database; each lane keeps its best 32
//     (sorted in registers) -> 128 candidates per query (>= 2k for k <= 64).
//  3. finalize: one wave per query rescoring its candidates in exact FP64
//     (sum of (a-b)^2 in dimension order, no fused multiply-add: the reference's
//     squared_euclidean), ordering by (d^2, index), and checking a certificate:
//     every non-candidate has approx d^2 >= tau_stripe, so if the k-th exact d^2
//     is below the smallest exact d^2 a non-candidate can have, the set is exact.
//     That bound uses |approx - exact| <= c (|x||q| + |x|^2 + |q|^2) together
//     with d >= | |x| - |q| |, so it scales with the query's own norm rather than
//     the largest norm in the data. Rows that fail are rescanned exactly by the
//     fallback kernel.
//  4. output: k neighbours per row sorted by index, distance = sqrt(d^2).
#include <algorithm>
#include <cfloat>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.h"

namespace snapgpu {
namespace knn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kDim = 32;      // augmented, padded row width
constexpr int kMaxData = 30;  // data dimensions supported
constexpr int kList = 32;     // candidates kept per lane (stripe)
constexpr int kCand = 4 * kList;  // candidates per query
constexpr int kMaxK = 100;

enum Mode : int32_t { kModeFP32 = 0, kModeBF16 = 1 };

// Error constants c of the approximate d^2 (see the certificate above):
// |approx - exact| <= c * (|x||q| + |x|^2 + |q|^2).
// BF16 (three-term split, six products): the omitted products and the split
// remainders are <= ~4.1 * 2^-24 per term, fp32 accumulation of 192 terms adds
// <= 192 * 2^-24, doubled by d^2 = |q|^2 - 2s, plus fp32 rounding of the inputs
// and of |q|^2: c = 3e-5.
// FP32: 32 sequential FMAs (<= 32 * 2^-24), doubled, plus input rounding: c = 1e-5.
constexpr double kEpsBF16 = 3.0e-5;
constexpr double kEpsFP32 = 1.0e-5;

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// ---------------------------------------------------------------------------
// 1. prepare
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) prepare_kernel(const double* __restrict__ X, int64_t n,
                                                      int d, int64_t ld, int64_t n_pad, float* xf,
                                                      float* qf, float* qnorm,
                                                      unsigned int* max_norm_bits) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int c = threadIdx.x & 31;
  if (row >= n_pad) return;
  double v = 0.0;
  if (row < n && c < d) v = X[row * ld + c];
  // |x|^2 in fp64 (order-independent enough: only used inside the error bound)
  double sq = v * v;
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 32);
  double xv = v, qv = v;
  if (c == kMaxData) {
    xv = row < n ? sq : 0.0;
    qv = -0.5;
  } else if (c > kMaxData) {
    xv = 0.0;
    qv = 0.0;
  }
  if (row >= n) qv = 0.0;
  const int64_t o = row * kDim + c;
  xf[o] = (float)xv;
  qf[o] = (float)qv;
  if (c == 0) {
    qnorm[row] = (float)sq;
    if (row < n) atomicMax(max_norm_bits, __float_as_uint((float)sqrt(sq) * 1.0000001f));
  }
}

// ---------------------------------------------------------------------------
// 2. candidates
// ---------------------------------------------------------------------------
struct CandParams {
  const float *xf, *qf, *qnorm;
  int64_t n, n_pad;
  const int32_t* rows;  // optional list of query rows (null = all)
  int64_t n_rows;
  int32_t* cand;        // [n_rows][kCand]
  float* tau;           // [n_rows][4]
};

// Inserts (s, p) into a descending list kept in registers.
__device__ __forceinline__ void list_insert(float (&L)[kList], int32_t (&I)[kList], float s, int32_t p) {
#pragma unroll
  for (int r = kList - 1; r > 0; --r) {
    const bool up = s > L[r - 1];
    const bool here = s > L[r];
    L[r] = up ? L[r - 1] : (here ? s : L[r]);
    I[r] = up ? I[r - 1] : (here ? p : I[r]);
  }
  if (s > L[0]) {
    L[0] = s;
    I[0] = p;
  }
}

// ---------------------------------------------------------------------------
// 2b. spatial order + per-workgroup centres (BF16 path)
// ---------------------------------------------------------------------------
// The BF16 error bound scales with the squared norms of the (shifted) vectors,
// and chunks can only be skipped when they are spatially compact. The points
// are therefore ordered as the leaves of a balanced kd-tree: the leaves are
// runs of 128 consecutive positions (the last may be shorter); a node covering
// leaves [a, b) is split at leaf a + (b - a + 1) / 2 after its points are
// sorted along the dimension of largest extent within the node. One radix sort
// per level, keyed by (node, coordinate). Every vector is then shifted by its
// workgroup's centre, so the norms (and the bound) are comparable to
// within-group distances instead of distances from the origin.
constexpr int kWgWaves = 8;   // waves per workgroup in the centred kernel
constexpr int kWgQueries = kWgWaves * 16;
constexpr int kChunk = 128;   // database points staged in LDS per step
static_assert(kChunk == kWgQueries, "database chunk w must hold the queries of workgroup w");
constexpr int kMaxLevels = 26;    // kd-tree depth bound (2^26 leaves >> INT32_MAX / 128)
constexpr int kSpreadSlice = 8192;  // positions per block in the extent reduction

__device__ __forceinline__ unsigned int f2ord(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned int o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Node (at depth `level`) of leaf `leaf` among nl leaves: the path bits.
__device__ __forceinline__ uint32_t kd_node(int64_t leaf, int64_t nl, int level) {
  int64_t a = 0, b = nl;
  uint32_t id = 0;
  for (int l = 0; l < level; ++l) {
    const int64_t m = a + (b - a + 1) / 2;
    const bool right = leaf >= m;
    id = (id << 1) | (right ? 1u : 0u);
    if (right) a = m; else b = m;
  }
  return id;
}

__global__ void kd_mm_init_kernel(unsigned int* mm, int64_t count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) mm[t] = (t & 1) ? 0u : 0xFFFFFFFFu;
}

// mm[node][c] = {min, max} (ordered-float bits) over the node's points.
// Block (node i, slice s) covers positions node_begin[i] + s * kSpreadSlice ...
__global__ void __launch_bounds__(256) kd_extent_kernel(const double* X, int d, int64_t ld,
                                                        const uint32_t* order, const int64_t* nodes,
                                                        int slices, unsigned int* mm) {
  __shared__ unsigned int lmin[kDim], lmax[kDim];
  const int64_t i = blockIdx.x / slices;
  const int s = blockIdx.x % slices;
  const int64_t b0 = nodes[2 * i], b1 = nodes[2 * i + 1];
  const int64_t p0 = b0 + (int64_t)s * kSpreadSlice, p1 = min(b1, p0 + kSpreadSlice);
  if (threadIdx.x < kDim) {
    lmin[threadIdx.x] = 0xFFFFFFFFu;
    lmax[threadIdx.x] = 0u;
  }
  __syncthreads();
  if (p0 < p1) {
    const int c = threadIdx.x & 31;
    unsigned int mn = 0xFFFFFFFFu, mx = 0u;
    if (c < d) {
      for (int64_t p = p0 + (threadIdx.x >> 5); p < p1; p += blockDim.x / 32) {
        const unsigned int o = f2ord((float)X[(int64_t)order[p] * ld + c]);
        mn = min(mn, o);
        mx = max(mx, o);
      }
      atomicMin(&lmin[c], mn);
      atomicMax(&lmax[c], mx);
    }
  }
  __syncthreads();
  if (p0 < p1 && threadIdx.x < d) {
    atomicMin(&mm[(i * kDim + threadIdx.x) * 2], lmin[threadIdx.x]);
    atomicMax(&mm[(i * kDim + threadIdx.x) * 2 + 1], lmax[threadIdx.x]);
  }
}

// keys[p] = (node << 32) | ordered bits of the point's coordinate along the
// node's widest dimension; vals[p] = order[p].
__global__ void kd_key_kernel(const double* X, int64_t n, int d, int64_t ld, const uint32_t* order,
                              int64_t nl, int level, const unsigned int* mm, uint64_t* keys,
                              uint32_t* vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t node = kd_node(p / kChunk, nl, level);
  int best = 0;
  float ext = -1.f;
  for (int c = 0; c < d; ++c) {
    const unsigned int lo = mm[((int64_t)node * kDim + c) * 2], hi = mm[((int64_t)node * kDim + c) * 2 + 1];
    const float e = hi >= lo ? ord2f(hi) - ord2f(lo) : 0.f;
    if (e > ext) {
      ext = e;
      best = c;
    }
  }
  const uint32_t row = order[p];
  keys[p] = ((uint64_t)node << 32) | f2ord((float)X[(int64_t)row * ld + best]);
  vals[p] = row;
}

__global__ void iota_u32_kernel(uint32_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

// centre[w][c] = mean over the workgroup's queries of X[row][c] (FP64 sum, fixed order)
__global__ void centre_kernel(const double* X, int64_t n, int d, int64_t ld, const uint32_t* order,
                              float* centre) {
  const int64_t w = blockIdx.x;
  const int c = threadIdx.x;  // 32 threads
  double s = 0.0;
  int cnt = 0;
  for (int q = 0; q < kWgQueries; ++q) {
    const int64_t qi = w * kWgQueries + q;
    if (qi >= n) break;
    if (c < d) s += X[(int64_t)order[qi] * ld + c];
    ++cnt;
  }
  centre[w * kDim + c] = (c < d && cnt > 0) ? (float)(s / cnt) : 0.f;
}

// xs[sp][c] = X[order[sp]][c] (c < d, zero elsewhere): the rows in sorted order
__global__ void permute_kernel(const double* X, int64_t n, int d, int64_t ld, const uint32_t* order,
                               int64_t n_rows_pad, double* xs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t sp = t >> 5;
  const int c = (int)(t & 31);
  if (sp >= n_rows_pad) return;
  xs[t] = (sp < n && c < d) ? X[(int64_t)order[sp] * ld + c] : 0.0;
}

// radius[w] >= max over the chunk's points of |x - centre[w]| (FP64, rounded up)
__global__ void __launch_bounds__(kChunk) radius_kernel(const double* xs, int64_t n, const float* centre,
                                                        double* radius) {
  __shared__ double red[kChunk / 64];
  const int64_t w = blockIdx.x;
  const int64_t sp = w * kChunk + threadIdx.x;
  double d2 = 0.0;
  if (sp < n) {
    for (int c = 0; c < kDim; ++c) {
      const double v = xs[sp * kDim + c] - (double)centre[w * kDim + c];
      d2 += v * v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < kChunk / 64; ++i) m = fmax(m, red[i]);
    radius[w] = sqrt(m) * (1.0 + 1e-12) + 1e-300;
  }
}

struct CentredParams {
  const double* xs;      // FP64 rows in sorted order, 32 wide
  const float* centre;   // [n_wg][32]
  const double* radius;  // [n_wg]
  const uint32_t* order; // sorted position -> row
  int64_t n, n_wg;
  int d;
  int prune;             // skip chunks that cannot hold a candidate
  int32_t* cand;         // [n][kCand] by sorted position
  float* tau;            // [n][4]
  float* qnorm_c;        // [n] |q - centre|^2 by sorted position
  unsigned long long* visited;  // chunks processed (statistics)
};

// Visiting order of the database chunks for workgroup w: w, w+1, w-1, w+2,
// w-2, ... (near chunks first in the spatial sort); -1 past either end.
__device__ __forceinline__ int64_t visit_chunk(int64_t v, int64_t w, int64_t nwg) {
  const int64_t half = (v + 1) >> 1;
  const int64_t k = (v & 1) ? w + half : w - half;
  return (k >= 0 && k < nwg) ? k : -1;
}

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = to_bf16(x);
  const float r1 = x - (float)h;
  m = to_bf16(r1);
  l = to_bf16(r1 - (float)m);
}

// Workgroup = 8 waves = 128 consecutive sorted queries sharing one centre.
// Database chunks (128 consecutive sorted points, so chunk w holds the
// workgroup's own queries) are visited nearest-index first. Before each batch
// of 64 visits the workgroup takes T = the largest, over its queries, of the
// smallest stripe threshold (approximate d^2 of the 32nd best); a chunk k is
// skipped when (|c_w - c_k| - r_w - r_k)^2 > T, a lower bound of the exact d^2
// between any of its points and any of the queries (triangle inequality on the
// bounding spheres). Thresholds only shrink, so every skipped point has exact
// d^2 > the final threshold: the certificate of the finalize step holds for it
// as for a point that was seen and not kept. Processed chunks are shifted by
// the centre, given their augmented coordinate |x - c|^2, split into three
// bf16 terms and staged in LDS; every wave then runs its 16 x 16 MFMA tiles.
// Accepted candidates of the candidate kernels wait in a per-lane queue of kQueue
// entries and enter the sorted list together. list_insert runs for the whole
// wave whenever one lane inserts, so inserting from the queues costs one pass
// per queued entry of the fullest lane instead of one pass per (tile, value) at
// which any of the 64 lanes accepted. While entries wait, the acceptance
// threshold is the list's older (lower) minimum, which only lets a few more
// candidates in; list_insert keeps the best kList whatever the order, and a
// point turned away was below a minimum the final list exceeds.
constexpr int kQueue = 12;

struct CandQueue {
  float v[kQueue];
  int32_t p[kQueue];
  int n;
};

__device__ __forceinline__ void queue_flush(CandQueue& q, float (&L)[kList], int32_t (&I)[kList]) {
#pragma unroll
  for (int t = 0; t < kQueue; ++t) {
    const bool has = t < q.n;
    if (__ballot(has)) {
      if (has) list_insert(L, I, q.v[t], q.p[t]);
    }
  }
  q.n = 0;
}

// Room for the next `room` appends: flush when some lane could overflow.
__device__ __forceinline__ void queue_reserve(CandQueue& q, int room, float (&L)[kList], int32_t (&I)[kList]) {
  if (__ballot(q.n > kQueue - room)) queue_flush(q, L, I);
}

__device__ __forceinline__ void queue_append(CandQueue& q, bool ins, float s, int32_t p) {
  if (ins) {
#pragma unroll
    for (int t = 0; t < kQueue; ++t) {
      if (q.n == t) {
        q.v[t] = s;
        q.p[t] = p;
      }
    }
    ++q.n;
  }
}

__global__ void __launch_bounds__(kWgWaves * 64) cand_centred_kernel(CentredParams P) {
  __shared__ __align__(16) __bf16 As[3][kChunk][kDim];
  __shared__ float cen[kDim];
  __shared__ float red[kWgWaves];
  __shared__ unsigned long long vmask;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int64_t wg = blockIdx.x;
  if (tid < kDim) cen[tid] = P.centre[wg * kDim + tid];
  __syncthreads();
  const int64_t qi = wg * kWgQueries + wv * 16 + j;  // sorted position
  const bool qvalid = qi < P.n;
  // query operand: (q - c, -1/2 at dim 30), split in three bf16 terms
  bf16x8 bq_h, bq_m, bq_l;
  float part = 0.f;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int c = g * 8 + e;
    float v = 0.f;
    // centre in FP64, then round: the rounding error is relative to the shifted value
    if (qvalid && c < P.d) v = (float)(P.xs[qi * kDim + c] - (double)cen[c]);
    if (c == kMaxData) v = -0.5f;
    if (c < P.d) part = fmaf(v, v, part);
    __bf16 h, m, l;
    split3(v, h, m, l);
    bq_h[e] = h;
    bq_m[e] = m;
    bq_l[e] = l;
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  const float nqc = part;  // |q - c|^2 (fp32)

  float L[kList];
  int32_t I[kList];
#pragma unroll
  for (int r = 0; r < kList; ++r) {
    L[r] = -INFINITY;
    I[r] = -1;
  }
  CandQueue Q;
  Q.n = 0;
  const double rw = P.radius[wg];
  unsigned long long visited = 0;
  // staging roles: thread t -> point t/4, dims (t%4)*8 .. +7
  const int sp = tid >> 2, sq = tid & 3;
  // the first batch is the workgroup's own chunk alone, so that the thresholds
  // are finite when the other chunks are tested
  for (int64_t v0 = 0, bs = 1; v0 < 2 * P.n_wg; v0 += bs, bs = 64) {
    // T: the weakest threshold any query of the workgroup still accepts (the
    // lists made exact first)
    queue_flush(Q, L, I);
    float tq = INFINITY;
    {
      float t = nqc - 2.0f * L[kList - 1];
      t = fminf(t, __shfl_xor(t, 16, 64));
      t = fminf(t, __shfl_xor(t, 32, 64));
      if (!qvalid) t = -INFINITY;
      for (int o = 1; o < 16; o <<= 1) t = fmaxf(t, __shfl_xor(t, o, 64));
      tq = t;
    }
    __syncthreads();  // previous batch done with vmask and As
    if (lane == 0) red[wv] = tq;
    __syncthreads();
    if (wv == 0) {
      float T = red[0];
      for (int i = 1; i < kWgWaves; ++i) T = fmaxf(T, red[i]);
      const int64_t k = lane < bs ? visit_chunk(v0 + lane, wg, P.n_wg) : -1;
      bool take = k >= 0;
      if (take && P.prune && k != wg && T < INFINITY) {
        double d2 = 0.0;
        for (int c = 0; c < kDim; ++c) {
          const double dc = (double)cen[c] - (double)P.centre[k * kDim + c];
          d2 += dc * dc;
        }
        const double lb = sqrt(d2) * (1.0 - 1e-12) - rw - P.radius[k];
        take = !(lb > 0.0 && lb * lb > (double)T);
      }
      const unsigned long long m = __ballot(take);
      if (lane == 0) vmask = m;
    }
    __syncthreads();
    unsigned long long m = vmask;
    // FP64 coordinates of this thread's point and 8 dimensions of the next chunk
    // to stage, loaded while the current chunk's tiles run (xs is zero-padded
    // to whole chunks, so the loads need no bounds test)
    double xr[8];
#define SG_PREFETCH(mask)                                                              \
  {                                                                                    \
    const int64_t cp = visit_chunk(v0 + __ffsll(mask) - 1, wg, P.n_wg) * kChunk;      \
    const double2* src = reinterpret_cast<const double2*>(P.xs + (cp + sp) * kDim + sq * 8); \
    _Pragma("unroll") for (int e = 0; e < 4; ++e) {                                    \
      const double2 t2 = src[e];                                                       \
      xr[2 * e] = t2.x;                                                                \
      xr[2 * e + 1] = t2.y;                                                            \
    }                                                                                  \
  }
    if (m != 0) SG_PREFETCH(m)
    while (m != 0) {
      const int bit = __ffsll(m) - 1;
      m &= m - 1;
      const int64_t c0 = visit_chunk(v0 + bit, wg, P.n_wg) * kChunk;
      ++visited;
      __syncthreads();
      {
        const int64_t p = c0 + sp;
        float v[8];
        float ps = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int c = sq * 8 + e;
          v[e] = (p < P.n && c < P.d) ? (float)(xr[e] - (double)cen[c]) : 0.f;
          ps = fmaf(v[e], v[e], ps);
        }
        ps += __shfl_xor(ps, 1, 64);
        ps += __shfl_xor(ps, 2, 64);
        if (sq == 3) v[6] = ps;  // dim 30 = |x - c|^2
        bf16x8 th, tm, tl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          __bf16 h, mm, l;
          split3(v[e], h, mm, l);
          th[e] = h;
          tm[e] = mm;
          tl[e] = l;
        }
        *reinterpret_cast<bf16x8*>(&As[0][sp][sq * 8]) = th;
        *reinterpret_cast<bf16x8*>(&As[1][sp][sq * 8]) = tm;
        *reinterpret_cast<bf16x8*>(&As[2][sp][sq * 8]) = tl;
      }
      __syncthreads();
      if (m != 0) SG_PREFETCH(m)
#pragma unroll 2
      for (int tt = 0; tt < kChunk / 16; ++tt) {
        const int row = tt * 16 + j;
        const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&As[0][row][g * 8]);
        const bf16x8 am = *reinterpret_cast<const bf16x8*>(&As[1][row][g * 8]);
        const bf16x8 al = *reinterpret_cast<const bf16x8*>(&As[2][row][g * 8]);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bq_h, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bq_m, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_l, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bq_h, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_m, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_h, acc, 0, 0, 0);
        queue_reserve(Q, 4, L, I);
        const float th = L[kList - 1];
        const int64_t pbase = c0 + tt * 16 + g * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float sv = acc[u];
          const int64_t p = pbase + u;
          queue_append(Q, sv > th && p < P.n, sv, (int32_t)p);
        }
      }
    }
  }
#undef SG_PREFETCH
  queue_flush(Q, L, I);
  if (tid == 0 && P.visited) atomicAdd(P.visited, visited);
  if (!qvalid) return;
  int32_t* out = P.cand + qi * kCand + g * kList;
#pragma unroll
  for (int r = 0; r < kList; ++r) out[r] = I[r] >= 0 ? (int32_t)P.order[I[r]] : -1;
  P.tau[qi * 4 + g] = (L[kList - 1] == -INFINITY) ? INFINITY : nqc - 2.0f * L[kList - 1];
  if (g == 0) P.qnorm_c[qi] = nqc;
}

__global__ void __launch_bounds__(256) cand_kernel(CandParams P) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int j = lane & 15;   // query within the wave
  const int g = lane >> 4;   // stripe / k-group
  const int64_t qi = wave * 16 + j;  // index into the query list
  const bool qvalid = qi < P.n_rows;
  const int64_t q = qvalid ? (P.rows ? (int64_t)P.rows[qi] : qi) : 0;
  if (wave * 16 >= P.n_rows) return;

  float L[kList];
  int32_t I[kList];
#pragma unroll
  for (int r = 0; r < kList; ++r) {
    L[r] = -INFINITY;
    I[r] = -1;
  }
  const int64_t ntiles = P.n_pad / 16;
  CandQueue Q;
  Q.n = 0;

  {
    float qv[kDim];
#pragma unroll
    for (int c = 0; c < kDim; ++c) qv[c] = P.qf[q * kDim + c];
    for (int64_t t = 0; t < ntiles; ++t) {
      queue_reserve(Q, 4, L, I);
      const float th = L[kList - 1];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t p = t * 16 + g * 4 + v;
        const float4* xr = reinterpret_cast<const float4*>(P.xf + p * kDim);
        float s = 0.f;
#pragma unroll
        for (int c4 = 0; c4 < kDim / 4; ++c4) {
          const float4 x4 = xr[c4];
          s = fmaf(x4.x, qv[c4 * 4 + 0], s);
          s = fmaf(x4.y, qv[c4 * 4 + 1], s);
          s = fmaf(x4.z, qv[c4 * 4 + 2], s);
          s = fmaf(x4.w, qv[c4 * 4 + 3], s);
        }
        queue_append(Q, s > th && p < P.n, s, (int32_t)p);
      }
    }
  }
  queue_flush(Q, L, I);
  if (!qvalid) return;
  int32_t* out = P.cand + qi * kCand + g * kList;
#pragma unroll
  for (int r = 0; r < kList; ++r) out[r] = I[r];
  // stripe bound on the approximate d^2 of every point this lane did not keep
  const float nq = P.qnorm[q];
  P.tau[qi * 4 + g] = (L[kList - 1] == -INFINITY) ? INFINITY : nq - 2.0f * L[kList - 1];
}

// ---------------------------------------------------------------------------
// 3. finalize (exact FP64 rescoring + certificate)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ __forceinline__ double exact_d2(const double* a, const double* b, int d) {
  double acc = 0.0;
  for (int i = 0; i < d; ++i) {
    const double t = a[i] - b[i];
    acc = acc + t * t;
  }
  return acc;
}

__device__ __forceinline__ bool key_less(double da, int32_t ia, double db, int32_t ib) {
  return da < db || (da == db && ia < ib);
}

// Bitonic sort of 128 (d2, idx) pairs held as slot 0 (elements 0..63) and slot 1.
__device__ __forceinline__ void sort128(double (&dd)[2], int32_t (&ii)[2], int lane) {
#pragma unroll
  for (int size = 2; size <= 128; size <<= 1) {
#pragma unroll
    for (int jj = size >> 1; jj > 0; jj >>= 1) {
      if (jj == 64) {
        const bool sw = key_less(dd[1], ii[1], dd[0], ii[0]);
        const double td = dd[0];
        const int32_t ti = ii[0];
        dd[0] = sw ? dd[1] : dd[0];
        ii[0] = sw ? ii[1] : ii[0];
        dd[1] = sw ? td : dd[1];
        ii[1] = sw ? ti : ii[1];
      } else {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const int e = r * 64 + lane;
          const double od = __shfl_xor(dd[r], jj, 64);
          const int32_t oi = __shfl_xor(ii[r], jj, 64);
          const bool asc = (e & size) == 0;
          const bool low = (e & jj) == 0;
          const bool mine_less = key_less(dd[r], ii[r], od, oi);
          const bool keep_mine = (low == asc) ? mine_less : !mine_less;
          dd[r] = keep_mine ? dd[r] : od;
          ii[r] = keep_mine ? ii[r] : oi;
        }
      }
    }
  }
}

struct FinParams {
  const double* X;
  int64_t n, ld;
  int d, k;
  const int32_t* rows;  // null = identity
  int64_t n_rows;
  const int32_t* cand;
  const float* tau;
  const float* qnorm;
  const float* qnorm_qi;  // if set: |q - centre|^2 by list position (centred BF16 path)
  const unsigned int* max_norm_bits;
  double eps_c;
  int force_fallback;
  int32_t* nbr;     // [n][k]
  double* dist;     // [n][k]
  int32_t* fail_rows;
  unsigned int* fail_count;
};

// Writes the k best (ascending (d2, idx)) entries, re-sorted by neighbour index.
__device__ void write_row(const FinParams& P, int64_t q, double (&dd)[2], int32_t (&ii)[2], int lane,
                          double* lds_d, int32_t* lds_i) {
  // elements 0..k-1 are the answer; place them by index rank
  const int k = P.k;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int e = r * 64 + lane;
    if (e < k) {
      lds_d[e] = dd[r];
      lds_i[e] = ii[r];
    }
  }
  wave_sync();
  for (int e = lane; e < k; e += 64) {
    const int32_t me = lds_i[e];
    int rank = 0;
    for (int o = 0; o < k; ++o) rank += lds_i[o] < me;
    P.nbr[q * k + rank] = me;
    P.dist[q * k + rank] = sqrt(lds_d[e]);
  }
  wave_sync();
}

// Smallest exact d^2 a point not kept by a stripe with threshold tau can have.
// Such a point x has approx d^2 >= tau, so exact d^2 >= tau - c(t r + t^2 + r^2)
// with t = |x|, r = |q|, and also exact d^2 >= (t - r)^2. Minimising the larger
// of the two over t in [0, M]: below r both fall, so t = r gives tau - 3c r^2;
// above r the first falls and the second rises, so the minimum is where they
// cross (larger root of (1+c)t^2 + (c-2) r t + (1+c) r^2 - tau = 0), capped at M.
__device__ double nonkept_lower_bound(double tau, double nq, double M, double c) {
  const double r = sqrt(nq);
  double best = tau - 3.0 * c * nq;  // t <= r
  const double a = 1.0 + c, bq = (c - 2.0) * r, cq = (1.0 + c) * nq - tau;
  const double disc = bq * bq - 4.0 * a * cq;
  double t = r;
  if (disc >= 0.0) t = (-bq + sqrt(disc)) / (2.0 * a);
  if (t < r) t = r;
  double cross;
  if (t > M) {
    cross = tau - c * (M * r + M * M + nq);  // beyond the largest norm: first bound at M
  } else {
    cross = (t - r) * (t - r);
  }
  best = fmin(best, cross);
  // guard against rounding in the bound itself
  const double mm = M < INFINITY ? M * M : 0.0;  // no cap when the norms are unbounded
  return best - 1e-7 * fabs(tau) - 4.0 * DBL_EPSILON * (nq + mm);
}

__global__ void __launch_bounds__(256) finalize_kernel(FinParams P) {
  __shared__ double lds_d[4][128];
  __shared__ int32_t lds_i[4][128];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t qi = (int64_t)blockIdx.x * 4 + wid;
  if (qi >= P.n_rows) return;
  const int64_t q = P.rows ? (int64_t)P.rows[qi] : qi;
  const double* xq = P.X + q * P.ld;
  double dd[2];
  int32_t ii[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int32_t c = P.cand[qi * kCand + r * 64 + lane];
    if (c < 0 || c == q) {
      dd[r] = INFINITY;
      ii[r] = INT32_MAX;
    } else {
      dd[r] = exact_d2(xq, P.X + (int64_t)c * P.ld, P.d);
      ii[r] = c;
    }
  }
  sort128(dd, ii, lane);
  // number of real candidates and the k-th exact distance
  const int nvalid = popc64(__ballot(ii[0] != INT32_MAX)) + popc64(__ballot(ii[1] != INT32_MAX));
  const int k = P.k;
  const double dk = __shfl(k - 1 < 64 ? dd[0] : dd[1], (k - 1) & 63, 64);
  float tmin = P.tau[qi * 4 + 0];
  for (int g = 1; g < 4; ++g) tmin = fminf(tmin, P.tau[qi * 4 + g]);
  const double M = (double)__uint_as_float(*P.max_norm_bits);
  const double nq = P.qnorm_qi ? (double)P.qnorm_qi[qi] : (double)P.qnorm[q];
  const double lb = nonkept_lower_bound(tmin, nq, P.qnorm_qi ? INFINITY : M, P.eps_c);
  bool ok = nvalid >= k && (tmin == INFINITY || lb > dk);
  if (P.force_fallback) ok = false;
  if (!ok) {
    if (lane == 0) {
      const unsigned int slot = atomicAdd(P.fail_count, 1u);
      P.fail_rows[slot] = (int32_t)q;
    }
    return;
  }
  write_row(P, q, dd, ii, lane, lds_d[wid], lds_i[wid]);
}

// ---------------------------------------------------------------------------
// Exact FP64 fallback for rows whose certificate failed: one wave per row, each
// lane keeps its best k (lexicographic (d2, idx)) in LDS, then a k-round merge.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) fallback_kernel(FinParams P, const int32_t* rows,
                                                      const unsigned int* count) {
  __shared__ double ld_d[kMaxK][64];
  __shared__ int32_t ld_i[kMaxK][64];
  __shared__ double out_d[128];
  __shared__ int32_t out_i[128];
  const int lane = threadIdx.x;
  const int k = P.k;
  const unsigned int nfail = *count;
  for (unsigned int f = blockIdx.x; f < nfail; f += gridDim.x) {
    const int64_t q = rows[f];
    const double* xq = P.X + q * P.ld;
    int cnt = 0;
    for (int r = 0; r < k; ++r) {
      ld_d[r][lane] = INFINITY;
      ld_i[r][lane] = INT32_MAX;
    }
    for (int64_t p = lane; p < P.n; p += 64) {
      if (p == q) continue;
      const double d2 = exact_d2(xq, P.X + p * P.ld, P.d);
      const int32_t pi = (int32_t)p;
      if (!key_less(d2, pi, ld_d[k - 1][lane], ld_i[k - 1][lane])) continue;
      int r = k - 1;
      while (r > 0 && key_less(d2, pi, ld_d[r - 1][lane], ld_i[r - 1][lane])) {
        ld_d[r][lane] = ld_d[r - 1][lane];
        ld_i[r][lane] = ld_i[r - 1][lane];
        --r;
      }
      ld_d[r][lane] = d2;
      ld_i[r][lane] = pi;
      ++cnt;
    }
    (void)cnt;
    // merge: k rounds of a wave-wide lexicographic argmin over the list heads
    int head = 0;
    for (int r = 0; r < k; ++r) {
      double hd = head < k ? ld_d[head][lane] : INFINITY;
      int32_t hi = head < k ? ld_i[head][lane] : INT32_MAX;
      double bd = hd;
      int32_t bi = hi;
      int bl = lane;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double od = __shfl_xor(bd, o, 64);
        const int32_t oi = __shfl_xor(bi, o, 64);
        const int ol = __shfl_xor(bl, o, 64);
        if (key_less(od, oi, bd, bi)) {
          bd = od;
          bi = oi;
          bl = ol;
        }
      }
      if (lane == bl) ++head;
      if (lane == 0) {
        out_d[r] = bd;
        out_i[r] = bi;
      }
    }
    wave_sync();
    for (int e = lane; e < k; e += 64) {
      const int32_t me = out_i[e];
      int rank = 0;
      for (int o = 0; o < k; ++o) rank += out_i[o] < me;
      P.nbr[q * k + rank] = me;
      P.dist[q * k + rank] = sqrt(out_d[e]);
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Second tier for the rows whose certificate failed: exact FP64 distances, with
// the database split into S segments (grid.y) so that a handful of rows still
// fills the GPU. Each (16 queries, segment) wave keeps 4 x 32 candidates; the
// keys are -fl32(d^2).
// ---------------------------------------------------------------------------
struct SegParams {
  const double* X;
  int64_t ld, n;
  int d;
  const int32_t* rows;
  int64_t n_rows;
  int64_t seg_size;  // multiple of 16
  int S;
  int32_t* cand;  // [n_rows][S][kCand]
  float* tau;     // [n_rows][S][4]  (d^2 bound per stripe)
};

// Each wave streams its segment in tiles of 16 points: the 16 x 32 FP64 tile is
// loaded coalesced (8 values per lane, the next tile's loaded while the current
// one is scored) and shared through the wave's LDS slice; lane (query j,
// stripe g) scores points 4g .. 4g + 3 of the tile.
__global__ void __launch_bounds__(256) exact_seg_kernel(SegParams P) {
  __shared__ double tile[4][16][kDim + 1];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int seg = blockIdx.y;
  const int j = lane & 15, g = lane >> 4;
  const int64_t qi = wave * 16 + j;
  if (wave * 16 >= P.n_rows) return;
  const bool qvalid = qi < P.n_rows;
  const int64_t q = qvalid ? (int64_t)P.rows[qi] : 0;
  double qv[kMaxData];
#pragma unroll
  for (int c = 0; c < kMaxData; ++c) qv[c] = c < P.d ? P.X[q * P.ld + c] : 0.0;
  float L[kList];
  int32_t I[kList];
#pragma unroll
  for (int r = 0; r < kList; ++r) {
    L[r] = -INFINITY;
    I[r] = -1;
  }
  const int64_t p0 = (int64_t)seg * P.seg_size;
  const int64_t p1 = min(P.n, p0 + P.seg_size);
  // element k = lane + 64 i of a tile: point k / 32, dimension k % 32; clamped
  // reads (values of points past p1 or dimensions past d are never used)
  double xn[8];
  auto load_tile = [&](int64_t t0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = lane + 64 * i;
      const int64_t p = min(t0 + (k >> 5), P.n - 1);
      const int c = min(k & 31, P.d - 1);
      xn[i] = P.X[p * P.ld + c];
    }
  };
  if (p0 < p1) load_tile(p0);
  for (int64_t t0 = p0; t0 < p1; t0 += 16) {
    wave_sync();  // previous tile's reads are done
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = lane + 64 * i;
      tile[wv][k >> 5][k & 31] = xn[i];
    }
    wave_sync();
    if (t0 + 16 < p1) load_tile(t0 + 16);
    const float th = L[kList - 1];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int pt = g * 4 + v;
      const int64_t p = t0 + pt;
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < kMaxData; ++c) {
        if (c < P.d) {
          const double df = tile[wv][pt][c] - qv[c];
          acc = fma(df, df, acc);
        }
      }
      const float sv = p < p1 ? -(float)acc : -INFINITY;
      const bool ins = sv > th;
      if (__ballot(ins)) {
        if (ins) list_insert(L, I, sv, (int32_t)p);
      }
    }
  }
  if (!qvalid) return;
  const int64_t slot = qi * P.S + seg;
  int32_t* out = P.cand + slot * kCand + g * kList;
#pragma unroll
  for (int r = 0; r < kList; ++r) out[r] = I[r];
  P.tau[slot * 4 + g] = (L[kList - 1] == -INFINITY) ? INFINITY : -L[kList - 1];
}

// One 256-thread block per row: exact rescoring of the S x 128 candidates,
// bitonic sort by (d^2, index) in LDS, certificate, output or hand-off.
constexpr int kMaxSeg = 32;
__global__ void __launch_bounds__(256) finalize_multi_kernel(FinParams P, SegParams Q,
                                                             int32_t* fail_rows, unsigned int* fail_count) {
  __shared__ double sd[kMaxSeg * kCand];
  __shared__ int32_t si[kMaxSeg * kCand];
  __shared__ float stau[256];
  const int64_t qi = blockIdx.x;
  if (qi >= Q.n_rows) return;
  const int64_t q = Q.rows[qi];
  const int C = Q.S * kCand;
  int Cp = 1;
  while (Cp < C) Cp <<= 1;
  const double* xq = P.X + q * P.ld;
  for (int e = threadIdx.x; e < Cp; e += blockDim.x) {
    double dv = INFINITY;
    int32_t iv = INT32_MAX;
    if (e < C) {
      const int32_t c = Q.cand[qi * (int64_t)C + e];
      if (c >= 0 && c != q) {
        dv = exact_d2(xq, P.X + (int64_t)c * P.ld, P.d);
        iv = c;
      }
    }
    sd[e] = dv;
    si[e] = iv;
  }
  float tm = INFINITY;
  for (int e = threadIdx.x; e < Q.S * 4; e += blockDim.x) tm = fminf(tm, Q.tau[qi * (int64_t)Q.S * 4 + e]);
  stau[threadIdx.x] = tm;
  __syncthreads();
  for (int size = 2; size <= Cp; size <<= 1) {
    for (int jj = size >> 1; jj > 0; jj >>= 1) {
      for (int e = threadIdx.x; e < Cp; e += blockDim.x) {
        const int o = e ^ jj;
        if (o > e) {
          const bool asc = (e & size) == 0;
          const bool gt = key_less(sd[o], si[o], sd[e], si[e]);  // element e larger than o
          if (gt == asc) {
            const double td = sd[e];
            const int32_t ti = si[e];
            sd[e] = sd[o];
            si[e] = si[o];
            sd[o] = td;
            si[o] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    float t = INFINITY;
    for (int e = 0; e < (int)blockDim.x; ++e) t = fminf(t, stau[e]);
    stau[0] = t;
  }
  __syncthreads();
  const int k = P.k;
  const float tmin = stau[0];
  const bool enough = si[k - 1] != INT32_MAX;
  const double lb = (double)tmin * (1.0 - 0x1p-20) - 64.0 * DBL_MIN;
  const bool ok = enough && (tmin == INFINITY || lb > sd[k - 1]);
  if (!ok) {
    if (threadIdx.x == 0) {
      const unsigned int slot = atomicAdd(fail_count, 1u);
      fail_rows[slot] = (int32_t)q;
    }
    return;
  }
  for (int e = threadIdx.x; e < k; e += blockDim.x) {
    const int32_t me = si[e];
    int rank = 0;
    for (int o = 0; o < k; ++o) rank += si[o] < me;
    P.nbr[q * k + rank] = me;
    P.dist[q * k + rank] = sqrt(sd[e]);
  }
}

// ---------------------------------------------------------------------------
// 5. wide path (d > 30 or k > 100, beyond the candidate lists above): exact
// FP64 distances from a block of queries to every point, in the oracle's
// arithmetic (dimension order, no fused multiply-add), keyed by their bit
// patterns (non-negative doubles order like their bits; the query itself gets
// the +inf key), then one stable radix sort per query: equal distances keep
// index order, so the first k entries are the exact neighbours with ties broken
// by index. A second stable sort puts each row's k neighbours in index order.
// O(n^2 d) arithmetic; the points are read transposed, so a block's 256 lanes
// load consecutive points of one dimension, and each lane serves kWideQ queries.
// ---------------------------------------------------------------------------
constexpr int kWideQ = 8;                     // queries per block
constexpr int64_t kWideItems = 1ll << 27;     // (query, point) pairs per sort

__global__ void transpose_kernel(const double* __restrict__ X, int64_t n, int d, int64_t ld,
                                 double* __restrict__ Xt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int64_t c = t / n, p = t - c * n;
  Xt[t] = X[p * ld + c];
}

__global__ void __launch_bounds__(256) wide_dist_kernel(const double* __restrict__ X, const double* __restrict__ Xt,
                                                        int64_t n, int d, int64_t ld, int64_t q0, int64_t nq,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t qb = (int64_t)blockIdx.y * kWideQ;
  if (p >= n) return;
  double acc[kWideQ];
  const double* xq[kWideQ];
#pragma unroll
  for (int j = 0; j < kWideQ; ++j) {
    acc[j] = 0.0;
    xq[j] = X + (q0 + min(qb + j, nq - 1)) * ld;  // rows past the block repeat its last query
  }
  for (int c = 0; c < d; ++c) {
    const double x = Xt[(int64_t)c * n + p];
#pragma unroll
    for (int j = 0; j < kWideQ; ++j) {
      const double t = xq[j][c] - x;
      acc[j] = acc[j] + t * t;
    }
  }
#pragma unroll
  for (int j = 0; j < kWideQ; ++j) {
    const int64_t qi = qb + j;
    if (qi >= nq) break;
    const int64_t o = qi * n + p;
    keys[o] = (q0 + qi == p) ? 0x7FF0000000000000ull : (uint64_t)__double_as_longlong(acc[j]);
    vals[o] = (uint32_t)p;
  }
}

__global__ void iota_offsets_kernel(int64_t* off, int64_t count, int64_t step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= count) off[i] = i * step;
}

// the k nearest of each sorted row: (neighbour index, d^2 bits) for the second sort
__global__ void wide_take_kernel(const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ svals, int64_t n,
                                 int64_t nq, int k, uint32_t* __restrict__ idx, uint64_t* __restrict__ d2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq * k) return;
  const int64_t qi = t / k, r = t - qi * k;
  idx[t] = svals[qi * n + r];
  d2[t] = skeys[qi * n + r];
}

__global__ void wide_write_kernel(const uint32_t* __restrict__ idx, const uint64_t* __restrict__ d2, int64_t q0,
                                  int64_t nq, int k, int32_t* __restrict__ nbr, double* __restrict__ dist) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq * k) return;
  nbr[q0 * k + t] = (int32_t)idx[t];
  dist[q0 * k + t] = sqrt(__longlong_as_double((long long)d2[t]));
}

struct WideWs {
  double* xt;
  uint64_t *k0, *k1, *tk0, *tk1;
  uint32_t *v0, *v1, *ti0, *ti1;
  int64_t *off, *off_k;
  void* tmp;
  size_t tmp_bytes;
  int64_t block_q;
};

static int64_t wide_block_queries(int64_t n) {
  return std::max<int64_t>(1, std::min<int64_t>(kWideItems / n, n));
}

static int64_t wide_layout(int64_t n, int d, int k, WideWs* w, char* base) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  const int64_t bq = wide_block_queries(n);
  const int64_t items = bq * n, kitems = bq * k;
  WideWs t;
  t.block_q = bq;
  t.xt = (double*)take(n * d * 8);
  t.k0 = (uint64_t*)take(items * 8);
  t.k1 = (uint64_t*)take(items * 8);
  t.v0 = (uint32_t*)take(items * 4);
  t.v1 = (uint32_t*)take(items * 4);
  t.tk0 = (uint64_t*)take(kitems * 8);
  t.tk1 = (uint64_t*)take(kitems * 8);
  t.ti0 = (uint32_t*)take(kitems * 4);
  t.ti1 = (uint32_t*)take(kitems * 4);
  t.off = (int64_t*)take((bq + 1) * 8);
  t.off_k = (int64_t*)take((bq + 1) * 8);
  size_t a = 0, b = 0;
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                    (uint32_t*)nullptr, (uint32_t*)nullptr, (int)items, (int)bq,
                                                    (int64_t*)nullptr, (int64_t*)nullptr, 0, 63);
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (uint64_t*)nullptr, (uint64_t*)nullptr, (int)kitems, (int)bq,
                                                    (int64_t*)nullptr, (int64_t*)nullptr, 0, 32);
  t.tmp_bytes = std::max(a, b);
  t.tmp = take((int64_t)t.tmp_bytes + 256);
  if (w) *w = t;
  return off;
}

static bool wide_needed(int d, int k) { return d > kMaxData || k > kMaxK; }

static int wide_knn(const double* X, int64_t n, int d, int64_t ld, int k, int32_t* nbr, double* dist, char* base,
                    hipStream_t st) {
  WideWs w;
  wide_layout(n, d, k, &w, base);
  transpose_kernel<<<(unsigned)ceil_div(n * d, 256), 256, 0, st>>>(X, n, d, ld, w.xt);
  SG_LAUNCH_CHECK();
  for (int64_t q0 = 0; q0 < n; q0 += w.block_q) {
    const int64_t nq = std::min<int64_t>(w.block_q, n - q0);
    const dim3 grid((unsigned)ceil_div(n, 256), (unsigned)ceil_div(nq, kWideQ));
    wide_dist_kernel<<<grid, 256, 0, st>>>(X, w.xt, n, d, ld, q0, nq, w.k0, w.v0);
    iota_offsets_kernel<<<(unsigned)ceil_div(nq + 1, 256), 256, 0, st>>>(w.off, nq, n);
    iota_offsets_kernel<<<(unsigned)ceil_div(nq + 1, 256), 256, 0, st>>>(w.off_k, nq, k);
    SG_LAUNCH_CHECK();
    size_t tb = w.tmp_bytes;
    // keys below 2^63 (non-negative doubles and the +inf key): 63 bits
    SG_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(w.tmp, tb, w.k0, w.k1, w.v0, w.v1, (int)(nq * n), (int)nq,
                                                       w.off, w.off + 1, 0, 63, st));
    wide_take_kernel<<<(unsigned)ceil_div(nq * k, 256), 256, 0, st>>>(w.k1, w.v1, n, nq, k, w.ti0, w.tk0);
    SG_LAUNCH_CHECK();
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < n) ++bits;
    tb = w.tmp_bytes;
    SG_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(w.tmp, tb, w.ti0, w.ti1, w.tk0, w.tk1, (int)(nq * k),
                                                       (int)nq, w.off_k, w.off_k + 1, 0, bits, st));
    wide_write_kernel<<<(unsigned)ceil_div(nq * k, 256), 256, 0, st>>>(w.ti1, w.tk1, q0, nq, k, nbr, dist);
    SG_LAUNCH_CHECK();
  }
  return SNAPGPU_OK;
}

struct Workspace {
  float *xf, *qf, *qnorm, *tau;
  int32_t *cand, *fail_rows, *fail_rows2;
  unsigned int *max_norm_bits, *fail_count, *fail_count2;
  float *centre, *qnorm_c, *tau2;
  int32_t* cand2;
  uint64_t *keys, *keys2;
  uint32_t *order0, *order;
  unsigned int* mm;
  int64_t* kd_nodes;
  double *xs, *radius;
  unsigned long long* visited;
  void* sort_tmp;
  size_t sort_bytes;
};

static int64_t layout(int64_t n, Workspace* w, char* base) {
  const int64_t n_pad = ((n + 15) / 16) * 16;
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  Workspace t;
  t.xf = (float*)take(n_pad * kDim * 4);
  t.qf = (float*)take(n_pad * kDim * 4);
  t.qnorm = (float*)take(n_pad * 4);
  t.tau = (float*)take(n * 4 * 4);
  t.cand = (int32_t*)take(n * kCand * 4);
  t.fail_rows = (int32_t*)take(n * 4);
  t.fail_rows2 = (int32_t*)take(n * 4);
  t.max_norm_bits = (unsigned int*)take(4);
  t.fail_count = (unsigned int*)take(4);
  t.fail_count2 = (unsigned int*)take(4);
  t.centre = (float*)take(ceil_div(n, kWgQueries) * kDim * 4);
  const int64_t t2rows = std::max<int64_t>(n, 2048);  // rows x segments of the second tier
  t.cand2 = (int32_t*)take(t2rows * kCand * 4);
  t.tau2 = (float*)take(t2rows * 4 * 4);
  t.qnorm_c = (float*)take(n * 4);
  t.keys = (uint64_t*)take(n * 8);
  t.keys2 = (uint64_t*)take(n * 8);
  t.order0 = (uint32_t*)take(n * 4);
  t.order = (uint32_t*)take(n * 4);
  const int64_t nl = ceil_div(n, kChunk);
  t.mm = (unsigned int*)take(nl * kDim * 2 * 4);
  t.kd_nodes = (int64_t*)take((4 * nl + 64) * 2 * 8);
  t.xs = (double*)take(ceil_div(n, kChunk) * kChunk * kDim * 8);
  t.radius = (double*)take(ceil_div(n, kChunk) * 8);
  t.visited = (unsigned long long*)take(8);
  size_t sb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sb, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (int)std::min<int64_t>(n, INT32_MAX));
  t.sort_bytes = sb;
  t.sort_tmp = take((int64_t)sb + 256);
  if (w) *w = t;
  return off;
}

}  // namespace knn
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::knn;

extern "C" {

int64_t snapgpu_knn_workspace_bytes(int64_t n, int32_t d, int32_t k) {
  if (n < 0) return 0;
  if (wide_needed(d, k)) return wide_layout(n, d, std::max(k, 1), nullptr, nullptr) + 256;
  return layout(n, nullptr, nullptr) + 256;
}

int snapgpu_knn(const double* d_X, int64_t n, int32_t d, int64_t ld, int32_t k, int32_t mode,
                int32_t* d_nbr, double* d_dist, void* d_workspace, int64_t workspace_bytes,
                int32_t flags, int64_t* n_fallback, int64_t* n_exact, int64_t* n_chunk_visits,
                void* stream) {
  clear_error();
  SG_ARG(n >= 2, "need at least 2 points");
  SG_ARG(n < INT32_MAX, "too many points");
  SG_ARG(d >= 1, "dimension must be >= 1");
  SG_ARG(ld >= d, "ld < d");
  SG_ARG(k >= 1 && k <= n - 1, "k must be in [1, n-1]");
  SG_ARG(mode == kModeFP32 || mode == kModeBF16, "mode must be 0 (fp32) or 1 (bf16)");
  SG_ARG(d_X && d_nbr && d_dist && d_workspace, "null pointer");
  SG_ARG(workspace_bytes >= snapgpu_knn_workspace_bytes(n, d, k), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* base = (char*)(((uintptr_t)d_workspace + 255) & ~(uintptr_t)255);
  if (wide_needed(d, k)) {
    const int rc = wide_knn(d_X, n, d, ld, k, d_nbr, d_dist, base, st);
    if (rc) return rc;
    if (n_fallback) *n_fallback = 0;
    if (n_exact) *n_exact = n;
    if (n_chunk_visits) *n_chunk_visits = 0;
    return SNAPGPU_OK;
  }
  Workspace w;
  layout(n, &w, base);
  const int64_t n_pad = ((n + 15) / 16) * 16;
  SG_HIP(hipMemsetAsync(w.max_norm_bits, 0, 4, st));
  SG_HIP(hipMemsetAsync(w.fail_count, 0, 4, st));
  prepare_kernel<<<(unsigned)ceil_div(n_pad * 32, 256), 256, 0, st>>>(
      d_X, n, d, ld, n_pad, w.xf, w.qf, w.qnorm, w.max_norm_bits);
  SG_LAUNCH_CHECK();
  CandParams cp{w.xf, w.qf, w.qnorm, n, n_pad, nullptr, n, w.cand, w.tau};
  FinParams fp{d_X, n, ld, d, k, nullptr, n, w.cand, w.tau, w.qnorm, nullptr, w.max_norm_bits,
               mode == kModeBF16 ? kEpsBF16 : kEpsFP32, (flags & 1) ? 1 : 0,
               d_nbr, d_dist, w.fail_rows, w.fail_count};
  if (mode == kModeBF16) {
    // spatial order of the queries, per-workgroup centres, centred LDS-staged MFMA tiles
    // kd-tree order: leaves of 128 consecutive positions, split along the widest dimension
    const int64_t nl = ceil_div(n, kChunk);
    int levels = 0;
    while (((int64_t)1 << levels) < nl) ++levels;
    SG_ARG(levels <= kMaxLevels, "too many points for the kd-tree order");
    std::vector<int64_t> tab;
    std::vector<int64_t> level_off(levels + 1, 0);
    for (int l = 0; l < levels; ++l) {
      level_off[l] = (int64_t)tab.size() / 2;
      for (int64_t id = 0; id < ((int64_t)1 << l); ++id) {
        int64_t a = 0, b = nl;
        for (int q = l - 1; q >= 0; --q) {
          const int64_t m = a + (b - a + 1) / 2;
          if ((id >> q) & 1) a = m; else b = m;
        }
        tab.push_back(std::min(a * kChunk, n));
        tab.push_back(std::min(b * kChunk, n));
      }
    }
    if (!tab.empty())
      SG_HIP(hipMemcpyAsync(w.kd_nodes, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st));
    iota_u32_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(w.order, n);
    for (int l = 0; l < levels; ++l) {
      const int64_t nodes = (int64_t)1 << l;
      const int64_t max_size = ceil_div(nl, nodes) * kChunk;
      const int slices = (int)std::max<int64_t>(1, ceil_div(max_size, kSpreadSlice));
      kd_mm_init_kernel<<<(unsigned)ceil_div(nodes * kDim * 2, 256), 256, 0, st>>>(w.mm, nodes * kDim * 2);
      kd_extent_kernel<<<(unsigned)(nodes * slices), 256, 0, st>>>(d_X, d, ld, w.order,
                                                                  w.kd_nodes + 2 * level_off[l], slices, w.mm);
      kd_key_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(d_X, n, d, ld, w.order, nl, l, w.mm, w.keys,
                                                               w.order0);
      SG_LAUNCH_CHECK();
      size_t sb = w.sort_bytes;
      SG_HIP(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, sb, w.keys, w.keys2, w.order0, w.order, (int)n, 0,
                                                32 + l, st));
    }
    const int64_t n_wg = ceil_div(n, kWgQueries);
    centre_kernel<<<(unsigned)n_wg, 32, 0, st>>>(d_X, n, d, ld, w.order, w.centre);
    permute_kernel<<<(unsigned)ceil_div(n_wg * kChunk * kDim, 256), 256, 0, st>>>(d_X, n, d, ld, w.order,
                                                                                n_wg * kChunk, w.xs);
    radius_kernel<<<(unsigned)n_wg, kChunk, 0, st>>>(w.xs, n, w.centre, w.radius);
    SG_HIP(hipMemsetAsync(w.visited, 0, 8, st));
    CentredParams cc{w.xs, w.centre, w.radius, w.order, n, n_wg, d, (flags & 2) ? 0 : 1,
                     w.cand, w.tau, w.qnorm_c, w.visited};
    cand_centred_kernel<<<(unsigned)n_wg, kWgWaves * 64, 0, st>>>(cc);
    SG_LAUNCH_CHECK();
    fp.rows = reinterpret_cast<const int32_t*>(w.order);
    fp.qnorm_qi = w.qnorm_c;
  } else {
    const unsigned cblocks = (unsigned)ceil_div(ceil_div(n, 16), 4);
    cand_kernel<<<cblocks, 256, 0, st>>>(cp);
    SG_LAUNCH_CHECK();
  }
  finalize_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, st>>>(fp);
  SG_LAUNCH_CHECK();
  unsigned int nfail = 0, nfail2 = 0;
  SG_HIP(hipMemcpyAsync(&nfail, w.fail_count, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (nfail > 0 && !(flags & 1)) {
    // second tier: exact FP64 distances for the failed rows (database split into
    // S segments so that few rows still fill the GPU), then the certificate again
    SG_HIP(hipMemsetAsync(w.fail_count2, 0, 4, st));
    const int64_t qwaves = ceil_div((int64_t)nfail, 16);
    int S = (int)std::min<int64_t>(kMaxSeg, std::max<int64_t>(1, ceil_div(2048, qwaves)));
    while (S > 1 && (int64_t)nfail * S > std::max<int64_t>(n, 2048)) --S;
    const int64_t seg = ceil_div(ceil_div(n, S), 16) * 16;
    SegParams sp{d_X, ld, n, d, w.fail_rows, (int64_t)nfail, seg, S, w.cand2, w.tau2};
    exact_seg_kernel<<<dim3((unsigned)ceil_div(qwaves, 4), (unsigned)S), 256, 0, st>>>(sp);
    SG_LAUNCH_CHECK();
    finalize_multi_kernel<<<(unsigned)nfail, 256, 0, st>>>(fp, sp, w.fail_rows2, w.fail_count2);
    SG_LAUNCH_CHECK();
    SG_HIP(hipMemcpyAsync(&nfail2, w.fail_count2, 4, hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
    if (nfail2 > 0) {
      fallback_kernel<<<(unsigned)std::min<int64_t>(nfail2, 65535), 64, 0, st>>>(fp, w.fail_rows2,
                                                                                 w.fail_count2);
      SG_LAUNCH_CHECK();
    }
  } else if (nfail > 0) {
    fallback_kernel<<<(unsigned)std::min<int64_t>(nfail, 65535), 64, 0, st>>>(fp, w.fail_rows,
                                                                              w.fail_count);
    SG_LAUNCH_CHECK();
  }
  if (n_chunk_visits) {
    unsigned long long nv = 0;
    if (mode == kModeBF16) {
      SG_HIP(hipMemcpyAsync(&nv, w.visited, 8, hipMemcpyDeviceToHost, st));
      SG_HIP(hipStreamSynchronize(st));
    }
    *n_chunk_visits = (int64_t)nv;
  }
  if (n_fallback) *n_fallback = (int64_t)nfail;
  if (n_exact) *n_exact = (int64_t)(flags & 1 ? nfail : nfail2);
  return SNAPGPU_OK;
}

}  // extern "C"
