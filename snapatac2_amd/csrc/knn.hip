// Path C: exact k-nearest-neighbour graph (Euclidean, self excluded by index).
//
// Reference: snapatac2-core/src/utils/knn.rs:9-32 (kd-tree; neighbours in
// increasing squared Euclidean distance, index == i skipped, k taken, distance =
// sqrt) and :58-78 (each row sorted by neighbour index); binding
// snapatac2-python/src/knn.rs:7-16.
//
// Pipeline (all on the device):
//  1. prepare: FP64 rows -> augmented 32-wide rows x~ = (x, |x|^2, 0) for the
//     database and q~ = (x, -1/2, 0) for the queries, so s = x~.q~ = x.q - |x|^2/2
//     and d^2 = |q|^2 - 2 s. Stored as three bf16 terms (x = hi + mid + lo +
//     O(2^-24 x)) and as fp32.
//  2. candidates: a wave owns 16 queries and streams every database point in
//     tiles of 16. BF16 mode: six v_mfma_f32_16x16x32_bf16 per tile (every
//     product of the split terms down to order 2^-24, fp32 accumulation), so the
//     approximation error is at fp32 level even where d^2 is much smaller than
//     |q|^2 (spectral embeddings); FP32 mode: SIMT FMAs. Output
//     lane l holds query l%16 and 4 points of the tile, so the 4 lanes of a query
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
#include <cfloat>

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

enum Mode : int32_t { kModeFP32 = 0, kModeBF16 = 1, kModeExact = 2 };

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
                                                      int d, int64_t ld, int64_t n_pad,
                                                      __bf16* xh, __bf16* xm, __bf16* xl, __bf16* qh,
                                                      __bf16* qm, __bf16* ql, float* xf, float* qf,
                                                      float* qnorm, unsigned int* max_norm_bits) {
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
  const float xv32 = (float)xv, qv32 = (float)qv;
  // three-term split: each step keeps the next 8 significant bits
  const __bf16 xhv = to_bf16(xv32), qhv = to_bf16(qv32);
  const float xr1 = xv32 - (float)xhv, qr1 = qv32 - (float)qhv;
  const __bf16 xmv = to_bf16(xr1), qmv = to_bf16(qr1);
  xh[o] = xhv;
  xm[o] = xmv;
  xl[o] = to_bf16(xr1 - (float)xmv);
  qh[o] = qhv;
  qm[o] = qmv;
  ql[o] = to_bf16(qr1 - (float)qmv);
  xf[o] = xv32;
  qf[o] = qv32;
  if (c == 0) {
    qnorm[row] = (float)sq;
    if (row < n) atomicMax(max_norm_bits, __float_as_uint((float)sqrt(sq) * 1.0000001f));
  }
}

// ---------------------------------------------------------------------------
// 2. candidates
// ---------------------------------------------------------------------------
struct CandParams {
  const __bf16 *xh, *xm, *xl, *qh, *qm, *ql;
  const float *xf, *qf, *qnorm;
  const double* X;  // FP64 rows (exact mode)
  int64_t ld;
  int d;
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

template <int MODE>
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

  if (MODE == kModeBF16) {
    const bf16x8 bq_h = *reinterpret_cast<const bf16x8*>(P.qh + q * kDim + g * 8);
    const bf16x8 bq_m = *reinterpret_cast<const bf16x8*>(P.qm + q * kDim + g * 8);
    const bf16x8 bq_l = *reinterpret_cast<const bf16x8*>(P.ql + q * kDim + g * 8);
    const bf16x8* xh8 = reinterpret_cast<const bf16x8*>(P.xh);
    const bf16x8* xm8 = reinterpret_cast<const bf16x8*>(P.xm);
    const bf16x8* xl8 = reinterpret_cast<const bf16x8*>(P.xl);
    bf16x8 ah = xh8[(int64_t)j * 4 + g];
    bf16x8 am = xm8[(int64_t)j * 4 + g];
    bf16x8 al = xl8[(int64_t)j * 4 + g];
    for (int64_t t = 0; t < ntiles; ++t) {
      // smallest terms first
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bq_h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bq_m, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bq_h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_m, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bq_h, acc, 0, 0, 0);
      if (t + 1 < ntiles) {  // prefetch the next tile's A operands
        const int64_t row = (t + 1) * 16 + j;
        ah = xh8[row * 4 + g];
        am = xm8[row * 4 + g];
        al = xl8[row * 4 + g];
      }
      const float th = L[kList - 1];
      const int64_t pbase = t * 16 + g * 4;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float s = acc[v];
        const int64_t p = pbase + v;
        const bool ins = s > th && p < P.n;
        if (__ballot(ins)) {
          if (ins) list_insert(L, I, s, (int32_t)p);
        }
      }
    }
  } else if (MODE == kModeExact) {
    // second tier: exact FP64 differences; the key is -fl32(d^2)
    double qv[kMaxData];
#pragma unroll
    for (int c = 0; c < kMaxData; ++c) qv[c] = c < P.d ? P.X[q * P.ld + c] : 0.0;
    for (int64_t t = 0; t < ntiles; ++t) {
      const float th = L[kList - 1];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t p = t * 16 + g * 4 + v;
        float s = -INFINITY;
        if (p < P.n) {
          const double* xr = P.X + p * P.ld;
          double acc = 0.0;
          for (int c = 0; c < P.d; ++c) {
            const double df = xr[c] - qv[c];
            acc = fma(df, df, acc);
          }
          s = -(float)acc;
        }
        const bool ins = s > th;
        if (__ballot(ins)) {
          if (ins) list_insert(L, I, s, (int32_t)p);
        }
      }
    }
  } else {
    float qv[kDim];
#pragma unroll
    for (int c = 0; c < kDim; ++c) qv[c] = P.qf[q * kDim + c];
    for (int64_t t = 0; t < ntiles; ++t) {
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
        const bool ins = s > th && p < P.n;
        if (__ballot(ins)) {
          if (ins) list_insert(L, I, s, (int32_t)p);
        }
      }
    }
  }
  if (!qvalid) return;
  int32_t* out = P.cand + qi * kCand + g * kList;
#pragma unroll
  for (int r = 0; r < kList; ++r) out[r] = I[r];
  // stripe bound on the approximate d^2 of every point this lane did not keep
  if (MODE == kModeExact) {
    P.tau[qi * 4 + g] = (L[kList - 1] == -INFINITY) ? INFINITY : -L[kList - 1];
  } else {
    const float nq = P.qnorm[q];
    P.tau[qi * 4 + g] = (L[kList - 1] == -INFINITY) ? INFINITY : nq - 2.0f * L[kList - 1];
  }
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
  const unsigned int* max_norm_bits;
  double eps_c;
  int mode;
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
  return best - 1e-7 * fabs(tau) - 4.0 * DBL_EPSILON * (nq + M * M);
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
  const double nq = (double)P.qnorm[q];
  double lb;
  if (P.mode == kModeExact) {
    // keys are fl32 of the exact FP64 d^2 (FMA accumulation; within a few ulp of
    // the rescored value): non-kept points have exact d^2 >= tau (1 - 2^-20)
    lb = (double)tmin * (1.0 - 0x1p-20) - 64.0 * DBL_MIN;
  } else {
    lb = nonkept_lower_bound(tmin, nq, M, P.eps_c);
  }
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

struct Workspace {
  __bf16 *xh, *xm, *xl, *qh, *qm, *ql;
  float *xf, *qf, *qnorm, *tau;
  int32_t *cand, *fail_rows, *fail_rows2;
  unsigned int *max_norm_bits, *fail_count, *fail_count2;
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
  t.xh = (__bf16*)take(n_pad * kDim * 2);
  t.xm = (__bf16*)take(n_pad * kDim * 2);
  t.xl = (__bf16*)take(n_pad * kDim * 2);
  t.qh = (__bf16*)take(n_pad * kDim * 2);
  t.qm = (__bf16*)take(n_pad * kDim * 2);
  t.ql = (__bf16*)take(n_pad * kDim * 2);
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
  if (w) *w = t;
  return off;
}

}  // namespace knn
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::knn;

extern "C" {

int64_t snapgpu_knn_workspace_bytes(int64_t n, int32_t d, int32_t k) {
  (void)d;
  (void)k;
  if (n < 0) return 0;
  return layout(n, nullptr, nullptr) + 256;
}

int snapgpu_knn(const double* d_X, int64_t n, int32_t d, int64_t ld, int32_t k, int32_t mode,
                int32_t* d_nbr, double* d_dist, void* d_workspace, int64_t workspace_bytes,
                int32_t flags, int64_t* n_fallback, int64_t* n_exact, void* stream) {
  clear_error();
  SG_ARG(n >= 2, "need at least 2 points");
  SG_ARG(n < INT32_MAX, "too many points");
  SG_ARG(d >= 1 && d <= kMaxData, "dimension must be in [1, 30]");
  SG_ARG(ld >= d, "ld < d");
  SG_ARG(k >= 1 && k <= kMaxK && k <= n - 1, "k must be in [1, min(100, n-1)]");
  SG_ARG(mode == kModeFP32 || mode == kModeBF16, "mode must be 0 (fp32) or 1 (bf16)");
  SG_ARG(d_X && d_nbr && d_dist && d_workspace, "null pointer");
  SG_ARG(workspace_bytes >= snapgpu_knn_workspace_bytes(n, d, k), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* base = (char*)(((uintptr_t)d_workspace + 255) & ~(uintptr_t)255);
  Workspace w;
  layout(n, &w, base);
  const int64_t n_pad = ((n + 15) / 16) * 16;
  SG_HIP(hipMemsetAsync(w.max_norm_bits, 0, 4, st));
  SG_HIP(hipMemsetAsync(w.fail_count, 0, 4, st));
  prepare_kernel<<<(unsigned)ceil_div(n_pad * 32, 256), 256, 0, st>>>(
      d_X, n, d, ld, n_pad, w.xh, w.xm, w.xl, w.qh, w.qm, w.ql, w.xf, w.qf, w.qnorm, w.max_norm_bits);
  SG_LAUNCH_CHECK();
  CandParams cp{w.xh, w.xm, w.xl, w.qh, w.qm, w.ql, w.xf, w.qf, w.qnorm, d_X, ld, d,
                n, n_pad, nullptr, n, w.cand, w.tau};
  const unsigned cblocks = (unsigned)ceil_div(ceil_div(n, 16), 4);
  if (mode == kModeBF16) cand_kernel<kModeBF16><<<cblocks, 256, 0, st>>>(cp);
  else cand_kernel<kModeFP32><<<cblocks, 256, 0, st>>>(cp);
  SG_LAUNCH_CHECK();
  FinParams fp{d_X, n, ld, d, k, nullptr, n, w.cand, w.tau, w.qnorm, w.max_norm_bits,
               mode == kModeBF16 ? kEpsBF16 : kEpsFP32, mode, (flags & 1) ? 1 : 0,
               d_nbr, d_dist, w.fail_rows, w.fail_count};
  finalize_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, st>>>(fp);
  SG_LAUNCH_CHECK();
  unsigned int nfail = 0, nfail2 = 0;
  SG_HIP(hipMemcpyAsync(&nfail, w.fail_count, 4, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (nfail > 0 && !(flags & 1)) {
    // second tier: exact FP64 distances for the failed rows, then the certificate again
    SG_HIP(hipMemsetAsync(w.fail_count2, 0, 4, st));
    CandParams cp2 = cp;
    cp2.rows = w.fail_rows;
    cp2.n_rows = nfail;
    cand_kernel<kModeExact><<<(unsigned)ceil_div(ceil_div((int64_t)nfail, 16), 4), 256, 0, st>>>(cp2);
    SG_LAUNCH_CHECK();
    FinParams fp2 = fp;
    fp2.rows = w.fail_rows;
    fp2.n_rows = nfail;
    fp2.mode = kModeExact;
    fp2.fail_rows = w.fail_rows2;
    fp2.fail_count = w.fail_count2;
    finalize_kernel<<<(unsigned)ceil_div((int64_t)nfail, 4), 256, 0, st>>>(fp2);
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
  if (n_fallback) *n_fallback = (int64_t)nfail;
  if (n_exact) *n_exact = (int64_t)(flags & 1 ? nfail : nfail2);
  return SNAPGPU_OK;
}

}  // extern "C"
