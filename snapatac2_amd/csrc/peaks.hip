// Cell x peak count matrix from the per-cell fragment store.
//
// Reference: create_peak_matrix (snapatac2-core/src/feature_count/matrix.rs:95-171)
// with RegionCounter::insert / insert_fragment (feature_count/counter.rs:92-139)
// over the fragments rebuilt by pair_to_fragments / single_to_fragments
// (feature_count/data_iter.rs:27-123): paired fragments [start, start + size)
// with the inclusive size filter; single-end reads insert one base at the
// stored position. A region is hit by a query interval when they share a base
// on the same chromosome (half-open overlap).
//
// Steps: one lane per fragment finds its chromosome (binary search over the
// genome offsets), then the regions its insertions (or the whole fragment)
// overlap: regions of a chromosome are sorted by start with a running maximum
// of their ends, so a query walks down from the last region starting before
// its end while that maximum passes its start. A count pass sizes the output,
// an emit pass writes one (cell << rbits | region) key per increment (rbits =
// the bits of the largest column), a radix sort over the rbits + cell bits
// groups them, and the runs of equal keys are the CSR entries and their
// counts (columns ascend inside a row, as the reference's BTreeMap gives).
#include <hipcub/hipcub.hpp>

#include "common.h"

namespace snapgpu {
namespace peaks {

constexpr int kThreads = 256;

struct Index {
  const uint64_t* start;    // region starts, sorted per chromosome
  const uint64_t* end;      // region ends (exclusive)
  const uint64_t* pmax;     // running maximum of end inside the chromosome's range
  const uint32_t* orig;     // region id (column) in the input order
  const int64_t* chrom_off; // [n_chrom + 1] range of each chromosome's regions
  const uint64_t* base;     // [n_chrom + 1] genome offset of each chromosome
  int32_t n_chrom;
};

__device__ __forceinline__ int chrom_of(const Index& I, uint64_t p) {
  int a = 0, b = I.n_chrom;  // last chromosome with base <= p
  while (b - a > 1) {
    const int mid = (a + b) / 2;
    if (I.base[mid] <= p) a = mid;
    else b = mid;
  }
  return a;
}

// Regions of chromosome c overlapping [qs, qe); calls f(orig, region start, region end).
template <typename F>
__device__ __forceinline__ void for_hits(const Index& I, int c, uint64_t qs, uint64_t qe, F f) {
  const int64_t lo = I.chrom_off[c], hi = I.chrom_off[c + 1];
  int64_t a = lo, b = hi;  // first region with start >= qe
  while (a < b) {
    const int64_t mid = (a + b) / 2;
    if (I.start[mid] < qe) a = mid + 1;
    else b = mid;
  }
  for (int64_t i = a - 1; i >= lo && I.pmax[i] > qs; --i)
    if (I.end[i] > qs) f(I.orig[i], I.start[i], I.end[i]);
}

struct Frag {
  bool on;
  int c;
  uint64_t s, size;
};

__device__ __forceinline__ Frag frag_at(const Index& I, const uint32_t* pos, const uint32_t* val, int64_t f,
                                        int paired, uint32_t min_len, uint32_t max_len) {
  Frag q{false, 0, 0, 0};
  const uint64_t p = pos[f];
  if (p >= I.base[I.n_chrom]) return q;  // outside the genome
  q.c = chrom_of(I, p);
  q.s = p - I.base[q.c];
  if (paired) {
    q.size = val[f];
    if (q.size < min_len || q.size > max_len) return q;
  } else {
    q.size = 1;
  }
  q.on = true;
  return q;
}

// Visits every (region) increment of fragment f: `emit(orig)`.
template <typename F>
__device__ __forceinline__ void visit(const Index& I, const Frag& q, int paired, int strategy, F emit) {
  if (!q.on) return;
  if (!paired) {  // single-end: one insertion at the stored base (counter.rs:104-108)
    for_hits(I, q.c, q.s, q.s + 1, [&](uint32_t o, uint64_t, uint64_t) { emit(o); });
    return;
  }
  const uint64_t a = q.s, b = q.s + q.size - 1;  // the two insertion bases
  if (strategy == SNAPGPU_FRAGMENT) {
    for_hits(I, q.c, q.s, q.s + q.size, [&](uint32_t o, uint64_t, uint64_t) { emit(o); });
  } else if (strategy == SNAPGPU_INSERTION) {
    for_hits(I, q.c, a, a + 1, [&](uint32_t o, uint64_t, uint64_t) { emit(o); });
    for_hits(I, q.c, b, b + 1, [&](uint32_t o, uint64_t, uint64_t) { emit(o); });
  } else {  // paired-insertion: regions hit by either insertion, once (counter.rs:120-135)
    for_hits(I, q.c, a, a + 1, [&](uint32_t o, uint64_t, uint64_t) { emit(o); });
    for_hits(I, q.c, b, b + 1, [&](uint32_t o, uint64_t rs, uint64_t re) {
      if (!(rs <= a && a < re)) emit(o);
    });
  }
}

__global__ void count_kernel(Index I, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ val, int64_t n,
                             int paired, int strategy, uint32_t min_len, uint32_t max_len, int64_t* __restrict__ hits) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const Frag q = frag_at(I, pos, val, f, paired, min_len, max_len);
  int64_t k = 0;
  visit(I, q, paired, strategy, [&](uint32_t) { ++k; });
  hits[f] = k;
}

__global__ void emit_kernel(Index I, const int64_t* __restrict__ cell_ptr, int64_t n_cells,
                            const uint32_t* __restrict__ pos, const uint32_t* __restrict__ val, int64_t n, int paired,
                            int strategy, uint32_t min_len, uint32_t max_len, const int64_t* __restrict__ hits_incl,
                            int64_t total, int rbits, uint64_t* __restrict__ keys) {
  // rows of the block's first and last fragment bound the rows of all of its
  // fragments (the row of a fragment grows with its index); every thread
  // reaches the barrier before any returns
  __shared__ int64_t row_span[2];
  const int64_t f0 = (int64_t)blockIdx.x * blockDim.x;
  if (threadIdx.x < 2) {
    const int64_t g = threadIdx.x == 0 ? f0 : (f0 + blockDim.x < n ? f0 + blockDim.x : n) - 1;
    int64_t a = 0, b = n_cells;  // last row with cell_ptr[row] <= g
    while (b - a > 1) {
      const int64_t mid = (a + b) / 2;
      if (cell_ptr[mid] <= g) a = mid;
      else b = mid;
    }
    row_span[threadIdx.x] = a;
  }
  __syncthreads();
  const int64_t f = f0 + threadIdx.x;
  if (f >= n) return;
  const Frag q = frag_at(I, pos, val, f, paired, min_len, max_len);
  if (!q.on) return;
  int64_t lo = row_span[0], hi = row_span[1] + 1;  // last row with cell_ptr[row] <= f
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (cell_ptr[mid] <= f) lo = mid;
    else hi = mid;
  }
  int64_t o = f == 0 ? 0 : hits_incl[f - 1];
  const int64_t o_end = hits_incl[f];
  const uint64_t cell = (uint64_t)lo << rbits;
  visit(I, q, paired, strategy, [&](uint32_t r) {
    if (o < o_end && o < total) keys[o] = cell | r;
    ++o;
  });
}

// run starts of the sorted keys: flag[j] = 1 where a new key starts (the
// flags depend only on equality, so on neither key layout)
__global__ void run_flag_kernel(const uint64_t* __restrict__ keys, int64_t n, int64_t* __restrict__ flag) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  flag[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1 : 0;
}

// distinct keys, widened from (cell << rbits | region) to (cell << 32 | region)
__global__ void run_start_kernel(const uint64_t* __restrict__ keys, int64_t n, const int64_t* __restrict__ rid,
                                 int rbits, uint64_t* __restrict__ ukey, int64_t* __restrict__ ustart) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (j == 0 || rid[j] != rid[j - 1]) {
    const uint64_t k = keys[j];
    ukey[rid[j] - 1] = ((k >> rbits) << 32) | (k & ((1ull << rbits) - 1));
    ustart[rid[j] - 1] = j;
  }
}

__global__ void csr_entries_kernel(const uint64_t* __restrict__ ukey, const int64_t* __restrict__ ustart, int64_t nu,
                                   int64_t total, int32_t* __restrict__ idx, uint32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nu) return;
  idx[r] = (int32_t)(ukey[r] & 0xFFFFFFFFull);
  cnt[r] = (uint32_t)((r + 1 < nu ? ustart[r + 1] : total) - ustart[r]);
}

__global__ void csr_indptr_kernel(const uint64_t* __restrict__ ukey, int64_t nu, int64_t n_cells,
                                  int64_t* __restrict__ indptr) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n_cells) return;
  const uint64_t k = (uint64_t)c << 32;
  int64_t a = 0, b = nu;  // first key >= c << 32
  while (a < b) {
    const int64_t mid = (a + b) / 2;
    if (ukey[mid] < k) a = mid + 1;
    else b = mid;
  }
  indptr[c] = a;
}

// Gene level of the gene matrix: entry e of a cells x regions CSR becomes the
// key (row << gbits | group[column]) with its count (every group below 2^gbits).
__global__ void group_key_kernel(const int64_t* __restrict__ indptr, int64_t n_rows, const int32_t* __restrict__ idx,
                                 const uint32_t* __restrict__ val, int64_t nnz, const uint32_t* __restrict__ group,
                                 int gbits, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  // rows of the block's first and last entry bound the rows of all of its
  // entries; every thread reaches the barrier before any returns
  __shared__ int64_t row_span[2];
  const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
  if (threadIdx.x < 2) {
    const int64_t g = threadIdx.x == 0 ? e0 : (e0 + blockDim.x < nnz ? e0 + blockDim.x : nnz) - 1;
    int64_t a = 0, b = n_rows;  // last row with indptr[row] <= g
    while (b - a > 1) {
      const int64_t mid = (a + b) / 2;
      if (indptr[mid] <= g) a = mid;
      else b = mid;
    }
    row_span[threadIdx.x] = a;
  }
  __syncthreads();
  const int64_t e = e0 + threadIdx.x;
  if (e >= nnz) return;
  int64_t lo = row_span[0], hi = row_span[1] + 1;  // last row with indptr[row] <= e
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (indptr[mid] <= e) lo = mid;
    else hi = mid;
  }
  keys[e] = ((uint64_t)lo << gbits) | group[idx[e]];
  vals[e] = val[e];
}

// Largest count of each run of equal keys (GeneCount::get_values keeps the
// maximum over a gene's transcripts, counter.rs:270-283).
__global__ void run_max_kernel(const uint64_t* __restrict__ ukey, const int64_t* __restrict__ ustart, int64_t nu,
                               int64_t total, const uint32_t* __restrict__ vals, int32_t* __restrict__ idx,
                               uint32_t* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nu) return;
  const int64_t a = ustart[r], b = r + 1 < nu ? ustart[r + 1] : total;
  uint32_t m = 0;
  for (int64_t j = a; j < b; ++j) m = vals[j] > m ? vals[j] : m;
  idx[r] = (int32_t)(ukey[r] & 0xFFFFFFFFull);
  out[r] = m;
}

inline int grid(int64_t n) { return (int)ceil_div(n > 0 ? n : 1, kThreads); }

inline Index to_index(const snapgpu_peak_index* p) {
  return Index{p->start, p->end, p->pmax, p->orig, p->chrom_off, p->base, p->n_chrom};
}

}  // namespace peaks
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::peaks;

extern "C" {

int64_t snapgpu_peak_count_workspace_bytes(int64_t n_fragments) {
  return snapgpu_scan_workspace_bytes(n_fragments > 0 ? n_fragments : 1);
}

int snapgpu_peak_count(const snapgpu_peak_index* index, const uint32_t* d_pos, const uint32_t* d_val,
                       int64_t n_fragments, int32_t paired, int32_t strategy, uint32_t min_len, uint32_t max_len,
                       int64_t* d_hits, void* d_ws, int64_t ws_bytes, int64_t* h_total, void* stream) {
  clear_error();
  SG_ARG(index != nullptr && h_total != nullptr && n_fragments >= 0, "bad arguments");
  SG_ARG(index->n_chrom >= 1, "empty genome");
  SG_ARG(strategy == SNAPGPU_INSERTION || strategy == SNAPGPU_FRAGMENT || strategy == SNAPGPU_PIC,
         "strategy must be insertion, fragment or paired-insertion");
  SG_ARG(ws_bytes >= snapgpu_peak_count_workspace_bytes(n_fragments), "workspace too small");
  *h_total = 0;
  if (n_fragments == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  count_kernel<<<grid(n_fragments), kThreads, 0, st>>>(to_index(index), d_pos, d_val, n_fragments, paired, strategy,
                                                       min_len, max_len, d_hits);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(d_hits, n_fragments, d_ws, ws_bytes, stream);
  if (rc) return rc;
  SG_HIP(hipMemcpyAsync(h_total, d_hits + n_fragments - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  return SNAPGPU_OK;
}

int64_t snapgpu_peak_pairs_workspace_bytes(int64_t total) {
  const int64_t t = total > 0 ? total : 1;
  size_t tmp = 0;
  hipcub::DoubleBuffer<uint64_t> kb(nullptr, nullptr);
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, kb, (int)(t < INT32_MAX ? t : INT32_MAX), 0, 64);
  return 8 * t /* alt keys */ + 8 * t /* run flags / ids */ + 8 * t /* run starts */ +
         snapgpu_scan_workspace_bytes(t) + (int64_t)tmp + 6 * 256;
}

int snapgpu_peak_pairs(const snapgpu_peak_index* index, const int64_t* d_cell_ptr, int64_t n_cells,
                       const uint32_t* d_pos, const uint32_t* d_val, int64_t n_fragments, int32_t paired,
                       int32_t strategy, uint32_t min_len, uint32_t max_len, const int64_t* d_hits, int64_t total,
                       uint64_t* d_keys, void* d_ws, int64_t ws_bytes, int64_t* h_nnz, void* stream) {
  clear_error();
  SG_ARG(index != nullptr && h_nnz != nullptr && total >= 0 && n_cells >= 1 && n_cells < (1ll << 31), "bad arguments");
  SG_ARG(total < (int64_t)INT32_MAX, "at most 2^31 - 1 increments per call");
  SG_ARG(ws_bytes >= snapgpu_peak_pairs_workspace_bytes(total), "workspace too small");
  SG_ARG(index->n_columns >= 0, "negative column count");
  *h_nnz = 0;
  if (total == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  char* p = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  auto take = [&](int64_t bytes) {
    char* at = p;
    p += (bytes + 255) & ~(int64_t)255;
    return at;
  };
  uint64_t* alt = reinterpret_cast<uint64_t*>(take(8 * total));
  int64_t* rid = reinterpret_cast<int64_t*>(take(8 * total));
  int64_t* ustart = reinterpret_cast<int64_t*>(take(8 * total));
  void* sws = take(snapgpu_scan_workspace_bytes(total));
  void* tmp = p;
  size_t tmp_bytes = (size_t)(ws_bytes - (p - (char*)d_ws));
  // every column is below 2^rbits (32 when the caller gives no column count)
  int rbits = 32;
  if (index->n_columns > 0) {
    rbits = 1;
    while (rbits < 32 && ((uint64_t)(index->n_columns - 1) >> rbits) != 0) ++rbits;
  }
  emit_kernel<<<grid(n_fragments), kThreads, 0, st>>>(to_index(index), d_cell_ptr, n_cells, d_pos, d_val,
                                                      n_fragments, paired, strategy, min_len, max_len, d_hits, total,
                                                      rbits, d_keys);
  SG_LAUNCH_CHECK();
  int end_bit = rbits + 1;  // rbits region bits, then the bits of n_cells
  while (end_bit < 64 && ((uint64_t)n_cells >> (end_bit - rbits)) != 0) ++end_bit;
  hipcub::DoubleBuffer<uint64_t> kb(d_keys, alt);
  SG_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, kb, (int)total, 0, end_bit, st));
  if (kb.Current() != d_keys) SG_HIP(hipMemcpyAsync(d_keys, kb.Current(), 8 * total, hipMemcpyDeviceToDevice, st));
  run_flag_kernel<<<grid(total), kThreads, 0, st>>>(d_keys, total, rid);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(rid, total, sws, snapgpu_scan_workspace_bytes(total), stream);
  if (rc) return rc;
  // unique keys overwrite the alternate buffer; their run starts go to ustart
  run_start_kernel<<<grid(total), kThreads, 0, st>>>(d_keys, total, rid, rbits, alt, ustart);
  SG_LAUNCH_CHECK();
  SG_HIP(hipMemcpyAsync(h_nnz, rid + total - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  return SNAPGPU_OK;
}

int snapgpu_peak_csr(const void* d_ws, int64_t ws_bytes, int64_t total, int64_t nnz, int64_t n_cells,
                     int64_t* d_indptr, int32_t* d_idx, uint32_t* d_val, void* stream) {
  clear_error();
  SG_ARG(total >= 0 && nnz >= 0 && nnz <= total && n_cells >= 1 && d_indptr != nullptr, "bad arguments");
  SG_ARG(ws_bytes >= snapgpu_peak_pairs_workspace_bytes(total), "workspace too small");
  hipStream_t st = as_stream(stream);
  const char* p = (const char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  const int64_t step = (8 * (total > 0 ? total : 1) + 255) & ~(int64_t)255;
  const uint64_t* ukey = reinterpret_cast<const uint64_t*>(p);
  const int64_t* ustart = reinterpret_cast<const int64_t*>(p + 2 * step);
  if (nnz > 0) {
    csr_entries_kernel<<<grid(nnz), kThreads, 0, st>>>(ukey, ustart, nnz, total, d_idx, d_val);
    SG_LAUNCH_CHECK();
  }
  csr_indptr_kernel<<<grid(n_cells + 1), kThreads, 0, st>>>(ukey, nnz, n_cells, d_indptr);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_group_max_workspace_bytes(int64_t nnz) {
  const int64_t t = nnz > 0 ? nnz : 1;
  size_t tmp = 0;
  hipcub::DoubleBuffer<uint64_t> kb(nullptr, nullptr);
  hipcub::DoubleBuffer<uint32_t> vb(nullptr, nullptr);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kb, vb, (int)(t < INT32_MAX ? t : INT32_MAX), 0, 64);
  // unique keys, run starts, their values, keys x2, values x2, run ids
  return 8 * t * 5 + 4 * t * 3 + snapgpu_scan_workspace_bytes(t) + (int64_t)tmp + 10 * 256;
}

int snapgpu_group_max(const int64_t* d_indptr, int64_t n_rows, const int32_t* d_idx, const uint32_t* d_val,
                      int64_t nnz, const uint32_t* d_group, int32_t n_groups, void* d_ws, int64_t ws_bytes,
                      int64_t* h_nnz_out, void* stream) {
  clear_error();
  SG_ARG(n_rows >= 1 && n_rows < (1ll << 31) && nnz >= 0 && nnz < (int64_t)INT32_MAX && h_nnz_out != nullptr &&
             n_groups >= 0,
         "bad arguments");
  SG_ARG(ws_bytes >= snapgpu_group_max_workspace_bytes(nnz), "workspace too small");
  *h_nnz_out = 0;
  if (nnz == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  char* p = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  auto take = [&](int64_t bytes) {
    char* at = p;
    p += (bytes + 255) & ~(int64_t)255;
    return at;
  };
  uint64_t* ukey = reinterpret_cast<uint64_t*>(take(8 * nnz));
  int64_t* ustart = reinterpret_cast<int64_t*>(take(8 * nnz));
  uint32_t* uval = reinterpret_cast<uint32_t*>(take(4 * nnz));
  uint64_t* kA = reinterpret_cast<uint64_t*>(take(8 * nnz));
  uint64_t* kB = reinterpret_cast<uint64_t*>(take(8 * nnz));
  uint32_t* vA = reinterpret_cast<uint32_t*>(take(4 * nnz));
  uint32_t* vB = reinterpret_cast<uint32_t*>(take(4 * nnz));
  int64_t* rid = reinterpret_cast<int64_t*>(take(8 * nnz));
  void* sws = take(snapgpu_scan_workspace_bytes(nnz));
  void* tmp = p;
  size_t tmp_bytes = (size_t)(ws_bytes - (p - (char*)d_ws));
  // every group is below 2^gbits (32 when the caller gives no group count)
  int gbits = 32;
  if (n_groups > 0) {
    gbits = 1;
    while (gbits < 32 && ((uint64_t)(n_groups - 1) >> gbits) != 0) ++gbits;
  }
  group_key_kernel<<<grid(nnz), kThreads, 0, st>>>(d_indptr, n_rows, d_idx, d_val, nnz, d_group, gbits, kA, vA);
  SG_LAUNCH_CHECK();
  int end_bit = gbits + 1;  // gbits group bits, then the bits of n_rows
  while (end_bit < 64 && ((uint64_t)n_rows >> (end_bit - gbits)) != 0) ++end_bit;
  hipcub::DoubleBuffer<uint64_t> kb(kA, kB);
  hipcub::DoubleBuffer<uint32_t> vb(vA, vB);
  SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kb, vb, (int)nnz, 0, end_bit, st));
  run_flag_kernel<<<grid(nnz), kThreads, 0, st>>>(kb.Current(), nnz, rid);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(rid, nnz, sws, snapgpu_scan_workspace_bytes(nnz), stream);
  if (rc) return rc;
  run_start_kernel<<<grid(nnz), kThreads, 0, st>>>(kb.Current(), nnz, rid, gbits, ukey, ustart);
  SG_LAUNCH_CHECK();
  // the sorted values are kept next to the run starts for snapgpu_group_max_csr
  SG_HIP(hipMemcpyAsync(uval, vb.Current(), 4 * nnz, hipMemcpyDeviceToDevice, st));
  SG_HIP(hipMemcpyAsync(h_nnz_out, rid + nnz - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  return SNAPGPU_OK;
}

int snapgpu_group_max_csr(const void* d_ws, int64_t ws_bytes, int64_t nnz, int64_t nnz_out, int64_t n_rows,
                          int64_t* d_indptr, int32_t* d_idx, uint32_t* d_val, void* stream) {
  clear_error();
  SG_ARG(nnz >= 0 && nnz_out >= 0 && nnz_out <= nnz && n_rows >= 1 && d_indptr != nullptr, "bad arguments");
  SG_ARG(ws_bytes >= snapgpu_group_max_workspace_bytes(nnz), "workspace too small");
  hipStream_t st = as_stream(stream);
  const char* p = (const char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  const int64_t t = nnz > 0 ? nnz : 1;
  const int64_t s8 = (8 * t + 255) & ~(int64_t)255;
  const uint64_t* ukey = reinterpret_cast<const uint64_t*>(p);
  const int64_t* ustart = reinterpret_cast<const int64_t*>(p + s8);
  const uint32_t* uval = reinterpret_cast<const uint32_t*>(p + 2 * s8);
  if (nnz_out > 0) {
    run_max_kernel<<<grid(nnz_out), kThreads, 0, st>>>(ukey, ustart, nnz_out, nnz, uval, d_idx, d_val);
    SG_LAUNCH_CHECK();
  }
  csr_indptr_kernel<<<grid(n_rows + 1), kThreads, 0, st>>>(ukey, nnz_out, n_rows, d_indptr);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
