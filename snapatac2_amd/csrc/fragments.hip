// Fragment-file import on the device: the text of a fragment file (lines
// "chrom\tstart\tend\tbarcode[\tcount[\tstrand]]") becomes the per-cell fragment
// store (CSR: row = barcode, column = genome base index, value = length) plus
// the per-barcode QC counts.
//
// Reference:
//   import_fragments / make_arraydata / count_fragments
//     (snapatac2-core/src/preprocessing/import.rs:31-204);
//   Fragment::from_str (snapatac2-core/src/preprocessing/qc.rs:306-345);
//   FragmentQCBuilder (qc.rs:427-466);
//   import_fragments and shift_fragment (snapatac2-python/src/preprocessing.rs:87-183).
//
// Steps (all on one stream; the host reads three scalars back in between):
//   1. newline count per 64 KiB block, scan, newline positions -> line starts;
//   2. one lane per line parses its fields (tab separated), looks the
//      chromosome up in a hashed name table and writes per-line records;
//   3. the records with end > start are compacted;
//   4. stable radix sorts put them in (barcode group, position, value) order:
//      by (position, value) first, then by the run index of the barcode in the
//      file (input sorted by barcode) or by a 64-bit hash of the barcode
//      (unsorted input; equal hashes of different barcodes are reported);
//   5. groups are the runs of equal barcodes; QC sums per group by a
//      segmented wave reduction; whitelist and minimum-count filters; the
//      groups of unsorted input are ranked by their barcode bytes (LSD sort of
//      8-byte big-endian words), which is the reference's external-sort order;
//   6. fill: each fragment of a kept group goes to row offset + its rank.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "common.h"

namespace snapgpu {
namespace frag {

constexpr int kThreads = 256;
constexpr int kRounds = 16;                                     // 16-byte words per thread
constexpr int64_t kByteBlock = (int64_t)kThreads * kRounds * 16;  // 64 KiB of text per block
constexpr int64_t kLineBlock = (int64_t)kThreads * kRounds;       // 4096 records per block

constexpr uint64_t kFnvOffset = 1469598103934665603ull;
constexpr uint64_t kFnvPrime = 1099511628211ull;

// per-line flags
constexpr uint8_t F_KEEP = 1;     // a record with end != start (import.rs:66)
constexpr uint8_t F_GENOME = 2;   // chromosome in the genome (import.rs:168)
constexpr uint8_t F_MITO = 4;     // chromosome in the mitochondrial list
constexpr uint8_t F_STRAND = 8;   // a strand field is present (single-end record)
constexpr uint8_t F_REV = 16;     // strand '-'
constexpr uint8_t F_RANGE = 32;   // stored coordinate >= chromosome length (genome.rs:387-406)
constexpr uint8_t F_HUGE = 64;    // end - start > 2^32 - 1
constexpr uint8_t F_BIG = 128;    // +-(end - start) outside int32

// error codes (low byte of the error word; the line index is above it)
enum : uint32_t {
  E_START_MISSING = 1,
  E_END_MISSING = 2,
  E_BARCODE_MISSING = 3,
  E_START_BAD = 4,
  E_END_BAD = 5,
  E_COUNT_BAD = 6,
  E_STRAND_BAD = 7,
  E_BARCODE_NONE = 8,
  E_END_BEFORE_START = 9,
  E_BARCODE_LONG = 10,
  E_REVERSE_PAIRED = 11,
  E_SIZE_I32 = 12,
  E_RANGE = 13,
  E_SIZE_U32 = 14,
};

// Host-packed hashed name table (chromosomes, or a barcode whitelist):
// header | slots[cap] | entries[n] | name bytes.
struct TableHeader {
  uint32_t n, cap;
  uint64_t total;  // sum of lengths (chromosome table)
};
struct TableSlot {
  uint64_t hash;
  int32_t idx;  // -1 = empty
  int32_t pad;
};
struct TableEntry {
  uint64_t base, length;
  uint32_t name_off, name_len, flags, pad;
};
static_assert(sizeof(TableHeader) == 16 && sizeof(TableSlot) == 16 && sizeof(TableEntry) == 32, "layout");

struct Counters {
  unsigned long long err;       // min (line << 8 | code) over parse errors
  unsigned long long first;     // min (line << 1 | has_strand) over records
  unsigned long long n_rec;     // records (not comments, not empty)
  unsigned long long n_strand;  // records with a strand field
  unsigned long long bad;       // min (line << 8 | code) over conversion errors in whitelisted groups
  unsigned long long dup;       // min group index of a barcode seen in two runs
  unsigned int max_bc;          // longest barcode (bytes) over kept records
  unsigned int collision;       // two barcodes with one hash next to each other
  unsigned int any_bad;         // a whitelisted group holds a record count_fragments cannot convert
  unsigned int max_size;        // largest end - start (low 32 bits) over kept records
};

__host__ __device__ inline uint64_t fnv_step(uint64_t h, uint32_t c) { return (h ^ c) * kFnvPrime; }

__device__ __forceinline__ uint8_t load_byte(const uint8_t* t, int64_t p) { return t[p]; }

// 8 bytes at the aligned address a (little endian); bytes at or past n read as 0.
__device__ __forceinline__ uint64_t load_word(const uint8_t* t, int64_t a, int64_t n) {
  if (a + 8 <= n) return *reinterpret_cast<const uint64_t*>(t + a);
  uint64_t w = 0;
  for (int k = 0; k < 8 && a + k < n; ++k) w |= (uint64_t)t[a + k] << (8 * k);
  return w;
}

__device__ bool bytes_equal(const uint8_t* a, const uint8_t* b, uint32_t len) {
  for (uint32_t k = 0; k < len; ++k)
    if (a[k] != b[k]) return false;
  return true;
}

__device__ int table_find(const uint8_t* tab, uint64_t h, const uint8_t* s, uint32_t len) {
  const TableHeader* H = reinterpret_cast<const TableHeader*>(tab);
  const uint32_t cap = H->cap, n = H->n;
  const TableSlot* S = reinterpret_cast<const TableSlot*>(tab + sizeof(TableHeader));
  const TableEntry* E = reinterpret_cast<const TableEntry*>(S + cap);
  const uint8_t* names = reinterpret_cast<const uint8_t*>(E + n);
  uint32_t i = (uint32_t)h & (cap - 1);
  for (uint32_t probe = 0; probe < cap; ++probe, i = (i + 1) & (cap - 1)) {
    const TableSlot sl = S[i];
    if (sl.idx < 0) return -1;
    if (sl.hash == h) {
      const TableEntry e = E[sl.idx];
      if (e.name_len == len && bytes_equal(names + e.name_off, s, len)) return sl.idx;
    }
  }
  return -1;
}

__device__ __forceinline__ const TableEntry& table_entry(const uint8_t* tab, int idx) {
  const TableHeader* H = reinterpret_cast<const TableHeader*>(tab);
  const TableSlot* S = reinterpret_cast<const TableSlot*>(tab + sizeof(TableHeader));
  return reinterpret_cast<const TableEntry*>(S + H->cap)[idx];
}

// Number of bytes equal to c in x.
__device__ __forceinline__ int count_byte(uint32_t x, uint32_t pat) {
  const uint32_t v = x ^ pat;
  uint32_t y = (v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
  y = ~(y | v | 0x7F7F7F7Fu);
  return __popc(y);
}

// 16-bit mask of the bytes equal to '\n' in a 16-byte word (bit k = byte k).
__device__ __forceinline__ uint32_t newline_mask(uint4 w) {
  uint32_t m = 0;
  const uint32_t x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t v = x[q] ^ 0x0A0A0A0Au;
    uint32_t y = (v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    y = ~(y | v | 0x7F7F7F7Fu);  // 0x80 in each byte equal to '\n'
    m |= (((y >> 7) & 1u) | ((y >> 14) & 2u) | ((y >> 21) & 4u) | ((y >> 28) & 8u)) << (4 * q);
  }
  return m;
}

// The 16 bytes of word index wi (bytes past n_bytes read as 0, not '\n').
__device__ __forceinline__ uint4 load16(const uint8_t* t, int64_t wi, int64_t n) {
  const int64_t off = wi * 16;
  if (off + 16 <= n) return *reinterpret_cast<const uint4*>(t + off);
  uint32_t x[4] = {0, 0, 0, 0};
  for (int k = 0; k < 16 && off + k < n; ++k) x[k >> 2] |= (uint32_t)t[off + k] << (8 * (k & 3));
  return make_uint4(x[0], x[1], x[2], x[3]);
}

template <typename T>
__device__ T block_sum(T v, T* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T s = 0;
  for (int k = 0; k < kThreads / 64; ++k) s += lds[k];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// 1. lines
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kThreads) nl_count_kernel(const uint8_t* __restrict__ text, int64_t n,
                                                            int64_t* __restrict__ blk) {
  __shared__ int64_t red[kThreads / 64];
  const int64_t w0 = (int64_t)blockIdx.x * kThreads * kRounds;
  int64_t c = 0;
  for (int r = 0; r < kRounds; ++r) {
    const int64_t wi = w0 + (int64_t)r * kThreads + threadIdx.x;
    if (wi * 16 < n) {
      const uint4 w = load16(text, wi, n);
      c += count_byte(w.x, 0x0A0A0A0Au) + count_byte(w.y, 0x0A0A0A0Au) + count_byte(w.z, 0x0A0A0A0Au) +
           count_byte(w.w, 0x0A0A0A0Au);
    }
  }
  c = block_sum<int64_t>(c, red);
  if (threadIdx.x == 0) blk[blockIdx.x] = c;
}

// line_start[0] = 0 and line_start[k + 1] = (position of newline k) + 1.
// blk_incl: inclusive scan of the per-block newline counts.
__global__ void __launch_bounds__(kThreads) nl_index_kernel(const uint8_t* __restrict__ text, int64_t n,
                                                            const int64_t* __restrict__ blk_incl,
                                                            int64_t* __restrict__ line_start) {
  __shared__ int64_t wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t w0 = (int64_t)blockIdx.x * kThreads * kRounds;
  int64_t base = blockIdx.x == 0 ? 0 : blk_incl[blockIdx.x - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) line_start[0] = 0;
  for (int r = 0; r < kRounds; ++r) {
    const int64_t wi = w0 + (int64_t)r * kThreads + threadIdx.x;
    uint32_t m = 0;
    if (wi * 16 < n) m = newline_mask(load16(text, wi, n));
    const uint64_t c = (uint64_t)__popc(m);
    const uint64_t incl = wave_incl_scan_u64(c, lane);
    if (lane == 63) wsum[wid] = (int64_t)incl;
    __syncthreads();
    int64_t before = 0, total = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
      const int64_t s = wsum[k];
      before += k < wid ? s : 0;
      total += s;
    }
    int64_t o = base + before + (int64_t)(incl - c);
    while (m) {
      const int k = __ffs(m) - 1;
      m &= m - 1;
      line_start[1 + o] = wi * 16 + k + 1;
      ++o;
    }
    base += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 2. parse
// ---------------------------------------------------------------------------

struct ParseOut {
  uint32_t* pos;
  uint32_t* size;
  uint32_t* cnt;
  uint8_t* flags;
  uint64_t* bc;   // barcode offset << 8 | length
  uint64_t* bch;  // FNV-1a hash of the barcode
};

__device__ __forceinline__ uint64_t sat_add(uint64_t a, int64_t d) {
  if (d >= 0) {
    const uint64_t r = a + (uint64_t)d;
    return r < a ? ~0ull : r;
  }
  const uint64_t m = (uint64_t)(-(d + 1)) + 1;  // |d| without overflow
  return a < m ? 0 : a - m;
}

struct LineStats {
  uint32_t n_rec, n_strand, n_keep, max_bc, max_size;
  unsigned long long first;  // min (line << 1 | has_strand) over records
};

// Parses line i and writes its record; accumulates the block statistics.
__device__ __forceinline__ void parse_line(const uint8_t* __restrict__ text, int64_t n_bytes,
                                           const int64_t* __restrict__ line_start, int64_t i,
                                           const uint8_t* __restrict__ chroms, int64_t shift_left,
                                           int64_t shift_right, const ParseOut& out, Counters* cnt,
                                           LineStats& ls) {
  bool is_rec = false, has_strand = false;
  uint32_t err = 0, bc_len_keep = 0;
  {
    const int64_t s = line_start[i];
    int64_t e = line_start[i + 1] - 1;
    if (e > n_bytes) e = n_bytes;
    if (e > s && load_byte(text, e - 1) == '\r') --e;
    uint8_t fl = 0;
    uint32_t pos = 0, size = 0, count = 1;
    uint64_t bcw = 0, bch = 0;
    const bool comment = e > s && load_byte(text, s) == '#';
    if (e > s && !comment) {
      is_rec = true;
      // current field
      uint32_t field = 0, flen = 0, c0 = 0;
      uint64_t num = 0, h = kFnvOffset;
      bool nondigit = false;
      int64_t fstart = s;
      // committed fields
      uint64_t chrom_h = 0, bc_h = 0, start = 0, end = 0;
      int64_t chrom_off = s, bc_off = s;
      uint32_t chrom_len = 0, bc_len = 0, bc_c0 = 0;
      bool start_bad = false, end_bad = false, count_bad = false, strand_bad = false;
      uint32_t strand = 0;  // 0 none, 1 '+', 2 '-'
      int64_t a = s & ~(int64_t)7;
      uint64_t word = load_word(text, a, n_bytes);
      for (int64_t p = s; p <= e; ++p) {
        uint32_t c = '\t';
        if (p < e) {
          if ((p & ~(int64_t)7) != a) {
            a = p & ~(int64_t)7;
            word = load_word(text, a, n_bytes);
          }
          c = (uint32_t)(word >> (8 * (p & 7))) & 0xFFu;
        }
        if (c == '\t') {
          const bool bad = nondigit || flen == 0;
          switch (field) {
            case 0: chrom_h = h; chrom_off = fstart; chrom_len = flen; break;
            case 1: start = num; start_bad = bad; break;
            case 2: end = num; end_bad = bad; break;
            case 3: bc_h = h; bc_off = fstart; bc_len = flen; bc_c0 = c0; break;
            case 4:
              if (flen == 1 && c0 == '.') count = 1;
              else { count = (uint32_t)num; count_bad = bad || num > 0xFFFFFFFFull; }
              break;
            case 5:
              if (flen == 1 && c0 == '.') strand = 0;
              else if (flen == 1 && c0 == '+') strand = 1;
              else if (flen == 1 && c0 == '-') strand = 2;
              else strand_bad = true;
              break;
            default: break;
          }
          ++field;
          flen = 0;
          num = 0;
          nondigit = false;
          h = kFnvOffset;
          fstart = p + 1;
        } else {
          const uint32_t d = c - '0';
          const bool dig = d < 10;
          const bool ovf = num > 1844674407370955161ull || (num == 1844674407370955161ull && d > 5);
          nondigit |= !dig || ovf;
          num = num * 10 + (dig ? d : 0);
          h = fnv_step(h, c);
          c0 = flen == 0 ? c : c0;
          ++flen;
        }
      }
      has_strand = strand != 0;  // a '.' strand field is no strand (qc.rs:338-344)
      if (field < 2) err = E_START_MISSING;
      else if (start_bad) err = E_START_BAD;
      else if (field < 3) err = E_END_MISSING;
      else if (end_bad) err = E_END_BAD;
      else if (field < 4) err = E_BARCODE_MISSING;
      else if (count_bad) err = E_COUNT_BAD;
      else if (strand_bad) err = E_STRAND_BAD;
      if (err == 0) {
        // shift_fragment (preprocessing.rs:173-183)
        const uint64_t st = shift_left != 0 ? sat_add(start, shift_left) : start;
        uint64_t en = end;
        if (has_strand && shift_left != 0) en = sat_add(end, shift_left);
        if (!has_strand && shift_right != 0) en = sat_add(end, shift_right);
        if (en < st) {
          err = E_END_BEFORE_START;
        } else if (en > st) {
          fl |= F_KEEP;
          if (bc_len == 1 && bc_c0 == '.') err = E_BARCODE_NONE;
          else if (bc_len > 255) err = E_BARCODE_LONG;
          bc_len_keep = bc_len;
          if (has_strand) fl |= F_STRAND;
          if (strand == 2) fl |= F_REV;
          const uint64_t sz = en - st;
          size = (uint32_t)sz;
          if (sz > 0xFFFFFFFFull) fl |= F_HUGE;
          if (sz > (strand == 2 ? 2147483648ull : 2147483647ull)) fl |= F_BIG;
          const int ci = table_find(chroms, chrom_h, text + chrom_off, chrom_len);
          if (ci >= 0) {
            const TableEntry& ce = table_entry(chroms, ci);
            fl |= F_GENOME;
            if (ce.flags & 1u) fl |= F_MITO;
            const uint64_t coord = strand == 2 ? en - 1 : st;
            if (coord >= ce.length) fl |= F_RANGE;
            else pos = (uint32_t)(ce.base + coord);
          }
          bcw = ((uint64_t)bc_off << 8) | (bc_len > 255 ? 255u : bc_len);
          bch = bc_h;
        }
      }
    }
    out.pos[i] = pos;
    out.size[i] = size;
    out.cnt[i] = count;
    out.flags[i] = err ? 0 : fl;
    out.bc[i] = bcw;
    out.bch[i] = bch;
    if (err) atomicMin(&cnt->err, ((unsigned long long)i << 8) | err);
    if (!err && (fl & F_KEEP)) {
      ++ls.n_keep;
      ls.max_size = size > ls.max_size ? size : ls.max_size;
    }
  }
  if (is_rec) {
    ++ls.n_rec;
    ls.n_strand += has_strand ? 1 : 0;
    const unsigned long long f = ((unsigned long long)i << 1) | (has_strand ? 1ull : 0ull);
    ls.first = f < ls.first ? f : ls.first;
  }
  ls.max_bc = bc_len_keep > ls.max_bc ? bc_len_keep : ls.max_bc;
}

// kLineBlock lines per block (lane t of round r takes line b0 + r * 256 + t);
// blk[block] = records kept. The block statistics go to the counters with one
// atomic each, so the counters see a few tens of thousands of atomics.
__global__ void __launch_bounds__(kThreads) parse_kernel(const uint8_t* __restrict__ text, int64_t n_bytes,
                                                         const int64_t* __restrict__ line_start, int64_t n_lines,
                                                         const uint8_t* __restrict__ chroms, int64_t shift_left,
                                                         int64_t shift_right, ParseOut out, Counters* cnt,
                                                         int64_t* __restrict__ blk) {
  __shared__ unsigned long long red[6][kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t b0 = (int64_t)blockIdx.x * kLineBlock;
  LineStats ls{0, 0, 0, 0, 0, ~0ull};
  for (int r = 0; r < kRounds; ++r) {
    const int64_t i = b0 + (int64_t)r * kThreads + threadIdx.x;
    if (i < n_lines) parse_line(text, n_bytes, line_start, i, chroms, shift_left, shift_right, out, cnt, ls);
  }
  unsigned long long v[6] = {ls.n_rec, ls.n_strand, ls.n_keep, ls.max_bc, ls.first, ls.max_size};
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] += __shfl_xor(v[q], d, 64);
    const unsigned long long m = __shfl_xor(v[3], d, 64), f = __shfl_xor(v[4], d, 64),
                             z = __shfl_xor(v[5], d, 64);
    v[3] = m > v[3] ? m : v[3];
    v[4] = f < v[4] ? f : v[4];
    v[5] = z > v[5] ? z : v[5];
  }
  if (lane == 0)
    for (int q = 0; q < 6; ++q) red[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) {
      for (int q = 0; q < 3; ++q) v[q] += red[q][w];
      v[3] = red[3][w] > v[3] ? red[3][w] : v[3];
      v[4] = red[4][w] < v[4] ? red[4][w] : v[4];
      v[5] = red[5][w] > v[5] ? red[5][w] : v[5];
    }
    blk[blockIdx.x] = (int64_t)v[2];
    if (v[0]) atomicAdd(&cnt->n_rec, v[0]);
    if (v[1]) atomicAdd(&cnt->n_strand, v[1]);
    if (v[3] && v[3] > *reinterpret_cast<volatile unsigned int*>(&cnt->max_bc))
      atomicMax(&cnt->max_bc, (unsigned int)v[3]);
    if (v[4] != ~0ull && v[4] < *reinterpret_cast<volatile unsigned long long*>(&cnt->first))
      atomicMin(&cnt->first, v[4]);
    if (v[5] && v[5] > *reinterpret_cast<volatile unsigned int*>(&cnt->max_size))
      atomicMax(&cnt->max_size, (unsigned int)v[5]);
  }
}

// ---------------------------------------------------------------------------
// 3. compaction of the kept records (file order)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kThreads) keep_write_kernel(const uint8_t* __restrict__ flags, int64_t n,
                                                              const int64_t* __restrict__ blk_incl,
                                                              uint32_t* __restrict__ kept) {
  __shared__ int64_t wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t b0 = (int64_t)blockIdx.x * kLineBlock;
  int64_t base = blockIdx.x == 0 ? 0 : blk_incl[blockIdx.x - 1];
  for (int r = 0; r < kRounds; ++r) {
    const int64_t i = b0 + (int64_t)r * kThreads + threadIdx.x;
    const bool k = i < n && (flags[i] & F_KEEP);
    const uint64_t bal = ballot64(k);
    if (lane == 0) wsum[wid] = popc64(bal);
    __syncthreads();
    int64_t before = 0, total = 0;
    for (int q = 0; q < kThreads / 64; ++q) {
      const int64_t s = wsum[q];
      before += q < wid ? s : 0;
      total += s;
    }
    if (k) kept[base + before + popc64(bal & lanemask_lt(lane))] = (uint32_t)i;
    base += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 4. sort keys and barcode comparisons
// ---------------------------------------------------------------------------

// 8 bytes starting at byte p (any alignment) from two aligned loads.
__device__ __forceinline__ uint64_t load_u8x8(const uint8_t* t, int64_t p, int64_t n) {
  const int64_t a = p & ~(int64_t)7;
  const int sh = (int)(p & 7) * 8;
  const uint64_t lo = load_word(t, a, n);
  if (sh == 0) return lo;
  return (lo >> sh) | (load_word(t, a + 8, n) << (64 - sh));
}

// Equality of the len bytes at offsets a and b, eight at a time.
__device__ __forceinline__ bool text_equal(const uint8_t* t, int64_t n, int64_t a, int64_t b, uint32_t len) {
  for (uint32_t k = 0; k < len; k += 8) {
    const uint64_t x = load_u8x8(t, a + k, n) ^ load_u8x8(t, b + k, n);
    const uint32_t rem = len - k;
    const uint64_t mask = rem >= 8 ? ~0ull : ((1ull << (8 * rem)) - 1);
    if (x & mask) return false;
  }
  return true;
}

__device__ __forceinline__ bool same_barcode(const uint8_t* text, int64_t n, const uint64_t* bc, const uint64_t* bch,
                                             uint32_t la, uint32_t lb, bool* collision) {
  if (bch[la] != bch[lb]) return false;
  const uint64_t a = bc[la], b = bc[lb];
  const bool eq = (a & 255) == (b & 255) && text_equal(text, n, (int64_t)(a >> 8), (int64_t)(b >> 8), (uint32_t)(a & 255));
  if (!eq) *collision = true;
  return eq;
}

__global__ void iota_kernel(uint32_t* __restrict__ v, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) v[j] = (uint32_t)j;
}

// (position, value) key of the record at sorted slot j, in pos_bits + vbits
// bits: position above the value, the value offset by max_size for
// single-end records so that -size .. +size order as integers. Records outside
// the genome get `outside` (above every genome key) and sort last inside their
// group.
__global__ void pv_key_kernel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kept,
                              const uint8_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                              const uint32_t* __restrict__ size, int64_t n, int paired, int vbits,
                              uint32_t max_size, uint64_t outside, uint64_t* __restrict__ key) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t line = kept[perm[j]];
  const uint8_t f = flags[line];
  uint64_t k = outside;
  if (f & F_GENOME) {
    const uint64_t sz = size[line];
    const uint64_t v = paired ? sz : ((f & F_REV) ? (uint64_t)max_size - sz : (uint64_t)max_size + sz);
    k = ((uint64_t)pos[line] << vbits) | v;
  }
  key[j] = k;
}

// Boundaries of equal-barcode runs over kept records in file order (input
// sorted by barcode): run[v] = 1 where a new run starts.
__global__ void run_bound_kernel(const uint8_t* __restrict__ text, int64_t n_bytes, const uint32_t* __restrict__ kept,
                                 const uint64_t* __restrict__ bc, const uint64_t* __restrict__ bch, int64_t n,
                                 int64_t* __restrict__ run) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  bool col = false;
  run[v] = (v == 0 || !same_barcode(text, n_bytes, bc, bch, kept[v - 1], kept[v], &col)) ? 1 : 0;
}

// Second sort key: the run index (sorted input) or the barcode hash.
__global__ void group_key_kernel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kept,
                                 const int64_t* __restrict__ run_id, const uint64_t* __restrict__ bch, int64_t n,
                                 uint64_t* __restrict__ key) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t v = perm[j];
  key[j] = run_id ? (uint64_t)run_id[v] : bch[kept[v]];
}

// Group boundaries over the final order: run changes (sorted input; run ids
// by kept record) or barcode changes (hash-sorted order, where two different
// barcodes with one hash next to each other are flagged).
__global__ void group_bound_kernel(const uint8_t* __restrict__ text, int64_t n_bytes, const uint32_t* __restrict__ perm,
                                   const uint32_t* __restrict__ kept, const uint64_t* __restrict__ bc,
                                   const uint64_t* __restrict__ bch, const int64_t* __restrict__ run, int64_t n,
                                   int64_t* __restrict__ gid, Counters* cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  bool b = j == 0, col = false;
  if (!b) {
    if (run) b = run[perm[j - 1]] != run[perm[j]];
    else b = !same_barcode(text, n_bytes, bc, bch, kept[perm[j - 1]], kept[perm[j]], &col);
  }
  gid[j] = b ? 1 : 0;
  if (col) atomicOr(&cnt->collision, 1u);
}

// Per-group QC sums (FragmentQCBuilder::update over the in-genome records,
// qc.rs:442-449), group start slot and first record. gid: inclusive scan of
// the boundaries (1-based group ids). Groups are contiguous in slot order, so
// a segmented wave scan leaves one atomic per group piece.
__global__ void __launch_bounds__(kThreads) group_stats_kernel(
    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kept, const uint8_t* __restrict__ flags,
    const uint32_t* __restrict__ cnt_in, const int64_t* __restrict__ gid, int64_t n,
    int64_t* __restrict__ gstart, uint32_t* __restrict__ gline, unsigned long long* __restrict__ g_total,
    unsigned long long* __restrict__ g_mito, unsigned long long* __restrict__ g_uniq,
    unsigned int* __restrict__ g_bad) {
  const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool act = j < n;
  int64_t g = -1;
  uint64_t tot = 0, mito = 0, uniq = 0;
  uint32_t bad = 0;  // conversion flags of in-genome records
  if (act) {
    const uint32_t line = kept[perm[j]];
    g = gid[j] - 1;
    const uint8_t f = flags[line];
    if (f & F_GENOME) {
      tot = cnt_in[line];
      mito = (f & F_MITO) ? 1 : 0;
      uniq = 1 - mito;
      bad = f & (F_RANGE | F_HUGE | F_REV | F_BIG);
    }
    if (j == 0 || gid[j - 1] != gid[j]) {
      gstart[g] = j;
      gline[g] = line;
    }
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int src = lane >= d ? lane - d : lane;
    const int64_t og = (int64_t)shfl_u64((uint64_t)g, src);
    const uint64_t ot = shfl_u64(tot, src), om = shfl_u64(mito, src), ou = shfl_u64(uniq, src);
    const uint32_t ob = __shfl(bad, src, 64);
    if (lane >= d && og == g) {
      tot += ot;
      mito += om;
      uniq += ou;
      bad |= ob;
    }
  }
  const int64_t ng = (int64_t)shfl_u64((uint64_t)g, lane < 63 ? lane + 1 : lane);
  const bool tail = act && (lane == 63 || ng != g);
  if (tail) {
    if (tot) atomicAdd(&g_total[g], (unsigned long long)tot);
    if (mito) atomicAdd(&g_mito[g], (unsigned long long)mito);
    if (uniq) atomicAdd(&g_uniq[g], (unsigned long long)uniq);
    if (bad) atomicOr(&g_bad[g], bad);
  }
}

// Whether a whitelisted group holds a record that count_fragments cannot
// convert (then bad_kernel names the first such line).
__global__ void bad_group_kernel(const unsigned int* __restrict__ g_bad, const uint8_t* __restrict__ wl,
                                 int64_t n_groups, int paired, Counters* cnt) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  const uint32_t mask = F_RANGE | F_HUGE | (paired ? F_REV : F_BIG);
  if (wl[g] && (g_bad[g] & mask)) atomicOr(&cnt->any_bad, 1u);
}

// Whitelist membership of each group's barcode (import.rs:74).
__global__ void whitelist_kernel(const uint8_t* __restrict__ text, const uint32_t* __restrict__ gline,
                                 const uint64_t* __restrict__ bc, const uint64_t* __restrict__ bch,
                                 int64_t n_groups, const uint8_t* __restrict__ wl_table, uint8_t* __restrict__ wl) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  uint8_t w = 1;
  if (wl_table) {
    const uint32_t line = gline[g];
    const uint64_t b = bc[line];
    w = table_find(wl_table, bch[line], text + (b >> 8), (uint32_t)(b & 255)) >= 0 ? 1 : 0;
  }
  wl[g] = w;
}

__global__ void group_hash_kernel(const uint32_t* __restrict__ gline, const uint64_t* __restrict__ bch,
                                  int64_t n_groups, uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  key[g] = bch[gline[g]];
  idx[g] = (uint32_t)g;
}

// A barcode in two runs of sorted input (import.rs:139-141), among
// whitelisted groups: equal hashes next to each other after sorting by hash.
__global__ void dup_kernel(const uint8_t* __restrict__ text, const uint64_t* __restrict__ key,
                           const uint32_t* __restrict__ idx, const uint32_t* __restrict__ gline,
                           const uint64_t* __restrict__ bc, const uint8_t* __restrict__ wl, int64_t n_groups,
                           Counters* cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0 || r >= n_groups || key[r] != key[r - 1]) return;
  // scan back over the equal-hash run for the same barcode
  const uint32_t g = idx[r];
  if (!wl[g]) return;
  const uint64_t b = bc[gline[g]];
  for (int64_t q = r - 1; q >= 0 && key[q] == key[r]; --q) {
    const uint32_t h = idx[q];
    if (!wl[h]) continue;
    const uint64_t c = bc[gline[h]];
    if ((b & 255) == (c & 255) && bytes_equal(text + (b >> 8), text + (c >> 8), (uint32_t)(b & 255))) {
      atomicMin(&cnt->dup, (unsigned long long)(g > h ? g : h));
      return;
    }
  }
}

// First record of a whitelisted group that count_fragments cannot convert
// (import.rs:173-198): coordinate outside the chromosome, or a size that does
// not fit the value type (u32 paired, i32 single-end).
__global__ void bad_kernel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kept,
                           const uint8_t* __restrict__ flags, const int64_t* __restrict__ gid,
                           const uint8_t* __restrict__ wl, int64_t n, int paired, Counters* cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t line = kept[perm[j]];
  const uint8_t f = flags[line];
  if (!(f & F_GENOME) || !wl[gid[j] - 1]) return;
  uint32_t code = 0;
  if (f & F_RANGE) code = E_RANGE;
  else if (f & F_HUGE) code = E_SIZE_U32;
  else if (paired && (f & F_REV)) code = E_REVERSE_PAIRED;
  else if (!paired && (f & F_BIG)) code = E_SIZE_I32;
  if (code) atomicMin(&cnt->bad, ((unsigned long long)line << 8) | code);
}

// Word w (8 bytes, big endian, zero padded) of each group's barcode, for the
// lexicographic LSD sort of the groups of unsorted input.
__global__ void word_key_kernel(const uint8_t* __restrict__ text, const uint32_t* __restrict__ order,
                                const uint32_t* __restrict__ gline, const uint64_t* __restrict__ bc,
                                int64_t n_groups, int w, uint64_t* __restrict__ key) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_groups) return;
  const uint64_t b = bc[gline[order[r]]];
  const uint8_t* s = text + (b >> 8);
  const int len = (int)(b & 255);
  uint64_t k = 0;
  for (int q = 0; q < 8; ++q) {
    const int p = 8 * w + q;
    k = (k << 8) | (p < len ? s[p] : 0u);
  }
  key[r] = k;
}

// keep flag of the group at rank r (whitelisted and at least min unique
// fragments, import.rs:142-148), for the row scan.
__global__ void keep_rank_kernel(const uint32_t* __restrict__ order, const uint8_t* __restrict__ wl,
                                 const unsigned long long* __restrict__ g_uniq, int64_t n_groups, uint64_t min_frag,
                                 int64_t* __restrict__ kr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_groups) return;
  const uint32_t g = order ? order[r] : (uint32_t)r;
  kr[r] = (wl[g] && g_uniq[g] >= min_frag) ? 1 : 0;
}

// rowid[g] (-1 = dropped), rowgroup[row], row length (in-genome records).
__global__ void rows_kernel(const uint32_t* __restrict__ order, const int64_t* __restrict__ kr_incl,
                            const unsigned long long* __restrict__ g_uniq,
                            const unsigned long long* __restrict__ g_mito, int64_t n_groups,
                            int64_t* __restrict__ rowid, uint32_t* __restrict__ rowgroup,
                            int64_t* __restrict__ rowlen) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_groups) return;
  const uint32_t g = order ? order[r] : (uint32_t)r;
  const int64_t inc = kr_incl[r], prev = r == 0 ? 0 : kr_incl[r - 1];
  if (inc != prev) {
    const int64_t row = inc - 1;
    rowid[g] = row;
    rowgroup[row] = g;
    rowlen[row] = (int64_t)(g_uniq[g] + g_mito[g]);
  } else {
    rowid[g] = -1;
  }
}

// ---------------------------------------------------------------------------
// 6. fill
// ---------------------------------------------------------------------------

__global__ void fill_kernel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ kept,
                            const uint8_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                            const uint32_t* __restrict__ size, const int64_t* __restrict__ gid,
                            const int64_t* __restrict__ gstart, const int64_t* __restrict__ rowid,
                            const int64_t* __restrict__ row_incl, int64_t n, int paired, uint32_t* __restrict__ out_pos,
                            uint32_t* __restrict__ out_val) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t line = kept[perm[j]];
  const uint8_t f = flags[line];
  if (!(f & F_GENOME)) return;
  const int64_t g = gid[j] - 1;
  const int64_t row = rowid[g];
  if (row < 0) return;
  const int64_t o = (row == 0 ? 0 : row_incl[row - 1]) + (j - gstart[g]);
  const uint32_t sz = size[line];
  out_pos[o] = pos[line];
  out_val[o] = paired ? sz : ((f & F_REV) ? (uint32_t)(0u - sz) : sz);
}

__global__ void rows_out_kernel(const uint8_t* __restrict__ text, const uint32_t* __restrict__ rowgroup,
                                const uint32_t* __restrict__ gline, const uint64_t* __restrict__ bc,
                                const int64_t* __restrict__ row_incl, const unsigned long long* __restrict__ g_uniq,
                                const unsigned long long* __restrict__ g_total,
                                const unsigned long long* __restrict__ g_mito, int64_t n_rows,
                                int64_t* __restrict__ cell_ptr, unsigned long long* __restrict__ qc,
                                uint8_t* __restrict__ names, int width) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const uint32_t g = rowgroup[r];
  if (r == 0) cell_ptr[0] = 0;
  cell_ptr[r + 1] = row_incl[r];
  if (qc) {
    qc[3 * r] = g_uniq[g];
    qc[3 * r + 1] = g_total[g];
    qc[3 * r + 2] = g_mito[g];
  }
  if (names) {
    const uint64_t b = bc[gline[g]];
    const uint8_t* s = text + (b >> 8);
    const int len = (int)(b & 255);
    uint8_t* o = names + r * (int64_t)width;
    for (int q = 0; q < width; ++q) o[q] = q < len ? s[q] : 0;
  }
}

// ---------------------------------------------------------------------------
// export: fragments -> text lines (Fragment's Display, qc.rs:286-303, as
// written by export_fragments, snapatac2-core/src/export.rs:90-106)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int dec_digits(uint64_t x) {
  int d = 1;
  while (x >= 10) {
    x /= 10;
    ++d;
  }
  return d;
}

__device__ __forceinline__ void put_dec(uint8_t* o, uint64_t x, int d) {
  for (int k = d - 1; k >= 0; --k) {
    o[k] = (uint8_t)('0' + x % 10);
    x /= 10;
  }
}

struct FmtLine {
  uint32_t chrom;
  uint64_t start, end;
  uint32_t strand;  // 0 none, 1 '+', 2 '-'
  int64_t row;
};

__device__ FmtLine fmt_line(int64_t f, const int64_t* __restrict__ cell_ptr, int64_t n_cells,
                            const uint32_t* __restrict__ pos, const uint32_t* __restrict__ val, int paired,
                            const uint8_t* __restrict__ chroms) {
  FmtLine L{};
  int64_t lo = 0, hi = n_cells;  // last row with cell_ptr[row] <= f
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (cell_ptr[mid] <= f) lo = mid;
    else hi = mid;
  }
  L.row = lo;
  const TableHeader* H = reinterpret_cast<const TableHeader*>(chroms);
  const uint32_t n = H->n;
  const uint32_t p = pos[f];
  uint32_t a = 0, b = n;  // last chromosome with base <= p
  while (b - a > 1) {
    const uint32_t mid = (a + b) / 2;
    if (table_entry(chroms, mid).base <= p) a = mid;
    else b = mid;
  }
  L.chrom = a;
  const uint64_t coord = p - table_entry(chroms, a).base;
  if (paired) {
    L.start = coord;
    L.end = coord + val[f];
  } else {
    const int32_t v = (int32_t)val[f];
    if (v >= 0) {
      L.start = coord;
      L.end = coord + (uint64_t)v;
      L.strand = 1;
    } else {
      const uint64_t m = (uint64_t)(-(int64_t)v);
      L.end = coord + 1;
      L.start = L.end >= m ? L.end - m : 0;
      L.strand = 2;
    }
  }
  return L;
}

__device__ __forceinline__ int name_len(const uint8_t* nm, int width) {
  int l = 0;
  while (l < width && nm[l] != 0) ++l;
  return l;
}

__global__ void fmt_len_kernel(const int64_t* __restrict__ cell_ptr, int64_t n_cells, const uint32_t* __restrict__ pos,
                               const uint32_t* __restrict__ val, int paired, const uint8_t* __restrict__ chroms,
                               const uint8_t* __restrict__ names, int width, const int64_t* __restrict__ order,
                               int64_t n, int64_t* __restrict__ line_end) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (reinterpret_cast<const TableHeader*>(chroms)->n == 0) {
    line_end[k] = 0;
    return;
  }
  const int64_t f = order ? order[k] : k;
  const FmtLine L = fmt_line(f, cell_ptr, n_cells, pos, val, paired, chroms);
  const int64_t len = table_entry(chroms, L.chrom).name_len + 1 + dec_digits(L.start) + 1 + dec_digits(L.end) + 1 +
                      name_len(names + L.row * (int64_t)width, width) + 2 + (L.strand ? 2 : 0) + 1;
  line_end[k] = len;
}

__global__ void fmt_write_kernel(const int64_t* __restrict__ cell_ptr, int64_t n_cells,
                                 const uint32_t* __restrict__ pos, const uint32_t* __restrict__ val, int paired,
                                 const uint8_t* __restrict__ chroms, const uint8_t* __restrict__ names, int width,
                                 const int64_t* __restrict__ order, int64_t n, const int64_t* __restrict__ line_end,
                                 uint8_t* __restrict__ text) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const TableHeader* H = reinterpret_cast<const TableHeader*>(chroms);
  if (H->n == 0) return;
  const int64_t f = order ? order[k] : k;
  const FmtLine L = fmt_line(f, cell_ptr, n_cells, pos, val, paired, chroms);
  const int64_t begin = k == 0 ? 0 : line_end[k - 1];
  const TableEntry& ce = table_entry(chroms, L.chrom);
  const uint8_t* nm = names + L.row * (int64_t)width;
  const int nl = name_len(nm, width);
  const int64_t len = ce.name_len + 1 + dec_digits(L.start) + 1 + dec_digits(L.end) + 1 + nl + 2 +
                      (L.strand ? 2 : 0) + 1;
  if (begin + len != line_end[k]) return;  // line ends from another call: write nothing
  uint8_t* o = text + begin;
  const uint8_t* cn = reinterpret_cast<const uint8_t*>(
                          reinterpret_cast<const TableEntry*>(chroms + sizeof(TableHeader) + H->cap * sizeof(TableSlot)) +
                          H->n) +
                      ce.name_off;
  for (uint32_t q = 0; q < ce.name_len; ++q) *o++ = cn[q];
  *o++ = '\t';
  int d = dec_digits(L.start);
  put_dec(o, L.start, d);
  o += d;
  *o++ = '\t';
  d = dec_digits(L.end);
  put_dec(o, L.end, d);
  o += d;
  *o++ = '\t';
  for (int q = 0; q < nl; ++q) *o++ = nm[q];
  *o++ = '\t';
  *o++ = '1';
  if (L.strand) {
    *o++ = '\t';
    *o++ = L.strand == 1 ? '+' : '-';
  }
  *o++ = '\n';
}

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------

struct Layout {
  int64_t blk, scan, line_start, pos, size, cnt, flags, bc, bch, kept, permA, permB, keyA, keyB, gid, run,
      gstart, gline, gtot, gmito, guniq, gbad, wl, ordA, ordB, gkA, gkB, kr, rowid, rowgroup, rowlen, counters, tmp,
      total;
  size_t tmp_bytes;
};

inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

Layout make_layout(int64_t n_bytes, int64_t L) {
  Layout o{};
  const int64_t nb = ceil_div(n_bytes > 0 ? n_bytes : 1, kByteBlock);
  const int64_t lb = ceil_div(L > 0 ? L : 1, kLineBlock);
  const int64_t Lc = L > 0 ? L : 1;
  int64_t p = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = p;
    p += align256(bytes);
    return at;
  };
  o.blk = take(8 * (nb > lb ? nb : lb));
  int64_t sw = snapgpu_scan_workspace_bytes(nb);
  const int64_t sw2 = snapgpu_scan_workspace_bytes(Lc + 1);
  if (sw2 > sw) sw = sw2;
  o.scan = take(sw);
  o.line_start = take(8 * (Lc + 1));
  o.pos = take(4 * Lc);
  o.size = take(4 * Lc);
  o.cnt = take(4 * Lc);
  o.flags = take(Lc);
  o.bc = take(8 * Lc);
  o.bch = take(8 * Lc);
  o.kept = take(4 * Lc);
  o.permA = take(4 * Lc);
  o.permB = take(4 * Lc);
  o.keyA = take(8 * Lc);
  o.keyB = take(8 * Lc);
  o.gid = take(8 * Lc);
  o.run = take(8 * Lc);
  o.gstart = take(8 * Lc);
  o.gline = take(4 * Lc);
  o.gtot = take(8 * Lc);
  o.gmito = take(8 * Lc);
  o.guniq = take(8 * Lc);
  o.gbad = take(4 * Lc);
  o.wl = take(Lc);
  o.ordA = take(4 * Lc);
  o.ordB = take(4 * Lc);
  o.gkA = take(8 * Lc);
  o.gkB = take(8 * Lc);
  o.kr = take(8 * Lc);
  o.rowid = take(8 * Lc);
  o.rowgroup = take(4 * Lc);
  o.rowlen = take(8 * Lc);
  o.counters = take(sizeof(Counters));
  size_t tmp = 0;
  hipcub::DoubleBuffer<uint64_t> kb(nullptr, nullptr);
  hipcub::DoubleBuffer<uint32_t> vb(nullptr, nullptr);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kb, vb, (int)(Lc < INT32_MAX ? Lc : INT32_MAX), 0, 64);
  o.tmp_bytes = tmp;
  o.tmp = take((int64_t)tmp);
  o.total = p + 256;
  return o;
}

template <typename T>
inline T* at(void* ws, int64_t off) {
  char* base = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
  return reinterpret_cast<T*>(base + off);
}

inline int grid(int64_t n, int t = kThreads) { return (int)ceil_div(n > 0 ? n : 1, t); }

const char* error_text(uint32_t code) {
  switch (code) {
    case E_START_MISSING: return "missing start position";
    case E_END_MISSING: return "missing end position";
    case E_BARCODE_MISSING: return "missing name (barcode)";
    case E_START_BAD: return "invalid start position";
    case E_END_BAD: return "invalid end position";
    case E_COUNT_BAD: return "invalid count (score) field";
    case E_STRAND_BAD: return "invalid strand";
    case E_BARCODE_NONE: return "the barcode field is '.'";
    case E_END_BEFORE_START: return "end position before start position";
    case E_BARCODE_LONG: return "barcode longer than 255 bytes";
    case E_REVERSE_PAIRED: return "a '-' strand record in a paired-end file: cannot convert its negative size to u32";
    case E_SIZE_I32: return "fragment size does not fit i32";
    case E_RANGE: return "position outside its chromosome";
    case E_SIZE_U32: return "fragment size does not fit u32";
    default: return "invalid record";
  }
}

int data_error(unsigned long long word) {
  const uint64_t line = word >> 8;
  set_error("line " + std::to_string(line + 1) + ": " + error_text((uint32_t)(word & 255)));
  return SNAPGPU_ERR_DATA;
}

uint64_t host_hash(const char* s, size_t len) {
  uint64_t h = kFnvOffset;
  for (size_t k = 0; k < len; ++k) h = fnv_step(h, (uint8_t)s[k]);
  return h;
}

}  // namespace frag
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::frag;

extern "C" {

int64_t snapgpu_frag_table_bytes(const char* const* names, int32_t n) {
  if (n < 0 || (n > 0 && names == nullptr)) return -1;
  uint32_t cap = 16;
  while (cap < 2u * (uint32_t)n) cap <<= 1;
  int64_t bytes = (int64_t)sizeof(TableHeader) + (int64_t)cap * sizeof(TableSlot) + (int64_t)n * sizeof(TableEntry);
  for (int32_t k = 0; k < n; ++k) bytes += (int64_t)strlen(names[k]);
  return bytes;
}

int snapgpu_frag_table_pack(const char* const* names, const uint64_t* lengths, const uint8_t* mito, int32_t n,
                            int32_t genome, void* out, int64_t out_bytes) {
  clear_error();
  SG_ARG(n >= 0 && n < (1 << 24), "table size out of range");
  SG_ARG(n == 0 || names != nullptr, "null names");
  SG_ARG(!genome || n == 0 || lengths != nullptr, "null lengths");
  const int64_t need = snapgpu_frag_table_bytes(names, n);
  SG_ARG(out != nullptr && out_bytes >= need, "output buffer too small");
  uint32_t cap = 16;
  while (cap < 2u * (uint32_t)n) cap <<= 1;
  uint8_t* o = static_cast<uint8_t*>(out);
  memset(o, 0, (size_t)need);
  TableHeader* H = reinterpret_cast<TableHeader*>(o);
  TableSlot* S = reinterpret_cast<TableSlot*>(o + sizeof(TableHeader));
  TableEntry* E = reinterpret_cast<TableEntry*>(S + cap);
  char* blob = reinterpret_cast<char*>(E + n);
  for (uint32_t k = 0; k < cap; ++k) S[k].idx = -1;
  uint64_t base = 0, name_off = 0;
  for (int32_t k = 0; k < n; ++k) {
    const size_t len = strlen(names[k]);
    SG_ARG(len < (1u << 30), "name too long");
    E[k].base = base;
    E[k].length = genome ? lengths[k] : 0;
    E[k].name_off = (uint32_t)name_off;
    E[k].name_len = (uint32_t)len;
    E[k].flags = (mito && mito[k]) ? 1u : 0u;
    memcpy(blob + name_off, names[k], len);
    name_off += len;
    if (genome) {
      base += lengths[k];
      SG_ARG(base <= 0xFFFFFFFFull, "genome longer than 2^32 - 1 bases: positions are stored as uint32");
    }
    const uint64_t h = host_hash(names[k], len);
    uint32_t i = (uint32_t)h & (cap - 1);
    while (S[i].idx >= 0) {
      const TableEntry& e = E[S[i].idx];
      SG_ARG(!(S[i].hash == h && e.name_len == len && memcmp(blob + e.name_off, names[k], len) == 0),
             std::string("duplicate name: ") + names[k]);
      i = (i + 1) & (cap - 1);
    }
    S[i].hash = h;
    S[i].idx = k;
  }
  H->n = (uint32_t)n;
  H->cap = cap;
  H->total = base;
  return SNAPGPU_OK;
}

int64_t snapgpu_frag_count_workspace_bytes(int64_t n_bytes) {
  const int64_t nb = ceil_div(n_bytes > 0 ? n_bytes : 1, kByteBlock);
  return align256(8 * nb) + snapgpu_scan_workspace_bytes(nb) + 512;
}

int snapgpu_frag_count_lines(const uint8_t* d_text, int64_t n_bytes, void* d_ws, int64_t ws_bytes,
                             int64_t* h_n_lines, void* stream) {
  clear_error();
  SG_ARG(n_bytes >= 0 && h_n_lines != nullptr, "bad arguments");
  SG_ARG(n_bytes == 0 || ((uintptr_t)d_text & 15) == 0, "text must be 16-byte aligned");
  SG_ARG(ws_bytes >= snapgpu_frag_count_workspace_bytes(n_bytes), "workspace too small");
  *h_n_lines = 0;
  if (n_bytes == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  const int64_t nb = ceil_div(n_bytes, kByteBlock);
  int64_t* blk = at<int64_t>(d_ws, 0);
  void* sws = at<char>(d_ws, align256(8 * nb));
  nl_count_kernel<<<(unsigned)nb, kThreads, 0, st>>>(d_text, n_bytes, blk);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(blk, nb, sws, snapgpu_scan_workspace_bytes(nb), stream);
  if (rc) return rc;
  int64_t n_nl = 0;
  uint8_t last = 0;
  SG_HIP(hipMemcpyAsync(&n_nl, blk + nb - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&last, d_text + n_bytes - 1, 1, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  *h_n_lines = n_nl + (last != '\n' ? 1 : 0);
  return SNAPGPU_OK;
}

int64_t snapgpu_frag_workspace_bytes(int64_t n_bytes, int64_t n_lines) {
  if (n_bytes < 0 || n_lines < 0) return -1;
  return make_layout(n_bytes, n_lines).total;
}

int snapgpu_frag_index(const uint8_t* d_text, int64_t n_bytes, int64_t n_lines, const void* d_chroms,
                       const void* d_whitelist, const snapgpu_frag_options* opt, void* d_ws, int64_t ws_bytes,
                       snapgpu_frag_summary* h_sum, void* stream) {
  clear_error();
  SG_ARG(n_bytes >= 0 && n_lines >= 0 && opt != nullptr && h_sum != nullptr && d_chroms != nullptr,
         "bad arguments");
  SG_ARG(n_bytes == 0 || ((uintptr_t)d_text & 15) == 0, "text must be 16-byte aligned");
  SG_ARG(n_lines <= n_bytes, "more lines than bytes");
  SG_ARG(n_lines < (int64_t)INT32_MAX, "at most 2^31 - 2 lines per call");
  const Layout Lo = make_layout(n_bytes, n_lines);
  SG_ARG(ws_bytes >= Lo.total, "workspace too small");
  *h_sum = snapgpu_frag_summary{};
  h_sum->n_lines = n_lines;
  h_sum->paired = 0;
  if (n_lines == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  void* w = d_ws;
  int64_t* blk = at<int64_t>(w, Lo.blk);
  void* sws = at<char>(w, Lo.scan);
  const int64_t sws_bytes = Lo.line_start - Lo.scan;
  int64_t* line_start = at<int64_t>(w, Lo.line_start);
  Counters* C = at<Counters>(w, Lo.counters);
  const uint8_t* chroms = static_cast<const uint8_t*>(d_chroms);

  // 1. line starts (the sentinel after a last line without '\n' is n_bytes + 1)
  const int64_t nb = ceil_div(n_bytes, kByteBlock);
  nl_count_kernel<<<(unsigned)nb, kThreads, 0, st>>>(d_text, n_bytes, blk);
  SG_LAUNCH_CHECK();
  int rc = snapgpu_scan_inclusive_i64(blk, nb, sws, sws_bytes, stream);
  if (rc) return rc;
  // line_start holds n_lines + 1 entries: check n_lines before writing them
  int64_t n_nl = 0;
  uint8_t last = 0;
  SG_HIP(hipMemcpyAsync(&n_nl, blk + nb - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&last, d_text + n_bytes - 1, 1, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  SG_ARG(n_lines == n_nl + (last != '\n' ? 1 : 0), "n_lines does not match the text");
  nl_index_kernel<<<(unsigned)nb, kThreads, 0, st>>>(d_text, n_bytes, blk, line_start);
  SG_LAUNCH_CHECK();
  const int64_t sentinel = n_bytes + 1;  // end of a last line without '\n'
  if (last != '\n') SG_HIP(hipMemcpyAsync(line_start + n_lines, &sentinel, 8, hipMemcpyHostToDevice, st));

  // 2. parse
  Counters c0{};
  c0.err = ~0ull;
  c0.first = ~0ull;
  c0.bad = ~0ull;
  c0.dup = ~0ull;
  SG_HIP(hipMemcpyAsync(C, &c0, sizeof(Counters), hipMemcpyHostToDevice, st));
  ParseOut po{at<uint32_t>(w, Lo.pos), at<uint32_t>(w, Lo.size), at<uint32_t>(w, Lo.cnt), at<uint8_t>(w, Lo.flags),
              at<uint64_t>(w, Lo.bc), at<uint64_t>(w, Lo.bch)};
  const int64_t lb = ceil_div(n_lines, kLineBlock);
  parse_kernel<<<(unsigned)lb, kThreads, 0, st>>>(d_text, n_bytes, line_start, n_lines, chroms, opt->shift_left,
                                                  opt->shift_right, po, C, blk);
  SG_LAUNCH_CHECK();

  // 3. compaction
  rc = snapgpu_scan_inclusive_i64(blk, lb, sws, sws_bytes, stream);
  if (rc) return rc;
  uint32_t* kept = at<uint32_t>(w, Lo.kept);
  keep_write_kernel<<<(unsigned)lb, kThreads, 0, st>>>(po.flags, n_lines, blk, kept);
  SG_LAUNCH_CHECK();
  Counters ch{};
  int64_t K = 0;
  SG_HIP(hipMemcpyAsync(&ch, C, sizeof(Counters), hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&K, blk + lb - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (ch.err != ~0ull) return data_error(ch.err);
  h_sum->n_records = (int64_t)ch.n_rec;
  h_sum->n_kept = K;
  h_sum->name_width = (int32_t)ch.max_bc;
  // paired-end unless the first record carries a strand (import.rs:46-55);
  // for unsorted input the first record is that of the smallest barcode, so
  // a file mixing both kinds is rejected there
  int paired = 0;
  if (ch.n_rec > 0) {
    if (opt->sorted_by_barcode) {
      paired = (ch.first & 1) ? 0 : 1;
    } else {
      if (ch.n_strand != 0 && ch.n_strand != ch.n_rec) {
        set_error("records with and without a strand field in one unsorted file");
        return SNAPGPU_ERR_DATA;
      }
      paired = ch.n_strand == 0 ? 1 : 0;
    }
  }
  h_sum->paired = paired;
  if (K == 0) return SNAPGPU_OK;

  // 4. sorts: (position, value), then run index or barcode hash
  uint32_t* permA = at<uint32_t>(w, Lo.permA);
  uint32_t* permB = at<uint32_t>(w, Lo.permB);
  uint64_t* keyA = at<uint64_t>(w, Lo.keyA);
  uint64_t* keyB = at<uint64_t>(w, Lo.keyB);
  void* tmp = at<char>(w, Lo.tmp);
  size_t tmp_bytes = Lo.tmp_bytes;
  // key widths: positions below the genome length, values up to the largest size
  TableHeader th{};
  SG_HIP(hipMemcpyAsync(&th, d_chroms, sizeof(TableHeader), hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  auto bits_of = [](uint64_t x) {
    int b = 1;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
  };
  const uint32_t max_size = ch.max_size;
  const int vbits = bits_of(paired ? (uint64_t)max_size : 2ull * max_size);
  const int pos_bits = bits_of(th.total);
  SG_ARG(pos_bits + vbits <= 64, "position and value keys exceed 64 bits");
  const uint64_t outside = (uint64_t)th.total << vbits;  // above every genome key (pos < total)
  iota_kernel<<<grid(K), kThreads, 0, st>>>(permA, K);
  SG_LAUNCH_CHECK();
  pv_key_kernel<<<grid(K), kThreads, 0, st>>>(permA, kept, po.flags, po.pos, po.size, K, paired, vbits, max_size,
                                              outside, keyA);
  SG_LAUNCH_CHECK();
  hipcub::DoubleBuffer<uint64_t> kb(keyA, keyB);
  hipcub::DoubleBuffer<uint32_t> vb(permA, permB);
  SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kb, vb, (int)K, 0, pos_bits + vbits, st));
  int64_t* run = at<int64_t>(w, Lo.run);
  int end_bit = 64;
  if (opt->sorted_by_barcode) {
    run_bound_kernel<<<grid(K), kThreads, 0, st>>>(d_text, n_bytes, kept, po.bc, po.bch, K, run);
    SG_LAUNCH_CHECK();
    rc = snapgpu_scan_inclusive_i64(run, K, sws, sws_bytes, stream);
    if (rc) return rc;
    int64_t n_runs = 0;
    SG_HIP(hipMemcpyAsync(&n_runs, run + K - 1, 8, hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
    end_bit = 1;
    while (end_bit < 64 && (n_runs >> end_bit) != 0) ++end_bit;
  }
  uint64_t* kcur = kb.Current();
  uint64_t* kalt = kb.Alternate();
  group_key_kernel<<<grid(K), kThreads, 0, st>>>(vb.Current(), kept, opt->sorted_by_barcode ? run : nullptr,
                                                 po.bch, K, kcur);
  SG_LAUNCH_CHECK();
  (void)kalt;
  SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kb, vb, (int)K, 0, end_bit, st));
  if (vb.Current() != permA) SG_HIP(hipMemcpyAsync(permA, vb.Current(), 4 * K, hipMemcpyDeviceToDevice, st));
  const uint32_t* perm = permA;

  // 5. groups
  int64_t* gid = at<int64_t>(w, Lo.gid);
  group_bound_kernel<<<grid(K), kThreads, 0, st>>>(d_text, n_bytes, perm, kept, po.bc, po.bch,
                                                   opt->sorted_by_barcode ? run : nullptr, K, gid, C);
  SG_LAUNCH_CHECK();
  rc = snapgpu_scan_inclusive_i64(gid, K, sws, sws_bytes, stream);
  if (rc) return rc;
  int64_t G = 0;
  SG_HIP(hipMemcpyAsync(&G, gid + K - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipMemcpyAsync(&ch, C, sizeof(Counters), hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (!opt->sorted_by_barcode && ch.collision) {
    set_error("two different barcodes share a 64-bit hash; cannot group this unsorted file");
    return SNAPGPU_ERR_CAPACITY;
  }
  h_sum->n_groups = G;
  int64_t* gstart = at<int64_t>(w, Lo.gstart);
  uint32_t* gline = at<uint32_t>(w, Lo.gline);
  auto* gtot = at<unsigned long long>(w, Lo.gtot);
  auto* gmito = at<unsigned long long>(w, Lo.gmito);
  auto* guniq = at<unsigned long long>(w, Lo.guniq);
  SG_HIP(hipMemsetAsync(gtot, 0, 8 * G, st));
  SG_HIP(hipMemsetAsync(gmito, 0, 8 * G, st));
  SG_HIP(hipMemsetAsync(guniq, 0, 8 * G, st));
  auto* gbad = at<unsigned int>(w, Lo.gbad);
  SG_HIP(hipMemsetAsync(gbad, 0, 4 * G, st));
  group_stats_kernel<<<grid(K), kThreads, 0, st>>>(perm, kept, po.flags, po.cnt, gid, K, gstart, gline, gtot, gmito,
                                                   guniq, gbad);
  SG_LAUNCH_CHECK();
  uint8_t* wl = at<uint8_t>(w, Lo.wl);
  whitelist_kernel<<<grid(G), kThreads, 0, st>>>(d_text, gline, po.bc, po.bch, G,
                                                 static_cast<const uint8_t*>(d_whitelist), wl);
  SG_LAUNCH_CHECK();
  uint32_t* ordA = at<uint32_t>(w, Lo.ordA);
  uint32_t* ordB = at<uint32_t>(w, Lo.ordB);
  uint64_t* gkA = at<uint64_t>(w, Lo.gkA);
  uint64_t* gkB = at<uint64_t>(w, Lo.gkB);
  if (opt->sorted_by_barcode && G > 1) {
    group_hash_kernel<<<grid(G), kThreads, 0, st>>>(gline, po.bch, G, gkA, ordA);
    SG_LAUNCH_CHECK();
    hipcub::DoubleBuffer<uint64_t> gk(gkA, gkB);
    hipcub::DoubleBuffer<uint32_t> go(ordA, ordB);
    SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, gk, go, (int)G, 0, 64, st));
    dup_kernel<<<grid(G), kThreads, 0, st>>>(d_text, gk.Current(), go.Current(), gline, po.bc, wl, G, C);
    SG_LAUNCH_CHECK();
  }
  bad_group_kernel<<<grid(G), kThreads, 0, st>>>(gbad, wl, G, paired, C);
  SG_LAUNCH_CHECK();
  SG_HIP(hipMemcpyAsync(&ch, C, sizeof(Counters), hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (ch.any_bad) {
    bad_kernel<<<grid(K), kThreads, 0, st>>>(perm, kept, po.flags, gid, wl, K, paired, C);
    SG_LAUNCH_CHECK();
    SG_HIP(hipMemcpyAsync(&ch, C, sizeof(Counters), hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
  }
  if (ch.dup != ~0ull) {
    set_error("a barcode appears in two separate runs: please sort the fragment file by barcodes (group " +
              std::to_string(ch.dup) + ")");
    return SNAPGPU_ERR_DATA;
  }
  if (ch.bad != ~0ull) return data_error(ch.bad);

  // group order: file order (sorted input) or barcode bytes (unsorted input)
  const uint32_t* order = nullptr;
  if (!opt->sorted_by_barcode) {
    iota_kernel<<<grid(G), kThreads, 0, st>>>(ordA, G);
    SG_LAUNCH_CHECK();
    hipcub::DoubleBuffer<uint64_t> gk(gkA, gkB);
    hipcub::DoubleBuffer<uint32_t> go(ordA, ordB);
    const int words = (int)ceil_div(ch.max_bc > 0 ? ch.max_bc : 1, 8);
    for (int wd = words - 1; wd >= 0; --wd) {
      word_key_kernel<<<grid(G), kThreads, 0, st>>>(d_text, go.Current(), gline, po.bc, G, wd, gk.Current());
      SG_LAUNCH_CHECK();
      SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, gk, go, (int)G, 0, 64, st));
    }
    if (go.Current() != ordA) SG_HIP(hipMemcpyAsync(ordA, go.Current(), 4 * G, hipMemcpyDeviceToDevice, st));
    order = ordA;
  }
  int64_t* kr = at<int64_t>(w, Lo.kr);
  keep_rank_kernel<<<grid(G), kThreads, 0, st>>>(order, wl, guniq, G, opt->min_num_fragments, kr);
  SG_LAUNCH_CHECK();
  rc = snapgpu_scan_inclusive_i64(kr, G, sws, sws_bytes, stream);
  if (rc) return rc;
  int64_t* rowid = at<int64_t>(w, Lo.rowid);
  uint32_t* rowgroup = at<uint32_t>(w, Lo.rowgroup);
  int64_t* rowlen = at<int64_t>(w, Lo.rowlen);
  rows_kernel<<<grid(G), kThreads, 0, st>>>(order, kr, guniq, gmito, G, rowid, rowgroup, rowlen);
  SG_LAUNCH_CHECK();
  int64_t R = 0;
  SG_HIP(hipMemcpyAsync(&R, kr + G - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  int64_t nnz = 0;
  if (R > 0) {
    rc = snapgpu_scan_inclusive_i64(rowlen, R, sws, sws_bytes, stream);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(&nnz, rowlen + R - 1, 8, hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
  }
  h_sum->n_rows = R;
  h_sum->nnz = nnz;
  return SNAPGPU_OK;
}

int snapgpu_frag_fill(const uint8_t* d_text, int64_t n_bytes, const void* d_ws, int64_t ws_bytes,
                      const snapgpu_frag_summary* h_sum, int64_t* d_cell_ptr, uint32_t* d_pos, uint32_t* d_val,
                      uint64_t* d_qc, uint8_t* d_names, int32_t name_width, void* stream) {
  clear_error();
  SG_ARG(h_sum != nullptr && d_cell_ptr != nullptr, "bad arguments");
  const Layout Lo = make_layout(n_bytes, h_sum->n_lines);
  SG_ARG(ws_bytes >= Lo.total, "workspace too small");
  SG_ARG(d_names == nullptr || name_width >= h_sum->name_width, "name_width below the longest barcode");
  SG_ARG(h_sum->nnz == 0 || (d_pos != nullptr && d_val != nullptr), "null outputs");
  hipStream_t st = as_stream(stream);
  const int64_t R = h_sum->n_rows, K = h_sum->n_kept;
  if (R == 0) {
    const int64_t zero = 0;
    SG_HIP(hipMemcpyAsync(d_cell_ptr, &zero, 8, hipMemcpyHostToDevice, st));
    SG_HIP(hipStreamSynchronize(st));
    return SNAPGPU_OK;
  }
  void* w = const_cast<void*>(d_ws);
  fill_kernel<<<grid(K), kThreads, 0, st>>>(at<uint32_t>(w, Lo.permA), at<uint32_t>(w, Lo.kept),
                                            at<uint8_t>(w, Lo.flags), at<uint32_t>(w, Lo.pos),
                                            at<uint32_t>(w, Lo.size), at<int64_t>(w, Lo.gid),
                                            at<int64_t>(w, Lo.gstart), at<int64_t>(w, Lo.rowid),
                                            at<int64_t>(w, Lo.rowlen), K, h_sum->paired, d_pos, d_val);
  SG_LAUNCH_CHECK();
  rows_out_kernel<<<grid(R), kThreads, 0, st>>>(
      d_text, at<uint32_t>(w, Lo.rowgroup), at<uint32_t>(w, Lo.gline), at<uint64_t>(w, Lo.bc),
      at<int64_t>(w, Lo.rowlen), at<unsigned long long>(w, Lo.guniq), at<unsigned long long>(w, Lo.gtot),
      at<unsigned long long>(w, Lo.gmito), R, d_cell_ptr, reinterpret_cast<unsigned long long*>(d_qc), d_names,
      name_width);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_frag_format_workspace_bytes(int64_t n_fragments) {
  return snapgpu_scan_workspace_bytes(n_fragments > 0 ? n_fragments : 1);
}

int snapgpu_frag_format(const int64_t* d_cell_ptr, int64_t n_cells, const uint32_t* d_pos, const uint32_t* d_val,
                        int64_t n_fragments, int32_t paired, const void* d_chroms, const uint8_t* d_names,
                        int32_t name_width, const int64_t* d_order, int64_t* d_line_end, uint8_t* d_text,
                        int64_t text_bytes, void* d_ws, int64_t ws_bytes, int64_t* h_text_bytes, void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && n_fragments >= 0 && d_chroms != nullptr && h_text_bytes != nullptr, "bad arguments");
  SG_ARG(name_width > 0 && d_names != nullptr, "names required");
  SG_ARG(ws_bytes >= snapgpu_frag_format_workspace_bytes(n_fragments), "workspace too small");
  *h_text_bytes = 0;
  if (n_fragments == 0) return SNAPGPU_OK;
  SG_ARG(n_cells > 0 && d_cell_ptr && d_pos && d_val && d_line_end, "null inputs");
  hipStream_t st = as_stream(stream);
  const uint8_t* chroms = static_cast<const uint8_t*>(d_chroms);
  if (d_text == nullptr) {
    fmt_len_kernel<<<grid(n_fragments), kThreads, 0, st>>>(d_cell_ptr, n_cells, d_pos, d_val, paired, chroms,
                                                           d_names, name_width, d_order, n_fragments, d_line_end);
    SG_LAUNCH_CHECK();
    int rc = snapgpu_scan_inclusive_i64(d_line_end, n_fragments, d_ws, ws_bytes, stream);
    if (rc) return rc;
  }
  int64_t total = 0;
  SG_HIP(hipMemcpyAsync(&total, d_line_end + n_fragments - 1, 8, hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  *h_text_bytes = total;
  if (d_text != nullptr) {
    SG_ARG(text_bytes >= total, "text buffer too small");
    fmt_write_kernel<<<grid(n_fragments), kThreads, 0, st>>>(d_cell_ptr, n_cells, d_pos, d_val, paired, chroms,
                                                             d_names, name_width, d_order, n_fragments, d_line_end,
                                                             d_text);
    SG_LAUNCH_CHECK();
  }
  return SNAPGPU_OK;
}

}  // extern "C"
