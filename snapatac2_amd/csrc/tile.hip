// Path A: fragments -> cell x bin CSR (the tile matrix).
//
// Reference semantics (snapatac2-core/src/feature_count/data_iter.rs):
//   gen_mat_pair (:346-415): per fragment (chrom, start) = get_position(col),
//     end = start + len - 1 (:371); skipped if the chromosome is excluded or the
//     length is outside [min, max] (:372-374); b1 = bin(start), b2 = bin(end);
//     Insertion: +1 b1, +1 b2 (:379-386); Fragment: +1 on b1..=b2 (:387-394);
//     PIC: +1 b1, +1 b2 if b2 != b1 (:395-406). Per row the counts are kept in a
//     BTreeMap, so the row is emitted with ascending bins and no zeros.
//   gen_mat_single (:312-344): +1 at the bin of the stored position.
//   get_position_rev panics when the position is past the chromosome end
//     (genome.rs:397-399); here that is reported as SNAPGPU_ERR_DATA.
//
// Design (no sort of the whole row, no hash map): one wavefront owns one cell
// and walks its fragments 64 at a time. Rows are sorted by start position
// (import.rs:208), so start bins never decrease and every event of a later
// fragment lies at or above the start bin of that fragment: once a tile's
// lower bound T for all future events is known, every merged (bin, count) below
// T is final and is streamed out; bins >= T wait in a small LDS "pending" list
// (at most max-span-in-bins + 1 entries). Within a tile the <= 128 new events
// are put in order (directly by rank when the fragments' bins do not
// interleave, the usual case, else by a register bitonic network over the
// wave), run-length encoded with ballots, and merged with the pending list by
// rank (binary search in LDS). The same code runs twice: pass 1 only counts, pass 2 writes.
#include <hipcub/hipcub.hpp>
#include <algorithm>

#include "common.h"

namespace snapgpu {
namespace tile {

constexpr uint32_t kInf = 0xFFFFFFFFu;

// Packed chromosome record (16 bytes). Entry n_chrom is a sentinel whose
// `start` is the genome length.
struct ChromRec {
  uint64_t start;    // first genome base index of the chromosome
  uint32_t bin_off;  // kept: first output bin; excluded: first bin of the next kept chromosome
  uint32_t kept;     // 1 = counted, 0 = excluded
};
static_assert(sizeof(ChromRec) == 16, "ChromRec layout");

struct Params {
  const uint64_t* cell_ptr;
  const uint32_t* pos;
  const uint32_t* len;  // null for single-end input
  const uint32_t* order;  // processing order of the cells (null = 0, 1, ...)
  int64_t n_cells;
  const ChromRec* table;
  int32_t n_chrom;
  uint32_t bin_size;
  uint64_t bin_magic;  // floor(2^64 / bin_size) + 1, for exact division of 32-bit offsets
  int32_t strategy;
  uint32_t min_len, max_len;
  int64_t* row_nnz;        // pass 1
  const int64_t* indptr;   // pass 2
  void* out_idx;           // pass 2
  uint32_t* out_val;       // pass 2
  int32_t* status;         // [0]=data error code, [1]=cell lo, [2]=cell hi, [3]=detail, [4]=capacity flag
};

enum : int32_t {
  kErrPosPastGenome = 1,
  kErrEndPastChrom = 2,
  kErrZeroLength = 3,
  kErrFillMismatch = 5,
  kErrSpan = 6,
  kErrUnsorted = 7,
};

// "fragment" counting expands every covered bin; longer spans are rejected.
constexpr uint32_t kMaxFragmentSpanBins = 1u << 20;

__device__ __forceinline__ uint32_t div_bin(uint32_t off, const Params& p) {
  return p.bin_size == 1 ? off : (uint32_t)__umul64hi((uint64_t)off, p.bin_magic);
}

__device__ void report(const Params& p, int32_t code, int64_t cell, int32_t detail) {
  if (atomicCAS(&p.status[0], 0, code) == 0) {
    p.status[1] = (int32_t)(cell & 0xFFFFFFFF);
    p.status[2] = (int32_t)(cell >> 32);
    p.status[3] = detail;
  }
}

// Chromosome containing genome index `pos` (< genome length): largest c with
// start[c] <= pos. Equivalent to GenomeBaseIndex::get_position (genome.rs:419-434).
__device__ __forceinline__ int find_chrom(const ChromRec* tab, int n_chrom, uint64_t pos, int hint) {
  if (tab[hint].start <= pos && pos < tab[hint + 1].start) return hint;
  int lo = 0, hi = n_chrom;  // answer in [lo, hi)
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (tab[mid].start <= pos) lo = mid; else hi = mid;
  }
  return lo;
}

// Lower bound on every output bin of fragments starting at or after genome index `pos`.
__device__ __forceinline__ uint32_t lower_bound_bin(const ChromRec* tab, int n_chrom, uint64_t pos,
                                                    int hint, const Params& p) {
  if (pos >= tab[n_chrom].start) return kInf;  // invalid position; reported when it is processed
  int c = find_chrom(tab, n_chrom, pos, hint);
  if (!tab[c].kept) return tab[c].bin_off;
  return tab[c].bin_off + div_bin((uint32_t)(pos - tab[c].start), p);
}

// Per-wave LDS scratch.
template <int CAP>
struct WaveLds {
  uint32_t pbin[CAP];        // pending bins (sorted, distinct)
  uint32_t pcnt[CAP];
  uint32_t nbin[128];        // run starts (bins) of one round's new events
  uint32_t mbin[CAP + 128];  // merge buffer
  uint32_t mcnt[CAP + 128];
  uint64_t fscan[64];        // fragment strategy: inclusive scan of events per lane
  uint32_t fb1[64];          // fragment strategy: first bin per lane
};

template <int CAP, bool FILL, typename IdxT>
struct CellState {
  WaveLds<CAP>* L;
  int np;            // pending entries
  int64_t emitted;   // entries streamed out so far
  int64_t out_base;  // FILL: indptr[cell]
  int64_t out_end;   // FILL: indptr[cell+1]
  bool overflow;
};

template <int CAP, bool FILL, typename IdxT>
__device__ __forceinline__ void emit(CellState<CAP, FILL, IdxT>& s, const Params& p, int64_t slot,
                                     uint32_t bin, uint32_t cnt) {
  if (FILL) {
    int64_t o = s.out_base + slot;
    if (o < s.out_end) {
      static_cast<IdxT*>(p.out_idx)[o] = (IdxT)bin;
      p.out_val[o] = cnt;
    }
  }
}

// Bitonic sort of 128 keys held as k[0] (elements 0..63) and k[1] (64..127).
__device__ __forceinline__ void bitonic128(uint32_t (&k)[2], int lane) {
#pragma unroll
  for (int size = 2; size <= 128; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      if (j == 64) {
        uint32_t a = k[0], b = k[1];
        k[0] = a < b ? a : b;
        k[1] = a < b ? b : a;
      } else {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const int e = r * 64 + lane;
          uint32_t x = k[r];
          uint32_t y = __shfl_xor(x, j, 64);
          const bool asc = (e & size) == 0;
          const bool low = (e & j) == 0;
          uint32_t mn = x < y ? x : y, mx = x < y ? y : x;
          k[r] = (low == asc) ? mn : mx;
        }
      }
    }
  }
}

// Number of entries below key (STRICT) or at or below key in a sorted LDS
// array of n <= CAP entries (CAP a power of two): a fixed search of
// log2(CAP) + 1 steps, the same for every lane, so no divergent loop.
template <bool STRICT, int CAP>
__device__ __forceinline__ int lds_count(const uint32_t* arr, int n, uint32_t key) {
  static_assert((CAP & (CAP - 1)) == 0, "CAP must be a power of two");
  int lo = 0;
#pragma unroll
  for (int step = CAP; step > 0; step >>= 1) {
    if (lo + step <= n) {
      const uint32_t v = arr[lo + step - 1];
      if (STRICT ? v < key : v <= key) lo += step;
    }
  }
  return lo;
}

// Fast paths for the per-fragment event pairs (k[0] = first event of lane i's
// fragment, k[1] = its second event or kInf; kInf in k[0] = no event). Start
// bins never decrease along a row and the second event is >= the first, so the
// pair order (lane 0 first, lane 0 second, lane 1 first, ...) is sorted unless
// some fragment ends in a later bin than where the next fragment starts. When
// it is sorted, the events are placed at their ranks. Otherwise (several
// fragments at one locus, most tiles of real and generated rows) the first
// events still ascend along the lanes and the second events usually do too:
// then the two ascending lists are merged by rank (a first event before a
// second event of the same bin). Either way the events end sorted (elements
// 0..63 in k[0], 64..127 in k[1], kInf last) and true is returned; when neither
// list order holds, k is unchanged and false is returned (bitonic sort).
template <int CAP>
__device__ __forceinline__ bool presort_pairs(uint32_t (&k)[2], int lane, WaveLds<CAP>* L) {
  const bool v0 = k[0] != kInf, v1 = k[1] != kInf;
  const uint64_t m0 = ballot64(v0), m1 = ballot64(v1);
  const uint64_t later = m0 & lanemask_gt(lane);
  const int nl = later ? lowbit64(later) : lane;
  const uint32_t nxt = __shfl(k[0], nl, 64);
  const uint32_t last = v1 ? k[1] : k[0];
  if (ballot64(v0 && later != 0 && last > nxt) != 0) {
    const uint64_t later1 = m1 & lanemask_gt(lane);
    const uint32_t nxt1 = __shfl(k[1], later1 ? lowbit64(later1) : lane, 64);
    if (ballot64((v0 && later != 0 && k[0] > nxt) || (v1 && later1 != 0 && k[1] > nxt1)) != 0) return false;
    const uint64_t lt = lanemask_lt(lane);
    const int n0 = popc64(m0), n1 = popc64(m1);
    const int i0 = popc64(m0 & lt), i1 = popc64(m1 & lt);
    if (v0) L->nbin[i0] = k[0];
    if (v1) L->nbin[64 + i1] = k[1];
    wave_sync();
    const int r0 = v0 ? i0 + lds_count<true, 64>(L->nbin + 64, n1, k[0]) : 0;
    const int r1 = v1 ? i1 + lds_count<false, 64>(L->nbin, n0, k[1]) : 0;
    L->mbin[lane] = kInf;
    L->mbin[64 + lane] = kInf;
    wave_sync();
    if (v0) L->mbin[r0] = k[0];
    if (v1) L->mbin[r1] = k[1];
    wave_sync();
    k[0] = L->mbin[lane];
    k[1] = L->mbin[64 + lane];
    wave_sync();
    return true;
  }
  const uint64_t lt = lanemask_lt(lane);
  const int r0 = popc64(m0 & lt) + popc64(m1 & lt);
  L->nbin[lane] = kInf;
  L->nbin[64 + lane] = kInf;
  wave_sync();
  if (v0) L->nbin[r0] = k[0];
  if (v1) L->nbin[r0 + 1] = k[1];
  wave_sync();
  k[0] = L->nbin[lane];
  k[1] = L->nbin[64 + lane];
  wave_sync();
  return true;
}

// Merges one round of up to 128 new unit events with the pending list, streams out
// every merged bin below T and keeps the rest pending.
template <int CAP, bool FILL, typename IdxT>
__device__ void merge_round(CellState<CAP, FILL, IdxT>& s, const Params& p, uint32_t (&k)[2],
                            uint32_t T, int lane, bool sorted = false) {
  WaveLds<CAP>* L = s.L;
  const uint64_t any = ballot64(k[0] != kInf) | ballot64(k[1] != kInf);
  if (any == 0 && s.np == 0) return;

  if (!sorted) bitonic128(k, lane);

  // Run-length encode the sorted events (element order: k[0] lanes 0..63, then k[1]).
  const uint32_t prev0 = __shfl_up(k[0], 1, 64);
  const uint32_t last0 = __shfl(k[0], 63, 64);
  uint32_t prev1 = __shfl_up(k[1], 1, 64);
  if (lane == 0) prev1 = last0;
  const bool st0 = k[0] != kInf && (lane == 0 || k[0] != prev0);
  const bool st1 = k[1] != kInf && k[1] != prev1;
  const uint64_t S0 = ballot64(st0), S1 = ballot64(st1);
  const int nv = popc64(ballot64(k[0] != kInf)) + popc64(ballot64(k[1] != kInf));
  const int n0 = popc64(S0);
  const int nN = n0 + popc64(S1);
  const uint64_t lt = lanemask_lt(lane), gt = lanemask_gt(lane);
  int idx0 = popc64(S0 & lt), idx1 = n0 + popc64(S1 & lt);
  // run lengths (the counts) only matter when writing; counting needs the runs alone
  uint32_t cnt0 = 0, cnt1 = 0;
  if (FILL && st0) {
    uint64_t m = S0 & gt;
    int nxt = m ? lowbit64(m) : (S1 ? 64 + lowbit64(S1) : nv);
    cnt0 = (uint32_t)(nxt - lane);
  }
  if (FILL && st1) {
    uint64_t m = S1 & gt;
    int nxt = m ? 64 + lowbit64(m) : nv;
    cnt1 = (uint32_t)(nxt - (64 + lane));
  }

  if (s.np == 0) {
    // Fast path: the new runs are the merged list.
    const bool f0 = st0 && k[0] < T, f1 = st1 && k[1] < T;
    const uint64_t F0 = ballot64(f0), F1 = ballot64(f1);
    const int nf0 = popc64(F0);
    const int nf = nf0 + popc64(F1);
    if (f0) emit(s, p, s.emitted + popc64(F0 & lt), k[0], cnt0);
    if (f1) emit(s, p, s.emitted + nf0 + popc64(F1 & lt), k[1], cnt1);
    // pending: starts with key >= T, ranked after the finals
    const int np = nN - nf;
    if (np > CAP) { s.overflow = true; return; }
    if (st0 && !f0) { L->pbin[idx0 - nf] = k[0]; if (FILL) L->pcnt[idx0 - nf] = cnt0; }
    if (st1 && !f1) { L->pbin[idx1 - nf] = k[1]; if (FILL) L->pcnt[idx1 - nf] = cnt1; }
    s.emitted += nf;
    s.np = np;
    wave_sync();
    return;
  }

  // General path: merge runs N with pending P by rank into M.
  if (st0) L->nbin[idx0] = k[0];
  if (st1) L->nbin[idx1] = k[1];
  wave_sync();
  const int np = s.np;
  if (st0) {
    int r = idx0 + lds_count<false, CAP>(L->pbin, np, k[0]);
    L->mbin[r] = k[0];
    if (FILL) L->mcnt[r] = cnt0;
  }
  if (st1) {
    int r = idx1 + lds_count<false, CAP>(L->pbin, np, k[1]);
    L->mbin[r] = k[1];
    if (FILL) L->mcnt[r] = cnt1;
  }
  for (int j = lane; j < np; j += 64) {
    uint32_t b = L->pbin[j];
    int r = j + lds_count<true, 128>(L->nbin, nN, b);
    L->mbin[r] = b;
    if (FILL) L->mcnt[r] = L->pcnt[j];
  }
  wave_sync();
  // Equal bins appear as adjacent pairs (pending first); fold them and split at T.
  const int nM = np + nN;
  int run_final = 0, run_pend = 0;
  for (int base = 0; base < nM; base += 64) {
    const int sidx = base + lane;
    bool st = false, fin = false;
    uint32_t b = kInf, c = 0;
    if (sidx < nM) {
      b = L->mbin[sidx];
      st = sidx == 0 || L->mbin[sidx - 1] != b;
      if (st) {
        if (FILL) {
          c = L->mcnt[sidx];
          if (sidx + 1 < nM && L->mbin[sidx + 1] == b) c += L->mcnt[sidx + 1];
        }
        fin = b < T;
      }
    }
    const uint64_t F = ballot64(st && fin), P = ballot64(st && !fin);
    if (st && fin) emit(s, p, s.emitted + run_final + popc64(F & lt), b, c);
    const int pr = run_pend + popc64(P & lt);
    if (st && !fin && pr < CAP) {
      L->pbin[pr] = b;
      if (FILL) L->pcnt[pr] = c;
    }
    run_final += popc64(F);
    run_pend += popc64(P);
  }
  if (run_pend > CAP) { s.overflow = true; return; }
  s.emitted += run_final;
  s.np = run_pend;
  wave_sync();
}

template <int CAP, bool FILL, typename IdxT>
__device__ void process_cell(const Params& p, const ChromRec* tab, WaveLds<CAP>* L, int64_t cell,
                             int lane) {
  const int64_t f0 = (int64_t)p.cell_ptr[cell], f1 = (int64_t)p.cell_ptr[cell + 1];
  CellState<CAP, FILL, IdxT> s;
  s.L = L;
  s.np = 0;
  s.emitted = 0;
  s.overflow = false;
  s.out_base = FILL ? p.indptr[cell] : 0;
  s.out_end = FILL ? p.indptr[cell + 1] : 0;
  const int nc = p.n_chrom;
  const uint64_t genome = tab[nc].start;
  const bool paired = p.strategy != SNAPGPU_SINGLE;
  int hint = 0;

  // the next tile's positions and lengths are loaded one tile ahead (clamped
  // to the cell's last fragment, so every load is in bounds and unconditional)
  uint32_t ps_n = 0u, ln_n = 1u;
  uint32_t prev_last = 0u;  // position of the previous tile's last fragment
  if (f0 < f1) {
    const int64_t fc = min(f0 + lane, f1 - 1);
    ps_n = p.pos[fc];
    if (paired) ln_n = p.len[fc];
  }
  for (int64_t t0 = f0; t0 < f1; t0 += 64) {
    const int64_t f = t0 + lane;
    const bool has = f < f1;
    const uint32_t ps = has ? ps_n : 0u;
    const uint32_t ln = (has && paired) ? ln_n : 1u;
    {
      const int64_t fc = min(f + 64, f1 - 1);
      ps_n = p.pos[fc];
      if (paired) ln_n = p.len[fc];
    }
    // Positions must not decrease along the row (the final-bin bound T below
    // relies on it); a row out of order is reported, and the caller sorts it.
    {
      uint32_t prev = (uint32_t)__shfl_up((int)ps, 1, 64);
      if (lane == 0) prev = prev_last;
      if (ballot64(has && ps < prev) != 0) {
        if (lane == 0) report(p, kErrUnsorted, cell, 0);
        s.np = 0;
        break;
      }
      prev_last = (uint32_t)__shfl((int)ps, (int)min<int64_t>(63, f1 - 1 - t0), 64);
    }
    bool kept = false;
    uint32_t b1 = kInf, b2 = kInf;
    if (has) {
      if ((uint64_t)ps >= genome) {
        report(p, kErrPosPastGenome, cell, (int32_t)lane);
      } else {
        const int c = find_chrom(tab, nc, ps, hint);
        hint = c;  // per-lane hint; rows are sorted so it is usually right
        const ChromRec rc = tab[c];
        bool in_size = !paired || (ln >= p.min_len && ln <= p.max_len);
        if (rc.kept && in_size) {
          const uint32_t off = (uint32_t)(ps - rc.start);
          b1 = rc.bin_off + div_bin(off, p);
          if (paired) {
            const uint64_t chrom_len = tab[c + 1].start - rc.start;
            if (ln == 0) {
              report(p, kErrZeroLength, cell, c);
            } else if ((uint64_t)off + ln - 1 >= chrom_len) {
              report(p, kErrEndPastChrom, cell, c);
            } else {
              b2 = rc.bin_off + div_bin(off + ln - 1, p);
              kept = true;
            }
          } else {
            kept = true;
          }
        }
      }
    }
    // Lower bound T of every event of the fragments after this tile.
    uint32_t T = kInf;
    if (t0 + 64 < f1) {
      const uint32_t pn = (uint32_t)__shfl((int)ps_n, 0, 64);  // position of fragment t0 + 64
      T = lower_bound_bin(tab, nc, pn, hint, p);
    }

    if (p.strategy != SNAPGPU_FRAGMENT) {
      uint32_t k[2];
      k[0] = kept ? b1 : kInf;
      if (p.strategy == SNAPGPU_INSERTION) k[1] = kept ? b2 : kInf;
      else if (p.strategy == SNAPGPU_PIC) k[1] = (kept && b2 != b1) ? b2 : kInf;
      else k[1] = kInf;
      const bool sorted = presort_pairs(k, lane, L);
      merge_round(s, p, k, T, lane, sorted);
      if (s.overflow) break;
    } else {
      // Every bin b1..=b2 gets +1. The events of the tile are laid out lane by
      // lane (ascending start bins) and handled 128 at a time.
      if (kept && b2 - b1 >= kMaxFragmentSpanBins) {
        report(p, kErrSpan, cell, (int32_t)(b2 - b1));
        kept = false;
      }
      const uint64_t nev = kept ? (uint64_t)(b2 - b1) + 1 : 0;
      const uint64_t incl = wave_incl_scan_u64(nev, lane);
      const uint64_t total = shfl_u64(incl, 63);
      L->fscan[lane] = incl;
      L->fb1[lane] = kept ? b1 : kInf;
      wave_sync();
      const uint64_t keptmask = ballot64(kept);
      // At least one round per tile, so pending bins below T are flushed even
      // when every fragment of the tile is excluded or filtered.
      const uint64_t span = total > 0 ? total : 1;
      for (uint64_t g0 = 0; g0 < span; g0 += 128) {
        uint32_t k[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const uint64_t g = g0 + (uint64_t)(r * 64 + lane);
          k[r] = kInf;
          if (g < total) {
            int lo = 0, hi = 63;  // first lane with incl > g
            while (lo < hi) {
              int mid = (lo + hi) >> 1;
              if (L->fscan[mid] > g) hi = mid; else lo = mid + 1;
            }
            const uint64_t excl = lo == 0 ? 0 : L->fscan[lo - 1];
            k[r] = L->fb1[lo] + (uint32_t)(g - excl);
          }
        }
        uint32_t Tr = T;
        const uint64_t gn = g0 + 128;
        if (gn < total) {
          int lo = 0, hi = 63;
          while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (L->fscan[mid] > gn) hi = mid; else lo = mid + 1;
          }
          const uint64_t excl = lo == 0 ? 0 : L->fscan[lo - 1];
          uint32_t kn = L->fb1[lo] + (uint32_t)(gn - excl);
          const uint64_t later = keptmask & (lo == 63 ? 0ull : (~0ull << (lo + 1)));
          if (later) {
            uint32_t bl = L->fb1[lowbit64(later)];
            kn = kn < bl ? kn : bl;
          }
          Tr = Tr < kn ? Tr : kn;
        }
        merge_round(s, p, k, Tr, lane);
        if (s.overflow) break;
      }
      wave_sync();
      if (s.overflow) break;
    }
  }
  if (!s.overflow && s.np > 0) {
    // End of the row: everything still pending is final.
    uint32_t k[2] = {kInf, kInf};
    merge_round(s, p, k, kInf, lane);
  }
  if (s.overflow) {
    if (lane == 0) atomicOr(&p.status[4], 1);
    return;
  }
  if (!FILL) {
    if (lane == 0) p.row_nnz[cell] = s.emitted;
  } else {
    if (lane == 0 && s.emitted != s.out_end - s.out_base) report(p, kErrFillMismatch, cell, 0);
  }
}

template <int PS, int WAVES, bool FILL, typename IdxT>
__global__ void __launch_bounds__(WAVES * 64) tile_kernel(Params p) {
  constexpr int CAP = 64 * PS;
  extern __shared__ __align__(16) unsigned char smem[];
  WaveLds<CAP>* waves = reinterpret_cast<WaveLds<CAP>*>(smem);
  ChromRec* tab = reinterpret_cast<ChromRec*>(smem + sizeof(WaveLds<CAP>) * WAVES);
  for (int i = threadIdx.x; i <= p.n_chrom; i += blockDim.x) tab[i] = p.table[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t slot = (int64_t)blockIdx.x * WAVES + wid;
  if (slot >= p.n_cells) return;
  const int64_t cell = p.order ? (int64_t)p.order[slot] : slot;
  process_cell<CAP, FILL, IdxT>(p, tab, &waves[wid], cell, lane);
}

template <int PS, int WAVES, bool FILL, typename IdxT>
static int launch(const Params& p, hipStream_t st) {
  constexpr int CAP = 64 * PS;
  const size_t lds = sizeof(WaveLds<CAP>) * WAVES + sizeof(ChromRec) * (size_t)(p.n_chrom + 1);
  if (lds > 160 * 1024) {
    set_error("chromosome table too large for LDS");
    return SNAPGPU_ERR_CAPACITY;
  }
  const int64_t blocks = ceil_div(p.n_cells, WAVES);
  SG_ARG(blocks < (1ll << 31), "too many cells for one launch");
  hipLaunchKernelGGL((tile_kernel<PS, WAVES, FILL, IdxT>), dim3((unsigned)blocks), dim3(WAVES * 64),
                     lds, st, p);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

template <bool FILL, typename IdxT>
static int launch_level(const Params& p, int level, hipStream_t st) {
  switch (level) {
    case 0: return launch<1, 4, FILL, IdxT>(p, st);
    case 1: return launch<4, 4, FILL, IdxT>(p, st);
    case 2: return launch<16, 2, FILL, IdxT>(p, st);
    case 3: return launch<64, 1, FILL, IdxT>(p, st);
    default: set_error("invalid capacity level"); return SNAPGPU_ERR_ARG;
  }
}
constexpr int kMaxLevel = 3;

static int check_status(const int32_t* d_status, hipStream_t st, bool* overflow) {
  int32_t h[8];
  SG_HIP(hipMemcpyAsync(h, d_status, sizeof(h), hipMemcpyDeviceToHost, st));
  SG_HIP(hipStreamSynchronize(st));
  if (h[0] != 0) {
    const int64_t cell = (int64_t)(uint32_t)h[1] | ((int64_t)h[2] << 32);
    std::string msg;
    switch (h[0]) {
      case kErrPosPastGenome:
        msg = "fragment position is beyond the end of the genome (cell " + std::to_string(cell) + ")";
        break;
      case kErrEndPastChrom:
        msg = "Position is out of range for chromosome index " + std::to_string(h[3]) +
              " (fragment end past the chromosome end, cell " + std::to_string(cell) + ")";
        break;
      case kErrZeroLength:
        msg = "zero-length fragment in cell " + std::to_string(cell) +
              " (import_fragments drops these)";
        break;
      case kErrSpan:
        msg = "a fragment covers " + std::to_string(h[3]) +
              " bins (limit 2^20 for counting_strategy='fragment'); use a larger bin_size";
        break;
      case kErrUnsorted:
        set_error("the fragments of cell " + std::to_string(cell) +
                  " are not sorted by position; sort the rows (snapgpu_frag_sort_rows) and call again");
        return SNAPGPU_ERR_UNSORTED;
      case kErrFillMismatch:
        msg = "tile fill produced a row length different from the count pass (cell " +
              std::to_string(cell) + ")";
        break;
      default: msg = "tile kernel error " + std::to_string(h[0]);
    }
    set_error(msg);
    return SNAPGPU_ERR_DATA;
  }
  *overflow = h[4] != 0;
  return SNAPGPU_OK;
}

// Sort keys for the largest-first order: one wave walks one cell, so a dispatch
// lasts at least as long as its longest cell; started first, the long cells run
// beside the bulk instead of forming a tail.
__global__ void order_keys_kernel(const uint64_t* cell_ptr, int64_t n, uint32_t* keys, uint32_t* ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t len = cell_ptr[i + 1] - cell_ptr[i];
  keys[i] = ~(uint32_t)min(len, (uint64_t)0xFFFFFFFFu);  // ascending keys = descending sizes
  ids[i] = (uint32_t)i;
}

// (position, value) as one 64-bit key, so a key sort orders a row by position
// and then value; single-end rows (no values) sort by position alone.
__global__ void pack_sort_keys_kernel(const uint32_t* pos, const uint32_t* len, int64_t n, uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = ((uint64_t)pos[i] << 32) | (len ? len[i] : 0u);
}

__global__ void unpack_sort_keys_kernel(const uint64_t* keys, int64_t n, uint32_t* pos, uint32_t* len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  pos[i] = (uint32_t)(k >> 32);
  if (len) len[i] = (uint32_t)k;
}

static int make_params(Params& p, const uint64_t* cell_ptr, const uint32_t* pos, const uint32_t* len,
                       int64_t n_cells, const void* table, int32_t n_chrom, uint32_t bin_size,
                       int32_t strategy, uint32_t min_len, uint32_t max_len, int32_t* status) {
  SG_ARG(n_cells >= 0, "n_cells < 0");
  SG_ARG(n_chrom >= 1, "n_chrom < 1");
  SG_ARG(bin_size >= 1, "bin_size must be >= 1");
  SG_ARG(strategy >= 0 && strategy <= 3, "unknown counting strategy");
  SG_ARG(strategy == SNAPGPU_SINGLE || len != nullptr, "paired input needs lengths");
  SG_ARG(status != nullptr && table != nullptr && cell_ptr != nullptr, "null pointer");
  p = Params{};
  p.cell_ptr = cell_ptr;
  p.pos = pos;
  p.len = len;
  p.n_cells = n_cells;
  p.table = static_cast<const ChromRec*>(table);
  p.n_chrom = n_chrom;
  p.bin_size = bin_size;
  p.bin_magic = bin_size > 1 ? (~0ull / bin_size) + 1 : 0;
  p.strategy = strategy;
  p.min_len = min_len;
  p.max_len = max_len;
  p.status = status;
  return SNAPGPU_OK;
}

}  // namespace tile
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::tile;

extern "C" {

int64_t snapgpu_tile_table_bytes(int32_t n_chrom) {
  return n_chrom < 0 ? 0 : (int64_t)sizeof(ChromRec) * (n_chrom + 1);
}

int snapgpu_tile_table_pack(const uint64_t* h_chrom_len, const uint8_t* h_excluded, int32_t n_chrom,
                            uint32_t bin_size, void* h_out, uint64_t* n_bins) {
  clear_error();
  SG_ARG(n_chrom >= 1, "n_chrom < 1");
  SG_ARG(bin_size >= 1, "bin_size must be >= 1");
  SG_ARG(h_chrom_len && h_excluded && h_out && n_bins, "null pointer");
  ChromRec* t = static_cast<ChromRec*>(h_out);
  uint64_t acc = 0, bins = 0;
  for (int c = 0; c < n_chrom; ++c) {
    t[c].start = acc;
    t[c].kept = h_excluded[c] ? 0u : 1u;
    t[c].bin_off = (uint32_t)bins;
    if (t[c].kept) bins += (h_chrom_len[c] + bin_size - 1) / bin_size;
    acc += h_chrom_len[c];
    SG_ARG(bins < kInf, "number of bins does not fit in 32 bits");
  }
  SG_ARG(acc <= 0xFFFFFFFFull, "genome longer than 2^32 - 1 bases (positions are uint32)");
  t[n_chrom].start = acc;
  t[n_chrom].bin_off = (uint32_t)bins;
  t[n_chrom].kept = 0;
  // excluded chromosomes: first bin of the next kept chromosome (a lower bound)
  for (int c = n_chrom - 1; c >= 0; --c)
    if (!t[c].kept) t[c].bin_off = t[c + 1].bin_off;
  *n_bins = bins;
  return SNAPGPU_OK;
}

int64_t snapgpu_tile_order_workspace_bytes(int64_t n_cells) {
  if (n_cells <= 0) return 256;
  size_t tmp = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (int)std::min<int64_t>(n_cells, INT32_MAX));
  return 3 * ((n_cells * 4 + 255) / 256 * 256) + (int64_t)tmp + 256;
}

int snapgpu_tile_order(const uint64_t* d_cell_ptr, int64_t n_cells, uint32_t* d_order, void* d_ws,
                       int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && n_cells < (1ll << 31), "bad cell count");
  if (n_cells == 0) return SNAPGPU_OK;
  SG_ARG(d_cell_ptr && d_order && d_ws, "null pointer");
  SG_ARG(ws_bytes >= snapgpu_tile_order_workspace_bytes(n_cells), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  const int64_t arr = (n_cells * 4 + 255) / 256 * 256;
  uint32_t* keys = (uint32_t*)q;
  uint32_t* keys2 = (uint32_t*)(q + arr);
  uint32_t* ids = (uint32_t*)(q + 2 * arr);
  void* tmp = q + 3 * arr;
  size_t tmp_bytes = (size_t)(ws_bytes - (3 * arr + (q - (char*)d_ws)));
  order_keys_kernel<<<(unsigned)ceil_div(n_cells, 256), 256, 0, st>>>(d_cell_ptr, n_cells, keys, ids);
  SG_LAUNCH_CHECK();
  SG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, ids, d_order, (int)n_cells, 0, 32,
                                            st));
  return SNAPGPU_OK;
}

int64_t snapgpu_frag_sort_rows_workspace_bytes(int64_t n_cells, int64_t n_fragments) {
  const int64_t f = std::max<int64_t>(n_fragments, 1);
  size_t tmp = 0;
  (void)hipcub::DeviceSegmentedSort::SortKeys(nullptr, tmp, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                              (int)std::min<int64_t>(f, INT32_MAX),
                                              (int)std::min<int64_t>(std::max<int64_t>(n_cells, 1), INT32_MAX),
                                              (const uint64_t*)nullptr, (const uint64_t*)nullptr);
  return 2 * ((f * 8 + 255) / 256 * 256) + (int64_t)tmp + 256;
}

int snapgpu_frag_sort_rows(const uint64_t* d_cell_ptr, int64_t n_cells, int64_t n_fragments, uint32_t* d_pos,
                           uint32_t* d_len, void* d_ws, int64_t ws_bytes, void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && n_cells < INT32_MAX, "bad cell count");
  SG_ARG(n_fragments >= 0 && n_fragments < INT32_MAX, "n_fragments must be below 2^31 (sort chunks of rows)");
  if (n_cells == 0 || n_fragments == 0) return SNAPGPU_OK;
  SG_ARG(d_cell_ptr && d_pos && d_ws, "null pointer");
  SG_ARG(ws_bytes >= snapgpu_frag_sort_rows_workspace_bytes(n_cells, n_fragments), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* q = (char*)(((uintptr_t)d_ws + 255) & ~(uintptr_t)255);
  const int64_t arr = (n_fragments * 8 + 255) / 256 * 256;
  uint64_t* keys = (uint64_t*)q;
  uint64_t* keys2 = (uint64_t*)(q + arr);
  void* tmp = q + 2 * arr;
  size_t tmp_bytes = (size_t)(ws_bytes - (2 * arr + (q - (char*)d_ws)));
  const unsigned g = (unsigned)ceil_div(n_fragments, 256);
  pack_sort_keys_kernel<<<g, 256, 0, st>>>(d_pos, d_len, n_fragments, keys);
  SG_LAUNCH_CHECK();
  SG_HIP(hipcub::DeviceSegmentedSort::SortKeys(tmp, tmp_bytes, keys, keys2, (int)n_fragments, (int)n_cells,
                                               d_cell_ptr, d_cell_ptr + 1, st));
  unpack_sort_keys_kernel<<<g, 256, 0, st>>>(keys2, n_fragments, d_pos, d_len);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_tile_count(const uint64_t* d_cell_ptr, const uint32_t* d_pos, const uint32_t* d_len,
                       int64_t n_cells, const void* d_table, int32_t n_chrom, uint32_t bin_size,
                       int32_t strategy, uint32_t min_len, uint32_t max_len, int64_t* d_row_nnz,
                       const uint32_t* d_order, int32_t* d_status, int32_t* level, void* stream) {
  clear_error();
  Params p;
  int rc = make_params(p, d_cell_ptr, d_pos, d_len, n_cells, d_table, n_chrom, bin_size, strategy,
                       min_len, max_len, d_status);
  if (rc) return rc;
  p.order = d_order;
  SG_ARG(level != nullptr && *level >= 0 && *level <= kMaxLevel, "bad level");
  SG_ARG(d_row_nnz != nullptr || n_cells == 0, "null row_nnz");
  if (n_cells == 0) return SNAPGPU_OK;
  p.row_nnz = d_row_nnz;
  hipStream_t st = as_stream(stream);
  for (int lv = *level; lv <= kMaxLevel; ++lv) {
    SG_HIP(hipMemsetAsync(d_status, 0, 8 * sizeof(int32_t), st));
    rc = launch_level<false, int32_t>(p, lv, st);
    if (rc) return rc;
    bool overflow = false;
    rc = check_status(d_status, st, &overflow);
    if (rc) return rc;
    if (!overflow) {
      *level = lv;
      return SNAPGPU_OK;
    }
  }
  set_error("a fragment spans more than 4095 bins; use a larger bin_size");
  return SNAPGPU_ERR_CAPACITY;
}

int snapgpu_tile_fill(const uint64_t* d_cell_ptr, const uint32_t* d_pos, const uint32_t* d_len,
                      int64_t n_cells, const void* d_table, int32_t n_chrom, uint32_t bin_size,
                      int32_t strategy, uint32_t min_len, uint32_t max_len, int32_t level,
                      const int64_t* d_indptr, void* d_out_idx, int32_t index_bits,
                      uint32_t* d_out_val, const uint32_t* d_order, int32_t* d_status, void* stream) {
  clear_error();
  Params p;
  int rc = make_params(p, d_cell_ptr, d_pos, d_len, n_cells, d_table, n_chrom, bin_size, strategy,
                       min_len, max_len, d_status);
  if (rc) return rc;
  p.order = d_order;
  SG_ARG(level >= 0 && level <= kMaxLevel, "bad level");
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  if (n_cells == 0) return SNAPGPU_OK;
  SG_ARG(d_indptr != nullptr, "null indptr");
  p.indptr = d_indptr;
  p.out_idx = d_out_idx;
  p.out_val = d_out_val;
  hipStream_t st = as_stream(stream);
  SG_HIP(hipMemsetAsync(d_status, 0, 8 * sizeof(int32_t), st));
  rc = index_bits == 32 ? launch_level<true, int32_t>(p, level, st)
                        : launch_level<true, int64_t>(p, level, st);
  if (rc) return rc;
  bool overflow = false;
  rc = check_status(d_status, st, &overflow);
  if (rc) return rc;
  if (overflow) {
    set_error("tile fill overflowed at the level chosen by the count pass");
    return SNAPGPU_ERR_CAPACITY;
  }
  return SNAPGPU_OK;
}

}  // extern "C"
