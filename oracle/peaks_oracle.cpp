// ORACLE — test infrastructure only. Only tests/ and the cpu_baseline legs of
// bench.py may load this library, as the checker or the timed CPU baseline; the
// product (snapatac2_amd) never calls it.
//
// C++ restatement of the cell x peak matrix, the same rules as oracle/peaks.py
// (create_peak_matrix, snapatac2-core/src/feature_count/matrix.rs:95-171;
// RegionCounter, feature_count/counter.rs:92-139; pair_to_fragments /
// single_to_fragments, feature_count/data_iter.rs:27-123), with a different
// search than the GPU's: regions of a chromosome sorted by start, a query scans
// forward from the first region starting at or after (query start - longest
// region) while regions start before the query end. Per-cell std::map counts,
// rows in parallel (OpenMP).
//
// Parity pins: tests/test_peaks_oracle.py checks this library against the
// Python restatement, which is pinned by hand-computed counts there.
#include <stdint.h>

#include <algorithm>
#include <map>
#include <vector>

namespace {

struct Reg {
  uint64_t s, e;
  uint32_t col;
};

}  // namespace

extern "C" {

// regions: (chrom id, start, end, col) arrays of length n_reg (chrom id -1 =
// not in the genome); base[n_chrom + 1]; strategy 0 insertion, 1 fragment,
// 2 paired-insertion. Writes the row lengths into row_len[n_cells] and, when
// out_idx is not NULL, the entries (out offsets = scan of row_len).
int64_t oracle_peak_run(const int64_t* cell_ptr, int64_t n_cells, const uint32_t* pos, const int32_t* val, int paired,
                        const uint64_t* base, int32_t n_chrom, const int32_t* reg_chrom, const uint64_t* reg_s,
                        const uint64_t* reg_e, const uint32_t* reg_col, int64_t n_reg, int strategy, uint64_t min_len,
                        uint64_t max_len, int64_t* row_len, const int64_t* out_off, int32_t* out_idx,
                        uint32_t* out_val) {
  std::vector<std::vector<Reg>> by(n_chrom);
  std::vector<uint64_t> longest(n_chrom, 0);
  for (int64_t r = 0; r < n_reg; ++r) {
    if (reg_chrom[r] < 0) continue;
    by[reg_chrom[r]].push_back(Reg{reg_s[r], reg_e[r], reg_col[r]});
    longest[reg_chrom[r]] = std::max(longest[reg_chrom[r]], reg_e[r] > reg_s[r] ? reg_e[r] - reg_s[r] : 0);
  }
  for (auto& v : by) std::sort(v.begin(), v.end(), [](const Reg& a, const Reg& b) { return a.s < b.s; });
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t i = 0; i < n_cells; ++i) {
    std::map<uint32_t, uint32_t> counts;
    std::vector<uint32_t> hits;
    auto query = [&](int c, uint64_t qs, uint64_t qe, std::vector<uint32_t>& out) {
      const auto& v = by[c];
      const uint64_t from = qs > longest[c] ? qs - longest[c] : 0;
      auto it = std::lower_bound(v.begin(), v.end(), from, [](const Reg& a, uint64_t x) { return a.s < x; });
      for (; it != v.end() && it->s < qe; ++it)
        if (it->e > qs) out.push_back(it->col);
    };
    for (int64_t f = cell_ptr[i]; f < cell_ptr[i + 1]; ++f) {
      const uint64_t p = pos[f];
      if (p >= base[n_chrom]) continue;
      const int c = (int)(std::upper_bound(base, base + n_chrom + 1, p) - base) - 1;
      const uint64_t s = p - base[c];
      hits.clear();
      if (!paired) {
        query(c, s, s + 1, hits);
      } else {
        const uint64_t size = (uint32_t)val[f];
        if (size < min_len || size > max_len) continue;
        if (strategy == 1) {
          query(c, s, s + size, hits);
        } else {
          query(c, s, s + 1, hits);
          query(c, s + size - 1, s + size, hits);
          if (strategy == 2) {
            std::sort(hits.begin(), hits.end());
            hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
          }
        }
      }
      for (uint32_t h : hits) ++counts[h];
    }
    row_len[i] = (int64_t)counts.size();
    total += (int64_t)counts.size();
    if (out_idx) {
      int64_t o = out_off[i];
      for (auto& kv : counts) {
        out_idx[o] = (int32_t)kv.first;
        out_val[o] = kv.second;
        ++o;
      }
    }
  }
  return total;
}

}  // extern "C"
