// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker or
// the timed CPU baseline; the product (snapatac2_amd) never calls it.
//
// CPU restatement of the reference tile-matrix path:
//   GenomeBaseIndex           snapatac2-core/src/genome.rs:286-487
//     new                     :296-312  (cumulative base lengths, stored order)
//     with_step               :365-384  (ceil(len / step) bins per chromosome)
//     get_position_rev        :387-406  (panics when pos >= chromosome length)
//     get_position            :419-434  (binary search on the cumulative lengths)
//   FragmentData::get_gindex  snapatac2-core/src/feature_count/data_iter.rs:156-173
//                             (excluded chromosomes are dropped and bins renumbered)
//   gen_mat_single            data_iter.rs:312-344
//   gen_mat_pair              data_iter.rs:346-415 (per-row BTreeMap<usize, u32>)
//   CountingStrategy          snapatac2-core/src/feature_count/counter.rs:19-36
// Rows are processed in parallel (OpenMP), as the reference does with rayon.
//
// Parity pins: tests/test_tile_oracle.py (GenomeBaseIndex known answers from
// genome.rs:489-597, hand-derived cases, brute force, invariants from
// snapatac2-python/tests/test_tools.py:144-162 and test_pipeline.py:10-31).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace {

struct Index {
  std::vector<uint64_t> base_accum;    // cumulative base lengths
  std::vector<uint64_t> binned_accum;  // cumulative bin counts
  uint64_t step = 1;
};

Index make_index(const std::vector<uint64_t>& lens, uint64_t step) {
  Index ix;
  ix.step = step;
  uint64_t acc = 0, acc_bin = 0;
  for (uint64_t l : lens) {
    acc += l;
    acc_bin += (l + step - 1) / step;  // div_ceil
    ix.base_accum.push_back(acc);
    ix.binned_accum.push_back(acc_bin);
  }
  return ix;
}

// genome.rs:419-434, on the step-1 index: returns (chrom, offset) or chrom = -1.
bool get_position(const Index& ix, uint64_t i, int* chrom, uint64_t* off) {
  const auto& a = ix.binned_accum;
  auto it = std::lower_bound(a.begin(), a.end(), i);
  size_t j = it - a.begin();
  if (it != a.end() && *it == i) {  // Ok(j): start of chromosome j+1
    if (j + 1 >= a.size()) return false;
    *chrom = (int)(j + 1);
    *off = 0;
    return true;
  }
  if (j >= a.size()) return false;  // past the genome: the reference panics (unwrap)
  uint64_t prev = j == 0 ? 0 : a[j - 1];
  *chrom = (int)j;
  *off = (i - prev) * ix.step;
  return true;
}

// genome.rs:387-406 on the (filtered) output index.
bool get_position_rev(const Index& ix, int c, uint64_t pos, uint64_t* bin) {
  uint64_t size = c == 0 ? ix.base_accum[0] : ix.base_accum[c] - ix.base_accum[c - 1];
  if (pos >= size) return false;
  uint64_t p = pos / ix.step;
  *bin = c == 0 ? p : ix.binned_accum[c - 1] + p;
  return true;
}

struct Result {
  std::vector<int64_t> indptr;
  std::vector<int64_t> indices;
  std::vector<uint32_t> values;
  uint64_t n_bins = 0;
  std::string error;
};

}  // namespace

extern "C" {

// strategy: 0 insertion, 1 fragment, 2 paired-insertion, 3 single-end.
// min_len / max_len < 0 mean "no filter".
void* oracle_tile_run(const uint64_t* cell_ptr, const uint32_t* pos, const uint32_t* len,
                      int64_t n_cells, const uint64_t* chrom_len, const uint8_t* excluded,
                      int32_t n_chrom, uint64_t bin_size, int32_t strategy, int64_t min_len,
                      int64_t max_len) {
  Result* res = new Result();
  std::vector<uint64_t> all(chrom_len, chrom_len + n_chrom), kept_lens;
  std::vector<int> new_id(n_chrom, -1);
  for (int c = 0; c < n_chrom; ++c) {
    if (!excluded[c]) {
      new_id[c] = (int)kept_lens.size();
      kept_lens.push_back(all[c]);
    }
  }
  Index ori = make_index(all, 1);
  Index nix = make_index(kept_lens, bin_size);
  res->n_bins = nix.binned_accum.empty() ? 0 : nix.binned_accum.back();

  std::vector<std::vector<std::pair<uint64_t, uint32_t>>> rows(n_cells);
  std::vector<std::string> errs(n_cells);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t r = 0; r < n_cells; ++r) {
    std::map<uint64_t, uint32_t> count;
    for (uint64_t k = cell_ptr[r]; k < cell_ptr[r + 1]; ++k) {
      int c;
      uint64_t start;
      if (!get_position(ori, pos[k], &c, &start)) {
        errs[r] = "position past the genome";
        break;
      }
      if (strategy == 3) {
        if (excluded[c]) continue;
        uint64_t b;
        if (!get_position_rev(nix, new_id[c], start, &b)) { errs[r] = "out of range"; break; }
        count[b] += 1;
        continue;
      }
      uint64_t fs = len[k];
      if (excluded[c] || (min_len >= 0 && fs < (uint64_t)min_len) ||
          (max_len >= 0 && fs > (uint64_t)max_len))
        continue;
      if (fs == 0) { errs[r] = "zero length"; break; }
      uint64_t end = start + fs - 1;
      uint64_t b1, b2;
      if (!get_position_rev(nix, new_id[c], start, &b1) ||
          !get_position_rev(nix, new_id[c], end, &b2)) {
        errs[r] = "Position is out of range";
        break;
      }
      if (strategy == 0) {
        count[b1] += 1;
        count[b2] += 1;
      } else if (strategy == 1) {
        for (uint64_t b = b1; b <= b2; ++b) count[b] += 1;
      } else {
        count[b1] += 1;
        if (b2 != b1) count[b2] += 1;
      }
    }
    rows[r].assign(count.begin(), count.end());
  }
  for (int64_t r = 0; r < n_cells; ++r)
    if (!errs[r].empty()) {
      res->error = errs[r] + " (cell " + std::to_string(r) + ")";
      break;
    }
  res->indptr.resize(n_cells + 1);
  res->indptr[0] = 0;
  for (int64_t r = 0; r < n_cells; ++r) res->indptr[r + 1] = res->indptr[r] + (int64_t)rows[r].size();
  res->indices.resize(res->indptr[n_cells]);
  res->values.resize(res->indptr[n_cells]);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_cells; ++r) {
    int64_t o = res->indptr[r];
    for (auto& kv : rows[r]) {
      res->indices[o] = (int64_t)kv.first;
      res->values[o] = kv.second;
      ++o;
    }
  }
  return res;
}

const char* oracle_tile_error(void* h) { return static_cast<Result*>(h)->error.c_str(); }
int64_t oracle_tile_nnz(void* h) { return (int64_t) static_cast<Result*>(h)->indices.size(); }
uint64_t oracle_tile_n_bins(void* h) { return static_cast<Result*>(h)->n_bins; }

void oracle_tile_copy(void* h, int64_t* indptr, int64_t* indices, uint32_t* values) {
  Result* r = static_cast<Result*>(h);
  memcpy(indptr, r->indptr.data(), r->indptr.size() * sizeof(int64_t));
  memcpy(indices, r->indices.data(), r->indices.size() * sizeof(int64_t));
  memcpy(values, r->values.data(), r->values.size() * sizeof(uint32_t));
}

void oracle_tile_free(void* h) { delete static_cast<Result*>(h); }

}  // extern "C"
