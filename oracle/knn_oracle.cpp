// ORACLE — test infrastructure only (see oracle/__init__.py).
//
// Exact k-nearest-neighbour graph by FP64 brute force, restating
// snapatac2-core/src/utils/knn.rs:
//   :9-32  every other point ranked by squared Euclidean distance (kdtree 0.7's
//          squared_euclidean: sum over dimensions of (a-b)*(a-b), no fused
//          multiply-add: built with -ffp-contract=off), the query itself skipped
//          by index (:27), k taken (:28), distance = sqrt (:28);
//   :58-78 each row re-sorted by neighbour index.
// Equal distances are ordered by neighbour index (the kd-tree's own order among
// exact ties is not pinned by the reference; see the oracle's docs).
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <utility>
#include <vector>

extern "C" void oracle_knn(const double* X, int64_t n, int32_t d, int32_t k, int32_t* nbr,
                           double* dist) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < n; ++q) {
    std::vector<std::pair<double, int64_t>> all;
    all.reserve(n - 1);
    const double* a = X + q * d;
    for (int64_t p = 0; p < n; ++p) {
      if (p == q) continue;
      const double* b = X + p * d;
      double acc = 0.0;
      for (int c = 0; c < d; ++c) {
        double t = a[c] - b[c];
        acc = acc + t * t;
      }
      all.emplace_back(acc, p);
    }
    int kk = (int)std::min<int64_t>(k, n - 1);
    std::partial_sort(all.begin(), all.begin() + kk, all.end());
    std::vector<std::pair<int64_t, double>> row;
    for (int i = 0; i < kk; ++i) row.emplace_back(all[i].second, sqrt(all[i].first));
    std::sort(row.begin(), row.end());
    for (int i = 0; i < kk; ++i) {
      nbr[q * k + i] = (int32_t)row[i].first;
      dist[q * k + i] = row[i].second;
    }
  }
}
