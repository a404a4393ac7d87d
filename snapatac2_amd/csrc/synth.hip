// Seeded synthetic fragment generator (workload input, not a reference path).
//
// Produces the stored encoding of obsm['fragment_paired'] (import.rs:161-210):
// per cell, fragments as (genome base index, length), sorted by (pos, len).
// Every random number is a hash of (seed, cell, fragment, stream), so any range
// of cells can be generated independently and reproducibly.
//   fragments per cell ~ round(LogNormal(mu, sigma)) clipped to [min, max]
//   start: with prob. p_uniform uniform over the genome, otherwise the centre of
//          a peak drawn from the cell type's activity + round(N(0, peak_sd))
//   length: 60% round(Exp(mean 80)), 25% round(N(200,20)), 15% round(N(400,40)),
//          clipped to [1, 4095]; the fragment is shifted to lie inside its
//          chromosome.
#include <hipcub/hipcub.hpp>

#include "common.h"

namespace snapgpu {
namespace synth {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rng(uint64_t seed, uint64_t cell, uint64_t frag, uint32_t stream) {
  return mix64(seed ^ mix64(cell * 0x100000001B3ull ^ mix64(frag * 0x2545F4914F6CDD1Dull + stream)));
}

// Uniform in (0, 1).
__device__ __forceinline__ double unif(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ double normal(uint64_t h1, uint64_t h2) {
  const double u1 = unif(h1), u2 = unif(h2);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void counts_kernel(uint64_t seed, int64_t first_cell, int64_t n_cells, double mu,
                              double sigma, int64_t lo, int64_t hi, int64_t* counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cells) return;
  const uint64_t cell = (uint64_t)(first_cell + i);
  const double z = normal(rng(seed, cell, ~0ull, 1), rng(seed, cell, ~0ull, 2));
  double v = rint(exp(mu + sigma * z));
  int64_t n = (int64_t)fmin(fmax(v, (double)lo), (double)hi);
  counts[i] = n;
}

struct GenParams {
  uint64_t seed;
  int64_t first_cell, n_cells;
  const int64_t* cell_ptr;  // local row pointers (cell_ptr[0] == 0)
  const uint64_t* chrom_start;  // n_chrom + 1
  int32_t n_chrom;
  const uint32_t* peak_centre;
  int32_t n_peaks;
  const float* type_cdf;  // n_types x n_peaks, row-wise cumulative
  int32_t n_types;
  double p_uniform, peak_sd;
  uint64_t* keys;
};

__global__ void __launch_bounds__(256) gen_kernel(GenParams g) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t genome = g.chrom_start[g.n_chrom];
  for (int64_t i = wave; i < g.n_cells; i += nwaves) {
    const uint64_t cell = (uint64_t)(g.first_cell + i);
    const int type = (int)(unif(rng(g.seed, cell, ~0ull, 3)) * g.n_types);
    const float* cdf = g.type_cdf + (int64_t)type * g.n_peaks;
    const int64_t b = g.cell_ptr[i], e = g.cell_ptr[i + 1];
    for (int64_t f = b + lane; f < e; f += 64) {
      const uint64_t j = (uint64_t)(f - b);
      int64_t start;
      if (unif(rng(g.seed, cell, j, 10)) < g.p_uniform) {
        start = (int64_t)(unif(rng(g.seed, cell, j, 11)) * (double)genome);
      } else {
        const float u = (float)unif(rng(g.seed, cell, j, 12));
        int lo = 0, hi = g.n_peaks - 1;  // first peak with cdf > u
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        const double z = normal(rng(g.seed, cell, j, 13), rng(g.seed, cell, j, 14));
        start = (int64_t)g.peak_centre[lo] + (int64_t)rint(g.peak_sd * z);
      }
      const double um = unif(rng(g.seed, cell, j, 20));
      double lf;
      if (um < 0.60) lf = rint(-80.0 * log(unif(rng(g.seed, cell, j, 21))));
      else if (um < 0.85) lf = rint(200.0 + 20.0 * normal(rng(g.seed, cell, j, 22), rng(g.seed, cell, j, 23)));
      else lf = rint(400.0 + 40.0 * normal(rng(g.seed, cell, j, 24), rng(g.seed, cell, j, 25)));
      uint64_t len = (uint64_t)fmin(fmax(lf, 1.0), 4095.0);
      if (start < 0) start = 0;
      if ((uint64_t)start >= genome) start = (int64_t)genome - 1;
      int lo = 0, hi = g.n_chrom;  // chromosome containing start
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (g.chrom_start[mid] <= (uint64_t)start) lo = mid; else hi = mid;
      }
      const uint64_t cs = g.chrom_start[lo], clen = g.chrom_start[lo + 1] - cs;
      if (len > clen) len = clen;
      uint64_t off = (uint64_t)start - cs;
      if (off + len > clen) off = clen - len;
      const uint64_t pos = cs + off;
      g.keys[f] = ((uint64_t)i << 44) | (pos << 12) | len;
    }
  }
}

__global__ void unpack_kernel(const uint64_t* keys, int64_t n, uint32_t* pos, uint32_t* len) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t k = keys[i];
    pos[i] = (uint32_t)((k >> 12) & 0xFFFFFFFFull);
    len[i] = (uint32_t)(k & 0xFFFull);
  }
}

// Cell type of each cell: the same draw gen_kernel makes for the cell.
__global__ void types_kernel(uint64_t seed, int64_t first_cell, int64_t n_cells, int32_t n_types,
                             int32_t* types) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cells) return;
  types[i] = (int32_t)(unif(rng(seed, (uint64_t)(first_cell + i), ~0ull, 3)) * n_types);
}

}  // namespace synth
}  // namespace snapgpu

using namespace snapgpu;
using namespace snapgpu::synth;

extern "C" {

int snapgpu_synth_counts(uint64_t seed, int64_t first_cell, int64_t n_cells, double mu, double sigma,
                         int64_t min_frags, int64_t max_frags, int64_t* d_counts, void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && first_cell >= 0, "negative cell range");
  if (n_cells == 0) return SNAPGPU_OK;
  counts_kernel<<<(unsigned)ceil_div(n_cells, 256), 256, 0, as_stream(stream)>>>(
      seed, first_cell, n_cells, mu, sigma, min_frags, max_frags, d_counts);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int snapgpu_synth_cell_types(uint64_t seed, int64_t first_cell, int64_t n_cells, int32_t n_types,
                             int32_t* d_types, void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && first_cell >= 0 && n_types >= 1, "bad cell range or type count");
  if (n_cells == 0) return SNAPGPU_OK;
  types_kernel<<<(unsigned)ceil_div(n_cells, 256), 256, 0, as_stream(stream)>>>(seed, first_cell, n_cells,
                                                                              n_types, d_types);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

int64_t snapgpu_synth_workspace_bytes(int64_t n_frags) {
  size_t tmp = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (int)std::min<int64_t>(n_frags, INT32_MAX));
  // two key arrays + the sort's temporary storage
  return 2 * n_frags * (int64_t)sizeof(uint64_t) + (int64_t)tmp + 256;
}

int snapgpu_synth_fragments(uint64_t seed, int64_t first_cell, int64_t n_cells,
                            const int64_t* d_cell_ptr, int64_t n_frags,
                            const uint64_t* d_chrom_start, int32_t n_chrom,
                            const uint32_t* d_peak_centre, int32_t n_peaks, const float* d_type_cdf,
                            int32_t n_types, double p_uniform, double peak_sd, uint32_t* d_pos,
                            uint32_t* d_len, void* d_workspace, int64_t workspace_bytes,
                            void* stream) {
  clear_error();
  SG_ARG(n_cells >= 0 && n_cells < (1ll << 20), "a chunk holds at most 2^20 - 1 cells");
  SG_ARG(n_frags >= 0 && n_frags <= INT32_MAX, "a chunk holds at most 2^31 - 1 fragments");
  SG_ARG(n_peaks >= 1 && n_types >= 1 && n_chrom >= 1, "empty generator tables");
  SG_ARG(workspace_bytes >= snapgpu_synth_workspace_bytes(n_frags), "workspace too small");
  if (n_frags == 0) return SNAPGPU_OK;
  hipStream_t st = as_stream(stream);
  uint64_t* keys = static_cast<uint64_t*>(d_workspace);
  uint64_t* sorted = keys + n_frags;
  void* tmp = sorted + n_frags;
  size_t tmp_bytes = (size_t)(workspace_bytes - 2 * n_frags * (int64_t)sizeof(uint64_t));
  GenParams g{seed, first_cell, n_cells, d_cell_ptr, d_chrom_start, n_chrom, d_peak_centre,
              n_peaks, d_type_cdf, n_types, p_uniform, peak_sd, keys};
  const int64_t blocks = std::min<int64_t>(ceil_div(n_cells, 4), 65536);
  gen_kernel<<<(unsigned)blocks, 256, 0, st>>>(g);
  SG_LAUNCH_CHECK();
  int end_bit = 44;
  while (end_bit < 64 && (1ll << (end_bit - 44)) <= n_cells) ++end_bit;
  SG_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, (int)n_frags, 0, end_bit, st));
  unpack_kernel<<<(unsigned)std::min<int64_t>(ceil_div(n_frags, 256), 16384), 256, 0, st>>>(
      sorted, n_frags, d_pos, d_len);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
