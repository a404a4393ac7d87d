/*
 * libsnapgpu — C ABI of the MI355X (gfx950) kernels behind snapatac2_amd.
 *
 * Conventions
 *   - Every pointer named d_* is device memory; h_* is host memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - Every entry point returns 0 (SNAPGPU_OK) or a SNAPGPU_ERR_* code; the
 *     message of the last failure on the calling thread is returned by
 *     snapgpu_last_error().
 *   - The library allocates no device memory: callers pass outputs and, where
 *     an entry point needs scratch space, a workspace whose size is given by the
 *     matching *_workspace_bytes() query.
 *   - Kernels are deterministic: the same inputs give bitwise-equal outputs.
 */
#ifndef SNAPGPU_H
#define SNAPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum snapgpu_status {
  SNAPGPU_OK = 0,
  SNAPGPU_ERR_ARG = 1,      /* bad argument (shape, enum value, null pointer) */
  SNAPGPU_ERR_HIP = 2,      /* HIP runtime failure                            */
  SNAPGPU_ERR_DATA = 3,     /* input breaks an invariant the reference panics on */
  SNAPGPU_ERR_CAPACITY = 4, /* input exceeds a documented kernel limit        */
  SNAPGPU_ERR_UNSORTED = 5  /* a row of fragments is not sorted by position:
                               sort it (snapgpu_frag_sort_rows), call again  */
};

/* Counting strategies of the tile matrix.
 * Reference: snapatac2-core/src/feature_count/counter.rs:19-36 (CountingStrategy)
 * and data_iter.rs:312-415 (gen_mat_single / gen_mat_pair). */
enum snapgpu_strategy {
  SNAPGPU_INSERTION = 0, /* "insertion": +1 at both insertion bins            */
  SNAPGPU_FRAGMENT = 1,  /* "fragment": +1 at every bin the fragment covers    */
  SNAPGPU_PIC = 2,       /* "paired-insertion": second +1 only if a new bin    */
  SNAPGPU_SINGLE = 3     /* single-end reads (obsm['fragment_single'])         */
};

const char* snapgpu_last_error(void);
int snapgpu_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Prefix sums                                                              */
/* ------------------------------------------------------------------------ */

/* Bytes of workspace snapgpu_scan_inclusive_i64 needs for n elements. */
int64_t snapgpu_scan_workspace_bytes(int64_t n);

/* In-place inclusive prefix sum of d_data[0..n). Used to turn per-row counts
 * into CSR row pointers (the reference builds them serially in anndata-rs
 * to_csr_data, called at data_iter.rs:413). */
int snapgpu_scan_inclusive_i64(int64_t* d_data, int64_t n, void* d_workspace,
                               int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Path A: fragment -> cell x bin CSR ("tile matrix")                        */
/* Replaces mk_tile_matrix (snapatac2-python/src/preprocessing.rs:286-345)   */
/* -> create_tile_matrix (feature_count/matrix.rs:26-93) -> gen_mat_pair /   */
/* gen_mat_single (feature_count/data_iter.rs:312-415).                      */
/* ------------------------------------------------------------------------ */

/* Size in bytes of the packed chromosome table for n_chrom chromosomes. */
int64_t snapgpu_tile_table_bytes(int32_t n_chrom);

/* Host-only. Packs the chromosome table the tile kernels read.
 *   h_chrom_len[n_chrom]: lengths in the stored order of
 *     uns['reference_sequences'] (GenomeBaseIndex::new, genome.rs:296-312);
 *   h_excluded[n_chrom]: 1 = chromosome excluded from the output, which
 *     renumbers the bins of the others (FragmentData::get_gindex,
 *     data_iter.rs:156-173);
 * Writes the table to h_out (snapgpu_tile_table_bytes bytes; copy it to the
 * device before use) and the number of output bins to *n_bins
 * (GenomeBaseIndex::with_step + len, genome.rs:341-384). */
int snapgpu_tile_table_pack(const uint64_t* h_chrom_len, const uint8_t* h_excluded,
                            int32_t n_chrom, uint32_t bin_size, void* h_out,
                            uint64_t* n_bins);

/* Processing order for the two passes: the cells by decreasing fragment count
 * (a wave walks one cell, so long cells started late would trail the launch).
 * Same output in any order. d_order[n_cells] out; workspace from
 * snapgpu_tile_order_workspace_bytes. */
int64_t snapgpu_tile_order_workspace_bytes(int64_t n_cells);
int snapgpu_tile_order(const uint64_t* d_cell_ptr, int64_t n_cells, uint32_t* d_order, void* d_ws,
                       int64_t ws_bytes, void* stream);

/* Sorts the fragments of every row by (position, value) in place, the order
 * import_fragments stores them in (import.rs:208). The tile passes need rows
 * whose positions do not decrease and return SNAPGPU_ERR_UNSORTED (before any
 * output is used) when one does; the reference's per-row BTreeMap accepts any
 * order (data_iter.rs:362-411), so sorting and calling again gives its result.
 * d_cell_ptr[n_cells + 1]: row pointers into d_pos / d_len, all within
 * [0, n_fragments], n_fragments < 2^31; d_len may be NULL (positions only).
 * Workspace:
 * snapgpu_frag_sort_rows_workspace_bytes(n_cells, n_fragments). */
int64_t snapgpu_frag_sort_rows_workspace_bytes(int64_t n_cells, int64_t n_fragments);
int snapgpu_frag_sort_rows(const uint64_t* d_cell_ptr, int64_t n_cells, int64_t n_fragments,
                           uint32_t* d_pos, uint32_t* d_len, void* d_ws, int64_t ws_bytes,
                           void* stream);

/* Pass 1: number of stored (bin, count) entries of each cell.
 *   d_cell_ptr[n_cells+1]: fragment row pointers of obsm['fragment_paired'|
 *     'fragment_single'] (import.rs:117-159);
 *   d_pos[F]: 0-based genome base index of each fragment (import.rs:161-210);
 *   d_len[F]: fragment length (paired) or NULL (single-end, whose value is not
 *     used by gen_mat_single);
 *   min_len/max_len: inclusive size filter (0 / UINT32_MAX = none);
 *   d_row_nnz[n_cells]: output;
 *   d_order: processing order from snapgpu_tile_order, or NULL (cell order);
 *   d_status[8]: int32 device scratch for error reporting;
 *   *level: in = first capacity level to try (0), out = level that succeeded
 *     (pass it to snapgpu_tile_fill).
 * Rows must be sorted by position (SNAPGPU_ERR_UNSORTED otherwise).
 * Synchronises `stream` once per level tried. */
int snapgpu_tile_count(const uint64_t* d_cell_ptr, const uint32_t* d_pos,
                       const uint32_t* d_len, int64_t n_cells, const void* d_table,
                       int32_t n_chrom, uint32_t bin_size, int32_t strategy,
                       uint32_t min_len, uint32_t max_len, int64_t* d_row_nnz,
                       const uint32_t* d_order, int32_t* d_status, int32_t* level, void* stream);

/* Pass 2: writes the CSR columns (ascending per row) and counts.
 *   d_indptr[n_cells+1]: row pointers (scan of pass-1 counts);
 *   d_out_idx: int32 columns if index_bits == 32, int64 if 64;
 *   d_out_val: uint32 counts. */
int snapgpu_tile_fill(const uint64_t* d_cell_ptr, const uint32_t* d_pos,
                      const uint32_t* d_len, int64_t n_cells, const void* d_table,
                      int32_t n_chrom, uint32_t bin_size, int32_t strategy,
                      uint32_t min_len, uint32_t max_len, int32_t level,
                      const int64_t* d_indptr, void* d_out_idx, int32_t index_bits,
                      uint32_t* d_out_val, const uint32_t* d_order, int32_t* d_status,
                      void* stream);

/* ------------------------------------------------------------------------ */
/* Feature selection input                                                  */
/* ------------------------------------------------------------------------ */

/* d_out[j] += sum of the stored counts of column j (uint64; zero it first).
 * Integer sums, so the order of accumulation does not matter. Replaces the
 * chunked column sums of select_features
 * (snapatac2-python/python/snapatac2/preprocessing/_basic.py:1053-1055). */
int snapgpu_csr_col_sum_u32(const void* d_indices, int32_t index_bits, const uint32_t* d_values,
                            int64_t nnz, uint64_t* d_out, void* stream);

/* d_out[j] = sum of the stored counts of column j (uint64; written, not added
 * to) for a CSR matrix given by its row pointers. Same sums as
 * snapgpu_csr_col_sum_u32 over the entries (select_features' column sums,
 * snapatac2-python/python/snapatac2/preprocessing/_basic.py:1053-1055), with
 * one global atomic per (block of 3968 rows, non-zero column) instead of one
 * per entry when every row's column indices ascend; rows out of order are
 * detected and summed entry by entry instead. Indices must lie in
 * [0, n_cols). d_ws16: 16 bytes of device scratch. Synchronises `stream`. */
int snapgpu_csr_col_sum_rows_u32(const int64_t* d_indptr, int64_t n_rows, const void* d_indices,
                                 int32_t index_bits, const uint32_t* d_values, int64_t n_cols,
                                 uint64_t* d_out, void* d_ws16, void* stream);

/* ------------------------------------------------------------------------ */
/* Synthetic workload generator (no reference counterpart: it produces the   */
/* obsm['fragment_paired'] encoding of import.rs:161-210 from a seed).       */
/* ------------------------------------------------------------------------ */

/* Fragments per cell for cells first_cell .. first_cell+n_cells-1:
 * round(exp(mu + sigma*N(0,1))) clipped to [min_frags, max_frags]. */
int snapgpu_synth_counts(uint64_t seed, int64_t first_cell, int64_t n_cells, double mu,
                         double sigma, int64_t min_frags, int64_t max_frags, int64_t* d_counts,
                         void* stream);

/* Cell type (0 .. n_types-1) of cells first_cell .. first_cell+n_cells-1: the
 * type whose peak activity snapgpu_synth_fragments draws the cell's peak
 * fragments from (the generator's ground truth, for judging embeddings). */
int snapgpu_synth_cell_types(uint64_t seed, int64_t first_cell, int64_t n_cells, int32_t n_types,
                             int32_t* d_types, void* stream);

int64_t snapgpu_synth_workspace_bytes(int64_t n_frags);

/* Fragments of a chunk of cells (< 2^20 cells, < 2^31 fragments), sorted per
 * cell by (pos, len). d_cell_ptr: local row pointers of the chunk (from
 * snapgpu_synth_counts + scan); d_chrom_start[n_chrom+1]: cumulative
 * chromosome starts; d_type_cdf[n_types*n_peaks]: per-type cumulative peak
 * activity. */
int snapgpu_synth_fragments(uint64_t seed, int64_t first_cell, int64_t n_cells,
                            const int64_t* d_cell_ptr, int64_t n_frags,
                            const uint64_t* d_chrom_start, int32_t n_chrom,
                            const uint32_t* d_peak_centre, int32_t n_peaks,
                            const float* d_type_cdf, int32_t n_types, double p_uniform,
                            double peak_sd, uint32_t* d_pos, uint32_t* d_len,
                            void* d_workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Path C: exact k-nearest-neighbour graph                                   */
/* Replaces nearest_neighbour_graph (snapatac2-python/src/knn.rs:7-16 ->      */
/* snapatac2-core/src/utils/knn.rs:9-32, 58-78).                             */
/* ------------------------------------------------------------------------ */

int64_t snapgpu_knn_workspace_bytes(int64_t n, int32_t d, int32_t k);

/* d_X: n x d row-major FP64 (leading dimension ld), d >= 1, 1 <= k <= n-1.
 * For d <= 30 and k <= 100:
 * mode 0: FP32 SIMT candidate distances; mode 1: BF16 MFMA (hi/lo split, FP32
 * accumulation). Both rescore candidates in exact FP64 and prove the result
 * with a certificate (failing rows are recomputed exactly), so the neighbour
 * sets equal FP64 brute force with ties broken by index.
 * Outputs: d_nbr[n*k] (int32) and d_dist[n*k] (FP64 Euclidean distance), each
 * row sorted by neighbour index (knn.rs:67). Rows whose certificate fails are
 * screened again with exact FP64 distances (*n_fallback counts them) and, if
 * that certificate fails too, scanned exactly (*n_exact). Mode 1 sorts the
 * points spatially and skips 128-point chunks whose bounding sphere is provably
 * farther than every candidate threshold of a 128-query workgroup;
 * *n_chunk_visits (may be NULL) receives the (workgroup, chunk) pairs processed,
 * out of ceil(n/128)^2. flags bit 0: send every row straight to the exact scan
 * (testing); bit 1: visit every chunk (no skipping).
 * Wider rows (d > 30) or more neighbours (k > 100), in either mode: exact FP64
 * distances from every query to every point (O(n^2 d)) and a stable sort of
 * each query's (distance, index) pairs; *n_exact = n, the other counters 0.
 * Synchronises `stream`. */
int snapgpu_knn(const double* d_X, int64_t n, int32_t d, int64_t ld, int32_t k, int32_t mode,
                int32_t* d_nbr, double* d_dist, void* d_workspace, int64_t workspace_bytes,
                int32_t flags, int64_t* n_fallback, int64_t* n_exact, int64_t* n_chunk_visits,
                void* stream);

/* ------------------------------------------------------------------------ */
/* Path B: the normalized-cosine operator inside spectral()                  */
/* Replaces spectral_embedding / idf / normalize / spectral_mf               */
/* (snapatac2-python/src/embedding.rs:24-59, 131-191, 269-326) and the scipy  */
/* LinearOperator matvec it hands to eigsh (embedding.rs:155-171).           */
/* ------------------------------------------------------------------------ */

/* Column selection by feature mask (spectral_embedding's slice_axis,
 * embedding.rs:36-41): the selected columns keep their order.
 * snapgpu_col_select_table: d_cols = the k selected columns, ascending and
 * distinct, in [0, m); d_sel = 2 * ceil(m / 32) words: for word w, the mask of
 * selected columns among 32w .. 32w + 31 and the count of selected columns
 * below 32w. Entries with columns outside [0, m) are dropped.
 * select_count: d_row_nnz[i] = selected entries of row i.
 * select_fill: copies them, columns renumbered 0 .. k-1, into the CSR given by
 * d_new_indptr (the scan of the counts). */
int snapgpu_col_select_table(const int64_t* d_cols, int64_t k, int64_t m, uint32_t* d_sel,
                             void* stream);
int snapgpu_csr_select_count(const int64_t* d_indptr, const void* d_idx, int32_t index_bits,
                             int64_t n, const uint32_t* d_sel, int64_t m, int64_t* d_row_nnz,
                             void* stream);
int snapgpu_csr_select_fill(const int64_t* d_indptr, const void* d_idx, int32_t index_bits,
                            const uint32_t* d_val, int64_t n, const uint32_t* d_sel, int64_t m,
                            const int64_t* d_new_indptr, int32_t* d_new_idx, uint32_t* d_new_val,
                            void* stream);

/* Stable CSR -> CSC transpose (rows ascending within each column, so column
 * reductions add in the reference's CSR order, embedding.rs:139-144).
 * h_indptr: host copy of the row pointers; entries are sorted chunk_nnz at a
 * time. Workspace: snapgpu_csr_transpose_workspace_bytes(m, chunk_nnz). */
int64_t snapgpu_csr_transpose_workspace_bytes(int64_t m, int64_t chunk_nnz);
int snapgpu_csr_transpose(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_val,
                          int64_t n, int64_t m, int64_t nnz, const int64_t* h_indptr,
                          int64_t chunk_nnz, int64_t* d_colptr, int32_t* d_rowidx,
                          uint32_t* d_cval, void* d_ws, int64_t ws_bytes, void* stream);

/* IDF weights from the CSC column pointers (embedding.rs:269-286). d_ws16: 16
 * bytes of device scratch. Synchronises `stream`. */
int snapgpu_spectral_idf(const int64_t* d_colptr, int64_t m, int64_t n, double* d_w,
                         void* d_ws16, void* stream);

/* Row scales s = 1/||x o w|| (normalize, embedding.rs:314-326), column sums c
 * and degrees d = X~ c - 1 (spectral_mf, embedding.rs:133-153); writes
 * dinv = 1/d, r = s*sqrt(dinv) (FP64 and FP32) and w2 = w^2 (FP32).
 * value_type: 0 = uint32 counts, 1 = float32 values (bit patterns in d_cnt).
 * Workspace: snapgpu_spectral_prepare_workspace_bytes(m) (the (w_j, c_j) pairs
 * the degree pass gathers). Fails with SNAPGPU_ERR_DATA if a non-empty row has
 * d <= 0. Synchronises. */
int64_t snapgpu_spectral_prepare_workspace_bytes(int64_t m);
int snapgpu_spectral_prepare(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt,
                             int64_t n, const int64_t* d_colptr, const int32_t* d_rowidx,
                             const uint32_t* d_ccnt, int64_t m, const double* d_w, double* d_s,
                             double* d_c, double* d_dinv, double* d_r64, float* d_r32,
                             float* d_w2, int32_t value_type, int32_t* d_status, void* d_ws,
                             int64_t ws_bytes, void* stream);

/* Y = A V for a block of b <= 32 columns (FP64 in/out, FP32 sparse products):
 * A v = r o (X (w^2 o (X^T (r o v)))) - dinv o v, the eigen() matvec
 * X @ (v.T @ X).T - D * v of embedding.rs:160-162. */
int64_t snapgpu_spectral_apply_workspace_bytes(int64_t n, int64_t m);
int snapgpu_spectral_apply(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt,
                           int64_t n, const int64_t* d_colptr, const int32_t* d_rowidx,
                           const uint32_t* d_ccnt, int64_t m, const float* d_r32,
                           const float* d_w2, const double* d_dinv, const double* d_V,
                           int64_t ldv, int32_t b, double* d_Y, int64_t ldy, int32_t value_type,
                           void* d_ws, int64_t ws_bytes, void* stream);

/* Adjoint with respect to the feature weights (no counterpart in the
 * reference; it differentiates the operator above, embedding.rs:133-162 with
 * the weighting of embedding.rs:314-326): dw = d<G, A(w) V>/dw, including the
 * row-norm and degree dependence on w. d_s, d_dinv, d_r32, d_w2 are the
 * outputs of snapgpu_spectral_prepare for the same d_w. The derivative with
 * respect to V is A G (A is symmetric): call snapgpu_spectral_apply on G.
 * Workspace: snapgpu_spectral_vjp_workspace_bytes(n, m). */
int64_t snapgpu_spectral_vjp_workspace_bytes(int64_t n, int64_t m);
int snapgpu_spectral_vjp_w(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt,
                           int64_t n, const int64_t* d_colptr, const int32_t* d_rowidx,
                           const uint32_t* d_ccnt, int64_t m, const double* d_w,
                           const double* d_s, const double* d_dinv, const float* d_r32,
                           const float* d_w2, const double* d_V, int64_t ldv, const double* d_G,
                           int64_t ldg, int32_t b, double* d_dw, int32_t value_type, void* d_ws,
                           int64_t ws_bytes, void* stream);

/* Panel layout for large operators (the same products as snapgpu_spectral_apply).
 * The matrix (CSR of the cells or CSC of the features) is re-laid out
 * panel-major: panel p holds the entries whose column lies in
 * [p * 32768, (p + 1) * 32768), so each product gathers from an L2-sized slice
 * of the dense operand; work is split into chunks of ~4096 entries aligned to
 * (row, panel) segments ("pairs"), whose FP32 sums are then added per row in
 * panel order. Integer counts are packed with the column into 4 bytes:
 * `packed` = 2 for counts below 128 (the form the 16-column product gathers
 * with the fewest instructions; needs dense_rows <= 2^26 for its 32-bit
 * offsets, else the generic product runs), 1 for counts below 65536, 0 for
 * float values or larger counts (a separate value array). snapgpu_panel_sizes
 * returns the densest packing the values allow; snapgpu_panel_build accepts any
 * packing the values fit and returns SNAPGPU_ERR_ARG for one they do not.
 * Build: snapgpu_panel_sizes, then allocate, then snapgpu_panel_build with the
 * same workspace (sizes leaves its tables there).
 *   ent / val       n_entries uint32 (val only when not packed)
 *   chunk_start     n_chunks + 1 int64;   chunk_pair n_chunks + 1 int32
 *   chunk_panel     n_chunks int32
 *   row_pair_ptr    n_rows + 1 int64;     row_pair_ids n_pairs int32 */
typedef struct snapgpu_panel_layout {
  const uint32_t* ent;
  const uint32_t* val;
  const int64_t* chunk_start;
  const int32_t* chunk_pair;
  const int32_t* chunk_panel;
  int64_t n_chunks;
  const int64_t* row_pair_ptr;
  const int32_t* row_pair_ids;
  int64_t n_pairs;
  int64_t n_rows;
  int64_t dense_rows;
  int32_t packed;
  int32_t value_type;
} snapgpu_panel_layout;

int32_t snapgpu_panel_count(int64_t dense_rows);
int64_t snapgpu_panel_build_workspace_bytes(int64_t nrows, int64_t dense_rows, int64_t nnz);
int snapgpu_panel_sizes(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val,
                        int64_t nrows, int64_t dense_rows, int64_t nnz, int32_t value_type,
                        void* d_ws, int64_t ws_bytes, int64_t* n_pairs, int64_t* n_chunks,
                        int32_t* packed, void* stream);
int snapgpu_panel_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val,
                        int64_t nrows, int64_t dense_rows, int64_t nnz, int32_t packed, void* d_ws,
                        int64_t ws_bytes, uint32_t* d_ent, uint32_t* d_val_out,
                        int64_t* d_chunk_start, int32_t* d_chunk_pair, int32_t* d_chunk_panel,
                        int64_t n_chunks, int64_t* d_row_pair_ptr, int32_t* d_row_pair_ids,
                        void* stream);
/* Y = A V as snapgpu_spectral_apply, over the two panel layouts (csr: rows =
 * cells, dense_rows = features; csc: rows = features, dense_rows = cells). */
int64_t snapgpu_spectral_apply_panel_workspace_bytes(int64_t n, int64_t m, int64_t pairs_csr,
                                                     int64_t pairs_csc);
int snapgpu_spectral_apply_panel(const snapgpu_panel_layout* csr, const snapgpu_panel_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv,
                                 const double* d_V, int64_t ldv, int32_t b, double* d_Y,
                                 int64_t ldy, void* d_ws, int64_t ws_bytes, void* stream);

/* out = diag(row_scale) X D in FP64 over one panel layout, for a dense operand D
 * of dense_rows x 2 FP64 (row-major, 16-byte rows); X's values are squared when
 * square_values != 0; row_scale may be NULL (1). Each (row, panel) segment is
 * summed in entry order, then a row's segments in panel order, so results are
 * bitwise reproducible. out: n_rows x 2. No counterpart in the reference: the
 * FP64 products of this library's eigenpair checks and weight adjoint (the
 * CSR form is snapgpu_csr_spmm_scaled_f64), gathering from an L2-sized slice
 * of D at a time. Workspace: snapgpu_panel_spmm_f64_workspace_bytes(n_pairs). */
int64_t snapgpu_panel_spmm_f64_workspace_bytes(int64_t n_pairs);
int snapgpu_panel_spmm_f64(const snapgpu_panel_layout* layout, const double* d_dense, int32_t square_values,
                           const double* d_row_scale, double* d_out, void* d_ws, int64_t ws_bytes,
                           void* stream);

/* Window layout (the same products as snapgpu_spectral_apply, b <= 16, counts
 * below 256; replaces the matvec of the reference's eigensolver call,
 * snapatac2-core/src/embedding.rs:155-171 via the operator of :269-326). One
 * workgroup keeps the sums of a block of output rows in 2048 LDS slots and walks
 * the dense operand in windows of snapgpu_window_count's 1024 rows, so the
 * gathers of one window reuse L1; within a window a row's entries are cut into pieces of at most 16, each
 * with its own slot, and a row's slots are added in order at the end: fixed
 * summation order. Build: snapgpu_window_plan (slots per row, workgroups;
 * returns the block rows and row slots on the host), upload them, then
 * snapgpu_window_build.
 *   ent            n_entries uint32: column in window | slot << 10 | piece start
 *                  << 21 | count << 24
 *   bucket_start   n_blocks * n_windows + 1 int64 (bucket = (block, window))
 *   quad_off       n_blocks * n_windows * 257 uint32 (piece-aligned chunk of
 *                  each of the 256 quads of a workgroup, relative to its bucket)
 *   block_row      n_blocks + 1 int64 (first row of each workgroup)
 *   row_slot       n_rows uint32 (first slot | number of slots << 16) */
typedef struct snapgpu_window_layout {
  const uint32_t* ent;
  const int64_t* bucket_start;
  const uint32_t* quad_off;
  const int64_t* block_row;
  const uint32_t* row_slot;
  int64_t n_blocks;
  int64_t n_windows;
  int64_t n_rows;
  int64_t dense_rows;
  int64_t n_entries;
} snapgpu_window_layout;

int64_t snapgpu_window_count(int64_t dense_rows); /* windows of 1024 dense rows */
int64_t snapgpu_window_plan_workspace_bytes(int64_t nrows);
int snapgpu_window_plan(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val,
                        int64_t nrows, int64_t dense_rows, int64_t nnz, void* d_ws,
                        int64_t ws_bytes, int64_t* h_block_row, int64_t max_blocks,
                        int64_t* n_blocks, uint32_t* h_row_slot, void* stream);
int64_t snapgpu_window_build_workspace_bytes(int64_t n_blocks, int64_t dense_rows);
int snapgpu_window_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val,
                         int64_t nrows, int64_t dense_rows, int64_t nnz,
                         const int64_t* d_block_row, int64_t n_blocks, const uint32_t* d_row_slot,
                         void* d_ws, int64_t ws_bytes, uint32_t* d_ent, int64_t* d_bucket_start,
                         uint32_t* d_quad_off, void* stream);
/* Y = A V over the two window layouts (csr: rows = cells, dense_rows = features;
 * csc: rows = features, dense_rows = cells), b <= 16. */
int64_t snapgpu_spectral_apply_window_workspace_bytes(int64_t n, int64_t m);
int snapgpu_spectral_apply_window(const snapgpu_window_layout* csr, const snapgpu_window_layout* csc,
                                  const float* d_r32, const float* d_w2, const double* d_dinv,
                                  const double* d_V, int64_t ldv, int32_t b, double* d_Y,
                                  int64_t ldy, void* d_ws, int64_t ws_bytes, void* stream);
/* Y = A V with the CSC product over the panel layout (rows = features,
 * dense_rows = cells) and the CSR product over the window layout (rows = cells,
 * dense_rows = features), b <= 16: the faster layout for each product at a
 * million cells. The same matvec as snapgpu_spectral_apply (embedding.rs:155-171). */
int64_t snapgpu_spectral_apply_mixed_workspace_bytes(int64_t n, int64_t m, int64_t pairs_csc);
int snapgpu_spectral_apply_mixed(const snapgpu_window_layout* csr, const snapgpu_panel_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv,
                                 const double* d_V, int64_t ldv, int32_t b, double* d_Y,
                                 int64_t ldy, void* d_ws, int64_t ws_bytes, void* stream);
/* snapgpu_spectral_vjp_w with its two FP32 sparse products over the mixed layout (the CSC
 * panel layout and the CSR window layout of snapgpu_spectral_apply_mixed), for
 * integer counts and b <= 8; the FP64 passes read the CSR and CSC arrays as
 * above. d_pws: snapgpu_spectral_vjp_mixed_workspace_bytes(csc n_pairs) bytes for
 * the panel product's pair sums (the mixed apply's workspace is large enough). */
int64_t snapgpu_spectral_vjp_mixed_workspace_bytes(int64_t pairs_csc);
int snapgpu_spectral_vjp_w_mixed(const snapgpu_window_layout* csr_windows,
                                 const snapgpu_panel_layout* csc_panels, const int64_t* d_indptr,
                                 const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                                 const int64_t* d_colptr, const int32_t* d_rowidx,
                                 const uint32_t* d_ccnt, int64_t m, const double* d_w,
                                 const double* d_s, const double* d_dinv, const float* d_r32,
                                 const float* d_w2, const double* d_V, int64_t ldv, const double* d_G,
                                 int64_t ldg, int32_t b, double* d_dw, void* d_ws, int64_t ws_bytes,
                                 void* d_pws, int64_t pws_bytes, void* stream);

/* Block-tiled layout (an LDS-staged form of the same products for b <= 16,
 * slower than the panel layout at a million cells). One
 * workgroup owns a block of output rows (rows ordered by decreasing entry
 * count, blocks of about equal work) and keeps their sums in LDS; the dense
 * operand is swept in panels of snapgpu_block_panel_rows() rows, each read
 * once per block into LDS, from which every entry of the tile gathers its row.
 * A row's entries inside one panel are cut into pieces of at most 16, each
 * with its own accumulator slot; a tile stores its pieces by decreasing length,
 * entry i of each piece in row i, so each thread of the workgroup reads one
 * piece with loads coalesced across its wave, and every slot has one writer:
 * fixed summation order. Build: snapgpu_block_plan (row order, pieces, blocks;
 * its workspace keeps them), then allocate, then snapgpu_block_build.
 *   ent / val      n_entries uint32 (val only when not packed: counts >= 1024
 *                  or float values)
 *   tile_start     n_blocks * n_panels + 1 int64 (tile = (block, panel))
 *   tile_off       n_blocks * n_panels * 16 uint16 (end of row i of each tile)
 *   row0           n_blocks + 1 int32 (first out position of each block)
 *   out_row        n_rows int32 (row of each out position)
 *   vslot          n_rows int32 (first slot of each out position in its block)
 *   nslots         n_blocks int32 */
typedef struct snapgpu_block_layout {
  const uint32_t* ent;
  const uint32_t* val;
  const int64_t* tile_start;
  const uint16_t* tile_off;
  const int32_t* row0;
  const int32_t* out_row;
  const int32_t* vslot;
  const int32_t* nslots;
  int64_t n_blocks;
  int64_t n_panels;
  int64_t n_rows;
  int64_t dense_rows;
  int64_t n_entries;
  int32_t packed;
  int32_t value_type;
} snapgpu_block_layout;

int32_t snapgpu_block_panel_rows(void);
int64_t snapgpu_block_plan_workspace_bytes(int64_t nrows);
/* Orders the rows, sizes their pieces and forms about target_blocks blocks;
 * returns the block count, the size of the build's piece table and whether the
 * entries pack. Synchronises. */
int snapgpu_block_plan(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                       int64_t dense_rows, int64_t nnz, int32_t value_type, int64_t target_blocks, void* d_ws,
                       int64_t ws_bytes, int64_t* n_blocks, int64_t* table_size, int32_t* packed, void* stream);
int64_t snapgpu_block_build_workspace_bytes(int64_t table_size);
/* Fills the layout arrays from the plan (d_plan_ws as left by snapgpu_block_plan).
 * Synchronises. */
int snapgpu_block_build(const int64_t* d_ptr, const int32_t* d_idx, const uint32_t* d_val, int64_t nrows,
                        int64_t dense_rows, int64_t nnz, int32_t packed, const void* d_plan_ws, int64_t plan_ws_bytes,
                        int64_t n_blocks, int64_t table_size, void* d_ws, int64_t ws_bytes, uint32_t* d_ent,
                        uint32_t* d_val_out, int64_t* d_tile_start, uint16_t* d_tile_off, int32_t* d_row0,
                        int32_t* d_out_row, int32_t* d_vslot, int32_t* d_nslots, void* stream);
/* Y = A V as snapgpu_spectral_apply (b <= 16), over the two block layouts (csr:
 * rows = cells, dense_rows = features; csc: rows = features, dense_rows = cells). */
int64_t snapgpu_spectral_apply_block_workspace_bytes(int64_t n, int64_t m);
int snapgpu_spectral_apply_block(const snapgpu_block_layout* csr, const snapgpu_block_layout* csc,
                                 const float* d_r32, const float* d_w2, const double* d_dinv, const double* d_V,
                                 int64_t ldv, int32_t b, double* d_Y, int64_t ldy, void* d_ws, int64_t ws_bytes,
                                 void* stream);

/* FP64 tall-skinny products for the block Lanczos basis (the work ARPACK's
 * dsaupd does inside eigsh, embedding.rs:164). Row-major operands.
 * C[ka x kb] = A[n x ka]^T B[n x kb], fixed-order reduction; kb <= 32. */
int64_t snapgpu_gemm_tn_workspace_bytes(int64_t n, int32_t ka, int32_t kb);
int snapgpu_gemm_tn_f64(const double* d_A, int64_t lda, const double* d_B, int64_t ldb,
                        int64_t n, int32_t ka, int32_t kb, double* d_C, void* d_ws,
                        int64_t ws_bytes, void* stream);

/* Allocation-free form of snapgpu_gemm_tn_f64. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-chunk partial products go to a static device array of
 * SNAPGPU_GEMM_TN_SCRATCH_BYTES held by the library on every GPU it runs on
 * (until the process ends) instead of a caller workspace; the call allocates
 * nothing. It fails with SNAPGPU_ERR_CAPACITY when
 * snapgpu_gemm_tn_workspace_bytes(n, ka, kb) - 256 exceeds that size (at most
 * 512 row chunks x ka x kb doubles: ka * kb <= 8192 always fits). Every call on
 * one GPU shares that GPU's array, so calls that overlap in time on one GPU
 * (other streams, other threads) race; calls ordered on one stream do not.
 * Same result, bit for bit, as snapgpu_gemm_tn_f64. */
#define SNAPGPU_GEMM_TN_SCRATCH_BYTES (32ll << 20)
int snapgpu_gemm_tn_f64_noalloc(const double* d_A, int64_t lda, const double* d_B, int64_t ldb,
                                int64_t n, int32_t ka, int32_t kb, double* d_C, void* stream);
/* C[n x kb] = beta*C + alpha*A[n x ka] B[ka x kb]; kb <= 32. */
int snapgpu_gemm_nn_f64(const double* d_A, int64_t lda, const double* d_B, int64_t n, int32_t ka,
                        int32_t kb, double alpha, double beta, double* d_C, int64_t ldc,
                        void* stream);
/* Seeded uniform(-1/2, 1/2) values (Lanczos start block). */
int snapgpu_random_f64(uint64_t seed, int64_t count, double* d_out, void* stream);

/* ---- Fragment-file import --------------------------------------------------
 * Replaces internal.import_fragments (snapatac2-python/src/preprocessing.rs:87-171)
 * and its core, preprocessing::import_fragments / make_arraydata /
 * count_fragments (snapatac2-core/src/preprocessing/import.rs:31-204): the
 * text of a fragment file ("chrom\tstart\tend\tbarcode[\tcount[\tstrand]]"
 * lines, '#' comment lines skipped; Fragment::from_str, qc.rs:306-345),
 * already decompressed and on the device (16-byte aligned), becomes
 * obsm['fragment_paired'] (uint32 lengths) or obsm['fragment_single'] (int32
 * signed lengths) as a CSR with int64 row pointers and uint32 genome base
 * positions, plus the FragmentQC counts of every stored barcode (qc.rs:427-466).
 * Rows: barcodes in file order (input sorted by barcode) or in byte order
 * (unsorted input: the order of the reference's external sort), restricted to
 * the whitelist and to at least min_num_fragments unique fragments; entries
 * ascend by (position, value) inside a row.
 * Sequence: snapgpu_frag_table_pack (host) for the chromosomes and the
 * optional whitelist -> copy both to the device -> snapgpu_frag_count_lines ->
 * snapgpu_frag_index (workspace from snapgpu_frag_workspace_bytes) -> allocate
 * n_rows + 1 row pointers, nnz positions and values -> snapgpu_frag_fill.
 * Records the reference would panic on fail with SNAPGPU_ERR_DATA and a
 * message naming the 1-based line. */
typedef struct snapgpu_frag_options {
  int64_t shift_left;          /* added to start (and to end of single-end records) */
  int64_t shift_right;         /* added to end of paired-end records */
  uint64_t min_num_fragments;  /* rows need at least this many unique fragments */
  int32_t sorted_by_barcode;   /* 1: groups are runs of equal barcodes in file order */
  int32_t reserved;
} snapgpu_frag_options;

typedef struct snapgpu_frag_summary {
  int64_t n_lines;    /* lines of the text */
  int64_t n_records;  /* non-comment, non-empty lines */
  int64_t n_kept;     /* records with end != start */
  int64_t n_groups;   /* barcodes among them */
  int64_t n_rows;     /* barcodes stored */
  int64_t nnz;        /* fragments stored */
  int32_t paired;     /* 1: fragment_paired, 0: fragment_single */
  int32_t name_width; /* bytes of the longest barcode */
} snapgpu_frag_summary;

/* Host-only. A hashed name table: genome = 1 for the chromosomes (names in the
 * stored order of uns['reference_sequences'], lengths, mito[k] = 1 for the
 * mitochondrial names; the genome must be shorter than 2^32 bases), genome = 0
 * for a barcode whitelist (lengths and mito may be NULL). */
int64_t snapgpu_frag_table_bytes(const char* const* names, int32_t n);
int snapgpu_frag_table_pack(const char* const* names, const uint64_t* lengths, const uint8_t* mito, int32_t n,
                            int32_t genome, void* h_out, int64_t out_bytes);
/* Number of lines of the text (a last line without '\n' counts). Synchronises. */
int64_t snapgpu_frag_count_workspace_bytes(int64_t n_bytes);
int snapgpu_frag_count_lines(const uint8_t* d_text, int64_t n_bytes, void* d_ws, int64_t ws_bytes,
                             int64_t* h_n_lines, void* stream);
/* Parse, group, QC and filter; the workspace keeps the state snapgpu_frag_fill
 * reads. d_whitelist may be NULL. Synchronises. */
int64_t snapgpu_frag_workspace_bytes(int64_t n_bytes, int64_t n_lines);
int snapgpu_frag_index(const uint8_t* d_text, int64_t n_bytes, int64_t n_lines, const void* d_chroms,
                       const void* d_whitelist, const snapgpu_frag_options* opt, void* d_ws, int64_t ws_bytes,
                       snapgpu_frag_summary* h_summary, void* stream);
/* d_cell_ptr[n_rows + 1], d_pos[nnz], d_val[nnz]; d_qc[3 n_rows] (optional):
 * per row n_fragment (unique), total count, mitochondrial; d_names[n_rows x
 * name_width] (optional): barcodes, zero padded. */
int snapgpu_frag_fill(const uint8_t* d_text, int64_t n_bytes, const void* d_ws, int64_t ws_bytes,
                      const snapgpu_frag_summary* h_summary, int64_t* d_cell_ptr, uint32_t* d_pos,
                      uint32_t* d_val, uint64_t* d_qc, uint8_t* d_names, int32_t name_width, void* stream);

/* Export: fragments -> text lines "chrom\tstart\tend\tbarcode\t1[\tstrand]\n",
 * as export_fragments writes them (snapatac2-core/src/export.rs:90-106, with
 * Fragment's Display, qc.rs:286-303). Line k prints fragment d_order[k] (or k
 * if d_order is NULL); d_names[n_cells x name_width] holds the barcodes. Call
 * once with d_text = NULL (line ends into d_line_end[n_fragments], total bytes
 * to *h_text_bytes), then with the text buffer. Synchronises. */
int64_t snapgpu_frag_format_workspace_bytes(int64_t n_fragments);
int snapgpu_frag_format(const int64_t* d_cell_ptr, int64_t n_cells, const uint32_t* d_pos, const uint32_t* d_val,
                        int64_t n_fragments, int32_t paired, const void* d_chroms, const uint8_t* d_names,
                        int32_t name_width, const int64_t* d_order, int64_t* d_line_end, uint8_t* d_text,
                        int64_t text_bytes, void* d_ws, int64_t ws_bytes, int64_t* h_text_bytes, void* stream);

/* ---- Nystrom spectral embedding ----------------------------------------------
 * Replaces internal.spectral_embedding_nystrom (snapatac2-python/src/embedding.rs:60-131)
 * with nystrom (:193-268) and idf_from_chunks (:288-312). The host picks the
 * landmark cells; the landmark embedding is the ordinary spectral path above.
 * Rows d_rows[n_take] of a CSR (int32 columns, uint32 values): row pointers
 * d_out_indptr[n_take + 1] and *h_nnz first (synchronises), then the entries. */
int64_t snapgpu_csr_take_rows_workspace_bytes(int64_t n_take);
int snapgpu_csr_take_rows_count(const int64_t* d_indptr, int64_t n_rows, const int64_t* d_rows, int64_t n_take,
                                int64_t* d_out_indptr, void* d_ws, int64_t ws_bytes, int64_t* h_nnz, void* stream);
int snapgpu_csr_take_rows_fill(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_val,
                               const int64_t* d_rows, int64_t n_take, const int64_t* d_out_indptr,
                               int32_t* d_out_idx, uint32_t* d_out_val, void* stream);
/* Entries per column (document frequency of idf_from_chunks), uint64 d_out[m];
 * d_ws16: 16 bytes of device scratch. */
int snapgpu_csr_col_counts(const int64_t* d_indptr, int64_t n, const int32_t* d_idx, int64_t nnz, int64_t m,
                           uint64_t* d_out, int32_t* d_ws16, void* stream);
/* d_s[i] = 1 / ||x_i o w||_2, 0 for an empty row (normalize, embedding.rs:314-326). */
int snapgpu_csr_row_inv_norm(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                             const double* d_w, double* d_s, void* stream);
/* C[i, :] = row_scale[i] * sum_j count_ij col_scale[j] B[j, :] (FP64, row-major,
 * k columns; NULL scales mean 1): both sparse products of nystrom
 * (seed^T @ evecs on the landmark CSC, sample @ that on all cells). */
int snapgpu_csr_spmm_scaled_f64(const int64_t* d_indptr, const int32_t* d_idx, const uint32_t* d_cnt, int64_t n,
                                const double* d_row_scale, const double* d_col_scale, const double* d_B, int64_t ldb,
                                int32_t k, double* d_C, int64_t ldc, void* stream);
/* E = U diag(1/evals) with row i divided by sqrt(dinv[i]): nystrom's scaling of the
 * landmark eigenvectors (embedding.rs:203-214). Row-major FP64. */
int snapgpu_nystrom_landmark_scale(const double* d_U, int64_t ldu, int64_t n, int32_t k, const double* d_dinv,
                                   const double* d_evals, double* d_E, int64_t lde, void* stream);
/* nystrom's per-chunk normalisation of Q (n x k), chunks of chunk_rows rows:
 * t = colsum(Q_chunk) * evals; d = Q t; d <= 0 -> the chunk's smallest
 * positive d; Q /= sqrt(d). SNAPGPU_ERR_DATA if a chunk needs that minimum and
 * has no positive d. Synchronises. */
int64_t snapgpu_nystrom_normalize_workspace_bytes(int64_t n, int32_t k, int64_t chunk_rows);
int snapgpu_nystrom_normalize(double* d_Q, int64_t ldq, int64_t n, int32_t k, const double* d_evals,
                              int64_t chunk_rows, void* d_ws, int64_t ws_bytes, void* stream);

/* ---- Cell x peak matrix ---------------------------------------------------------
 * Replaces internal.mk_peak_matrix (snapatac2-python/src/preprocessing.rs:370-420) ->
 * create_peak_matrix (snapatac2-core/src/feature_count/matrix.rs:95-171) with
 * RegionCounter (feature_count/counter.rs:73-159) over the fragments of
 * obsm['fragment_paired'|'fragment_single'] (data_iter.rs:27-123).
 * The region index (device arrays, built on the host from the peak list):
 * regions grouped by chromosome in the stored chromosome order, sorted by start
 * inside a chromosome, pmax = running maximum of end inside the chromosome,
 * orig = the region's column (first-occurrence order of the distinct regions),
 * chrom_off[n_chrom + 1] = each chromosome's range, base[n_chrom + 1] = genome
 * offsets of the chromosomes, n_columns = the number of region columns (every
 * orig < n_columns; 0 when unknown, which sorts 32 region bits). Sequence: snapgpu_peak_count (increments per
 * fragment, scanned; total to *h_total) -> snapgpu_peak_pairs (keys sorted and
 * grouped; distinct (cell, region) pairs to *h_nnz) -> snapgpu_peak_csr. */
typedef struct snapgpu_peak_index {
  const uint64_t* start;
  const uint64_t* end;
  const uint64_t* pmax;
  const uint32_t* orig;
  const int64_t* chrom_off;
  const uint64_t* base;
  int32_t n_chrom;
  int32_t n_columns;
} snapgpu_peak_index;

int64_t snapgpu_peak_count_workspace_bytes(int64_t n_fragments);
int snapgpu_peak_count(const snapgpu_peak_index* index, const uint32_t* d_pos, const uint32_t* d_val,
                       int64_t n_fragments, int32_t paired, int32_t strategy, uint32_t min_len, uint32_t max_len,
                       int64_t* d_hits, void* d_ws, int64_t ws_bytes, int64_t* h_total, void* stream);
int64_t snapgpu_peak_pairs_workspace_bytes(int64_t total);
int snapgpu_peak_pairs(const snapgpu_peak_index* index, const int64_t* d_cell_ptr, int64_t n_cells,
                       const uint32_t* d_pos, const uint32_t* d_val, int64_t n_fragments, int32_t paired,
                       int32_t strategy, uint32_t min_len, uint32_t max_len, const int64_t* d_hits, int64_t total,
                       uint64_t* d_keys, void* d_ws, int64_t ws_bytes, int64_t* h_nnz, void* stream);
/* d_indptr[n_cells + 1], d_idx[nnz] (int32 region columns, ascending per row),
 * d_val[nnz] (uint32 counts), from the workspace of snapgpu_peak_pairs. */
int snapgpu_peak_csr(const void* d_ws, int64_t ws_bytes, int64_t total, int64_t nnz, int64_t n_cells,
                     int64_t* d_indptr, int32_t* d_idx, uint32_t* d_val, void* stream);

/* Gene level of the gene matrix (create_gene_matrix with id_type "gene",
 * feature_count/matrix.rs:173-273; GeneCount::get_values, counter.rs:270-283):
 * each row's entries grouped by d_group[column] (the gene of the region's
 * transcript), keeping the largest count of a group. snapgpu_group_max sorts
 * and groups (distinct (row, group) pairs to *h_nnz_out; synchronises), then
 * snapgpu_group_max_csr writes d_indptr[n_rows + 1], d_idx (groups, ascending
 * per row) and d_val from its workspace. n_groups: every d_group value is below
 * it (0 when unknown, which sorts 32 group bits). */
int64_t snapgpu_group_max_workspace_bytes(int64_t nnz);
int snapgpu_group_max(const int64_t* d_indptr, int64_t n_rows, const int32_t* d_idx, const uint32_t* d_val,
                      int64_t nnz, const uint32_t* d_group, int32_t n_groups, void* d_ws, int64_t ws_bytes,
                      int64_t* h_nnz_out, void* stream);
int snapgpu_group_max_csr(const void* d_ws, int64_t ws_bytes, int64_t nnz, int64_t nnz_out, int64_t n_rows,
                          int64_t* d_indptr, int32_t* d_idx, uint32_t* d_val, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SNAPGPU_H */
