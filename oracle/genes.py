"""ORACLE (test infrastructure only) for the gene matrix.

Restates create_gene_matrix (snapatac2-core/src/feature_count/matrix.rs:173-273):
promoter regions of Promoters::new (snapatac2-core/src/genome.rs:197-230), the
region counts of oracle/peaks.py (RegionCounter over the distinct regions; the
reference labels distinct region k with transcript k, TranscriptCount,
counter.rs:214-245), and for id_type 'gene' the largest count among a gene's
transcripts (GeneCount::get_values, counter.rs:270-283), genes in
first-occurrence order.

Pinned by the hand-computed case of tests/test_genes_oracle.py.
"""
from __future__ import annotations

import numpy as np

from .peaks import peak_matrix


def promoters(transcripts, upstream, downstream, include_gene_body):
    out = []
    for t in transcripts:
        left, right = t["left"] - 1, t["right"] - 1
        if t["strand"] == "+":
            out.append(f"{t['chrom']}:{max(left - upstream, 0)}-{downstream + (right if include_gene_body else left)}")
        else:
            out.append(f"{t['chrom']}:{max((left if include_gene_body else right) - downstream, 0)}-{right + upstream}")
    return out


def gene_matrix(cell_ptr, pos, val, paired, chrom_sizes, transcripts, id_type="gene", upstream=2000,
                downstream=0, include_gene_body=True, strategy="paired-insertion", min_len=None, max_len=None,
                region_counter=peak_matrix):
    """region_counter: oracle.peaks.peak_matrix, or peak_matrix_cpp (the C++
    restatement) for the timed CPU baseline."""
    regs = promoters(transcripts, upstream, downstream, include_gene_body)
    ip, ix, d = region_counter(cell_ptr, pos, val, paired, chrom_sizes, regs, strategy, min_len, max_len)
    if id_type == "transcript":
        return ip, ix, d, [t["transcript_id"] for t in transcripts]
    genes = {}
    for t in transcripts:
        genes.setdefault(t["gene_name"], len(genes))
    indptr, indices, data = [0], [], []
    for r in range(len(ip) - 1):
        best = {}
        for k, v in zip(ix[ip[r]:ip[r + 1]], d[ip[r]:ip[r + 1]]):
            g = genes[transcripts[int(k)]["gene_name"]]
            best[g] = max(best.get(g, 0), int(v))
        for g in sorted(best):
            indices.append(g)
            data.append(best[g])
        indptr.append(len(indices))
    return (np.asarray(indptr, dtype=np.int64), np.asarray(indices, dtype=np.int64),
            np.asarray(data, dtype=np.uint32), list(genes))
