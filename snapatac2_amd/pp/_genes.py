"""Drop-in for snapatac2.pp.make_gene_matrix (cell x gene activity), on the GPU.

Reference: snapatac2-python/python/snapatac2/preprocessing/_basic.py:707-827
(Python API), internal.mk_gene_matrix (snapatac2-python/src/preprocessing.rs:432-499),
read_transcripts_from_gtf / _gff and Promoters::new (snapatac2-core/src/genome.rs:
64-230), create_gene_matrix (feature_count/matrix.rs:173-273), TranscriptCount /
GeneCount (feature_count/counter.rs:160-290).

The annotation is read on the host (transcript records of a GTF or GFF file)
and turned into promoter regions; the region counts are the GPU peak-matrix
kernels; the gene level (the largest count among a gene's transcripts) is the
GPU group-max.
"""
from __future__ import annotations

import ctypes as C
import gzip
import re
from pathlib import Path

import numpy as np

from .. import _lib
from .._device import DeviceCSR
from ..anndata import AnnData
from ..genome import read_chrom_sizes
from ._basic import FRAGMENT_PAIRED, FRAGMENT_SINGLE, _as_device_fragments
from ._peaks import parse_regions, region_counts

_GTF_ATTR = re.compile(r'\s*([^\s"]+)\s+(?:"([^"]*)"|([^;\s]+))\s*')


def _attrs_gtf(s: str) -> dict:
    out = {}
    for part in s.strip().split(";"):
        if part.strip():
            m = _GTF_ATTR.fullmatch(part)
            if m:
                out.setdefault(m.group(1), m.group(2) if m.group(2) is not None else m.group(3))
    return out


def _attrs_gff(s: str) -> dict:
    out = {}
    for part in s.strip().split(";"):
        if "=" in part:
            k, v = part.split("=", 1)
            out.setdefault(k.strip(), v.strip())
    return out


def read_transcripts(path, transcript_name_key="transcript_name", transcript_id_key="transcript_id",
                     gene_name_key="gene_name", gene_id_key="gene_id"):
    """Transcript records (type 'transcript') of a GTF (or GFF: .gff / .gff3
    suffix before an optional .gz) as dicts with 1-based left/right
    (read_transcripts_from_gtf / _gff, genome.rs:64-190)."""
    p = Path(path)
    name = p.name[:-3] if p.name.endswith(".gz") else p.name
    gff = name.endswith((".gff", ".gff3"))
    opener = gzip.open if p.name.endswith(".gz") else open
    out = []
    with opener(p, "rt") as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            f = line.rstrip("\n").split("\t")
            if len(f) < 9:
                raise _lib.PanicException(f"failed to parse {'GFF' if gff else 'GTF'} line: {line.strip()}")
            if f[2] != "transcript":
                continue
            a = _attrs_gff(f[8]) if gff else _attrs_gtf(f[8])
            for key in (transcript_id_key, gene_name_key, gene_id_key):
                if key not in a:
                    raise _lib.PanicException(f"failed to find '{key}' in record: {line.strip()}")
            out.append(dict(chrom=f[0], left=int(f[3]), right=int(f[4]), strand=f[6],
                            transcript_id=a[transcript_id_key], gene_name=a[gene_name_key],
                            gene_id=a[gene_id_key], transcript_name=a.get(transcript_name_key)))
    return out


def promoter_regions(transcripts, upstream: int, downstream: int, include_gene_body: bool):
    """Promoters::new (genome.rs:197-230): "chrom:start-end" per transcript."""
    regs = []
    for t in transcripts:
        left, right = t["left"] - 1, t["right"] - 1
        if t["strand"] == "+":
            start, end = max(left - upstream, 0), downstream + (right if include_gene_body else left)
        elif t["strand"] == "-":
            start, end = max((left if include_gene_body else right) - downstream, 0), right + upstream
        else:
            raise _lib.PanicException(f"Miss strand information for {t['transcript_id']}")
        regs.append(f"{t['chrom']}:{start}-{end}")
    return regs


def _promoter_tuples(transcripts, upstream: int, downstream: int, include_gene_body: bool):
    """The distinct regions of promoter_regions as (chrom, start, end), as
    parse_regions reads them back from the strings (the strip of the whole
    string removes the chromosome's leading spaces); None when a region has a
    negative end, which the string parse rejects."""
    seen, regs = set(), []
    for t in transcripts:
        left, right = t["left"] - 1, t["right"] - 1
        if t["strand"] == "+":
            start, end = max(left - upstream, 0), downstream + (right if include_gene_body else left)
        elif t["strand"] == "-":
            start, end = max((left if include_gene_body else right) - downstream, 0), right + upstream
        else:
            raise _lib.PanicException(f"Miss strand information for {t['transcript_id']}")
        if end < 0:
            return None
        r = (str(t["chrom"]).lstrip(), start, end)
        if r not in seen:
            seen.add(r)
            regs.append(r)
    return regs


def gene_counts(frags, chrom_sizes, transcripts, id_type="gene", upstream=2000, downstream=0,
                include_gene_body=True, counting_strategy="paired-insertion", min_frag_size=None,
                max_frag_size=None):
    """(DeviceCSR, var_names, gene_name per transcript or None)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    st = _lib.stream_ptr()
    # counts per distinct region; the reference labels distinct region k with transcript k
    regs = _promoter_tuples(transcripts, upstream, downstream, include_gene_body)
    if regs is None:  # the string parse raises the reference's error
        regs = parse_regions(promoter_regions(transcripts, upstream, downstream, include_gene_body))[1]
    X = region_counts(frags, chrom_sizes, regs, counting_strategy, min_frag_size, max_frag_size)
    n_cells = X.shape[0]
    if id_type == "transcript":
        X = DeviceCSR(X.indptr, X.indices, X.data, (n_cells, len(transcripts)))
        return X, [t["transcript_id"] for t in transcripts], [t["gene_name"] for t in transcripts]
    if id_type != "gene":
        raise _lib.PanicException("id_type must be 'transcript' or 'gene'")
    genes = {}
    for t in transcripts:
        genes.setdefault(t["gene_name"], len(genes))
    group = np.asarray([genes[t["gene_name"]] for t in transcripts] or [0], dtype=np.uint32)
    dev = X.indptr.device
    d_group = torch.from_numpy(group.view(np.int32)).to(dev)
    nnz = X.nnz
    indptr = torch.zeros(n_cells + 1, dtype=torch.int64, device=dev)
    if nnz == 0 or n_cells == 0:
        return DeviceCSR(indptr, torch.empty(0, dtype=torch.int32, device=dev),
                         torch.empty(0, dtype=torch.int32, device=dev), (n_cells, len(genes))), list(genes), None
    wsb = int(L.snapgpu_group_max_workspace_bytes(nnz))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    out_nnz = C.c_int64(0)
    _lib.check(L.snapgpu_group_max(X.indptr.data_ptr(), n_cells, X.indices.data_ptr(), X.data.data_ptr(), nnz,
                                   d_group.data_ptr(), len(genes), ws.data_ptr(), wsb, C.byref(out_nnz), st))
    idx = torch.empty(max(out_nnz.value, 1), dtype=torch.int32, device=dev)
    val = torch.empty(max(out_nnz.value, 1), dtype=torch.int32, device=dev)
    _lib.check(L.snapgpu_group_max_csr(ws.data_ptr(), wsb, nnz, out_nnz.value, n_cells, indptr.data_ptr(),
                                       idx.data_ptr(), val.data_ptr(), st))
    return DeviceCSR(indptr, idx[:out_nnz.value], val[:out_nnz.value], (n_cells, len(genes))), list(genes), None


def make_gene_matrix(
    adata,
    gene_anno,
    *,
    inplace: bool = False,
    file: Path | None = None,
    backend="hdf5",
    chunk_size: int = 500,
    use_x: bool = False,
    id_type: str = "gene",
    upstream: int = 2000,
    downstream: int = 0,
    include_gene_body: bool = True,
    transcript_name_key: str = "transcript_name",
    transcript_id_key: str = "transcript_id",
    gene_name_key: str = "gene_name",
    gene_id_key: str = "gene_id",
    min_frag_size: int | None = None,
    max_frag_size: int | None = None,
    counting_strategy: str = "paired-insertion",
):
    """Cell by gene activity matrix (snapatac2.pp.make_gene_matrix, _basic.py:707-827).

    gene_anno: a GTF/GFF path, or an object with an `annotation` path (a
    Genome). `use_x` and backed output (`file`) are not provided here.
    """
    if hasattr(gene_anno, "annotation"):
        gene_anno = gene_anno.annotation
    if use_x:
        raise NotImplementedError("use_x=True (counts from .X) is not provided")
    if file is not None:
        raise NotImplementedError("backed AnnData output needs HDF5, which is not available here")
    if FRAGMENT_SINGLE in adata.obsm:
        frags = _as_device_fragments(adata.obsm[FRAGMENT_SINGLE], False)
    elif FRAGMENT_PAIRED in adata.obsm:
        frags = _as_device_fragments(adata.obsm[FRAGMENT_PAIRED], True)
    else:
        raise RuntimeError("No fragment data found in the anndata object")
    tx = read_transcripts(gene_anno, transcript_name_key, transcript_id_key, gene_name_key, gene_id_key)
    X, names, gene_names = gene_counts(frags, read_chrom_sizes(adata.uns), tx, id_type, upstream, downstream,
                                       include_gene_body, counting_strategy, min_frag_size, max_frag_size)
    import pandas as pd
    if inplace:
        target = adata
    else:
        target = AnnData(X=X, obs=adata.obs.copy() if hasattr(adata, "obs") else None)
    if isinstance(target, AnnData):
        target.set_n_vars(len(names), names)
        target.X = X
    else:
        target.X = X.to_scipy()
        target.var_names = pd.Index(names)
    if gene_names is not None:
        target.var = pd.DataFrame({"gene_name": gene_names}, index=pd.Index(names))
    return None if inplace else target
