"""Seeded synthetic scATAC fragment workloads, generated on the device.

Generator spec (all configurations, seed 0 by default):
  genome: hg38 chr1-22, X, Y in stored (lexicographic) order, 3,088,269,832 bp;
  50,000 peak centres uniform over the genome; 10 cell types, each with a
  Dirichlet(0.1) activity vector over the peaks; cell type uniform;
  fragments per cell ~ LogNormal(ln 5000, 0.8) rounded, clipped to [500, 100000];
  start: 30% uniform, 70% peak centre + round(N(0, 150));
  length: 60% Exp(mean 80), 25% N(200, 20), 15% N(400, 40), rounded, in [1, 4095].
The per-fragment draws come from a counter-based hash of (seed, cell, fragment),
so a cell's fragments do not depend on how the cells are chunked.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._device import DeviceFragments
from .anndata import AnnData
from .genome import reference_sequences, synthetic_genome

DEFAULT_CHUNK_FRAGS = 256 * 1024 * 1024


class GeneratorTables:
    def __init__(self, seed: int = 0, chrom_sizes: dict[str, int] | None = None,
                 n_peaks: int = 50_000, n_types: int = 10, alpha: float = 0.1):
        self.seed = int(seed)
        self.chrom_sizes = chrom_sizes or synthetic_genome()
        lens = np.asarray(list(self.chrom_sizes.values()), dtype=np.uint64)
        self.chrom_start = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        genome = int(self.chrom_start[-1])
        rng = np.random.Generator(np.random.PCG64(self.seed))
        self.peak_centre = rng.integers(0, genome, n_peaks, dtype=np.uint64).astype(np.uint32)
        act = rng.dirichlet(np.full(n_peaks, alpha), size=n_types)
        cdf = np.cumsum(act, axis=1)
        cdf /= cdf[:, -1:]
        cdf = cdf.astype(np.float32)
        cdf[:, -1] = 1.0
        self.type_cdf = np.ascontiguousarray(cdf)
        self.n_peaks, self.n_types = n_peaks, n_types


def fragment_counts(n_cells: int, seed: int = 0, first_cell: int = 0, device=None,
                    mu: float = math.log(5000.0), sigma: float = 0.8,
                    lo: int = 500, hi: int = 100_000):
    torch = _lib.require_gpu()
    dev = device or torch.device("cuda")
    out = torch.empty(n_cells, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().snapgpu_synth_counts(seed, first_cell, n_cells, mu, sigma, lo, hi,
                                               out.data_ptr(), _lib.stream_ptr()))
    return out


def cell_types(n_cells: int, seed: int = 0, first_cell: int = 0, n_types: int = 10, device=None):
    """The generator's cell type of each cell (int32 device tensor): the type
    whose peak activity the cell's peak fragments are drawn from."""
    torch = _lib.require_gpu()
    dev = device or torch.device("cuda")
    out = torch.empty(n_cells, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().snapgpu_synth_cell_types(seed, first_cell, n_cells, n_types, out.data_ptr(),
                                                   _lib.stream_ptr()))
    return out


def _scan_to_indptr(counts):
    torch = _lib.require_gpu()
    L = _lib.lib()
    n = counts.numel()
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=counts.device)
    indptr[1:] = counts
    ws_b = int(L.snapgpu_scan_workspace_bytes(n))
    ws = torch.empty(max(ws_b, 8), dtype=torch.uint8, device=counts.device)
    _lib.check(L.snapgpu_scan_inclusive_i64(indptr.data_ptr() + 8, n, ws.data_ptr(), ws_b,
                                            _lib.stream_ptr()))
    return indptr


def generate_chunk(tables: GeneratorTables, first_cell: int, n_cells: int, device=None,
                   p_uniform: float = 0.3, peak_sd: float = 150.0) -> DeviceFragments:
    """Fragments of cells first_cell .. first_cell + n_cells - 1 (row pointers start at 0)."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    dev = device or torch.device("cuda")
    counts = fragment_counts(n_cells, tables.seed, first_cell, dev)
    indptr = _scan_to_indptr(counts)
    nf = int(indptr[-1].item())
    pos = torch.empty(nf, dtype=torch.int32, device=dev)
    ln = torch.empty(nf, dtype=torch.int32, device=dev)
    ws_b = int(L.snapgpu_synth_workspace_bytes(nf))
    ws = torch.empty(ws_b, dtype=torch.uint8, device=dev)
    cs = torch.from_numpy(tables.chrom_start.view(np.int64)).to(dev)
    pc = torch.from_numpy(tables.peak_centre.view(np.int32)).to(dev)
    cdf = torch.from_numpy(tables.type_cdf).to(dev)
    _lib.check(L.snapgpu_synth_fragments(tables.seed, first_cell, n_cells, indptr.data_ptr(), nf,
                                         cs.data_ptr(), len(tables.chrom_sizes), pc.data_ptr(),
                                         tables.n_peaks, cdf.data_ptr(), tables.n_types,
                                         p_uniform, peak_sd, pos.data_ptr(), ln.data_ptr(),
                                         ws.data_ptr(), ws_b, _lib.stream_ptr()))
    del ws
    return DeviceFragments(indptr, pos, ln, int(tables.chrom_start[-1]), paired=True)


def iter_chunks(n_cells: int, seed: int = 0, max_chunk_frags: int = DEFAULT_CHUNK_FRAGS,
                tables: GeneratorTables | None = None, device=None):
    """Yields (first_cell, DeviceFragments) chunks of whole cells, each holding
    at most `max_chunk_frags` fragments (but at least one cell)."""
    torch = _lib.require_gpu()
    tables = tables or GeneratorTables(seed)
    counts = fragment_counts(n_cells, tables.seed, 0, device).cpu().numpy()
    c0 = 0
    cum = np.cumsum(counts)
    while c0 < n_cells:
        base = cum[c0 - 1] if c0 > 0 else 0
        c1 = int(np.searchsorted(cum, base + max_chunk_frags, side="right"))
        c1 = max(c1, c0 + 1)
        c1 = min(c1, n_cells, c0 + (1 << 20) - 1)
        yield c0, generate_chunk(tables, c0, c1 - c0, device)
        c0 = c1


def synthetic_anndata(n_cells: int, seed: int = 0, device=None) -> AnnData:
    """AnnData with obsm['fragment_paired'] (device resident) and
    uns['reference_sequences'] for `n_cells` synthetic cells."""
    tables = GeneratorTables(seed)
    frags = generate_chunk(tables, 0, n_cells, device)
    ad = AnnData(n_obs=n_cells, n_vars=0)
    ad.obsm["fragment_paired"] = frags
    ad.uns["reference_sequences"] = reference_sequences(tables.chrom_sizes)
    return ad
