"""snapatac2_amd: MI355X-native drop-ins for SnapATAC2's hot paths.

    snapatac2_amd.pp.import_fragments  fragment files -> the per-cell fragment store
    snapatac2_amd.pp.add_tile_matrix   fragments -> cell x bin CSR
    snapatac2_amd.pp.make_peak_matrix  fragments -> cell x peak CSR
    snapatac2_amd.pp.make_gene_matrix  fragments -> cell x gene CSR
    snapatac2_amd.pp.select_features   feature mask for spectral
    snapatac2_amd.tl.spectral          normalized-cosine Laplacian eigenmaps
                                       (full operator, or Nystrom with sample_size)
    snapatac2_amd.tl.tune_feature_weights  eigengap ascent through the operator's adjoint
    snapatac2_amd.pp.knn               exact k-nearest-neighbour graph

All compute runs in hand-written gfx950 HIP kernels (libsnapgpu.so).
"""
from .anndata import AnnData
from . import genome, pp, tl

__all__ = ["AnnData", "genome", "pp", "tl"]
