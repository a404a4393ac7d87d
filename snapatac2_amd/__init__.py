"""snapatac2_amd: MI355X-native drop-ins for three SnapATAC2 hot paths.

    snapatac2_amd.pp.add_tile_matrix   fragments -> cell x bin CSR
    snapatac2_amd.pp.select_features   feature mask for spectral
    snapatac2_amd.tl.spectral          normalized-cosine Laplacian eigenmaps
    snapatac2_amd.pp.knn               exact k-nearest-neighbour graph

All compute runs in hand-written gfx950 HIP kernels (libsnapgpu.so).
"""
from .anndata import AnnData
from . import genome, pp, tl

__all__ = ["AnnData", "genome", "pp", "tl"]
