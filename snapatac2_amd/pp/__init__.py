"""Preprocessing drop-ins (snapatac2.pp)."""
from ._basic import add_tile_matrix, make_tile_matrix, make_tile_matrix_chunked, select_features
from ._import import export_fragment_text, import_fragment_text, import_fragments
from ._knn import knn
from ._genes import make_gene_matrix
from ._peaks import make_peak_matrix

__all__ = ["add_tile_matrix", "select_features", "make_tile_matrix", "make_tile_matrix_chunked",
           "knn", "import_fragments", "import_fragment_text", "export_fragment_text", "make_peak_matrix",
           "make_gene_matrix"]
