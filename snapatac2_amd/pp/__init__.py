"""Preprocessing drop-ins (snapatac2.pp)."""
from ._basic import add_tile_matrix, select_features, make_tile_matrix
from ._knn import knn

__all__ = ["add_tile_matrix", "select_features", "make_tile_matrix", "knn"]
