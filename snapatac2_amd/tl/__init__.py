"""Tool drop-ins (snapatac2.tl) and the differentiable spectral operator."""
from ._adjoint import normalized_cosine, spectral_gap, spectral_gap_grad, tune_feature_weights
from ._embedding import spectral

__all__ = ["spectral", "normalized_cosine", "spectral_gap", "spectral_gap_grad", "tune_feature_weights"]
