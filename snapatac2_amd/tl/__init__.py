"""Tool drop-ins (snapatac2.tl)."""
from ._embedding import spectral

__all__ = ["spectral"]
