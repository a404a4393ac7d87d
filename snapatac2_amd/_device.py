"""Device-resident containers for the fragment store and CSR matrices.

Both keep their arrays as torch tensors on the GPU (torch is used for device
memory only; all compute goes through libsnapgpu). Unsigned 32-bit arrays are
held in int32 tensors and reinterpreted by the kernels.
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


def to_device_u32(a: np.ndarray, device):
    torch = _torch()
    a = np.ascontiguousarray(a)
    if a.dtype != np.uint32:
        a = a.astype(np.uint32)
    return torch.from_numpy(a.view(np.int32)).to(device)


def from_device_u32(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


class DeviceFragments:
    """obsm['fragment_paired'] (values = lengths) or obsm['fragment_single']
    stored on the device: row pointers int64 [n+1], positions uint32 [F],
    lengths uint32 [F] (paired) or signed int32 [F] (single; not read by the
    tile kernels). Layout follows import.rs:117-210."""

    def __init__(self, cell_ptr, pos, values, n_cols: int, paired: bool = True):
        self.cell_ptr, self.pos, self.values = cell_ptr, pos, values
        self.n_cols = int(n_cols)
        self.paired = paired

    @property
    def shape(self):
        return (self.cell_ptr.numel() - 1, self.n_cols)

    @property
    def n_fragments(self) -> int:
        return self.pos.numel()

    @property
    def device(self):
        return self.pos.device

    @classmethod
    def from_scipy(cls, m, paired: bool = True, device="cuda"):
        torch = _torch()
        indptr = torch.from_numpy(np.asarray(m.indptr, dtype=np.int64)).to(device)
        pos = to_device_u32(np.asarray(m.indices).astype(np.uint32, copy=False), device)
        if paired:
            vals = to_device_u32(np.asarray(m.data).astype(np.uint32, copy=False), device)
        else:
            vals = torch.from_numpy(np.asarray(m.data, dtype=np.int32)).to(device)
        return cls(indptr, pos, vals, m.shape[1], paired)

    def to_scipy(self):
        import scipy.sparse as sp
        data = from_device_u32(self.values) if self.paired else self.values.cpu().numpy()
        return sp.csr_matrix((data, from_device_u32(self.pos), self.cell_ptr.cpu().numpy()),
                             shape=self.shape)


class DeviceCSR:
    """A CSR matrix on the device: indptr int64 [n+1], indices int32/int64,
    data (uint32 counts in an int32 tensor, or float32)."""

    def __init__(self, indptr, indices, data, shape, data_dtype=np.uint32):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.shape = (int(shape[0]), int(shape[1]))
        self.data_dtype = np.dtype(data_dtype)

    @property
    def nnz(self) -> int:
        return self.indices.numel()

    @property
    def dtype(self):
        return self.data_dtype

    @property
    def device(self):
        return self.indptr.device

    def to_scipy(self):
        import scipy.sparse as sp
        d = self.data.cpu().numpy()
        if self.data_dtype == np.uint32:
            d = d.view(np.uint32)
        return sp.csr_matrix((d, self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                             shape=self.shape)

    def __repr__(self) -> str:
        return f"DeviceCSR(shape={self.shape}, nnz={self.nnz}, dtype={self.data_dtype})"
