"""Device state of the normalized-cosine operator used by spectral().

A = D^-1/2 (X~ X~^T - I) D^-1/2 with X~ = rownorm(X diag(w)) and
D = diag(X~ X~^T 1 - 1) (snapatac2-python/src/embedding.rs:131-171, 269-326).
The matrix stays as counts in CSR + CSC; see csrc/spectral.hip.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .._device import DeviceCSR
from ._lanczos import _Linalg

TRANSPOSE_CHUNK = 1 << 28  # entries sorted per transpose chunk
BLOCK_TARGET = 512  # blocks per direction of the block-tiled layout (about two per CU)
DEFAULT_TILED = "mixed"  # layout of large operators (CSC panels + CSR windows; panels where windows cannot hold the counts)
# the CSR windows pay off when they fill the chip: with fewer workgroups (about two per CU
# of an MI355X) the default takes the panel layout instead
MIXED_MIN_BLOCKS = 512


def _to_device_csr(X, torch) -> DeviceCSR:
    """A scipy/numpy matrix (or DeviceCSR) as a DeviceCSR with int32 columns."""
    if isinstance(X, DeviceCSR):
        return X
    import scipy.sparse as sp
    Xc = sp.csr_matrix(X)
    Xc.sort_indices()
    data = Xc.data
    if data.dtype.kind in "ui" and (data.size == 0 or data.min() >= 0) and \
            (data.size == 0 or data.max() < 2**32):
        vals = data.astype(np.uint32).view(np.int32)
        dt = np.uint32
    else:
        vals = data.astype(np.float32).view(np.int32)
        dt = np.float32
    dev = torch.device("cuda")
    return DeviceCSR(torch.from_numpy(Xc.indptr.astype(np.int64)).to(dev),
                     torch.from_numpy(Xc.indices.astype(np.int32)).to(dev),
                     torch.from_numpy(vals).to(dev), Xc.shape, dt)


def select_columns(X: DeviceCSR, cols: np.ndarray | None) -> DeviceCSR:
    """Columns `cols` (ascending original ids) of X, on the device."""
    torch = _lib.require_gpu()
    if cols is None:
        if X.indices.dtype == torch.int32:
            return X
        cols = np.arange(X.shape[1])
    L = _lib.lib()
    n, m_all = X.shape
    cols = np.asarray(cols, dtype=np.int64)
    if cols.size and (cols[0] < 0 or cols[-1] >= m_all or np.any(np.diff(cols) <= 0)):
        raise ValueError("cols must be distinct ascending column ids of X")
    d_cols = torch.from_numpy(cols).to(X.device)
    sel = torch.empty(2 * (-(-m_all // 32)), dtype=torch.int32, device=X.device)
    st = _lib.stream_ptr()
    _lib.check(L.snapgpu_col_select_table(d_cols.data_ptr(), cols.size, m_all, sel.data_ptr(), st))
    bits = 32 if X.indices.dtype == torch.int32 else 64
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=X.device)
    _lib.check(L.snapgpu_csr_select_count(X.indptr.data_ptr(), X.indices.data_ptr(), bits, n,
                                          sel.data_ptr(), m_all, indptr.data_ptr() + 8, st))
    ws_b = int(L.snapgpu_scan_workspace_bytes(n))
    ws = torch.empty(max(ws_b, 8), dtype=torch.uint8, device=X.device)
    _lib.check(L.snapgpu_scan_inclusive_i64(indptr.data_ptr() + 8, n, ws.data_ptr(), ws_b, st))
    nnz = int(indptr[-1].item())
    idx = torch.empty(nnz, dtype=torch.int32, device=X.device)
    val = torch.empty(nnz, dtype=torch.int32, device=X.device)
    _lib.check(L.snapgpu_csr_select_fill(X.indptr.data_ptr(), X.indices.data_ptr(), bits,
                                         X.data.data_ptr(), n, sel.data_ptr(), m_all, indptr.data_ptr(),
                                         idx.data_ptr(), val.data_ptr(), st))
    return DeviceCSR(indptr, idx, val, (n, len(cols)), X.data_dtype)


class SpectralOperator:
    """Holds CSR/CSC counts, weights, scales and degrees of A on the device."""

    # densest panel entry packing used (2: counts < 128, 1: counts < 65536,
    # 0: separate values); lower it to run the generic panel product
    panel_pack_max = 2

    def __init__(self, X: DeviceCSR, feature_weights=None, use_panels: bool | None = None,
                 layout: str | None = None):
        """layout: 'mixed' (the CSC product over L2-sized panels, the CSR
        product over 1024-row windows, b <= 16 per launch), 'panel' (both over
        panels, b <= 32), 'window' (both over windows, b <= 16), 'block'
        (LDS-tiled products, b <= 16) or 'csr' (row-parallel products); by
        default DEFAULT_TILED when either dense operand spans more than one
        32,768-row panel, else 'csr'. use_panels=True / False stands for the
        default tiled layout / 'csr'. The window layouts need integer counts
        below 256; other matrices get 'panel' instead."""
        torch = _lib.require_gpu()
        L = _lib.lib()
        self.torch = torch
        self.X = X
        n, m = X.shape
        self.n, self.m = n, m
        dev = X.device
        self.value_type = 1 if np.dtype(X.data_dtype) == np.float32 else 0
        st = _lib.stream_ptr()
        # CSC copy (deterministic stable transpose)
        h_indptr = X.indptr.cpu().numpy()
        chunk = int(min(max(X.nnz, 1), TRANSPOSE_CHUNK))
        ws_b = int(L.snapgpu_csr_transpose_workspace_bytes(m, chunk))
        ws = torch.empty(ws_b, dtype=torch.uint8, device=dev)
        self.colptr = torch.empty(m + 1, dtype=torch.int64, device=dev)
        self.rowidx = torch.empty(X.nnz, dtype=torch.int32, device=dev)
        self.cval = torch.empty(X.nnz, dtype=torch.int32, device=dev)
        _lib.check(L.snapgpu_csr_transpose(X.indptr.data_ptr(), X.indices.data_ptr(),
                                           X.data.data_ptr(), n, m, X.nnz, h_indptr.ctypes.data,
                                           chunk, self.colptr.data_ptr(), self.rowidx.data_ptr(),
                                           self.cval.data_ptr(), ws.data_ptr(), ws_b, st))
        del ws
        f64 = dict(dtype=torch.float64, device=dev)
        if feature_weights is None:
            self.w = torch.empty(m, **f64)
            scratch = torch.empty(16, dtype=torch.uint8, device=dev)
            _lib.check(L.snapgpu_spectral_idf(self.colptr.data_ptr(), m, n, self.w.data_ptr(),
                                              scratch.data_ptr(), st))
        else:
            self.w = torch.empty(m, **f64)
            self._load_weights(feature_weights)
        self.s = torch.empty(n, **f64)
        self.c = torch.empty(m, **f64)
        self.dinv = torch.empty(n, **f64)
        self.r64 = torch.empty(n, **f64)
        self.r32 = torch.empty(n, dtype=torch.float32, device=dev)
        self.w2 = torch.empty(m, dtype=torch.float32, device=dev)
        self._status = torch.zeros(8, dtype=torch.int32, device=dev)
        self._prep_ws_b = int(L.snapgpu_spectral_prepare_workspace_bytes(m))
        self._prep_ws = torch.empty(self._prep_ws_b, dtype=torch.uint8, device=dev)
        self._prepare()
        # panel layouts (dense operands larger than one L2-sized panel)
        pc = int(L.snapgpu_panel_count(m))
        pr = int(L.snapgpu_panel_count(n))
        chosen = layout is None
        if layout is None:
            tiled = (pc > 1 or pr > 1) if use_panels is None else bool(use_panels)
            layout = DEFAULT_TILED if tiled else "csr"
        if layout not in ("mixed", "window", "block", "panel", "csr"):
            raise ValueError("layout must be 'mixed', 'window', 'block', 'panel' or 'csr'")
        if layout in ("mixed", "window") and self.value_type != 0:
            layout = "panel"  # the window entries hold integer counts
        self.layout = layout
        self.use_panels = layout != "csr"
        self._panel_keep = []
        if layout in ("mixed", "window"):
            fits = True
            try:
                self.window_csr = self._build_window(X.indptr, X.indices, X.data, n, m, X.nnz)
                if layout == "window":
                    self.window_csc = self._build_window(self.colptr, self.rowidx, self.cval, m, n, X.nnz)
            except _lib.CapacityError:
                fits = False  # counts >= 256 or too many windows
            if fits and layout == "mixed" and chosen and self.window_csr.n_blocks < MIXED_MIN_BLOCKS:
                fits = False  # the default: too few window workgroups to fill the device
            if not fits:  # the panel layout takes them
                self._panel_keep = []
                self.window_csr = self.window_csc = None
                layout = self.layout = "panel"
            if layout == "window":
                self._pws_b = int(L.snapgpu_spectral_apply_window_workspace_bytes(n, m))
                self._pws = torch.empty(self._pws_b, dtype=torch.uint8, device=dev)
            elif layout == "mixed":
                self.panel_csc = self._build_panel(self.colptr, self.rowidx, self.cval, m, n, X.nnz)
                self._pws_b = int(L.snapgpu_spectral_apply_mixed_workspace_bytes(n, m, self.panel_csc.n_pairs))
                self._pws = torch.empty(self._pws_b, dtype=torch.uint8, device=dev)
        if layout == "panel":
            self.panel_csr = self._build_panel(X.indptr, X.indices, X.data, n, m, X.nnz)
            self.panel_csc = self._build_panel(self.colptr, self.rowidx, self.cval, m, n, X.nnz)
            self._pws_b = int(L.snapgpu_spectral_apply_panel_workspace_bytes(
                n, m, self.panel_csr.n_pairs, self.panel_csc.n_pairs))
            self._pws = torch.empty(self._pws_b, dtype=torch.uint8, device=dev)
        elif layout == "block":
            self.block_csr = self._build_block(X.indptr, X.indices, X.data, n, m, X.nnz)
            self.block_csc = self._build_block(self.colptr, self.rowidx, self.cval, m, n, X.nnz)
            self._pws_b = int(L.snapgpu_spectral_apply_block_workspace_bytes(n, m))
            self._pws = torch.empty(self._pws_b, dtype=torch.uint8, device=dev)
        self._ws_b = int(L.snapgpu_spectral_apply_workspace_bytes(n, m))
        self._ws = torch.empty(self._ws_b, dtype=torch.uint8, device=dev)
        self.n_applies = 0
        self.n_columns = 0
        self.events = None  # optional list collecting (start, end) events per application
        self._vjp_ws = None

    def _build_panel(self, ptr, idx, val, nrows: int, dense_rows: int, nnz: int):
        """Panel-major copy of one storage direction (see snapgpu_panel_build)."""
        torch = self.torch
        L = _lib.lib()
        dev = ptr.device
        st = _lib.stream_ptr()
        ws_b = int(L.snapgpu_panel_build_workspace_bytes(nrows, dense_rows, nnz))
        ws = torch.empty(ws_b, dtype=torch.uint8, device=dev)
        K, NC, packed = C.c_int64(), C.c_int64(), C.c_int32()
        _lib.check(L.snapgpu_panel_sizes(ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows,
                                         nnz, self.value_type, ws.data_ptr(), ws_b, C.byref(K),
                                         C.byref(NC), C.byref(packed), st))
        packed = C.c_int32(min(packed.value, self.panel_pack_max))
        i32 = dict(dtype=torch.int32, device=dev)
        ent = torch.empty(max(nnz, 1), **i32)
        val_out = None if packed.value else torch.empty(max(nnz, 1), **i32)
        cstart = torch.empty(NC.value + 1, dtype=torch.int64, device=dev)
        cpair = torch.empty(NC.value + 1, **i32)
        cpanel = torch.empty(max(NC.value, 1), **i32)
        rp_ptr = torch.empty(nrows + 1, dtype=torch.int64, device=dev)
        rp_ids = torch.empty(max(K.value, 1), **i32)
        _lib.check(L.snapgpu_panel_build(
            ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows, nnz, packed.value,
            ws.data_ptr(), ws_b, ent.data_ptr(), val_out.data_ptr() if val_out is not None else None,
            cstart.data_ptr(), cpair.data_ptr(), cpanel.data_ptr(), NC.value, rp_ptr.data_ptr(),
            rp_ids.data_ptr(), st))
        self._panel_keep.append((ent, val_out, cstart, cpair, cpanel, rp_ptr, rp_ids))
        return _lib.PanelLayout(ent.data_ptr(), val_out.data_ptr() if val_out is not None else None,
                                cstart.data_ptr(), cpair.data_ptr(), cpanel.data_ptr(), NC.value,
                                rp_ptr.data_ptr(), rp_ids.data_ptr(), K.value, nrows, dense_rows,
                                packed.value, self.value_type)

    def _build_window(self, ptr, idx, val, nrows: int, dense_rows: int, nnz: int):
        """Window layout of one storage direction (see snapgpu_window_plan/build)."""
        torch = self.torch
        L = _lib.lib()
        dev = ptr.device
        st = _lib.stream_ptr()
        ws_b = int(L.snapgpu_window_plan_workspace_bytes(nrows))
        ws = torch.empty(ws_b, dtype=torch.uint8, device=dev)
        max_blocks = nrows + 1
        h_block_row = np.zeros(max_blocks + 1, dtype=np.int64)
        h_row_slot = np.zeros(nrows, dtype=np.uint32)
        nb = C.c_int64()
        _lib.check(L.snapgpu_window_plan(ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows, nnz,
                                         ws.data_ptr(), ws_b, h_block_row.ctypes.data, max_blocks, C.byref(nb),
                                         h_row_slot.ctypes.data, st))
        del ws
        nb = nb.value
        block_row = torch.from_numpy(h_block_row[:nb + 1]).to(dev)
        row_slot = torch.from_numpy(h_row_slot.view(np.int32)).to(dev)
        nw = int(L.snapgpu_window_count(dense_rows))
        bws_b = int(L.snapgpu_window_build_workspace_bytes(nb, dense_rows))
        bws = torch.empty(bws_b, dtype=torch.uint8, device=dev)
        ent = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        bstart = torch.empty(nb * nw + 1, dtype=torch.int64, device=dev)
        qoff = torch.empty(nb * nw * 257, dtype=torch.int32, device=dev)
        _lib.check(L.snapgpu_window_build(ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows, nnz,
                                          block_row.data_ptr(), nb, row_slot.data_ptr(), bws.data_ptr(), bws_b,
                                          ent.data_ptr(), bstart.data_ptr(), qoff.data_ptr(), st))
        self._panel_keep.append((ent, bstart, qoff, block_row, row_slot))
        return _lib.WindowLayout(ent.data_ptr(), bstart.data_ptr(), qoff.data_ptr(), block_row.data_ptr(),
                                 row_slot.data_ptr(), nb, nw, nrows, dense_rows, nnz)

    def _build_block(self, ptr, idx, val, nrows: int, dense_rows: int, nnz: int,
                     target_blocks: int = BLOCK_TARGET):
        """Block-tiled copy of one storage direction (see snapgpu_block_plan/build)."""
        torch = self.torch
        L = _lib.lib()
        dev = ptr.device
        st = _lib.stream_ptr()
        pws_b = int(L.snapgpu_block_plan_workspace_bytes(nrows))
        pws = torch.empty(pws_b, dtype=torch.uint8, device=dev)
        nb, ts, packed = C.c_int64(), C.c_int64(), C.c_int32()
        _lib.check(L.snapgpu_block_plan(ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows, nnz,
                                        self.value_type, target_blocks, pws.data_ptr(), pws_b, C.byref(nb),
                                        C.byref(ts), C.byref(packed), st))
        P = -(-dense_rows // int(L.snapgpu_block_panel_rows()))
        i32 = dict(dtype=torch.int32, device=dev)
        ent = torch.empty(max(nnz, 1), **i32)
        val_out = None if packed.value else torch.empty(max(nnz, 1), **i32)
        tile_start = torch.empty(nb.value * P + 1, dtype=torch.int64, device=dev)
        tile_off = torch.empty(max(nb.value * P, 1) * 16, dtype=torch.int16, device=dev)
        row0 = torch.empty(nb.value + 1, **i32)
        out_row = torch.empty(nrows, **i32)
        vslot = torch.empty(nrows, **i32)
        nslots = torch.empty(max(nb.value, 1), **i32)
        bws_b = int(L.snapgpu_block_build_workspace_bytes(ts.value))
        bws = torch.empty(bws_b, dtype=torch.uint8, device=dev)
        _lib.check(L.snapgpu_block_build(
            ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), nrows, dense_rows, nnz, packed.value, pws.data_ptr(),
            pws_b, nb.value, ts.value, bws.data_ptr(), bws_b, ent.data_ptr(), _lib.ptr(val_out),
            tile_start.data_ptr(), tile_off.data_ptr(), row0.data_ptr(), out_row.data_ptr(), vslot.data_ptr(),
            nslots.data_ptr(), st))
        del pws, bws
        self._panel_keep.append((ent, val_out, tile_start, row0, out_row, vslot, nslots, tile_off))
        return _lib.BlockLayout(ent.data_ptr(), _lib.ptr(val_out), tile_start.data_ptr(), tile_off.data_ptr(),
                                row0.data_ptr(),
                                out_row.data_ptr(), vslot.data_ptr(), nslots.data_ptr(), nb.value, P, nrows,
                                dense_rows, nnz, packed.value, self.value_type)

    def _load_weights(self, w) -> None:
        torch = self.torch
        if isinstance(w, torch.Tensor):
            w = w.detach().to(device=self.w.device, dtype=torch.float64).reshape(-1)
        else:
            w = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64).reshape(-1)).to(self.w.device)
        if w.numel() != self.m:
            raise ValueError(f"feature_weights has {w.numel()} entries, expected {self.m}")
        self.w.copy_(w)

    def _prepare(self) -> None:
        """Row scales, column sums, degrees and r, w^2 for the current weights."""
        L = _lib.lib()
        X = self.X
        _lib.check(L.snapgpu_spectral_prepare(
            X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), self.n,
            self.colptr.data_ptr(), self.rowidx.data_ptr(), self.cval.data_ptr(), self.m,
            self.w.data_ptr(), self.s.data_ptr(), self.c.data_ptr(), self.dinv.data_ptr(),
            self.r64.data_ptr(), self.r32.data_ptr(), self.w2.data_ptr(), self.value_type,
            self._status.data_ptr(), self._prep_ws.data_ptr(), self._prep_ws_b, _lib.stream_ptr()))

    def set_weights(self, w) -> None:
        """Replace the feature weights and recompute scales and degrees on the device."""
        self._load_weights(w)
        self._prepare()

    def vjp_w(self, V, G):
        """dw = d<G, A(w) V>/dw (FP64 device vector of length m) for FP64 n x b blocks."""
        torch = self.torch
        L = _lib.lib()
        X = self.X
        if V.shape != G.shape or V.shape[0] != self.n:
            raise ValueError("V and G must both be n x b")
        if V.stride(1) != 1 or G.stride(1) != 1:
            raise ValueError("V and G must be row-major")
        if self._vjp_ws is None:
            self._vjp_ws_b = int(L.snapgpu_spectral_vjp_workspace_bytes(self.n, self.m))
            self._vjp_ws = torch.empty(self._vjp_ws_b, dtype=torch.uint8, device=self.w.device)
        dw = torch.zeros(self.m, dtype=torch.float64, device=self.w.device)
        mixed = self.layout == "mixed"
        step = 8 if mixed else 32  # the mixed products take V and G side by side in 16 columns
        part = torch.empty_like(dw) if V.shape[1] > step else dw
        for c0 in range(0, V.shape[1], step):
            b = min(step, V.shape[1] - c0)
            Vc, Gc = V[:, c0:c0 + b], G[:, c0:c0 + b]
            if mixed:
                _lib.check(L.snapgpu_spectral_vjp_w_mixed(
                    C.byref(self.window_csr), C.byref(self.panel_csc),
                    X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), self.n,
                    self.colptr.data_ptr(), self.rowidx.data_ptr(), self.cval.data_ptr(), self.m,
                    self.w.data_ptr(), self.s.data_ptr(), self.dinv.data_ptr(), self.r32.data_ptr(),
                    self.w2.data_ptr(), Vc.data_ptr(), Vc.stride(0), Gc.data_ptr(), Gc.stride(0), b,
                    part.data_ptr(), self._vjp_ws.data_ptr(), self._vjp_ws_b, self._pws.data_ptr(),
                    self._pws_b, _lib.stream_ptr()))
                if part is not dw:
                    dw += part
                continue
            _lib.check(L.snapgpu_spectral_vjp_w(
                X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), self.n,
                self.colptr.data_ptr(), self.rowidx.data_ptr(), self.cval.data_ptr(), self.m,
                self.w.data_ptr(), self.s.data_ptr(), self.dinv.data_ptr(), self.r32.data_ptr(),
                self.w2.data_ptr(), Vc.data_ptr(), Vc.stride(0), Gc.data_ptr(), Gc.stride(0), b,
                part.data_ptr(), self.value_type, self._vjp_ws.data_ptr(), self._vjp_ws_b,
                _lib.stream_ptr()))
            if part is not dw:
                dw += part
        return dw

    def apply(self, V, Y, b: int | None = None) -> None:
        """Y[:, :b] = A V[:, :b]; V, Y row-major FP64 device tensors (any row stride)."""
        L = _lib.lib()
        X = self.X
        b = V.shape[1] if b is None else b
        wmax = 16 if self.layout in ("block", "window", "mixed") else 32
        if b > wmax:
            for c0 in range(0, b, wmax):
                self.apply(V[:, c0:c0 + wmax], Y[:, c0:c0 + wmax])
            return
        ev = None
        if self.events is not None:
            ev = (self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True))
            ev[0].record()
        if self.layout == "mixed":
            _lib.check(L.snapgpu_spectral_apply_mixed(
                C.byref(self.window_csr), C.byref(self.panel_csc), self.r32.data_ptr(),
                self.w2.data_ptr(), self.dinv.data_ptr(), V.data_ptr(), V.stride(0), b, Y.data_ptr(),
                Y.stride(0), self._pws.data_ptr(), self._pws_b, _lib.stream_ptr()))
        elif self.layout == "window":
            _lib.check(L.snapgpu_spectral_apply_window(
                C.byref(self.window_csr), C.byref(self.window_csc), self.r32.data_ptr(),
                self.w2.data_ptr(), self.dinv.data_ptr(), V.data_ptr(), V.stride(0), b, Y.data_ptr(),
                Y.stride(0), self._pws.data_ptr(), self._pws_b, _lib.stream_ptr()))
        elif self.layout == "block":
            _lib.check(L.snapgpu_spectral_apply_block(
                C.byref(self.block_csr), C.byref(self.block_csc), self.r32.data_ptr(),
                self.w2.data_ptr(), self.dinv.data_ptr(), V.data_ptr(), V.stride(0), b, Y.data_ptr(),
                Y.stride(0), self._pws.data_ptr(), self._pws_b, _lib.stream_ptr()))
        elif self.layout == "panel":
            _lib.check(L.snapgpu_spectral_apply_panel(
                C.byref(self.panel_csr), C.byref(self.panel_csc), self.r32.data_ptr(),
                self.w2.data_ptr(), self.dinv.data_ptr(), V.data_ptr(), V.stride(0), b, Y.data_ptr(),
                Y.stride(0), self._pws.data_ptr(), self._pws_b, _lib.stream_ptr()))
        else:
            _lib.check(L.snapgpu_spectral_apply(
                X.indptr.data_ptr(), X.indices.data_ptr(), X.data.data_ptr(), self.n,
                self.colptr.data_ptr(), self.rowidx.data_ptr(), self.cval.data_ptr(), self.m,
                self.r32.data_ptr(), self.w2.data_ptr(), self.dinv.data_ptr(), V.data_ptr(),
                V.stride(0), b, Y.data_ptr(), Y.stride(0), self.value_type, self._ws.data_ptr(),
                self._ws_b, _lib.stream_ptr()))
        if ev is not None:
            ev[1].record()
            self.events.append(ev)
        self.n_applies += 1
        self.n_columns += b

    def matmat(self, V):
        Y = self.torch.empty_like(V)
        self.apply(V, Y)
        return Y

    @property
    def degree(self):
        """d = X~ X~^T 1 - 1 (what spectral_mf returns for Nystrom, embedding.rs:189-190)."""
        return 1.0 / self.dinv

    def apply_f64(self, U):
        """A U for U (n x b FP64) with both sparse products in FP64
        (snapgpu_csr_spmm_scaled_f64 on the CSC, then on the CSR), for checks of
        eigenvector estimates that the FP32 products' rounding (about 1e-7 of
        |lambda_1|) would blur. None for float-valued matrices (the FP64
        products read integer counts)."""
        if self.value_type != 0:
            return None
        torch = self.torch
        L = _lib.lib()
        st = _lib.stream_ptr()
        U = U.to(torch.float64).contiguous()
        b = U.shape[1]
        w2 = (self.w * self.w).contiguous()
        # r o U up front: the CSC product then gathers one row per entry, not also r
        rU = (self.r64[:, None] * U).contiguous()
        Z = torch.empty((self.m, b), dtype=torch.float64, device=U.device)
        _lib.check(L.snapgpu_csr_spmm_scaled_f64(self.colptr.data_ptr(), self.rowidx.data_ptr(),
                                                 self.cval.data_ptr(), self.m, w2.data_ptr(),
                                                 0, rU.data_ptr(), b, b, Z.data_ptr(), b, st))
        Y = torch.empty_like(U)
        _lib.check(L.snapgpu_csr_spmm_scaled_f64(self.X.indptr.data_ptr(), self.X.indices.data_ptr(),
                                                 self.X.data.data_ptr(), self.n, self.r64.data_ptr(), 0,
                                                 Z.data_ptr(), b, b, Y.data_ptr(), b, st))
        return Y - self.dinv[:, None] * U

    def gram_f64(self, U):
        """U^T A U (b x b, FP64) with its sparse product in FP64. Since
        A = R X W^2 X^T R - D^-1, U^T A U = P^T P - U^T D^-1 U with
        P = W X^T (R U): one sparse product (the CSC: over its panel layout,
        snapgpu_panel_spmm_f64, when the operator has one, else
        snapgpu_csr_spmm_scaled_f64, with row scales w) and two tall-skinny FP64
        products, where apply_f64 takes two sparse products. b <= 32. None for
        float-valued matrices."""
        if self.value_type != 0:
            return None
        torch = self.torch
        L = _lib.lib()
        U = U.to(torch.float64).contiguous()
        b = U.shape[1]
        rU = (self.r64[:, None] * U).contiguous()
        P = torch.empty((self.m, b), dtype=torch.float64, device=U.device)
        if self.layout in ("panel", "mixed"):
            # two columns at a time over the CSC panel layout (L2-sized slices of rU)
            ws_b = int(L.snapgpu_panel_spmm_f64_workspace_bytes(self.panel_csc.n_pairs))
            if self._pws_b >= ws_b:
                ws, ws_b = self._pws, self._pws_b
            else:
                ws = torch.empty(ws_b, dtype=torch.uint8, device=U.device)
            if b % 2:
                rU = torch.nn.functional.pad(rU, (0, 1))
            P2 = torch.empty((self.m, 2), dtype=torch.float64, device=U.device)
            for c0 in range(0, b, 2):
                D = rU[:, c0:c0 + 2].contiguous()
                _lib.check(L.snapgpu_panel_spmm_f64(C.byref(self.panel_csc), D.data_ptr(), 0, self.w.data_ptr(),
                                                    P2.data_ptr(), ws.data_ptr(), ws_b, _lib.stream_ptr()))
                P[:, c0:c0 + 2] = P2[:, :min(2, b - c0)]
        else:
            _lib.check(L.snapgpu_csr_spmm_scaled_f64(self.colptr.data_ptr(), self.rowidx.data_ptr(),
                                                     self.cval.data_ptr(), self.m, self.w.data_ptr(), 0,
                                                     rU.data_ptr(), b, b, P.data_ptr(), b, _lib.stream_ptr()))
        PP = _Linalg(torch, self.m, b, U.device).tn(P, P).clone()
        dU = (self.dinv[:, None] * U).contiguous()
        return PP - _Linalg(torch, self.n, b, U.device).tn(U, dU)

    def rayleigh_f64(self, U):
        """Rayleigh quotients u^T A u of the columns of U with the FP64 products
        of apply_f64 (None for float-valued matrices)."""
        AU = self.apply_f64(U)
        return None if AU is None else (U.to(AU.dtype) * AU).sum(0)
