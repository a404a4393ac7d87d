"""ctypes binding of libsnapgpu.so (the C ABI declared in include/snapgpu.h).

The library is the only compute path: if it is missing, or no ROCm GPU is
present, every operation raises instead of falling back to CPU code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libsnapgpu.so"

OK, ERR_ARG, ERR_HIP, ERR_DATA, ERR_CAPACITY, ERR_UNSORTED = 0, 1, 2, 3, 4, 5

_p = C.c_void_p
_i32, _i64, _u32, _u64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "snapgpu_last_error": (C.c_char_p, []),
    "snapgpu_abi_version": (C.c_int, []),
    "snapgpu_scan_workspace_bytes": (_i64, [_i64]),
    "snapgpu_scan_inclusive_i64": (C.c_int, [_p, _i64, _p, _i64, _p]),
    "snapgpu_tile_table_bytes": (_i64, [_i32]),
    "snapgpu_tile_table_pack": (C.c_int, [_p, _p, _i32, _u32, _p, C.POINTER(_u64)]),
    "snapgpu_tile_order_workspace_bytes": (_i64, [_i64]),
    "snapgpu_tile_order": (C.c_int, [_p, _i64, _p, _p, _i64, _p]),
    "snapgpu_frag_sort_rows_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_frag_sort_rows": (C.c_int, [_p, _i64, _i64, _p, _p, _p, _i64, _p]),
    "snapgpu_tile_count": (C.c_int, [_p, _p, _p, _i64, _p, _i32, _u32, _i32, _u32, _u32, _p, _p, _p,
                                     C.POINTER(_i32), _p]),
    "snapgpu_tile_fill": (C.c_int, [_p, _p, _p, _i64, _p, _i32, _u32, _i32, _u32, _u32, _i32, _p,
                                    _p, _i32, _p, _p, _p, _p]),
    "snapgpu_csr_col_sum_u32": (C.c_int, [_p, _i32, _p, _i64, _p, _p]),
    "snapgpu_csr_col_sum_rows_u32": (C.c_int, [_p, _i64, _p, _i32, _p, _i64, _p, _p, _p]),
    "snapgpu_synth_counts": (C.c_int, [_u64, _i64, _i64, C.c_double, C.c_double, _i64, _i64, _p,
                                       _p]),
    "snapgpu_synth_cell_types": (C.c_int, [_u64, _i64, _i64, _i32, _p, _p]),
    "snapgpu_synth_workspace_bytes": (_i64, [_i64]),
    "snapgpu_synth_fragments": (C.c_int, [_u64, _i64, _i64, _p, _i64, _p, _i32, _p, _i32, _p,
                                          _i32, C.c_double, C.c_double, _p, _p, _p, _i64, _p]),
    "snapgpu_knn_workspace_bytes": (_i64, [_i64, _i32, _i32]),
    "snapgpu_knn": (C.c_int, [_p, _i64, _i32, _i64, _i32, _i32, _p, _p, _p, _i64, _i32,
                              C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), _p]),
    "snapgpu_col_select_table": (C.c_int, [_p, _i64, _i64, _p, _p]),
    "snapgpu_csr_select_count": (C.c_int, [_p, _p, _i32, _i64, _p, _i64, _p, _p]),
    "snapgpu_csr_select_fill": (C.c_int, [_p, _p, _i32, _p, _i64, _p, _i64, _p, _p, _p, _p]),
    "snapgpu_csr_transpose_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_csr_transpose": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _p, _i64, _p, _p, _p, _p,
                                        _i64, _p]),
    "snapgpu_spectral_idf": (C.c_int, [_p, _i64, _i64, _p, _p, _p]),
    "snapgpu_spectral_prepare_workspace_bytes": (_i64, [_i64]),
    "snapgpu_spectral_prepare": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p,
                                           _p, _p, _p, _i32, _p, _p, _i64, _p]),
    "snapgpu_spectral_apply_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_spectral_apply": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p,
                                         _i64, _i32, _p, _i64, _i32, _p, _i64, _p]),
    "snapgpu_spectral_vjp_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_spectral_vjp_w": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p,
                                         _p, _i64, _p, _i64, _i32, _p, _i32, _p, _i64, _p]),
    "snapgpu_panel_count": (_i32, [_i64]),
    "snapgpu_panel_build_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "snapgpu_panel_sizes": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _i32, _p, _i64,
                                      C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i32), _p]),
    "snapgpu_panel_build": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _i32, _p, _i64, _p, _p, _p, _p,
                                      _p, _i64, _p, _p, _p]),
    "snapgpu_spectral_apply_panel_workspace_bytes": (_i64, [_i64, _i64, _i64, _i64]),
    "snapgpu_spectral_apply_panel": (C.c_int, [_p, _p, _p, _p, _p, _p, _i64, _i32, _p, _i64, _p,
                                               _i64, _p]),
    "snapgpu_panel_spmm_f64_workspace_bytes": (_i64, [_i64]),
    "snapgpu_panel_spmm_f64": (C.c_int, [_p, _p, _i32, _p, _p, _p, _i64, _p]),
    "snapgpu_block_panel_rows": (_i32, []),
    "snapgpu_block_plan_workspace_bytes": (_i64, [_i64]),
    "snapgpu_block_plan": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _i32, _i64, _p, _i64, C.POINTER(_i64),
                                     C.POINTER(_i64), C.POINTER(_i32), _p]),
    "snapgpu_block_build_workspace_bytes": (_i64, [_i64]),
    "snapgpu_block_build": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _i32, _p, _i64, _i64, _i64, _p, _i64, _p, _p,
                                      _p, _p, _p, _p, _p, _p, _p]),
    "snapgpu_window_count": (_i64, [_i64]),
    "snapgpu_window_plan_workspace_bytes": (_i64, [_i64]),
    "snapgpu_window_plan": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _p, _i64, _p, _i64, C.POINTER(_i64), _p, _p]),
    "snapgpu_window_build_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_window_build": (C.c_int, [_p, _p, _p, _i64, _i64, _i64, _p, _i64, _p, _p, _i64, _p, _p, _p, _p]),
    "snapgpu_spectral_apply_window_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_spectral_apply_window": (C.c_int, [_p, _p, _p, _p, _p, _p, _i64, _i32, _p, _i64, _p, _i64, _p]),
    "snapgpu_spectral_apply_mixed_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "snapgpu_spectral_apply_mixed": (C.c_int, [_p, _p, _p, _p, _p, _p, _i64, _i32, _p, _i64, _p, _i64, _p]),
    "snapgpu_spectral_vjp_mixed_workspace_bytes": (_i64, [_i64]),
    "snapgpu_spectral_vjp_w_mixed": (C.c_int, [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p,
                                               _p, _i64, _p, _i64, _i32, _p, _p, _i64, _p, _i64, _p]),
    "snapgpu_spectral_apply_block_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_spectral_apply_block": (C.c_int, [_p, _p, _p, _p, _p, _p, _i64, _i32, _p, _i64, _p, _i64, _p]),
    "snapgpu_gemm_tn_workspace_bytes": (_i64, [_i64, _i32, _i32]),
    "snapgpu_gemm_tn_f64": (C.c_int, [_p, _i64, _p, _i64, _i64, _i32, _i32, _p, _p, _i64, _p]),
    "snapgpu_gemm_tn_f64_noalloc": (C.c_int, [_p, _i64, _p, _i64, _i64, _i32, _i32, _p, _p]),
    "snapgpu_gemm_nn_f64": (C.c_int, [_p, _i64, _p, _i64, _i32, _i32, C.c_double, C.c_double,
                                      _p, _i64, _p]),
    "snapgpu_random_f64": (C.c_int, [_u64, _i64, _p, _p]),
    "snapgpu_frag_table_bytes": (_i64, [C.POINTER(C.c_char_p), _i32]),
    "snapgpu_frag_table_pack": (C.c_int, [C.POINTER(C.c_char_p), _p, _p, _i32, _i32, _p, _i64]),
    "snapgpu_frag_count_workspace_bytes": (_i64, [_i64]),
    "snapgpu_frag_count_lines": (C.c_int, [_p, _i64, _p, _i64, C.POINTER(_i64), _p]),
    "snapgpu_frag_workspace_bytes": (_i64, [_i64, _i64]),
    "snapgpu_frag_index": (C.c_int, [_p, _i64, _i64, _p, _p, _p, _p, _i64, _p, _p]),
    "snapgpu_frag_fill": (C.c_int, [_p, _i64, _p, _i64, _p, _p, _p, _p, _p, _p, _i32, _p]),
    "snapgpu_csr_take_rows_workspace_bytes": (_i64, [_i64]),
    "snapgpu_csr_take_rows_count": (C.c_int, [_p, _i64, _p, _i64, _p, _p, _i64, C.POINTER(_i64), _p]),
    "snapgpu_csr_take_rows_fill": (C.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _p]),
    "snapgpu_csr_col_counts": (C.c_int, [_p, _i64, _p, _i64, _i64, _p, _p, _p]),
    "snapgpu_csr_row_inv_norm": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p]),
    "snapgpu_csr_spmm_scaled_f64": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _i32, _p, _i64, _p]),
    "snapgpu_nystrom_landmark_scale": (C.c_int, [_p, _i64, _i64, _i32, _p, _p, _p, _i64, _p]),
    "snapgpu_nystrom_normalize_workspace_bytes": (_i64, [_i64, _i32, _i64]),
    "snapgpu_nystrom_normalize": (C.c_int, [_p, _i64, _i64, _i32, _p, _i64, _p, _i64, _p]),
    "snapgpu_peak_count_workspace_bytes": (_i64, [_i64]),
    "snapgpu_peak_count": (C.c_int, [_p, _p, _p, _i64, _i32, _i32, _u32, _u32, _p, _p, _i64,
                                     C.POINTER(_i64), _p]),
    "snapgpu_peak_pairs_workspace_bytes": (_i64, [_i64]),
    "snapgpu_peak_pairs": (C.c_int, [_p, _p, _i64, _p, _p, _i64, _i32, _i32, _u32, _u32, _p, _i64, _p, _p,
                                     _i64, C.POINTER(_i64), _p]),
    "snapgpu_peak_csr": (C.c_int, [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p]),
    "snapgpu_group_max_workspace_bytes": (_i64, [_i64]),
    "snapgpu_group_max": (C.c_int, [_p, _i64, _p, _p, _i64, _p, _i32, _p, _i64, C.POINTER(_i64), _p]),
    "snapgpu_group_max_csr": (C.c_int, [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p]),
    "snapgpu_frag_format_workspace_bytes": (_i64, [_i64]),
    "snapgpu_frag_format": (C.c_int, [_p, _i64, _p, _p, _i64, _i32, _p, _p, _i32, _p, _p, _p, _i64,
                                      _p, _i64, C.POINTER(_i64), _p]),
}


class PanelLayout(C.Structure):
    """snapgpu_panel_layout (include/snapgpu.h): device pointers and sizes."""
    _fields_ = [("ent", _p), ("val", _p), ("chunk_start", _p), ("chunk_pair", _p),
                ("chunk_panel", _p), ("n_chunks", _i64), ("row_pair_ptr", _p),
                ("row_pair_ids", _p), ("n_pairs", _i64), ("n_rows", _i64), ("dense_rows", _i64),
                ("packed", _i32), ("value_type", _i32)]


class BlockLayout(C.Structure):
    """snapgpu_block_layout (include/snapgpu.h): device pointers and sizes."""
    _fields_ = [("ent", _p), ("val", _p), ("tile_start", _p), ("tile_off", _p), ("row0", _p), ("out_row", _p),
                ("vslot", _p),
                ("nslots", _p), ("n_blocks", _i64), ("n_panels", _i64), ("n_rows", _i64), ("dense_rows", _i64),
                ("n_entries", _i64), ("packed", _i32), ("value_type", _i32)]


class WindowLayout(C.Structure):
    """snapgpu_window_layout (include/snapgpu.h): device pointers and sizes."""
    _fields_ = [("ent", _p), ("bucket_start", _p), ("quad_off", _p), ("block_row", _p), ("row_slot", _p),
                ("n_blocks", _i64), ("n_windows", _i64), ("n_rows", _i64), ("dense_rows", _i64),
                ("n_entries", _i64)]


class PeakIndex(C.Structure):
    """snapgpu_peak_index (include/snapgpu.h): device pointers, the chromosome and column counts."""
    _fields_ = [("start", _p), ("end", _p), ("pmax", _p), ("orig", _p), ("chrom_off", _p), ("base", _p),
                ("n_chrom", _i32), ("n_columns", _i32)]


class FragOptions(C.Structure):
    """snapgpu_frag_options (include/snapgpu.h)."""
    _fields_ = [("shift_left", _i64), ("shift_right", _i64), ("min_num_fragments", _u64),
                ("sorted_by_barcode", _i32), ("reserved", _i32)]


class FragSummary(C.Structure):
    """snapgpu_frag_summary (include/snapgpu.h)."""
    _fields_ = [("n_lines", _i64), ("n_records", _i64), ("n_kept", _i64), ("n_groups", _i64),
                ("n_rows", _i64), ("nnz", _i64), ("paired", _i32), ("name_width", _i32)]


class SnapGpuError(RuntimeError):
    """A failure reported by libsnapgpu."""


class PanicException(SnapGpuError):
    """Raised where the reference Rust code panics on invalid data."""


class CapacityError(SnapGpuError):
    """A fixed capacity of a kernel or layout was exceeded (SNAPGPU_ERR_CAPACITY)."""


_lib = None
_handle = None


class _Trace:
    """Per-entry-point GPU time: each library call that returns a status code is
    bracketed by two HIP events on the calling thread's current stream (the
    stream every call here launches on), and summary() adds up the time between
    them once the device is idle."""

    def __init__(self):
        import threading
        self._lock = threading.Lock()
        self._pending = []

    def run(self, name, fn, args):
        import torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = fn(*args)
        e1.record()
        with self._lock:
            self._pending.append((name, e0, e1))
        return rc

    def summary(self) -> dict:
        """{entry point: {"calls": n, "ms": GPU milliseconds}}, longest first."""
        out: dict = {}
        if not self._pending:
            return out
        import torch
        torch.cuda.synchronize()
        with self._lock:
            for name, e0, e1 in self._pending:
                rec = out.setdefault(name, {"calls": 0, "ms": 0.0})
                rec["calls"] += 1
                rec["ms"] += e0.elapsed_time(e1)
        return dict(sorted(out.items(), key=lambda kv: -kv[1]["ms"]))


class _TracedLib:
    """The library with its status-returning entry points timed by a _Trace."""

    def __init__(self, handle, trace):
        self._handle, self._trace = handle, trace

    def __getattr__(self, name):
        fn = getattr(self._handle, name)
        res, args = SIGNATURES.get(name, (None, []))
        if res is not C.c_int or not args:
            return fn  # sizes, versions and the error text are not timed

        def call(*args):
            return self._trace.run(name, fn, args)
        return call


_trace = None


def enable_trace() -> None:
    """Time every library call from now on (what SNAPGPU_TRACE does at load)."""
    global _trace, _lib
    lib()
    if _trace is None:
        _trace = _Trace()
        _lib = _TracedLib(_handle, _trace)


def disable_trace() -> dict:
    """Stop timing; returns the summary of the calls timed so far."""
    global _trace, _lib
    out = trace_summary()
    _trace = None
    _lib = _handle
    return out


def trace_summary() -> dict:
    return _trace.summary() if _trace is not None else {}


def _trace_at_exit(dest: str) -> None:
    import json
    import sys
    rows = trace_summary()
    if dest in ("1", "stderr"):
        for name, rec in rows.items():
            print(f"{rec['ms']:12.3f} ms {rec['calls']:8d}  {name}", file=sys.stderr)
    else:
        with open(dest, "w") as f:
            json.dump(rows, f, indent=1)


def lib() -> C.CDLL:
    """The loaded library. With SNAPGPU_TRACE set (1: a table on stderr at exit;
    a path: JSON written there at exit), every status-returning call is timed
    on the GPU with HIP events (enable_trace / trace_summary for the same from
    code)."""
    global _lib, _handle
    if _handle is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m snapatac2_amd.build` "
                "(there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _handle = _lib = handle
        import os
        dest = os.environ.get("SNAPGPU_TRACE")
        if dest:
            import atexit
            enable_trace()
            atexit.register(_trace_at_exit, dest)
    return _lib


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = lib().snapgpu_last_error().decode(errors="replace")
    if rc == ERR_ARG:
        raise ValueError(msg)
    if rc == ERR_DATA:
        raise PanicException(msg)
    if rc == ERR_CAPACITY:
        raise CapacityError(msg)
    raise SnapGpuError(msg)


def require_gpu():
    """Returns torch and the current device; raises when no ROCm GPU is visible."""
    import torch
    if not torch.cuda.is_available():
        raise SnapGpuError("snapatac2_amd needs a ROCm GPU (MI355X / gfx950); none is visible")
    return torch


def stream_ptr(torch_mod=None) -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int:
    """Device address of a torch tensor (0 for None)."""
    return 0 if t is None else t.data_ptr()
