"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/snapgpu.h declares; the ctypes table covers them all."""
import ctypes
import os
import re

from snapatac2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(snapgpu_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 8
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert declared_symbols() == set(_lib.SIGNATURES)


def test_host_only_entry_points():
    L = _lib.lib()
    assert L.snapgpu_abi_version() >= 1
    import numpy as np
    lens = np.asarray([1000, 600, 100], dtype=np.uint64)
    excl = np.asarray([0, 0, 1], dtype=np.uint8)
    buf = np.zeros(L.snapgpu_tile_table_bytes(3), dtype=np.uint8)
    nb = ctypes.c_uint64()
    assert L.snapgpu_tile_table_pack(lens.ctypes.data, excl.ctypes.data, 3, 100,
                                     buf.ctypes.data, ctypes.byref(nb)) == 0
    assert nb.value == 16
    assert L.snapgpu_tile_table_pack(lens.ctypes.data, excl.ctypes.data, 3, 0,
                                     buf.ctypes.data, ctypes.byref(nb)) == _lib.ERR_ARG
    assert b"bin_size" in L.snapgpu_last_error()


def test_gemm_tn_noalloc_rejects_beyond_its_scratch():
    """The allocation-free basis product checks its size on the host, before any
    HIP call: 512 chunks x 320 x 32 doubles (40 MiB) exceed the 32 MiB array."""
    L = _lib.lib()
    rc = L.snapgpu_gemm_tn_f64_noalloc(None, 320, None, 32, 512 * 4096, 320, 32, None, None)
    assert rc == _lib.ERR_CAPACITY
    msg = L.snapgpu_last_error()
    assert b"capacity" in msg and b"33554432" in msg
    assert L.snapgpu_gemm_tn_f64_noalloc(None, 4, None, 40, 10, 4, 40, None, None) == _lib.ERR_ARG


def test_prepare_checks_its_workspace_on_the_host():
    """The preparation needs m (w_j, c_j) pairs of workspace (16 bytes each, plus
    alignment); a missing or short workspace is refused before any HIP call."""
    L = _lib.lib()
    m = 1000
    need = int(L.snapgpu_spectral_prepare_workspace_bytes(m))
    assert need >= 16 * m
    args = [None, None, None, 10, None, None, None, m] + [None] * 7 + [0, None]
    assert L.snapgpu_spectral_prepare(*args, None, need, None) == _lib.ERR_ARG
    assert b"workspace" in L.snapgpu_last_error()
    assert L.snapgpu_spectral_prepare(*args, 1 << 20, need - 1, None) == _lib.ERR_ARG
    assert b"workspace" in L.snapgpu_last_error()


def test_probe_library_is_separate():
    """The measurement aids live in libsnapgpu_probe.so, not in the product ABI."""
    from snapatac2_amd.build import PROBE_LIB
    probe = ctypes.CDLL(str(PROBE_LIB))
    for name in ("snapgpu_probe_last_error", "snapgpu_probe_copy_bytes", "snapgpu_probe_gather"):
        assert hasattr(probe, name)
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in ("snapgpu_copy_bytes", "snapgpu_gather_probe", "snapgpu_probe_gather"):
        assert not hasattr(lib, name)


def test_trace_switch_without_calls():
    """Turning the call trace on and off (no GPU call made): size queries and the
    version pass through untimed, and the plain library comes back."""
    plain = _lib.lib()
    _lib.enable_trace()
    try:
        traced = _lib.lib()
        assert traced is not plain
        assert traced.snapgpu_abi_version() == plain.snapgpu_abi_version()
        assert traced.snapgpu_scan_workspace_bytes(1000) == plain.snapgpu_scan_workspace_bytes(1000)
        assert _lib.trace_summary() == {}
    finally:
        assert _lib.disable_trace() == {}
    assert _lib.lib() is plain
