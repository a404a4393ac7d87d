"""SNAPGPU_TRACE's timing of library calls (snapatac2_amd._lib): every
status-returning entry point called while tracing is on shows up with its call
count and the GPU time between HIP events recorded around it; queries of sizes
and the error text are not timed; turning tracing off restores the plain library."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def test_trace_times_library_calls(torch_gpu):
    from snapatac2_amd import _lib
    from snapatac2_amd.pp._basic import _column_sums
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = sp.random(2000, 3000, density=0.01, format="csr", random_state=np.random.default_rng(1),
                  data_rvs=lambda k: np.ones(k)).astype(np.uint32)
    X = X + sp.csr_matrix((np.ones(2000, dtype=np.uint32), (np.arange(2000), np.arange(2000) % 3000)),
                          shape=(2000, 3000))
    plain = _lib.lib()
    _lib.enable_trace()
    try:
        assert _lib.lib() is not plain
        count = _column_sums(sp.csr_matrix(X, dtype=np.uint32))
        op = SpectralOperator(_to_device_csr(sp.csr_matrix(X, dtype=np.uint32), torch_gpu))
        op.matmat(torch_gpu.randn(2000, 4, dtype=torch_gpu.float64, device="cuda"))
        s = _lib.trace_summary()
    finally:
        out = _lib.disable_trace()
    assert _lib.lib() is plain and _lib.trace_summary() == {}
    assert np.array_equal(count, np.ravel(X.sum(axis=0)).astype(np.float64))
    for name in ("snapgpu_csr_col_sum_rows_u32", "snapgpu_spectral_prepare", "snapgpu_spectral_apply"):
        assert s[name]["calls"] >= 1 and s[name]["ms"] >= 0.0, name
    assert "snapgpu_spectral_apply_workspace_bytes" not in s and "snapgpu_last_error" not in s
    assert set(out) == set(s)
