"""snapgpu_gemm_tn_f64_noalloc (the allocation-free basis product) against
snapgpu_gemm_tn_f64 with a caller workspace: bit-identical results, on two calls
in a row, the second with a smaller input (the static scratch array keeps the
first call's partial products, which the second must not pick up), and both
within FP64 rounding of torch's A^T B."""
import pytest

from snapatac2_amd import _lib

pytestmark = pytest.mark.gpu


def _with_ws(torch, A, B):
    L = _lib.lib()
    n, ka = A.shape
    kb = B.shape[1]
    ws_b = int(L.snapgpu_gemm_tn_workspace_bytes(n, ka, kb))
    ws = torch.empty(ws_b, dtype=torch.uint8, device=A.device)
    C = torch.full((ka, kb), float("nan"), dtype=torch.float64, device=A.device)
    _lib.check(L.snapgpu_gemm_tn_f64(A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0), n, ka, kb,
                                     C.data_ptr(), ws.data_ptr(), ws_b, _lib.stream_ptr()))
    return C


def _noalloc(torch, A, B):
    L = _lib.lib()
    n, ka = A.shape
    kb = B.shape[1]
    C = torch.full((ka, kb), float("nan"), dtype=torch.float64, device=A.device)
    _lib.check(L.snapgpu_gemm_tn_f64_noalloc(A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0), n,
                                             ka, kb, C.data_ptr(), _lib.stream_ptr()))
    return C


# (n, ka, kb): a Lanczos-sized product that fills the scratch with 512 chunks,
# then smaller ones with fewer chunks and narrower partials
CALLS = [(2_100_000, 256, 32), (300_000, 40, 16), (5_000, 7, 3)]


def test_noalloc_matches_workspace_entry_point(torch_gpu):
    torch = torch_gpu
    gen = torch.Generator(device="cuda").manual_seed(11)
    for n, ka, kb in CALLS:
        print(f"no-alloc gemm_tn n={n} ka={ka} kb={kb}", flush=True)
        A = torch.randn(n, ka, dtype=torch.float64, device="cuda", generator=gen)
        B = torch.randn(n, kb, dtype=torch.float64, device="cuda", generator=gen)
        c0 = _noalloc(torch, A, B)
        c1 = _with_ws(torch, A, B)
        torch.cuda.synchronize()
        assert torch.equal(c0, c1), float((c0 - c1).abs().max())
        ref = A.T @ B
        bound = 1e-13 * (A.abs().T @ B.abs()) + 1e-300
        assert bool(((c0 - ref).abs() <= bound).all())
        del A, B


def test_noalloc_strided_views(torch_gpu):
    """Column views with a leading dimension above the width and an odd base
    offset (the scalar-load path), twice in a row."""
    torch = torch_gpu
    gen = torch.Generator(device="cuda").manual_seed(12)
    for n, ka, kb in [(70_001, 129, 7), (1_000, 17, 17)]:
        Ab = torch.randn(n, ka + 3, dtype=torch.float64, device="cuda", generator=gen)
        Bb = torch.randn(n, kb + 2, dtype=torch.float64, device="cuda", generator=gen)
        A, B = Ab[:, 1:1 + ka], Bb[:, 1:1 + kb]
        assert torch.equal(_noalloc(torch, A, B), _with_ws(torch, A, B))
