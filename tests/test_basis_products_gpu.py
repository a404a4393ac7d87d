"""FP64 tall-skinny products of the block Lanczos basis (FP64 MFMA kernels)
against torch's FP64 matmul on the same device tensors.

C = A^T B (snapgpu_gemm_tn_f64) and C = beta C + alpha A B
(snapgpu_gemm_nn_f64) over column counts that hit every tile width, partial
column groups, odd leading dimensions and base pointers off 16-byte
alignment (the scalar-load path), and more columns than one LDS slab of B^T.
Tolerance: |C - C_torch| <= 1e-13 * (|A|^T |B|) + 1e-300 elementwise — both
sides are FP64 sums of the same products in different orders.
"""
import pytest

from snapatac2_amd import _lib

pytestmark = pytest.mark.gpu


def _tn(torch, A, B):
    L = _lib.lib()
    n, ka = A.shape
    kb = B.shape[1]
    ws_b = int(L.snapgpu_gemm_tn_workspace_bytes(n, ka, kb))
    ws = torch.empty(ws_b, dtype=torch.uint8, device=A.device)
    C = torch.full((ka, kb), float("nan"), dtype=torch.float64, device=A.device)
    _lib.check(L.snapgpu_gemm_tn_f64(A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0), n, ka, kb,
                                     C.data_ptr(), ws.data_ptr(), ws_b, _lib.stream_ptr()))
    return C


def _nn(torch, A, M, C, alpha, beta):
    L = _lib.lib()
    ka, kb = M.shape
    _lib.check(L.snapgpu_gemm_nn_f64(A.data_ptr(), A.stride(0), M.data_ptr(), A.shape[0], ka, kb,
                                     float(alpha), float(beta), C.data_ptr(), C.stride(0),
                                     _lib.stream_ptr()))


def _views(torch, n, ka, kb, pad, shift, gen):
    """A (n x ka) and B (n x kb) as column views of wider buffers: leading
    dimension ka + pad, base pointer shifted by `shift` doubles."""
    Abuf = torch.randn(n, ka + pad + shift, dtype=torch.float64, device="cuda", generator=gen)
    Bbuf = torch.randn(n, kb + pad + shift, dtype=torch.float64, device="cuda", generator=gen)
    return Abuf[:, shift:shift + ka], Bbuf[:, shift:shift + kb]


CASES = [
    # n, ka, kb, pad, shift
    (1, 1, 1, 0, 0),
    (5, 3, 5, 0, 0),
    (1000, 16, 16, 0, 0),
    (1000, 17, 17, 1, 1),
    (4097, 40, 16, 0, 0),
    (70001, 100, 32, 3, 0),
    (70001, 129, 7, 0, 1),
    (200000, 288, 16, 48, 0),
    (100003, 300, 16, 36, 0),
    (50000, 320, 32, 16, 0),
    (30000, 400, 16, 0, 0),
    (30000, 401, 20, 1, 1),
    (9000, 1000, 16, 2, 0),
]


@pytest.mark.parametrize("n,ka,kb,pad,shift", CASES)
def test_gemm_tn_matches_torch(torch_gpu, n, ka, kb, pad, shift):
    torch = torch_gpu
    gen = torch.Generator(device="cuda").manual_seed(n + ka + kb)
    A, B = _views(torch, n, ka, kb, pad, shift, gen)
    C = _tn(torch, A, B)
    ref = A.T @ B
    bound = 1e-13 * (A.abs().T @ B.abs()) + 1e-300
    assert bool(((C - ref).abs() <= bound).all()), float(((C - ref).abs() / bound).max())


@pytest.mark.parametrize("n,ka,kb,pad,shift", CASES)
def test_gemm_nn_matches_torch(torch_gpu, n, ka, kb, pad, shift):
    torch = torch_gpu
    gen = torch.Generator(device="cuda").manual_seed(7 * n + ka + kb)
    A, _ = _views(torch, n, ka, kb, pad, shift, gen)
    M = torch.randn(ka, kb, dtype=torch.float64, device="cuda", generator=gen)
    Cbuf = torch.randn(n, kb + pad, dtype=torch.float64, device="cuda", generator=gen)
    C = Cbuf[:, :kb]
    C0 = C.clone()
    tail = Cbuf[:, kb:].clone()
    for alpha, beta in ((1.0, 0.0), (-1.0, 1.0), (0.5, -2.0)):
        C.copy_(C0)
        _nn(torch, A, M, C, alpha, beta)
        ref = beta * C0 + alpha * (A @ M)
        bound = 1e-13 * (abs(beta) * C0.abs() + abs(alpha) * (A.abs() @ M.abs())) + 1e-300
        assert bool(((C - ref).abs() <= bound).all()), (alpha, beta,
                                                        float(((C - ref).abs() / bound).max()))
    # columns of the buffer beyond kb are untouched
    assert torch.equal(Cbuf[:, kb:], tail)


def test_gemm_tn_repeatable(torch_gpu):
    """Fixed summation order: two calls give identical bits."""
    torch = torch_gpu
    gen = torch.Generator(device="cuda").manual_seed(3)
    A, B = _views(torch, 300000, 288, 16, 48, 0, gen)
    C1 = _tn(torch, A, B)
    C2 = _tn(torch, A, B)
    assert torch.equal(C1, C2)
