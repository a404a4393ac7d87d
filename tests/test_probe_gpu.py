"""bench.py's gather-rate probe (libsnapgpu_probe.so, a measurement aid outside
libsnapgpu): every index is gathered once by the 16 lanes of its group."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from snapatac2_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count", [1, 255, 256, 4096 * 64 * 4 + 17, 3_000_001])
def test_probe_covers_every_index(torch_gpu, count):
    torch = torch_gpu
    L = bench.probe_lib()
    rows = 1000
    gen = torch.Generator(device="cuda").manual_seed(count)
    idx = torch.randint(0, rows, (count,), dtype=torch.int32, device="cuda", generator=gen)
    dense = torch.arange(rows * 16, dtype=torch.float32, device="cuda") % 3.0
    out = torch.empty(4096 * 256, dtype=torch.float32, device="cuda")
    bench.probe_check(L, L.snapgpu_probe_gather(idx.data_ptr(), count, dense.data_ptr(), rows, out.data_ptr(),
                                                out.numel(), _lib.stream_ptr()))
    expect = dense.view(rows, 16).double()[idx.long()].sum()
    assert abs(out.double().sum().item() - expect.item()) <= 1e-6 * max(1.0, expect.item())


def test_probe_rejects_short_output(torch_gpu):
    torch = torch_gpu
    L = bench.probe_lib()
    dense = torch.zeros(16, dtype=torch.float32, device="cuda")
    out = torch.empty(10, dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="4096"):
        bench.probe_check(L, L.snapgpu_probe_gather(0, 0, dense.data_ptr(), 1, out.data_ptr(), 10,
                                                    _lib.stream_ptr()))


def test_copy_probe_copies(torch_gpu):
    torch = torch_gpu
    L = bench.probe_lib()
    src = torch.randint(0, 255, (1 << 20,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros_like(src)
    bench.probe_check(L, L.snapgpu_probe_copy_bytes(src.data_ptr(), dst.data_ptr(), src.numel(),
                                                    _lib.stream_ptr()))
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
