"""The gather-rate probe that bench.py uses as the ceiling of the operator's
sparse products: every index is gathered once by the 16 lanes of its group."""
import pytest

from snapatac2_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count", [1, 255, 256, 4096 * 64 * 4 + 17, 3_000_001])
def test_probe_covers_every_index(torch_gpu, count):
    torch = torch_gpu
    rows = 1000
    gen = torch.Generator(device="cuda").manual_seed(count)
    idx = torch.randint(0, rows, (count,), dtype=torch.int32, device="cuda", generator=gen)
    dense = torch.arange(rows * 16, dtype=torch.float32, device="cuda") % 3.0
    out = torch.empty(4096 * 256, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().snapgpu_gather_probe(idx.data_ptr(), count, dense.data_ptr(), rows,
                                               out.data_ptr(), out.numel(), _lib.stream_ptr()))
    expect = dense.view(rows, 16).double()[idx.long()].sum()
    assert abs(out.double().sum().item() - expect.item()) <= 1e-6 * max(1.0, expect.item())


def test_probe_rejects_short_output(torch_gpu):
    torch = torch_gpu
    dense = torch.zeros(16, dtype=torch.float32, device="cuda")
    out = torch.empty(10, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().snapgpu_gather_probe(0, 0, dense.data_ptr(), 1, out.data_ptr(), 10,
                                                   _lib.stream_ptr()))
