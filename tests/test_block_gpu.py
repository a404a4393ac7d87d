"""Block-tiled operator layout on the GPU (csrc/block.hip).

- The layout arrays that snapgpu_block_plan/build write equal the CPU model's
  (tests/_block_model.py): row order, blocks, first slots, tile starts and every
  packed entry.
- Y = A V through the block layout against the FP64 oracle (scipy restatement
  of embedding.rs:133-171) and against the row-parallel FP32 path, for b = 1 .. 16,
  packed and unpacked counts, float values, empty rows and columns, one and many
  panels on both sides; bitwise reproducible.
Tolerance: max |Y - Y_ref| <= 1e-5 max |Y_ref| (FP32 sparse products)."""
import numpy as np
import pytest
import scipy.sparse as sp

import _block_model as bm
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _counts(n, m, density, seed):
    rng = np.random.default_rng(seed)
    X = sp.random(n, m, density=density, format="csr", random_state=seed,
                  data_rvs=lambda k: rng.integers(1, 4, k)).astype(np.uint32)
    X = X + sp.csr_matrix((np.ones(n, dtype=np.uint32), (np.arange(n), rng.integers(0, 20, n))), shape=(n, m))
    return sp.csr_matrix(X, dtype=np.uint32)


def test_layout_arrays_equal_the_model(torch_gpu):
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(2500, 4100, 0.02, 3).tolil()
    X[9, :3000] = 3
    X[11, :] = 0
    X = sp.csr_matrix(X, dtype=np.uint32)
    X.eliminate_zeros()
    op = SpectralOperator(_to_device_csr(X, torch), layout="block")
    L = op.block_csr
    mod = bm.build(X, X.shape[1], 512)
    keep = dict(zip(("ent", "val", "tile_start", "row0", "out_row", "vslot", "nslots", "tile_off"),
                    op._panel_keep[0]))
    assert L.n_blocks == mod["nb"] and L.n_panels == mod["P"] and L.packed == 1
    assert np.array_equal(keep["out_row"].cpu().numpy(), mod["order"])
    assert np.array_equal(keep["row0"].cpu().numpy(), mod["row0"])
    assert np.array_equal(keep["vslot"].cpu().numpy(), mod["vslot"])
    assert np.array_equal(keep["nslots"].cpu().numpy()[:mod["nb"]], mod["nslots"])
    assert np.array_equal(keep["tile_start"].cpu().numpy(), mod["tile_start"])
    assert np.array_equal(keep["tile_off"].cpu().numpy().view(np.uint16).reshape(-1, 16)[:mod["nb"] * mod["P"]],
                          mod["tile_off"])
    packed = mod["ent"] | (mod["val"] << 10)
    assert np.array_equal(keep["ent"].cpu().numpy().view(np.uint32)[:X.nnz], packed.astype(np.uint32))


@pytest.mark.parametrize("kind", ["small", "multi_panel", "float", "large_counts", "empty_rows"])
def test_block_operator_vs_oracle(torch_gpu, kind):
    torch = torch_gpu
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    n, m, dens = {"small": (700, 900, 0.05), "multi_panel": (9000, 7000, 0.01)}.get(kind, (6000, 5000, 0.01))
    X = _counts(n, m, dens, 21)
    if kind == "float":
        X = X.astype(np.float32)
        X.data *= np.float32(0.37)
    elif kind == "large_counts":
        X.data[X.indptr[0]:X.indptr[40]] *= 3000  # counts >= 1024: the unpacked encoding
    elif kind == "empty_rows":
        X = X.tolil()
        X[100:140, :] = 0
        X[:, 2000:2600] = 0
        X = sp.csr_matrix(X, dtype=np.uint32)
        X.eliminate_zeros()
    Xd = _to_device_csr(X, torch)
    op = SpectralOperator(Xd, layout="block")
    ref_op = SpectralOperator(Xd, layout="csr")
    assert op.block_csr.packed == (kind in ("small", "multi_panel", "empty_rows"))
    Xh, dinv, _, _ = osp.operator_parts(X)
    rng = np.random.default_rng(3)
    for b in (16, 1, 7, 12, 20):
        V = rng.standard_normal((n, b))
        Vt = torch.from_numpy(V).cuda()
        Y = op.matmat(Vt).cpu().numpy()
        ref = Xh @ (Xh.T @ V) - dinv[:, None] * V
        scale = np.abs(ref).max()
        assert np.abs(Y - ref).max() / scale < 1e-5, (kind, b)
        assert np.abs(Y - ref_op.matmat(Vt).cpu().numpy()).max() / scale < 1e-5, (kind, b)
    Vt = torch.from_numpy(rng.standard_normal((n, 16))).cuda()
    assert torch.equal(op.matmat(Vt), op.matmat(Vt))


def test_block_operator_in_the_eigensolver(torch_gpu):
    """30 eigenpairs through the block layout match those through the panel
    layout (both FP32 products inside the FP64 block Lanczos)."""
    torch = torch_gpu
    from snapatac2_amd.tl._lanczos import eigsh_lm
    from snapatac2_amd.tl._operator import SpectralOperator, _to_device_csr
    X = _counts(40000, 36000, 0.004, 5)
    Xd = _to_device_csr(X, torch)
    ev_b, U_b, _ = eigsh_lm(SpectralOperator(Xd, layout="block"), 30, seed=0)
    ev_p, U_p, _ = eigsh_lm(SpectralOperator(Xd, layout="panel"), 30, seed=0)
    np.testing.assert_allclose(ev_b, ev_p, rtol=1e-6, atol=1e-9)
