"""The CPU model of the block-tiled layout (tests/_block_model.py) places every
entry exactly once, keeps every slot and gather index in range and gives the
sparse product, on matrices that exercise multi-piece segments, empty rows,
one-panel and many-panel shapes (no GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

import _block_model as bm


@pytest.mark.parametrize("n,m,dens,tb", [(300, 5000, 0.02, 8), (50, 70000, 0.05, 4), (2000, 900, 0.2, 16)])
def test_model_layout_and_product(n, m, dens, tb):
    rng = np.random.default_rng(n)
    X = sp.random(n, m, density=dens, format="lil", random_state=n, data_rvs=lambda k: rng.integers(1, 5, k))
    X[5, :] = 0
    X[7, :min(m, 3000)] = 2  # segments of up to 1024 entries: 64 pieces
    X = sp.csr_matrix(X)
    X.eliminate_zeros()
    lay = bm.build(X, m, tb)
    assert lay["nslots"].max() <= bm.kBC
    D = rng.standard_normal((lay["P"] * bm.kBR, 3))
    np.testing.assert_allclose(bm.product(lay, D, n), X @ D[:m], rtol=1e-12, atol=1e-12)
