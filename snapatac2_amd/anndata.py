"""A minimal in-memory AnnData container.

The `anndata` package is not a dependency. The drop-in functions duck-type
their `adata` argument (they only use .X, .obs, .var, .obsm, .obsp, .uns,
.n_obs, .n_vars, .shape), so a real anndata.AnnData works as well as this class.
"""
from __future__ import annotations

import numpy as np


class AnnData:
    def __init__(self, X=None, obs=None, var=None, obsm=None, obsp=None, uns=None,
                 n_obs: int | None = None, n_vars: int | None = None):
        import pandas as pd
        self.X = X
        if n_obs is None:
            n_obs = X.shape[0] if X is not None else (len(obs) if obs is not None else 0)
        if n_vars is None:
            n_vars = X.shape[1] if X is not None else (len(var) if var is not None else 0)
        self._n_obs, self._n_vars = int(n_obs), int(n_vars)
        self.obs = obs if obs is not None else pd.DataFrame(index=pd.RangeIndex(self._n_obs).astype(str))
        self.var = var if var is not None else pd.DataFrame(index=pd.RangeIndex(self._n_vars))
        self.obsm = dict(obsm or {})
        self.obsp = dict(obsp or {})
        self.uns = dict(uns or {})
        self._var_names = None
        self.isbacked = False

    @property
    def n_obs(self) -> int:
        return self._n_obs

    @property
    def n_vars(self) -> int:
        return self._n_vars

    @property
    def shape(self) -> tuple[int, int]:
        return (self._n_obs, self._n_vars)

    @property
    def obs_names(self):
        return self.obs.index

    @property
    def var_names(self):
        return self._var_names if self._var_names is not None else self.var.index

    def set_n_vars(self, n: int, var_names=None) -> None:
        import pandas as pd
        self._n_vars = int(n)
        self.var = pd.DataFrame(index=pd.RangeIndex(self._n_vars))
        self._var_names = var_names

    def chunked_X(self, chunk_size: int):
        """Yields (csr_chunk, start, end) like the reference AnnData.chunked_X."""
        X = self.X
        if hasattr(X, "to_scipy"):
            X = X.to_scipy()
        for s in range(0, self._n_obs, chunk_size):
            e = min(s + chunk_size, self._n_obs)
            yield X[s:e], s, e

    def __repr__(self) -> str:
        return (f"AnnData object with n_obs x n_vars = {self._n_obs} x {self._n_vars}\n"
                f"    obsm: {list(self.obsm)}\n    obsp: {list(self.obsp)}\n    uns: {list(self.uns)}")
