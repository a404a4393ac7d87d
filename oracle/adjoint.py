"""ORACLE (test infrastructure only) for path D, the derivative of the spectral operator.

The reference has no derivative of its operator; what is differentiated is the
forward map it builds in snapatac2-python/src/embedding.rs:131-171 with the
weighting of embedding.rs:314-326 (restated in oracle/spectral.py):

    Y = A(w) V,  A(w) = D^-1/2 (X~ X~^T - I) D^-1/2,
    X~ = rownorm(X diag(w)),  D = diag(X~ X~^T 1 - 1).

Three independent FP64 restatements of d<G, A(w) V>/dw:
  fd_grad            central finite differences of the forward (h = 1e-3)
  torch_sparse_grad  torch autograd through torch.sparse.mm (CPU, FP64)
  closed_form_grad   the closed form the device kernels evaluate
and, for full-size checks, the forward and the closed form through torch's
sparse products on any device (TorchCounts, apply_tensors,
closed_form_grad_tensors; torch.sparse autograd materialises a dense n x m
gradient there, so it is used at small sizes only), and for matrices too large
for those COO copies the forward by row chunks with torch gathers and
index_add_ (ChunkedCounts, with idf_from_df).

Pinned by tests/test_adjoint_oracle.py: the three (and the torch-tensor
forms) agree with each other, the
forward equals oracle/spectral.py's operator, dV = A G (symmetry), the scale
invariance <dw, w> = 0 (A(c w) = A(w): a dropped row-norm term breaks it), and
Hellmann-Feynman against finite differences of dense eigenvalues;
ChunkedCounts.apply equals the forward over several chunks, and idf_from_df
equals oracle/spectral.py's idf, including its all-equal and clamp cases.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import spectral as _sp


def apply(X, w, V) -> np.ndarray:
    """A(w) V in FP64 (empty rows map v to +v, as dinv = -1 there)."""
    Xh, dinv, _, _ = _sp.operator_parts(X, feature_weights=w)
    V = np.asarray(V, dtype=np.float64)
    return np.asarray(Xh @ (Xh.T @ V)) - dinv[:, None] * V


def objective(X, w, V, G) -> float:
    return float(np.sum(np.asarray(G, dtype=np.float64) * apply(X, w, V)))


def fd_grad(X, w, V, G, h: float = 1e-3, coords=None, directions=None):
    """Central differences of <G, A(w) V>: per coordinate (coords) and/or along the
    given directions (rows of `directions`). Returns (d_coords, d_directions)."""
    w = np.asarray(w, dtype=np.float64)
    dc = None
    if coords is not None:
        dc = np.empty(len(coords))
        for t, j in enumerate(coords):
            e = np.zeros_like(w)
            e[j] = h
            dc[t] = (objective(X, w + e, V, G) - objective(X, w - e, V, G)) / (2 * h)
    dd = None
    if directions is not None:
        dirs = np.atleast_2d(directions)
        dd = np.empty(dirs.shape[0])
        for t, p in enumerate(dirs):
            dd[t] = (objective(X, w + h * p, V, G) - objective(X, w - h * p, V, G)) / (2 * h)
    return dc, dd


def torch_sparse_grad(X, w, V, G):
    """(Y, dV, dw) by torch autograd through torch.sparse.mm, FP64 on the CPU."""
    import torch
    Xc = sp.csr_matrix(X, dtype=np.float64)
    Xc.sort_indices()
    coo = Xc.tocoo()
    f64 = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
    Y, dV, dw = torch_sparse_grad_tensors(
        torch.tensor(coo.row, dtype=torch.long), torch.tensor(coo.col, dtype=torch.long),
        f64(coo.data), Xc.shape, f64(w), f64(V), f64(G))
    return Y.numpy(), dV.numpy(), dw.numpy()


def torch_sparse_grad_tensors(rows, cols, vals, shape, w, V, G):
    """torch_sparse_grad on torch tensors (COO entries, FP64), on their device.
    Used at sizes where the CPU is too slow; torch's own sparse kernels, not
    this project's, do the arithmetic."""
    import torch
    n, m = shape
    dev = vals.device
    wt = w.detach().clone().requires_grad_(True)
    Vt = V.detach().clone().requires_grad_(True)
    vw = vals * wt[cols]
    nrm = torch.zeros(n, dtype=torch.float64, device=dev).index_add(0, rows, vw * vw)
    u = torch.where(nrm > 0, nrm.clamp_min(1e-300).rsqrt(), torch.zeros_like(nrm))
    Xn = torch.sparse_coo_tensor(torch.stack([rows, cols]), vw * u[rows], (n, m)).coalesce()
    c = torch.sparse.mm(Xn.t(), torch.ones(n, 1, dtype=torch.float64, device=dev))
    d = torch.sparse.mm(Xn, c).squeeze(1) - 1.0
    nonempty = torch.zeros(n, dtype=torch.bool, device=dev)
    nonempty[rows] = True
    a = torch.where(nonempty, d.abs().rsqrt(), torch.zeros_like(d))
    Y = a[:, None] * torch.sparse.mm(Xn, torch.sparse.mm(Xn.t(), a[:, None] * Vt)) - Vt / d[:, None]
    (Y * G).sum().backward()
    return Y.detach(), Vt.grad, wt.grad


class TorchCounts:
    """X, X^T, X o X and (X o X)^T as FP64 torch COO tensors on the device of
    `vals` (torch's own sparse kernels do the arithmetic, not this project's),
    for checks at sizes where scipy on the CPU is too slow."""

    def __init__(self, rows, cols, vals, shape):
        import torch
        self.torch = torch
        self.n, self.m = shape
        idx = torch.stack([rows, cols])
        self.X = torch.sparse_coo_tensor(idx, vals, shape).coalesce()
        self.XT = self.X.t().coalesce()
        self.X2 = torch.sparse_coo_tensor(idx, vals * vals, shape).coalesce()
        self.X2T = self.X2.t().coalesce()
        self.nonempty = torch.zeros(self.n, dtype=torch.bool, device=vals.device)
        self.nonempty[rows] = True

    def mm(self, A, B):
        """Sparse A times a dense vector or matrix B."""
        if B.dim() == 1:
            return self.torch.sparse.mm(A, B[:, None]).squeeze(1)
        return self.torch.sparse.mm(A, B)

    def scales(self, w):
        """(u, d, a): row scales, degrees and a = |d|^-1/2 (0 on empty rows)."""
        torch = self.torch
        nrm = self.mm(self.X2, w * w)
        u = torch.where(nrm > 0, nrm.clamp_min(1e-300).rsqrt(), torch.zeros_like(nrm))
        e = self.mm(self.XT, u)
        d = u * self.mm(self.X, w * w * e) - 1.0
        a = torch.where(self.nonempty, d.abs().rsqrt(), torch.zeros_like(d))
        return u, d, a, e


def apply_tensors(T: TorchCounts, w, V):
    """A(w) V in FP64 through torch's sparse products."""
    u, d, a, _ = T.scales(w)
    r = (u * a)[:, None]
    return r * T.mm(T.X, (w * w)[:, None] * T.mm(T.XT, r * V)) - V / d[:, None]


def closed_form_grad_tensors(T: TorchCounts, w, V, G):
    """closed_form_grad through torch's sparse products (same formulas)."""
    torch = T.torch
    u, d, a, e = T.scales(w)
    r = (u * a)[:, None]
    w2 = w * w
    P = T.mm(T.XT, r * V)
    Q = T.mm(T.XT, r * G)
    h = (P * Q).sum(1)
    alpha = (G * T.mm(T.X, w2[:, None] * P)).sum(1) + (V * T.mm(T.X, w2[:, None] * Q)).sum(1)
    beta = (G * V).sum(1)
    zero = torch.zeros_like(u)
    gamma = torch.where(T.nonempty, -0.5 * alpha * u * a ** 3 + beta / d ** 2, zero)
    phi = T.mm(T.XT, gamma * u)
    psi = T.mm(T.X, w2 * phi)
    kappa = torch.where(T.nonempty, alpha * a * u ** 3 + gamma * u ** 2 * (d + 1.0) + u ** 3 * psi, zero)
    sigma = T.mm(T.X2T, kappa)
    return w * (2.0 * h + 2.0 * e * phi - sigma)


def closed_form_grad(X, w, V, G) -> np.ndarray:
    """dw from the closed form (u = row scales, d = degrees, a = d^-1/2, r = u a):
        P = X^T (r o V), Q = X^T (r o G), h = rowdot(P, Q), e = X^T u
        alpha = rowdot(G, X (w^2 o P)) + rowdot(V, X (w^2 o Q)),  beta = rowdot(G, V)
        gamma = -alpha u d^-3/2 / 2 + beta d^-2,  phi = X^T (gamma o u)
        psi = X (w^2 o phi),  kappa = alpha a u^3 + gamma u^2 (d + 1) + u^3 psi
        dw = w o (2 h + 2 e o phi - (X o X)^T kappa)."""
    Xc = sp.csr_matrix(X, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    V = np.asarray(V, dtype=np.float64)
    G = np.asarray(G, dtype=np.float64)
    nrm = np.asarray(Xc.multiply(Xc) @ (w * w)).ravel()
    with np.errstate(divide="ignore"):
        u = np.where(nrm > 0, 1.0 / np.sqrt(np.where(nrm > 0, nrm, 1.0)), 0.0)
    e = np.asarray(Xc.T @ u).ravel()
    d = u * np.asarray(Xc @ (w * w * e)).ravel() - 1.0
    ne = u > 0
    a = np.where(ne, 1.0 / np.sqrt(np.abs(d)), 0.0)
    r = u * a
    P = np.asarray(Xc.T @ (r[:, None] * V))
    Q = np.asarray(Xc.T @ (r[:, None] * G))
    h = np.sum(P * Q, axis=1)
    w2 = w * w
    alpha = np.sum(G * np.asarray(Xc @ (w2[:, None] * P)), axis=1) + \
        np.sum(V * np.asarray(Xc @ (w2[:, None] * Q)), axis=1)
    beta = np.sum(G * V, axis=1)
    gamma = np.where(ne, -0.5 * alpha * u * a ** 3 + beta / d ** 2, 0.0)
    phi = np.asarray(Xc.T @ (gamma * u)).ravel()
    psi = np.asarray(Xc @ (w2 * phi)).ravel()
    kappa = np.where(ne, alpha * a * u ** 3 + gamma * u ** 2 * (d + 1.0) + u ** 3 * psi, 0.0)
    sigma = np.asarray(Xc.multiply(Xc).T @ kappa).ravel()
    return w * (2.0 * h + 2.0 * e * phi - sigma)


class ChunkedCounts:
    """X (a device CSR of integer counts: int64 row pointers, int32/int64
    columns, uint32 counts in an int32 tensor) for FP64 products at sizes where
    TorchCounts' COO copies do not fit: the entries are visited in row chunks,
    each turned into (row, column, value) tensors, and the products are torch
    gathers and index_add_ (torch's kernels, not this project's). Same formulas
    as TorchCounts (embedding.rs:133-162 with the weighting of :314-326)."""

    def __init__(self, indptr, indices, counts, shape, chunk_entries: int = 1 << 28):
        import torch
        self.torch = torch
        self.n, self.m = int(shape[0]), int(shape[1])
        self.indptr, self.indices, self.counts = indptr, indices, counts
        ip = indptr.cpu().numpy()
        bounds = [0]
        while bounds[-1] < self.n:
            r0 = bounds[-1]
            r1 = int(np.searchsorted(ip, ip[r0] + chunk_entries, side="right")) - 1
            bounds.append(min(self.n, max(r1, r0 + 1)))
        self.bounds = bounds
        self.nonempty = (indptr[1:] > indptr[:-1])

    def chunks(self, squared: bool = False):
        torch = self.torch
        for r0, r1 in zip(self.bounds[:-1], self.bounds[1:]):
            a, b = int(self.indptr[r0].item()), int(self.indptr[r1].item())
            lens = torch.diff(self.indptr[r0:r1 + 1])
            rows = torch.repeat_interleave(torch.arange(r0, r1, device=lens.device), lens)
            cols = self.indices[a:b].long()
            v = (self.counts[a:b].long() & 0xFFFFFFFF).double()
            yield rows, cols, (v * v if squared else v)

    def mm(self, B, squared: bool = False):
        """X B (or (X o X) B) for a dense m x k (or m-vector) B."""
        torch = self.torch
        vec = B.dim() == 1
        Bm = B[:, None] if vec else B
        out = torch.zeros((self.n, Bm.shape[1]), dtype=torch.float64, device=Bm.device)
        for rows, cols, v in self.chunks(squared):
            out.index_add_(0, rows, v[:, None] * Bm[cols])
        return out[:, 0] if vec else out

    def rmm(self, B, squared: bool = False):
        """X^T B for a dense n x k (or n-vector) B."""
        torch = self.torch
        vec = B.dim() == 1
        Bm = B[:, None] if vec else B
        out = torch.zeros((self.m, Bm.shape[1]), dtype=torch.float64, device=Bm.device)
        for rows, cols, v in self.chunks(squared):
            out.index_add_(0, cols, v[:, None] * Bm[rows])
        return out[:, 0] if vec else out

    def df(self):
        """Stored entries per column (idf's document frequency)."""
        torch = self.torch
        out = torch.zeros(self.m, dtype=torch.int64, device=self.indptr.device)
        for _, cols, _ in self.chunks():
            out += torch.bincount(cols, minlength=self.m)
        return out

    def col_sums(self):
        torch = self.torch
        out = torch.zeros(self.m, dtype=torch.float64, device=self.indptr.device)
        for _, cols, v in self.chunks():
            out.index_add_(0, cols, v)
        return out

    def scales(self, w):
        """(u, d, a): row scales 1/||x o w||, degrees d = X~ X~^T 1 - 1 and |d|^-1/2."""
        torch = self.torch
        nrm = self.mm(w * w, squared=True)
        u = torch.where(nrm > 0, nrm.clamp_min(1e-300).rsqrt(), torch.zeros_like(nrm))
        e = self.rmm(u)
        d = u * self.mm(w * w * e) - 1.0
        a = torch.where(self.nonempty, d.abs().rsqrt(), torch.zeros_like(d))
        return u, d, a

    def apply(self, w, V):
        """A(w) V in FP64."""
        u, d, a = self.scales(w)
        r = (u * a)[:, None]
        return r * self.mm((w * w)[:, None] * self.rmm(r * V)) - V / d[:, None]

    def scaled_csr(self, w):
        """(Xh, XhT, d): Xh = diag(u a) X diag(w) and its transpose as FP64 torch
        sparse CSR tensors (torch's own sparse kernels then apply them), with
        A(w) v = Xh (XhT v) - v / d. Needs the entries in memory three times over."""
        torch = self.torch
        u, d, a = self.scales(w)
        r = u * a
        rows, cols, vals = [], [], []
        for rr, cc, v in self.chunks():
            rows.append(rr)
            cols.append(cc)
            vals.append(v * r[rr] * w[cc])
        rows, cols, vals = torch.cat(rows), torch.cat(cols), torch.cat(vals)
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # torch marks sparse CSR as beta
            Xh = torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, (self.n, self.m)).coalesce().to_sparse_csr()
            XhT = torch.sparse_coo_tensor(torch.stack([cols, rows]), vals, (self.m, self.n)).coalesce().to_sparse_csr()
        return Xh, XhT, d


def idf_from_df(df, n: int):
    """idf (embedding.rs:269-286) from the document frequencies (torch, FP64):
    all equal -> 1; else df 0 -> 1, df n -> n - 1, w = ln(n / df)."""
    import torch
    if bool((df == df[0]).all()):
        return torch.ones(df.shape[0], dtype=torch.float64, device=df.device)
    d = df.double().clone()
    d[df == 0] = 1.0
    d[df == n] = n - 1.0
    return torch.log(n / d)
