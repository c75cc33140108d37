"""Local PCA on the device (wot/ot/util.py:240-255 compute_pca).

The reference stacks both days' cells (x: cells x genes), subtracts the
per-gene mean and fits ``sklearn.decomposition.PCA(n_components,
random_state=58951)`` on x^T.  So genes are the samples, and sklearn centers
each cell over the genes once more.  The cells' coordinates are
``components_.T`` (the right singular vectors V of the doubly-centered matrix),
scaled by ``singular_values_`` when the cost is built (ot_model.py:301, 245-247).

Here the double centering is one libwotk kernel pass (``wotk_pca_center``).
The top-k singular triplets come from the FP64 Gram matrix of the smaller side
(one rocBLAS GEMM) when that side has at most ``EXACT_MAX`` entries: its top-k
eigenpairs by Rayleigh-Ritz subspace iteration converged to residuals of
``RITZ_TOL`` times the largest eigenvalue (the full symmetric eigensolver when
that does not converge or k is not small).  Otherwise they come from a seeded block subspace
iteration (GEMMs plus thin QR), with ``POWER_ITERS`` iterations and
``OVERSAMPLE`` extra columns.  Signs follow one fixed convention: each
component's largest-magnitude cell coordinate is positive.  Squared distances
between cells, which are all the cost uses, do not depend on signs.  sklearn's
'auto' solver is randomized for large inputs; its answer differs from the exact
singular vectors by its own approximation error.
"""
import numpy as np
import scipy.sparse

from .. import _native
from . import device as _dev

EXACT_MAX = 8192
OVERSAMPLE = 10
POWER_ITERS = 10
SEED = 58951
RITZ_TOL = 1e-13            # Gram eigenpairs: residual / largest eigenvalue
RITZ_ABORT = 1e-6           # residual after 4 iterations above this: use the full solver
MAX_SUBSPACE_ITERS = 12


class DevicePCA:
    """The fitted attributes the reference reads from sklearn's PCA."""

    def __init__(self, V, S, gene_mean, cell_mean):
        self.V = V                   # device, cells x k (components_.T)
        self.S = S                   # device, k
        self.gene_mean = gene_mean   # device, genes (the reference's mean_shift)
        self.cell_mean = cell_mean   # device, cells
        self.n_components_ = int(S.shape[0])

    @property
    def singular_values_(self):
        return self.S.cpu().numpy()

    @property
    def components_(self):
        return self.V.T.cpu().numpy()

    @property
    def explained_variance_(self):
        g = self.gene_mean.shape[0]
        return (self.S ** 2 / max(g - 1, 1)).cpu().numpy()

    def coordinates(self):
        """V diag(S): the coordinates the cost is computed from (device)."""
        return self.V * self.S[None, :]


def _stack(torch, m1, m2, device):
    parts = []
    for m in (m1, m2):
        if isinstance(m, torch.Tensor):
            parts.append(m.to(device))
            continue
        if scipy.sparse.issparse(m):
            m = m.toarray()
        m = np.asarray(m)
        if m.dtype not in (np.float32, np.float64):
            m = m.astype(np.float64)
        parts.append(torch.from_numpy(np.ascontiguousarray(m)).to(device))
    dt = torch.float32 if all(p.dtype == torch.float32 for p in parts) else torch.float64
    return torch.cat([p.to(dt) for p in parts], 0).contiguous()


def center(x):
    """(A, gene_mean, cell_mean) for x (cells x genes, device float32/64)."""
    torch = _dev.torch_mod()
    n, g = x.shape
    L = _native.lib()
    A = torch.empty((n, g), dtype=torch.float64, device=x.device)
    mu = torch.empty(g, dtype=torch.float64, device=x.device)
    nu = torch.empty(n, dtype=torch.float64, device=x.device)
    ws = _dev.workspace(torch, L.wotk_pca_center_workspace_size(n, g), x.device)
    dt = _native.DTYPE_F64 if x.dtype == torch.float64 else _native.DTYPE_F32
    _native.check(L.wotk_pca_center(_dev._ptr(x), dt, x.stride(0), n, g, _dev._ptr(A), A.stride(0), _dev._ptr(mu),
                                    _dev._ptr(nu), _dev._ptr(ws), ws.numel(), _dev._stream(torch)))
    return A, mu, nu


def _gram_top(torch, M, k):
    """Top-k eigenpairs of the symmetric PSD Gram matrix M (device, FP64).

    Block subspace iteration with l = min(2k + 10, size) columns: Y = M Q, an
    orthonormal basis of span(Y) (symmetric orthonormalisation, twice), then
    Rayleigh-Ritz with the small l x l eigenproblems solved on the host (they
    are launch-bound on the device), until every top-k Ritz residual
    ||M v - lambda v|| is below RITZ_TOL * lambda_max.  Falls back to the full
    device eigensolver when l is not small next to the size, when the basis
    degenerates, when the residual after 4 iterations is above RITZ_ABORT
    (slow convergence: a flat spectrum past k), or after MAX_SUBSPACE_ITERS
    iterations.  Returns (eigenvalues
    descending, eigenvectors)."""
    size = M.shape[0]
    l = min(2 * k + 10, size)

    def full():
        w, Q = torch.linalg.eigh(M)
        idx = torch.argsort(w, descending=True)[:k]
        return w[idx], Q[:, idx]

    if 4 * l > size:
        return full()
    dev = M.device
    gen = torch.Generator(device=dev).manual_seed(SEED)
    Q = torch.randn((size, l), dtype=torch.float64, device=dev, generator=gen)

    def orthonormal(Y):
        for _ in range(2):
            G = (Y.T @ Y).cpu().numpy()
            w, V = np.linalg.eigh(0.5 * (G + G.T))
            if not w[0] > 1e-13 * w[-1]:
                return None
            Y = Y @ torch.from_numpy(V / np.sqrt(w)[None, :]).to(dev)
        return Y

    for it in range(MAX_SUBSPACE_ITERS):
        Y = orthonormal(M @ Q)
        if Y is None:
            return full()
        H = (Y.T @ (M @ Y)).cpu().numpy()
        w, Z = np.linalg.eigh(0.5 * (H + H.T))
        order = np.argsort(w)[::-1]
        w, Z = w[order], Z[:, order]
        Q = Y @ torch.from_numpy(np.ascontiguousarray(Z)).to(dev)
        if it % 4 == 3:
            wk = torch.from_numpy(w[:k].copy()).to(dev)
            Vk = Q[:, :k]
            rel_res = float(torch.linalg.vector_norm(M @ Vk - Vk * wk[None, :], dim=0).max()) / float(w[0])
            if rel_res <= RITZ_TOL:
                return wk, Vk
            if it == 3 and rel_res > RITZ_ABORT:
                break  # a flat spectrum past k: the full solver is faster than many more iterations
    return full()


def _top_exact(torch, A, k):
    n, g = A.shape
    if n <= g:
        w, V = _gram_top(torch, A @ A.T, k)   # cells Gram: eigenvectors are V
        S = torch.sqrt(torch.clamp(w, min=0.0))
    else:
        w, U = _gram_top(torch, A.T @ A, k)   # genes Gram: eigenvectors are U, V = A U / S
        S = torch.sqrt(torch.clamp(w, min=0.0))
        V = (A @ U) / torch.where(S > 0, S, torch.ones_like(S))[None, :]
    return V, S


def _top_subspace(torch, A, k):
    n, g = A.shape
    l = min(k + OVERSAMPLE, n, g)
    gen = torch.Generator(device=A.device).manual_seed(SEED)
    Om = torch.randn((g, l), dtype=torch.float64, device=A.device, generator=gen)
    Q, _ = torch.linalg.qr(A @ Om)
    for _ in range(POWER_ITERS):
        Z, _ = torch.linalg.qr(A.T @ Q)
        Q, _ = torch.linalg.qr(A @ Z)
    B = Q.T @ A
    Ub, Sb, _ = torch.linalg.svd(B, full_matrices=False)
    return (Q @ Ub)[:, :k], Sb[:k]


def fit(x, n_components, method='auto'):
    """DevicePCA of the stacked cells x (device tensor, cells x genes)."""
    torch = _dev.torch_mod()
    n, g = x.shape
    k = min(int(n_components), n)
    if k < 1 or k > min(n, g):
        raise ValueError('n_components=%d must be between 1 and min(n_samples, n_features)=%d'
                         % (n_components, min(n, g)))
    A, mu, nu = center(x)
    if method == 'auto':
        method = 'exact' if min(n, g) <= EXACT_MAX else 'subspace'
    if method == 'exact':
        V, S = _top_exact(torch, A, k)
    elif method == 'subspace':
        V, S = _top_subspace(torch, A, k)
    else:
        raise ValueError('method must be auto, exact or subspace')
    del A
    # sign convention: each component's largest-magnitude cell coordinate is positive
    piv = torch.argmax(V.abs(), dim=0)
    sgn = torch.sign(V[piv, torch.arange(k, device=V.device)])
    sgn = torch.where(sgn == 0, torch.ones_like(sgn), sgn)
    return DevicePCA(V * sgn[None, :], S, mu, nu)


def compute_pca(m1, m2, n_components, device='cuda', method='auto'):
    """Drop-in for wot.ot.compute_pca (util.py:240-255): returns (pca_1, pca_2,
    pca, mean_shift) with pca_1 / pca_2 the days' rows of components_.T
    (host float64), pca exposing singular_values_ / components_ and
    mean_shift the per-gene mean."""
    torch = _dev.torch_mod()
    x = _stack(torch, m1, m2, device)
    pca = fit(x, n_components, method)
    V = pca.V.cpu().numpy()
    n1 = m1.shape[0]
    return V[:n1], V[n1:], pca, pca.gene_mean.cpu().numpy()
