"""CPU oracle for local PCA (wot/ot/util.py:240-255).

TEST INFRASTRUCTURE ONLY (never imported by the product).

* ``reference_pca``     the reference's call as written: stack, subtract the
                        per-gene mean, sklearn PCA(n_components,
                        random_state=58951) fitted on x^T (svd_solver 'auto'
                        unless given), components_.T split by day
* ``double_centered``   the matrix the PCA decomposes (closed form)
* ``exact_coordinates`` numpy SVD of that matrix: V diag(S) and S, signs fixed
                        so each component's largest-magnitude entry is positive
* ``sqdist``            pairwise squared distances (sign-invariant comparisons)

Pinning (tests/test_oracle_pca.py): sklearn's exact solver against the numpy
SVD (fixes the centering: sklearn centers the columns of x^T, i.e. each cell),
centering invariants (zero row and column means), and the spectrum against
the eigenvalues of the Gram matrix.
"""
import numpy as np


def reference_pca(m1, m2, n_components, svd_solver='auto'):
    import sklearn.decomposition
    x = np.vstack([np.asarray(m1), np.asarray(m2)])
    mean_shift = x.mean(axis=0)
    x = x - mean_shift
    n_components = min(n_components, x.shape[0])
    pca = sklearn.decomposition.PCA(n_components=n_components, random_state=58951, svd_solver=svd_solver)
    pca.fit(x.T)
    comp = pca.components_.T
    n1 = np.asarray(m1).shape[0]
    return comp[:n1], comp[n1:], pca, mean_shift


def double_centered(m1, m2):
    x = np.vstack([np.asarray(m1, dtype=np.float64), np.asarray(m2, dtype=np.float64)])
    x = x - x.mean(axis=0)
    return x - x.mean(axis=1, keepdims=True)


def exact_coordinates(m1, m2, k):
    A = double_centered(m1, m2)
    U, S, Vt = np.linalg.svd(A, full_matrices=False)  # A = U S Vt: U are the cells' singular vectors
    V = U[:, :k]
    piv = np.argmax(np.abs(V), axis=0)
    sgn = np.sign(V[piv, np.arange(k)])
    sgn[sgn == 0] = 1.0
    V = V * sgn
    return V * S[:k], S[:k], V


def sqdist(Z):
    n = (Z * Z).sum(1)
    return np.maximum(n[:, None] + n[None, :] - 2.0 * Z @ Z.T, 0.0)
