"""PCA step before the cost (wot/ot/util.py:240-255), on the host with sklearn
exactly as the reference does it."""
import numpy as np
import scipy.sparse


def compute_pca(m1, m2, n_components):
    """Stack both days, centre per feature, fit ``PCA(n_components,
    random_state=58951)`` on the transposed matrix and split
    ``components_.T`` back into the two days.  Returns (pca_1, pca_2, pca,
    mean_shift); the caller scales by diag(singular_values_)."""
    import sklearn.decomposition
    a = m1.toarray() if scipy.sparse.issparse(m1) else np.asarray(m1)
    b = m2.toarray() if scipy.sparse.issparse(m2) else np.asarray(m2)
    x = np.vstack([a, b])
    mean_shift = x.mean(axis=0)
    x = x - mean_shift
    n_components = min(n_components, x.shape[0])
    pca = sklearn.decomposition.PCA(n_components=n_components, random_state=58951)
    pca.fit(x.T)
    comp = pca.components_.T
    return comp[:a.shape[0]], comp[a.shape[0]:a.shape[0] + b.shape[0]], pca, mean_shift
