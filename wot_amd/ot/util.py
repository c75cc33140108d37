"""PCA step before the cost (wot/ot/util.py:240-255): ``compute_pca`` runs on
the device (``wot_amd.ot.pca``: libwotk double centering, FP64 Gram / subspace
iteration) with the reference's signature and return values."""
from .pca import compute_pca  # noqa: F401
