"""AnnData container: the real ``anndata.AnnData`` when it is importable,
otherwise a minimal stand-in with the attributes the OT path touches
(X, obs, var, shape, boolean row selection, obs_names/var_names, copy)."""
import numpy as np
import pandas as pd

try:  # pragma: no cover - depends on the environment
    from anndata import AnnData  # noqa: F401
    HAVE_ANNDATA = True
except ImportError:  # pragma: no cover
    HAVE_ANNDATA = False

    class AnnData:
        def __init__(self, X=None, obs=None, var=None):
            self.X = X
            n, m = (X.shape if X is not None else (0, 0))
            self.obs = obs if obs is not None else pd.DataFrame(index=pd.RangeIndex(n).astype(str))
            self.var = var if var is not None else pd.DataFrame(index=pd.RangeIndex(m).astype(str))
            if len(self.obs) != n or len(self.var) != m:
                raise ValueError('obs/var lengths do not match X shape %s' % (X.shape,))

        @property
        def shape(self):
            return self.X.shape

        @property
        def obs_names(self):
            return self.obs.index

        @property
        def var_names(self):
            return self.var.index

        def __getitem__(self, key):
            rows, cols = key if isinstance(key, tuple) else (key, slice(None))
            rows = np.asarray(rows) if not isinstance(rows, slice) else rows
            cols = np.asarray(cols) if not isinstance(cols, slice) else cols
            X = self.X[rows][:, cols]
            return AnnData(X, self.obs.iloc[rows] if not isinstance(rows, slice) else self.obs[rows],
                           self.var.iloc[cols] if not isinstance(cols, slice) else self.var[cols])

        def copy(self):
            return AnnData(np.array(self.X, copy=True), self.obs.copy(), self.var.copy())
