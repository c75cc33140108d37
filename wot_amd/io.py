"""Transport-map files (the subset of wot/io/io.py:380-478 the OT driver and
TransportMapModel use).

'txt' (tab-separated, ``id`` index label) and 'npz' need nothing beyond
numpy/pandas; 'h5ad' is written and read through the HDF5 C library
(``wot_amd.h5ad``) in AnnData's on-disk layout, or by the anndata package when
it is installed.
"""
import numpy as np
import pandas as pd


def check_file_extension(name, output_format):
    """Append '.<format>' unless present (io.py:475-478)."""
    if not str(name).lower().endswith('.' + output_format):
        name += '.' + output_format
    return name


def write_dataset(ds, path, output_format='txt'):
    """Write an AnnData-like transport map (io.py:439-452)."""
    path = check_file_extension(str(path), output_format)
    if output_format == 'txt':
        pd.DataFrame(np.asarray(ds.X), index=ds.obs.index, columns=ds.var.index).to_csv(
            path, index_label='id', sep='\t', doublequote=False)
    elif output_format == 'h5ad':
        if hasattr(ds, 'write'):  # the anndata package
            ds.write(path)
        else:
            from . import h5ad
            h5ad.write_h5ad(path, np.asarray(ds.X), ds.obs, ds.var)
    elif output_format == 'npz':
        obs = {('obs_' + c): ds.obs[c].values for c in ds.obs.columns}
        np.savez(path, X=np.asarray(ds.X), obs_index=np.array([str(v) for v in ds.obs.index], dtype=str),
                 var_index=np.array([str(v) for v in ds.var.index], dtype=str), **obs)
    else:
        raise ValueError('Unknown file format')
    return path


def read_npz_tmap(path):
    """Read back a map written with output_format='npz' -> (X, obs DataFrame, var index)."""
    z = np.load(path, allow_pickle=False)
    obs = pd.DataFrame({k[4:]: z[k] for k in z.files if k.startswith('obs_') and k != 'obs_index'},
                       index=pd.Index(z['obs_index']))
    return z['X'], obs, pd.Index(z['var_index'])


def read_dataset(path):
    """Read a transport map written by write_dataset ('h5ad', 'txt' / 'csv'
    with an 'id' index column, or 'npz') as AnnData (the map formats of
    wot/io/io.py:380-430 read_dataset)."""
    from . import _anndata
    p = str(path)
    low = p.lower()
    if low.endswith('.h5ad'):
        from . import h5ad
        return h5ad.read_h5ad(p)
    if low.endswith('.npz'):
        X, obs, var = read_npz_tmap(p)
        return _anndata.AnnData(np.asarray(X, dtype=np.float64), obs, pd.DataFrame(index=var))
    if low.endswith('.txt') or low.endswith('.csv') or low.endswith('.tsv'):
        sep = ',' if low.endswith('.csv') else '\t'
        df = pd.read_csv(p, sep=sep, index_col=0)
        return _anndata.AnnData(df.values.astype(np.float64), pd.DataFrame(index=df.index.astype(str)),
                                pd.DataFrame(index=df.columns.astype(str)))
    raise ValueError('Unknown file format for %s (supported: .h5ad, .npz, .txt, .tsv, .csv)' % p)
