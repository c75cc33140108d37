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


# ---- cell / gene filters (wot/io/io.py:499-518 filter_adata) ---------------
def _set_file_ext(path):
    """Extension of a gene-set file, with '.gmt.txt' / '.grp.txt' / '.gmx.txt'
    read as the inner format (io.py:480-495)."""
    name = str(path).rsplit('/', 1)[-1].lower()
    parts = name.split('.')
    if len(parts) >= 3 and parts[-1] == 'txt' and parts[-2] in ('gmt', 'grp', 'gmx'):
        return parts[-2]
    return parts[-1] if len(parts) > 1 else ''


def read_set_ids(path):
    """Every id listed in a set file (the obs index of wot.io.read_sets(path),
    io.py:183-204): grp / txt (one id per line; lines starting with '#' or '>'
    skipped), gmt (tab-separated: name, description, members) or gmx (one set
    per column: name, description, then members).  Ids are kept in first-seen
    order, de-duplicated case-insensitively as the reference's reader does."""
    ext = _set_file_ext(path)
    ids = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    if ext in ('txt', 'grp'):
        ids = [ln.strip() for ln in lines if ln and ln[0] not in '#>' and ln.strip()]
    elif ext == 'gmt':
        for ln in lines:
            fields = ln.rstrip('\n').split('\t')
            if len(fields) > 2:
                ids.extend(v.strip() for v in fields[2:] if v.strip())
    elif ext == 'gmx':
        rows = [ln.split('\t') for ln in lines]
        for r in rows[2:]:
            ids.extend(v.strip() for v in r if v.strip())
    else:
        raise ValueError('Unknown file format "{}"'.format(ext))
    seen = {}
    for v in ids:
        seen.setdefault(v.lower(), v)
    return list(seen.values())


def _filter_mask(index, frame, spec):
    """Boolean keep-mask over ``index`` for one filter spec: an existing file
    path (its ids), a comma-separated id list, one boolean column name of
    ``frame``, or (beyond the reference) any non-string iterable of ids."""
    import os
    if isinstance(spec, str):
        if os.path.exists(spec):
            return np.asarray(index.isin(read_set_ids(spec)))
        names = spec.split(',')
        if len(names) == 1 and names[0] in frame:
            return np.asarray(frame[names[0]] == True)  # noqa: E712 (the column may hold 0/1)
        return np.asarray(index.isin(names))
    return np.asarray(index.isin(list(spec)))


def filter_adata(adata, obs_filter=None, var_filter=None):
    """Keep the cells (``obs_filter``) and genes (``var_filter``) a filter
    selects (io.py:499-518)."""
    if obs_filter is not None:
        adata = adata[_filter_mask(adata.obs.index, adata.obs, obs_filter)].copy()
    if var_filter is not None:
        adata = adata[:, _filter_mask(adata.var.index, adata.var, var_filter)].copy()
    return adata
