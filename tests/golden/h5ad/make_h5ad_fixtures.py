"""Write h5ad fixtures with h5py in AnnData's on-disk layout (encoding 0.1.0,
the layout anndata 0.7 writes for a dense transport map and for a CSR X), to
pin wot_amd.h5ad's reader against an independent HDF5 writer.

Run with a Python that has h5py (this image: the conda interpreter, in a
clean environment):

    env -i PATH=/usr/bin:/bin HOME=/tmp /opt/conda/bin/python3.9 tests/golden/h5ad/make_h5ad_fixtures.py

Contents are closed-form, so the test states the expected values directly.
"""
import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STR = h5py.string_dtype(encoding='utf-8')


def dataframe(g, index, columns):
    g.attrs['_index'] = '_index'
    g.attrs['column-order'] = np.array(list(columns), dtype=object) if columns else np.zeros(0)
    g.attrs['encoding-type'] = 'dataframe'
    g.attrs['encoding-version'] = '0.1.0'
    g.create_dataset('_index', data=np.array(index, dtype=object), dtype=STR)
    for k, v in columns.items():
        v = np.asarray(v)
        if v.dtype.kind in 'OU':
            g.create_dataset(k, data=v.astype(object), dtype=STR)
        else:
            g.create_dataset(k, data=v)


def empty_groups(f):
    for k in ('uns', 'obsm', 'varm', 'obsp', 'varp', 'layers'):
        f.create_group(k)


# dense transport map 4 x 3: X[i, j] = (i + 1) / (j + 2), growth columns g0, g1
with h5py.File(os.path.join(HERE, 'tmap_dense_h5py.h5ad'), 'w') as f:
    f.attrs['encoding-type'] = 'anndata'
    f.attrs['encoding-version'] = '0.1.0'
    X = np.array([[(i + 1) / (j + 2) for j in range(3)] for i in range(4)])
    f.create_dataset('X', data=X)
    dataframe(f.create_group('obs'), ['c1-t0', 'c2-t0', 'c3-t0', 'c4-t0'],
              {'g0': np.array([1.0, 2.0, 3.0, 4.0]), 'g1': np.array([0.5, 1.5, 2.5, 3.5]),
               'n': np.array([1, 2, 3, 4], dtype=np.int64), 'ok': np.array([True, False, True, True]),
               'label': np.array(['a', 'b', 'ä', 'd'], dtype=object)})
    dataframe(f.create_group('var'), ['c1-t1', 'c2-t1', 'c3-t1'], {})
    empty_groups(f)

# CSR X 3 x 4 (anndata's csr_matrix group), obs/var without the '_index' attribute
with h5py.File(os.path.join(HERE, 'tmap_csr_h5py.h5ad'), 'w') as f:
    g = f.create_group('X')
    g.attrs['encoding-type'] = 'csr_matrix'
    g.attrs['encoding-version'] = '0.1.0'
    g.attrs['shape'] = np.array([3, 4])
    g.create_dataset('data', data=np.array([1.0, 2.0, 3.0, 4.0]))
    g.create_dataset('indices', data=np.array([0, 3, 1, 2], dtype=np.int32))
    g.create_dataset('indptr', data=np.array([0, 2, 3, 4], dtype=np.int32))
    for key, ids in (('obs', ['r0', 'r1', 'r2']), ('var', ['k0', 'k1', 'k2', 'k3'])):
        grp = f.create_group(key)
        grp.create_dataset('index', data=np.array(ids, dtype=object), dtype=STR)
    empty_groups(f)
print('written')
