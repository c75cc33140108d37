"""h5ad transport maps through the HDF5 C library (wot_amd.h5ad): round
trips, fixtures written by h5py in AnnData's layout (an independent writer),
h5py reading our files when an h5py interpreter is available, and the
TransportMapModel index files (from_directory / to_json / from_json)."""
import json
import os
import subprocess

import numpy as np
import pandas as pd
import pytest

from wot_amd import _anndata, h5ad, io

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'h5ad')
H5PY_PYTHON = '/opt/conda/bin/python3.9'


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for dt in (np.float64, np.float32):
        X = rng.random((7, 5)).astype(dt)
        obs = pd.DataFrame({'g0': rng.random(7), 'g1': rng.random(7), 'n': np.arange(7),
                            'flag': rng.random(7) > 0.5, 'label': list('abcdefg')},
                           index=['cell-%d-ü' % i for i in range(7)])
        var = pd.DataFrame(index=['t%d' % j for j in range(5)])
        p = str(tmp_path / ('m_%s.h5ad' % np.dtype(dt).name))
        h5ad.write_h5ad(p, X, obs, var)
        a = h5ad.read_h5ad(p)
        assert a.X.dtype == dt and np.array_equal(a.X, X)
        assert list(a.obs.index) == list(obs.index) and list(a.var.index) == list(var.index)
        for c in obs.columns:
            assert list(a.obs[c]) == list(obs[c]), c
        assert a.obs['flag'].dtype == bool and a.obs['n'].dtype == np.int64
        assert list(h5ad.read_index(p, 'var')) == list(var.index)


def test_layout_attributes(tmp_path):
    p = str(tmp_path / 'm.h5ad')
    h5ad.write_h5ad(p, np.eye(3), pd.DataFrame({'g0': [1.0, 2, 3]}, index=list('abc')),
                    pd.DataFrame(index=list('xyz')))
    assert h5ad.read_attrs(p) == {'encoding-type': 'anndata', 'encoding-version': '0.1.0'}
    assert h5ad.read_attrs(p, 'obs') == {'encoding-type': 'dataframe', 'encoding-version': '0.1.0',
                                         '_index': '_index', 'column-order': ['g0']}
    assert h5ad.read_attrs(p, 'var')['column-order'] == []
    for k in ('X', 'obs', 'var', 'uns', 'obsm', 'varm', 'obsp', 'varp', 'layers'):
        assert h5ad.has_key(p, k), k


def test_reads_h5py_written_fixtures():
    a = h5ad.read_h5ad(os.path.join(GOLD, 'tmap_dense_h5py.h5ad'))
    want = np.array([[(i + 1) / (j + 2) for j in range(3)] for i in range(4)])
    assert np.array_equal(a.X, want)
    assert list(a.obs.index) == ['c1-t0', 'c2-t0', 'c3-t0', 'c4-t0']
    assert list(a.var.index) == ['c1-t1', 'c2-t1', 'c3-t1']
    assert list(a.obs.columns) == ['g0', 'g1', 'n', 'ok', 'label']
    assert list(a.obs['g1']) == [0.5, 1.5, 2.5, 3.5]
    assert list(a.obs['n']) == [1, 2, 3, 4]
    assert list(a.obs['ok']) == [True, False, True, True]
    assert list(a.obs['label']) == ['a', 'b', 'ä', 'd']
    b = h5ad.read_h5ad(os.path.join(GOLD, 'tmap_csr_h5py.h5ad'))
    assert np.array_equal(b.X, [[1, 0, 0, 2], [0, 3, 0, 0], [0, 0, 4, 0]])
    assert list(b.obs.index) == ['r0', 'r1', 'r2']  # no '_index' attribute: the 'index' dataset
    assert list(h5ad.read_index(os.path.join(GOLD, 'tmap_csr_h5py.h5ad'), 'var')) == ['k0', 'k1', 'k2', 'k3']


def _h5py_available():
    if not os.path.exists(H5PY_PYTHON):
        return False
    r = subprocess.run(['env', '-i', 'PATH=/usr/bin:/bin', 'HOME=/tmp', H5PY_PYTHON, '-c', 'import h5py'],
                       capture_output=True)
    return r.returncode == 0


@pytest.mark.skipif(not _h5py_available(), reason='no interpreter with h5py in this image')
def test_h5py_reads_our_files(tmp_path):
    p = str(tmp_path / 'm.h5ad')
    X = np.arange(12.0).reshape(4, 3) / 7
    obs = pd.DataFrame({'g0': [1.0, 2, 3, 4], 'ok': [True, False, False, True], 'label': list('wxyz')},
                       index=['a', 'b', 'c', 'ü'])
    h5ad.write_h5ad(p, X, obs, pd.DataFrame(index=['u', 'v', 'w']))
    script = '''
import h5py, json, sys
f = h5py.File(sys.argv[1], 'r')
o = f['obs']
out = {'root': {k: str(v) for k, v in f.attrs.items()}, 'keys': sorted(f.keys()),
       'X': f['X'][()].tolist(), 'X_dtype': str(f['X'].dtype),
       'obs_attrs': {k: (list(map(str, v)) if hasattr(v, 'tolist') else str(v)) for k, v in o.attrs.items()},
       'obs_index': [s.decode() for s in o[o.attrs['_index']][()]],
       'index_dtype': str(h5py.check_string_dtype(o['_index'].dtype)),
       'g0': o['g0'][()].tolist(), 'ok': o['ok'][()].tolist(), 'ok_dtype': str(o['ok'].dtype),
       'label': [s.decode() for s in o['label'][()]],
       'var_index': [s.decode() for s in f['var']['_index'][()]]}
print(json.dumps(out))
'''
    r = subprocess.run(['env', '-i', 'PATH=/usr/bin:/bin', 'HOME=/tmp', H5PY_PYTHON, '-c', script, p],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d['root'] == {'encoding-type': 'anndata', 'encoding-version': '0.1.0'}
    assert d['keys'] == ['X', 'layers', 'obs', 'obsm', 'obsp', 'uns', 'var', 'varm', 'varp']
    assert np.array_equal(np.array(d['X']), X) and d['X_dtype'] == 'float64'
    assert d['obs_attrs']['encoding-type'] == 'dataframe' and d['obs_attrs']['_index'] == '_index'
    assert d['obs_attrs']['column-order'] == ['g0', 'ok', 'label']
    assert d['obs_index'] == ['a', 'b', 'c', 'ü'] and "encoding='utf-8'" in d['index_dtype']
    assert d['g0'] == [1.0, 2.0, 3.0, 4.0] and d['label'] == list('wxyz')
    assert d['ok'] == [True, False, False, True] and d['ok_dtype'] == 'bool'
    assert d['var_index'] == ['u', 'v', 'w']


def _write_course(tmp_path, fmt='h5ad'):
    rng = np.random.default_rng(3)
    sizes = [4, 3, 5]
    ids = [['d%d_%d' % (t, i) for i in range(n)] for t, n in enumerate(sizes)]
    maps = []
    for t in range(2):
        X = rng.random((sizes[t], sizes[t + 1]))
        maps.append(X)
        ds = _anndata.AnnData(X, pd.DataFrame({'g0': np.ones(sizes[t])}, index=ids[t]),
                              pd.DataFrame(index=ids[t + 1]))
        io.write_dataset(ds, str(tmp_path / ('tmaps_%d_%d' % (t, t + 1))), output_format=fmt)
    return maps, ids


def test_write_dataset_h5ad_and_read_dataset(tmp_path):
    maps, ids = _write_course(tmp_path)
    ds = io.read_dataset(str(tmp_path / 'tmaps_0_1.h5ad'))
    assert np.array_equal(ds.X, maps[0]) and list(ds.obs.index) == ids[0] and list(ds.var.index) == ids[1]
    assert list(ds.obs['g0']) == [1.0] * 4


def test_model_from_directory_and_json(tmp_path):
    from wot_amd.tmap import TransportMapModel
    maps, ids = _write_course(tmp_path)
    m = TransportMapModel.from_directory(str(tmp_path / 'tmaps'))
    assert m.timepoints == [0.0, 1.0, 2.0]
    assert set(m.day_pairs) == {(0.0, 1.0), (1.0, 2.0)}
    assert list(m.meta.index) == ids[0] + ids[1] + ids[2]
    assert list(m.meta['day']) == [0.0] * 4 + [1.0] * 3 + [2.0] * 5
    cp = m.get_coupling(1.0, 2.0)  # host read of the path
    assert np.array_equal(cp.X, maps[1])
    idx = str(tmp_path / 'index.json')
    m.to_json(idx)
    m2 = TransportMapModel.from_json(idx)
    assert m2.timepoints == m.timepoints
    assert list(m2.meta.index) == list(m.meta.index)
    assert np.array_equal(m2.get_coupling(0.0, 1.0).X, maps[0])
    with pytest.raises(ValueError, match='No transport maps found'):
        TransportMapModel.from_directory(str(tmp_path / 'nothing'))
