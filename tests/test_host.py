"""Host-side logic that needs no GPU: configuration parsing
(wot/ot/initializer.py), transport-map output (wot/io/io.py:439-478), the
AnnData stand-in, and argument errors raised before any device work."""
import numpy as np
import pandas as pd
import pytest

from wot_amd._anndata import AnnData
from wot_amd.io import check_file_extension, read_npz_tmap, write_dataset
from wot_amd.ot import initializer as ini


def test_per_timepoint_config_is_averaged_onto_pairs():
    cfg = pd.DataFrame({'t': [2, 0, 1], 'epsilon': [0.3, 0.1, 0.2], 'lambda1': [3, 1, 2]})
    pairs = ini.parse_configuration(cfg)
    assert pairs == {(0.0, 1.0): {'epsilon': pytest.approx(0.15), 'lambda1': 1.5},
                     (1.0, 2.0): {'epsilon': pytest.approx(0.25), 'lambda1': 2.5}}


def test_per_timepair_config_keeps_only_pair_fields():
    cfg = pd.DataFrame({'t0': [0.0, 1.0], 't1': [1.0, 2.0], 'epsilon': [0.1, 0.2], 'tau': [5, 6]})
    pairs = ini.parse_configuration(cfg)
    assert pairs == {(0.0, 1.0): {'epsilon': 0.1}, (1.0, 2.0): {'epsilon': 0.2}}


def test_config_errors():
    with pytest.raises(ValueError):
        ini.parse_configuration(pd.DataFrame({'x': [1]}))
    with pytest.raises(ValueError):
        ini.parse_configuration({(0, 1): {}})
    with pytest.raises(ValueError):
        ini.parse_per_timepair_configuration({(0, 1, 2): {}})
    with pytest.raises(ValueError):
        ini.parse_per_timepair_configuration({(0, 1): 3})
    with pytest.raises(ValueError):
        ini.parse_configuration(3)
    assert ini.parse_configuration(None) is None


def test_inline_day_pairs_and_parameter_file(tmp_path):
    pairs = ini.parse_configuration('t0,t1,lambda1;0,1,2;1,2,3')
    assert pairs == {(0, 1): {'lambda1': 2}, (1, 2): {'lambda1': 3}}
    p = tmp_path / 'params.txt'
    p.write_text('epsilon\t0.07\nlambda2\t20\n')
    assert ini.parse_parameter_file(str(p)) == {'epsilon': 0.07, 'lambda2': 20}


def test_write_dataset_formats(tmp_path):
    a = AnnData(np.arange(6.0).reshape(2, 3), pd.DataFrame({'g0': [1.0, 2.0]}, index=['a', 'b']),
                pd.DataFrame(index=['x', 'y', 'z']))
    p = write_dataset(a, str(tmp_path / 'm'), 'npz')
    X, obs, var = read_npz_tmap(p)
    assert np.array_equal(X, a.X) and list(obs.index) == ['a', 'b'] and list(var) == ['x', 'y', 'z']
    np.testing.assert_array_equal(obs['g0'].values, [1.0, 2.0])
    t = write_dataset(a, str(tmp_path / 'm'), 'txt')
    back = pd.read_csv(t, sep='\t', index_col='id')
    assert back.shape == (2, 3) and list(back.columns) == ['x', 'y', 'z']
    with pytest.raises(ValueError):
        write_dataset(a, str(tmp_path / 'm'), 'xyz')
    assert check_file_extension('a/b', 'h5ad') == 'a/b.h5ad'
    assert check_file_extension('a/b.H5AD', 'h5ad') == 'a/b.H5AD'


def test_anndata_standin_selection():
    a = AnnData(np.arange(12.0).reshape(4, 3), pd.DataFrame({'day': [0, 0, 1, 1]}, index=list('abcd')),
                pd.DataFrame(index=list('xyz')))
    b = a[np.array([False, True, True, False]), :]
    assert b.shape == (2, 3) and list(b.obs_names) == ['b', 'c']
    assert np.array_equal(b.X, a.X[1:3])
