"""Host logic of OTModel.compute_all_transport_maps (no device needed: the
per-pair solve is stubbed): every pair's file is written, by the background
writer, with the stubbed map's content and ids; a write error surfaces from
the call; existing files are skipped when overwrite is off."""
import os

import numpy as np
import pandas as pd
import pytest

pytest.importorskip('torch')


def _model(tmp_path, monkeypatch, days=4, n=5):
    import wot_amd.ot as ot
    from wot_amd import _anndata
    rng = np.random.default_rng(0)
    X = rng.standard_normal((days * n, 3))
    obs = pd.DataFrame({'day': np.repeat(np.arange(days, dtype=float), n)},
                       index=['c%d' % i for i in range(days * n)])
    model = ot.OTModel(_anndata.AnnData(X, obs, pd.DataFrame(index=['g0', 'g1', 'g2'])), local_pca=0)

    def fake_map(t0, t1, covariate=None, cost_matrix=None):
        a = obs.index[obs['day'] == t0]
        b = obs.index[obs['day'] == t1]
        T = np.full((len(a), len(b)), t0 + 10 * t1)
        return _anndata.AnnData(T, pd.DataFrame(index=a), pd.DataFrame(index=b))
    monkeypatch.setattr(model, 'compute_transport_map', fake_map)
    return model


def test_every_pair_written_in_background(tmp_path, monkeypatch):
    from wot_amd.io import read_npz_tmap
    model = _model(tmp_path, monkeypatch)
    out = str(tmp_path / 'tm')
    model.compute_all_transport_maps(tmap_out=out, output_file_format='npz')
    for t in range(3):
        X, obs, var = read_npz_tmap(out + '_{}_{}.npz'.format(float(t), float(t + 1)))  # the reference's names
        assert X.shape == (5, 5) and np.all(X == t + 10 * (t + 1))
        assert list(obs.index) == ['c%d' % i for i in range(5 * t, 5 * t + 5)]


def test_write_error_surfaces(tmp_path, monkeypatch):
    import wot_amd.io as io
    model = _model(tmp_path, monkeypatch)

    def failing(ds, path, output_format='txt'):
        raise OSError('disk full')
    monkeypatch.setattr(io, 'write_dataset', failing)
    with pytest.raises(OSError, match='disk full'):
        model.compute_all_transport_maps(tmap_out=str(tmp_path / 'tm'), output_file_format='npz')


def test_existing_files_skipped(tmp_path, monkeypatch):
    model = _model(tmp_path, monkeypatch)
    out = str(tmp_path / 'tm')
    open(out + '_0.0_1.0.npz', 'w').close()
    model.compute_all_transport_maps(tmap_out=out, output_file_format='npz', overwrite=False)
    assert os.path.getsize(out + '_0.0_1.0.npz') == 0
    assert os.path.exists(out + '_1.0_2.0.npz') and os.path.exists(out + '_2.0_3.0.npz')
