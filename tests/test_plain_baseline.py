"""The plain-GPU baseline (wot_amd/plain_gpu.py, the reference solvers restated
in torch FP64) must compute what the reference computes: run on the CPU
device against the golden vectors written from the reference solver file, so
the speedups measured against it compare like with like."""
import glob
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from wot_amd import plain_gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
SV2 = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, 'sv2_*.npz')))
GAP = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, 'gap_*.npz')))


def _load(name):
    z = np.load(os.path.join(GOLD, name + '.npz'), allow_pickle=False)
    return {k: z[k] for k in z.files}


def _check(R, want):
    R = R.numpy()
    scale = np.abs(want).max()
    np.testing.assert_allclose(R, want, rtol=1e-8, atol=1e-13 * scale)


@pytest.mark.parametrize('name', SV2 + ['diag3_stablev2'])
def test_plain_stablev2_matches_reference_vectors(name):
    z = _load(name)
    R = plain_gpu.transport_stablev2(C=z['C'], G=z['G'], device='cpu', **json.loads(str(z['params'])))
    _check(R, z['R'])


@pytest.mark.parametrize('name', GAP + ['diag3_gap'])
def test_plain_duality_gap_matches_reference_vectors(name):
    z = _load(name)
    st = {}
    with np.errstate(all='ignore'):
        R = plain_gpu.optimal_transport_duality_gap(C=z['C'], G=z['G'], device='cpu', stats=st,
                                                    **json.loads(str(z['params'])))
    _check(R, z['R'])
    assert st['iterations'] > 0
