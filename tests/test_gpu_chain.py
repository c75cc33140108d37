"""FP32 MFMA GEMM chain (long-range couplings) against exact integer products
and the FP64 oracle chain."""
import numpy as np
import pandas as pd
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def tm():
    import torch
    assert torch.cuda.is_available()
    import wot_amd.tmap as tm
    return tm


def test_gemm_integer_exact(tm):
    """Small integers are exact in FP32: pins the MFMA operand/accumulator layout."""
    rng = np.random.default_rng(0)
    for (m, k, n) in [(4, 2, 5), (33, 17, 65), (128, 16, 128), (129, 31, 257), (300, 5, 7)]:
        A = rng.integers(-4, 5, (m, k)).astype(np.float32)
        B = rng.integers(-4, 5, (k, n)).astype(np.float32)
        C = tm.gemm_f32(A, B).cpu().numpy()
        assert np.array_equal(C, A.astype(np.int64) @ B.astype(np.int64)), (m, k, n)


def test_chain_integer_fixture_shapes(tm):
    """The shapes of the reference's tests/inputs/transport_maps (4x2, 2x5, 5x6, 6x3)."""
    rng = np.random.default_rng(2)
    maps = [rng.integers(0, 9, s).astype(np.float64) for s in [(4, 2), (2, 5), (5, 6), (6, 3)]]
    got = tm.chain_product([m.astype(np.float32) for m in maps]).cpu().numpy()
    assert np.array_equal(got, O.chain(maps))


def test_chain_of_plans_fp32_vs_fp64(tm):
    rng = np.random.default_rng(3)
    sizes = [700, 900, 650, 800, 1000, 750]
    maps = [rng.random((sizes[i], sizes[i + 1])) ** 8 / sizes[i + 1] for i in range(5)]
    ref = O.chain(maps)
    got = tm.chain_product([m.astype(np.float32) for m in maps]).cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    rn = got / got.sum(1, keepdims=True)
    rr = ref / ref.sum(1, keepdims=True)
    row_err = np.abs(rn - rr).max() / np.abs(rr).max()
    print('chain rel frob %.3e row-normalised max %.3e' % (rel, row_err))
    assert rel < 1e-5 and row_err < 1e-5


def test_glue_reorders_intermediate_cells(tm):
    from wot_amd._anndata import AnnData
    rng = np.random.default_rng(4)
    x0 = rng.random((3, 4))
    x1 = rng.random((4, 2))
    t0 = AnnData(x0, pd.DataFrame(index=['a', 'b', 'c']), pd.DataFrame(index=['p', 'q', 'r', 's']))
    perm = [2, 0, 3, 1]
    t1 = AnnData(x1[perm], pd.DataFrame(index=[['p', 'q', 'r', 's'][i] for i in perm]),
                 pd.DataFrame(index=['u', 'v']))
    out = tm.glue_transport_maps(t0, t1)
    np.testing.assert_allclose(out.X, x0 @ x1, rtol=1e-14)  # FP64, as the reference's np.dot
    assert list(out.obs.index) == ['a', 'b', 'c'] and list(out.var.index) == ['u', 'v']
    out32 = tm.glue_transport_maps(t0, t1, precision='float32')
    np.testing.assert_allclose(out32.X, x0 @ x1, rtol=1e-6)
    with pytest.raises(ValueError):
        tm.glue_transport_maps(t0, t1, precision='float16')
