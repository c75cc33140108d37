"""The C-ABI library loads and exports every symbol include/wotk.h declares
(no compute calls: this runs without a GPU)."""
import os
import re

import pytest

from wot_amd import _native

HEADER = os.path.join(os.path.dirname(__file__), '..', 'include', 'wotk.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(wotk_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ('wotk_transport_stablev2', 'wotk_transport_duality_gap', 'wotk_cost_sqeuclid',
              'wotk_median_f64', 'wotk_last_error'):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_native.SIGNATURES)


def test_abi_version_and_sizes():
    lib = _native.lib()
    assert lib.wotk_abi_version() == 1
    # workspace queries are pure host arithmetic
    assert lib.wotk_solver_workspace_size(2000, 2500) > 16 * (2000 + 2500)
    assert lib.wotk_cost_workspace_size(100, 200, 30, 0) >= 8 * 100 * 200
    assert lib.wotk_median_workspace_size(10) > 0


def test_errors_map_to_reference_exceptions():
    lib = _native.lib()
    # an invalid shape is rejected on the host before any device work
    rc = lib.wotk_transport_stablev2(None, 4, None, 4, 0, 5, None, 1.0, 50.0, 0.05, 10, 1e4, 1.0, 0, 50,
                                     None, 5, None, None, None, None, 0, None, None)
    assert rc == _native.WOTK_EINVAL
    with pytest.raises(ValueError):
        _native.check(rc)
    assert 'bad shape' in lib.wotk_last_error().decode()


def test_new_entry_points_reject_bad_arguments_on_the_host():
    """Argument checks of the map-walk, PCA, interpolation and EMD entry points
    run before any device work (no GPU needed)."""
    import ctypes
    lib = _native.lib()
    T = _native.TMap
    # a chain whose second map does not take the first one's output cells
    maps = (T * 2)(T(8, _native.DTYPE_F64, 0, 5, 4, 5, None, 0), T(8, _native.DTYPE_F64, 0, 3, 7, 3, None, 0))
    assert lib.wotk_tmap_workspace_size(maps, 2, _native.PUSH_FORWARD, 3) == 0
    outs = (ctypes.c_void_p * 2)(8, 8)
    rc = lib.wotk_tmap_propagate(maps, 2, _native.PUSH_FORWARD, 8, 3, 3, 1, outs, 3, 8, 1 << 20, None, None)
    assert rc == _native.WOTK_EINVAL and 'takes 7 cells' in lib.wotk_last_error().decode()
    one = (T * 1)(T(8, 7, 0, 5, 4, 5, None, 0))  # unknown dtype
    rc = lib.wotk_tmap_propagate(one, 1, _native.PUSH_FORWARD, 8, 3, 3, 1, outs, 3, 8, 1 << 20, None, None)
    assert rc == _native.WOTK_EINVAL and 'dtype' in lib.wotk_last_error().decode()
    one = (T * 1)(T(8, _native.DTYPE_F64, 0, 5, 4, 5, 8, 2))  # transposed copy with ld < rows
    assert lib.wotk_tmap_propagate(one, 1, _native.PULL_BACK, 8, 3, 3, 1, outs, 3, 8, 1 << 20, None,
                                   None) == _native.WOTK_EINVAL
    good = (T * 1)(T(8, _native.DTYPE_F64, 0, 5, 4, 5, None, 0))
    assert lib.wotk_tmap_workspace_size(good, 1, _native.PUSH_FORWARD, 20) > 0
    assert lib.wotk_tmap_transpose(8, 3, 4, 2, 2, 8, 2, None) == _native.WOTK_EINVAL
    assert lib.wotk_pca_center(8, 0, 2, 4, 3, None, 0, 8, 8, 8, 1 << 20, None) == _native.WOTK_EINVAL  # ldx < g
    assert lib.wotk_pca_center_workspace_size(100, 50) > 0
    assert lib.wotk_interp_weights(8, 2, 4, 3, 0.5, 8, 8, None) == _native.WOTK_EINVAL  # ld < J
    assert lib.wotk_interp_gather(8, 1, 8, 1, 4, 8, 3, 3, 3, 0.5, 8, 3, None) == _native.WOTK_EINVAL  # ld0 < d
    assert lib.wotk_rows_normalize_f64(8, 3, 4, 2, None) == _native.WOTK_EINVAL  # ld < k
    few = (ctypes.c_double * 2)()
    assert lib.wotk_probe_issue_mix(few, 2, 8, 1 << 20, None) == _native.WOTK_EINVAL  # needs 4 rates
    assert lib.wotk_probe_valu_rates(few, 2, 8, 1 << 20, None) == _native.WOTK_EINVAL  # needs 7 rates
    # adjoint: shape checks before any device work; workspace grows with the column splits
    assert lib.wotk_ot_adjoint(8, 0, 8, 0, 4, 4, 8, 8, 2, 8, 1.0, 8, 8, 1.0, 50.0, 0.05, 8, None, None, None, 8,
                               1 << 20, None) == _native.WOTK_EINVAL  # ldc < J
    # 100 x 5000, d = 30: 20 column splits of 256 columns, each with its own 100 x 30 FP64 partial V
    assert lib.wotk_adjoint_workspace_size(100, 5000, 30) >= 20 * 100 * 30 * 8


def test_emd_is_host_code():
    """The exact EMD takes host pointers and needs no device."""
    import numpy as np
    from wot_amd.ot.emd import emd2
    M = np.array([[0.0, 2.0], [1.0, 0.0]])
    assert emd2(np.array([0.5, 0.5]), np.array([0.5, 0.5]), M) == 0.0
    assert emd2(np.array([1.0, 0.0]), np.array([0.0, 1.0]), M) == 2.0
