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
