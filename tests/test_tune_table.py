"""Host side of the tile table and the autotuner's cache (no GPU needed)."""
import json

import pytest

from wot_amd import _native, tune


@pytest.fixture(autouse=True)
def clean_table():
    tune.reset()
    yield
    tune.reset()


def test_variants_default_first_and_distinct():
    for k in (tune.ROW_SWEEP, tune.BATCH_SWEEP, tune.BATCH_SHARED):
        v = tune.variants(k)
        assert len(v) == len(set(v)) >= 4
        assert all(r >= 1 and t % 64 == 0 for r, t in v)
    assert tune.variants(tune.ROW_SWEEP)[0] == (4, 256)
    assert tune.variants(tune.BATCH_SWEEP)[0] == (2, 256)


def test_bucket_arithmetic():
    lg = lambda n: max(0, int(n).bit_length() - 1)
    for rows, cols in [(1, 1), (15000, 15000), (333, 517), (2 ** 31, 7)]:
        assert tune.bucket(tune.ROW_SWEEP, rows, cols) == min(lg(rows), 31) * 32 + min(lg(cols), 31)
    assert tune.bucket(tune.BATCH_SWEEP, 0, 10000, 9) == 8 * 32 + 13
    assert tune.bucket(tune.BATCH_SWEEP, 0, 10000, 99) == 15 * 32 + 13   # combinations clamp to 16
    L = _native.lib()
    assert L.wotk_num_buckets(tune.ROW_SWEEP) == 1024 and L.wotk_num_buckets(tune.BATCH_SWEEP) == 512
    assert L.wotk_num_buckets(tune.BATCH_SHARED) == 512
    assert tune.bucket(tune.BATCH_SHARED, 0, 10000, 9) == tune.bucket(tune.BATCH_SWEEP, 0, 10000, 9)
    assert L.wotk_shape_bucket(7, 1, 1, 1) == -1


def test_set_get_reset():
    b = tune.bucket(tune.ROW_SWEEP, 15000, 15000)
    tune.set_tile(tune.ROW_SWEEP, b, (2, 512))
    assert tune.get_tile(tune.ROW_SWEEP, b) == (2, 512)
    assert tune.get_tile(tune.ROW_SWEEP, b + 1) == (4, 256)
    tune.set_tile(tune.BATCH_SWEEP, -1, (1, 512))
    assert all(tune.get_tile(tune.BATCH_SWEEP, k) == (1, 512) for k in (0, 100, 511))
    tune.set_tile(tune.ROW_SWEEP, b, None)
    assert tune.get_tile(tune.ROW_SWEEP, b) == (4, 256)


def test_rejects_unknown_tiles_and_buckets():
    with pytest.raises(ValueError, match='not a compiled variant'):
        tune.set_tile(tune.ROW_SWEEP, 0, (3, 256))
    with pytest.raises(ValueError, match='out of range'):
        tune.set_tile(tune.BATCH_SWEEP, 512, (2, 256))
    with pytest.raises(ValueError, match='unknown kernel'):
        tune.get_tile(5, 0)


def test_cache_loads_only_this_device_and_valid_tiles(tmp_path, monkeypatch):
    path = tmp_path / 'tiles.json'
    monkeypatch.setenv('WOTK_TILE_CACHE', str(path))
    monkeypatch.setattr(tune, 'device_key', lambda: 'devA/gfx950/abi1')
    b = tune.bucket(tune.ROW_SWEEP, 15000, 15000)
    entries = {
        'devA/gfx950/abi1|row_sweep|%d' % b: {'rows': 8, 'threads': 512},
        'devA/gfx950/abi1|row_sweep|%d' % (b + 1): {'rows': 3, 'threads': 256},   # not compiled: ignored
        'devB/gfx942/abi1|row_sweep|%d' % (b + 2): {'rows': 1, 'threads': 256},   # other device: ignored
        'devA/gfx950/abi1|batch_sweep|7': {'rows': 4, 'threads': 256},
    }
    path.write_text(json.dumps({'entries': entries}))
    assert tune.apply_persisted(force=True) == 2
    assert tune.get_tile(tune.ROW_SWEEP, b) == (8, 512)
    assert tune.get_tile(tune.ROW_SWEEP, b + 1) == (4, 256)
    assert tune.get_tile(tune.ROW_SWEEP, b + 2) == (4, 256)
    assert tune.get_tile(tune.BATCH_SWEEP, 7) == (4, 256)
    # persisting merges into the file
    tune._persist(tune.BATCH_SWEEP, [9], (1, 256), {(1, 256): 1.5, (2, 256): 2.0})
    data = json.loads(path.read_text())
    assert data['entries']['devA/gfx950/abi1|batch_sweep|9']['rows'] == 1
    assert len(data['entries']) == 5


def test_cache_off(monkeypatch):
    monkeypatch.setenv('WOTK_TILE_CACHE', 'off')
    assert tune.cache_path() is None
    assert tune.apply_persisted(force=True) == 0
