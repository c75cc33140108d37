"""Pins the transport-map oracle (oracle/tmap_oracle.py) and the host logic
of wot_amd.tmap that runs without a GPU (populations, time points, paths,
argument errors)."""
import os

import numpy as np
import pandas as pd
import pytest

from oracle import tmap_oracle as T
from wot_amd.tmap.population import Population, find_path, unique_timepoint

FIX = os.path.join(os.path.dirname(__file__), 'golden', 'transport_maps')


def fixture_maps():
    """The reference's integer maps t0->t1 (4x2), t1->t2 (2x5), t2->t3 (5x6), t3->t4 (6x3)."""
    return [pd.read_csv(os.path.join(FIX, 't%d_t%d.csv' % (d, d + 1)), index_col='id').values.astype(np.float64)
            for d in range(4)]


def test_fixture_push_forward_by_hand():
    maps = fixture_maps()
    p = np.array([[1.0, 0, 0, 0]])  # cell c1-t0
    one = T.push_forward(maps, p, 0, 1)
    np.testing.assert_allclose(one, [[2 / 5, 3 / 5]], rtol=0, atol=1e-15)
    # [0.4, 0.6] @ [[2..6], [7..11]] = [5, 6, 7, 8, 9], normalised by 35
    two = T.push_forward(maps, p, 0, 2)
    np.testing.assert_allclose(two, [[5 / 35, 6 / 35, 7 / 35, 8 / 35, 9 / 35]], rtol=0, atol=1e-15)
    raw = T.push_forward(maps, p, 0, 2, normalize=False)
    np.testing.assert_array_equal(raw, [[25.0, 30.0, 35.0, 40.0, 45.0]])


def test_fixture_pull_back_by_hand():
    maps = fixture_maps()
    q = np.array([[1.0, 0]])  # cell c1-t1: column 0 of t0->t1 = [2, 4, 6, 8]
    np.testing.assert_allclose(T.pull_back(maps, q, 1, 0), [[0.1, 0.2, 0.3, 0.4]], rtol=0, atol=1e-15)
    # c4-t2 pulled to day 0: t1_t2[:, 3] = [5, 10]; t0_t1 @ [5, 10] = [40, 70, 100, 130]
    q = np.array([[0, 0, 0, 1.0, 0]])
    np.testing.assert_array_equal(T.pull_back(maps, q, 2, 0, normalize=False), [[40.0, 70.0, 100.0, 130.0]])


def brute_push(Tm, p):
    out = np.zeros(Tm.shape[1])
    for j in range(Tm.shape[1]):
        for i in range(Tm.shape[0]):
            out[j] += p[i] * Tm[i, j]
    return out


def brute_pull(Tm, q):
    out = np.zeros(Tm.shape[0])
    for i in range(Tm.shape[0]):
        for j in range(Tm.shape[1]):
            out[i] += Tm[i, j] * q[j]
    return out


def test_brute_force_and_adjoint():
    rng = np.random.default_rng(4)
    sizes = [5, 7, 3, 6]
    maps = [rng.random((sizes[t], sizes[t + 1])) for t in range(3)]
    p = rng.random((2, 5))
    q = rng.random((2, 6))
    got = T.push_forward(maps, p, 0, 3, normalize=False)
    for k in range(2):
        ref = p[k]
        for m in maps:
            ref = brute_push(m, ref)
        np.testing.assert_allclose(got[k], ref, rtol=1e-13)
    back = T.pull_back(maps, q, 3, 0, normalize=False)
    for k in range(2):
        ref = q[k]
        for m in maps[::-1]:
            ref = brute_pull(m, ref)
        np.testing.assert_allclose(back[k], ref, rtol=1e-13)
    # <pull(q), p> == <q, push(p)>: a transposed map in either walk breaks this
    for k in range(2):
        assert np.dot(back[k], p[k]) == pytest.approx(np.dot(q[k], got[k]), rel=1e-13)


def test_normalised_walk_is_the_scaled_raw_walk():
    rng = np.random.default_rng(5)
    maps = [rng.random((40, 50)), rng.random((50, 30)), rng.random((30, 45))]
    p = rng.random((3, 40))
    raw = T.push_forward(maps, p, 0, 3, normalize=False)
    np.testing.assert_allclose(T.push_forward(maps, p, 0, 3), raw / raw.sum(1, keepdims=True), rtol=1e-13)
    q = rng.random((3, 45))
    raw = T.pull_back(maps, q, 3, 0, normalize=False)
    np.testing.assert_allclose(T.pull_back(maps, q, 3, 0), raw / raw.sum(1, keepdims=True), rtol=1e-13)


def test_binary_with_other():
    P = np.array([[0.5, 0, 0, 0.1], [0, 2.0, 0, 0]])
    B, other = T.binary_with_other(P)
    assert other
    np.testing.assert_array_equal(B, [[1, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]])
    B, other = T.binary_with_other(np.array([[1.0, 1.0]]))
    assert not other and B.shape == (1, 2)


def test_fates_rows_and_last_day():
    rng = np.random.default_rng(6)
    maps = [rng.random((4, 6)), rng.random((6, 5))]
    P = np.array([[1.0, 1, 0, 0, 0], [0, 0, 1, 0, 0]])  # 'Other' = cells 3, 4
    X = T.fates(maps, P, 2)
    assert X.shape == (4 + 6 + 5, 3)
    np.testing.assert_allclose(X.sum(1), 1.0, rtol=1e-14)
    np.testing.assert_array_equal(X[10:], [[1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1]])
    # day-1 cell i: mass into each set / total mass
    M = maps[1]
    sets = np.array([[1, 1, 0, 0, 0], [0, 0, 1, 0, 0], [0, 0, 0, 1, 1]], float)
    np.testing.assert_allclose(X[4:10], (M @ sets.T) / M.sum(1, keepdims=True), rtol=1e-13)


def test_transition_table_identity_maps_count_overlaps():
    n = 8
    maps = [np.eye(n), np.eye(n)]
    start = np.array([[1.0, 1, 1, 0, 0, 0, 0, 0], [0, 0, 0, 1, 1, 0, 0, 0]])
    end = np.array([[1.0, 0, 0, 1, 0, 0, 0, 0], [0, 1, 1, 0, 1, 1, 0, 0]])
    tt = T.transition_table(maps, start, 0, end, 2)
    # rows: start sets + Other (cells 5-7); columns: end sets + Other (cells 6, 7)
    S = np.vstack([start, [0, 0, 0, 0, 0, 1, 1, 1]])
    E = np.vstack([end, [0, 0, 0, 0, 0, 0, 1, 1]])
    overlap = S @ E.T
    np.testing.assert_allclose(tt, overlap / overlap.sum(), rtol=1e-15)
    assert tt.sum() == pytest.approx(1.0)


def test_trajectories_compose_push_and_pull():
    rng = np.random.default_rng(7)
    sizes = [5, 6, 4, 7]
    maps = [rng.random((sizes[t], sizes[t + 1])) for t in range(3)]
    P = rng.random((2, 4))
    X = T.trajectories(maps, P, 2)
    assert X.shape == (sum(sizes), 2)
    Pn = P / P.sum(1, keepdims=True)
    np.testing.assert_allclose(X[11:15].T, Pn, rtol=1e-15)
    np.testing.assert_allclose(X[0:5].T, T.pull_back(maps, Pn, 2, 0), rtol=1e-13)
    np.testing.assert_allclose(X[15:].T, T.push_forward(maps, Pn, 2, 3), rtol=1e-13)
    np.testing.assert_allclose(X.sum(0), 4.0, rtol=1e-13)  # each day's block sums to 1


def test_census_by_hand():
    sets = np.array([[1, 0], [1, 1], [0, 1]], float)
    c = T.census(np.array([[1.0, 1.0, 2.0], [0, 0, 0]]), sets)
    np.testing.assert_allclose(c, [[0.5, 0.75], [0, 0]])


# ---------------------------------------------------------------------------
# host logic of wot_amd.tmap (no device work)
# ---------------------------------------------------------------------------
def test_population_copy_semantics():
    a = Population(1, [0.0, 2.0, 0.0, 1.0], 'a')
    b = Population(1, [0.0, 0.0, 0.0, 3.0], 'b')
    out = Population.copy(a, b, normalize=False, add_missing=True)
    assert [p.name for p in out] == ['a', 'b', 'Other']
    np.testing.assert_array_equal(out[0].p, [0, 1, 0, 1])
    np.testing.assert_array_equal(out[2].p, [1, 0, 1, 0])
    out = Population.copy(a, normalize=True)
    np.testing.assert_allclose(out[0].p, [0, 2 / 3, 0, 1 / 3])
    np.testing.assert_array_equal(a.p, [0, 2, 0, 1])  # the input is untouched
    assert Population.get_missing_population(Population(0, [1.0, 1.0])) is None


def test_unique_timepoint_errors():
    assert unique_timepoint(Population(2, [1]), Population(2, [0])) == 2
    with pytest.raises(ValueError, match='not from the same day'):
        unique_timepoint(Population(1, [1]), Population(2, [1]))
    with pytest.raises(ValueError, match='No cells found'):
        unique_timepoint()


def test_find_path():
    tp = [0, 1, 2, 3]
    assert find_path(0, 3, None, tp) == [(0, 1), (1, 2), (2, 3)]
    # the longest chain of available pairs wins: 0->1->2->3 over 0->2->3 or 0->3
    assert find_path(0, 3, [(0, 1), (1, 2), (2, 3), (0, 2), (0, 3)], tp) == [(0, 1), (1, 2), (2, 3)]
    assert find_path(0, 3, [(0, 2), (2, 3), (0, 3)], tp) == [(0, 2), (2, 3)]
    # equally long: a day keeps the predecessor found first (ascending days)
    assert find_path(0, 3, [(0, 1), (1, 3), (0, 2), (2, 3)], tp) == [(0, 1), (1, 3)]
    with pytest.raises(ValueError, match='not reachable'):
        find_path(0, 3, [(0, 1), (2, 3)], tp)
    with pytest.raises(ValueError, match='must be present'):
        find_path(0, 9, None, tp)


def test_model_argument_errors_before_device_work():
    from wot_amd.tmap import TransportMapModel
    meta = pd.DataFrame({'day': [0.0, 0, 1, 1, 2]}, index=['a', 'b', 'c', 'd', 'e'])
    m = TransportMapModel({}, meta)
    assert m.timepoints == [0.0, 1.0, 2.0]
    assert m.day_pairs == [(0.0, 1.0), (1.0, 2.0)]
    last = Population(2.0, [1.0])
    first = Population(0.0, [1.0, 0.0])
    with pytest.raises(ValueError, match='No further timepoints'):
        m.push_forward(last)
    with pytest.raises(ValueError, match='before source'):
        m.push_forward(Population(1.0, [1.0, 1.0]), to_time=0.0)
    with pytest.raises(ValueError, match='No previous timepoints'):
        m.pull_back(first)
    with pytest.raises(ValueError, match='after source'):
        m.pull_back(Population(1.0, [1.0, 1.0]), to_time=2.0)
    with pytest.raises(ValueError):
        m.push_forward(Population(7.0, [1.0]))  # day not in the time points
    assert m.can_push_forward(first) and not m.can_pull_back(first)
    assert not m.can_push_forward(last) and m.can_pull_back(last)
    pops = m.population_from_ids(['a', 'zz'], ['zz'], at_time=0)
    np.testing.assert_array_equal(pops[0].p, [1, 0])
    assert pops[1] is None
    sets = m.population_from_cell_sets({'s1': ['d'], 's2': ['q']}, at_time=1)
    assert len(sets) == 1 and sets[0].name == 's1' and list(sets[0].p) == [0, 1]
    with pytest.raises(ValueError, match='Day pair'):
        m.get_coupling(0.0, 5.0)
    with pytest.raises(ValueError, match='No transport map'):
        m.get_coupling(0.0, 1.0)
