"""FP64 CPU oracle for populations through transport maps.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``wot_amd``) imports this
module; ``tests/`` and ``bench.py``'s ``cpu_baseline`` leg use it as the
checker / CPU baseline.

Plain numpy restatement of wot's TransportMapModel walks over a list of maps
(``maps[t]`` = the map from day t to day t+1, rows = cells of day t).
Populations are in the reference's layout: one row per population
(``np.vstack([pop.p ...])``).

* ``push_forward``      wot/tmap/transport_map_model.py:235-299
* ``pull_back``         wot/tmap/transport_map_model.py:301-365
* ``binary_with_other`` wot/population.py:31-69 (copy(normalize=False, add_missing=True))
* ``fates``             wot/tmap/transport_map_model.py:40-69
* ``transition_table``  wot/tmap/transport_map_model.py:71-102
* ``trajectories``      wot/tmap/transport_map_model.py:105-142
* ``census``            wot/tmap/transport_map_model.py:576-615 (all cells in the set matrix)

Pinning (tests/test_oracle_tmap.py): hand-computed values on the reference's
integer fixture maps (tests/golden/transport_maps, copied from the reference's
tests/inputs/transport_maps), brute-force loops, the adjoint identity
<pull(q), p> = <q, push(p)>, linearity of the normalised walk, identity maps
(transition table = set overlaps), and row sums of fates.
"""
import numpy as np


def _normalize_rows(P):
    # (p.T / np.sum(p, axis=1)).T  (transport_map_model.py:291-292, 357-358)
    return P / P.sum(axis=1, keepdims=True)


def push_forward(maps, P, i, j, normalize=True):
    """Populations P (k x cells of day i) pushed to day j >= i (:284-293)."""
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    for t in range(i, j):
        P = P @ np.asarray(maps[t], dtype=np.float64)
        if normalize:
            P = _normalize_rows(P)
    return P


def pull_back(maps, P, i, j, normalize=True):
    """Populations P (k x cells of day i) pulled back to day j <= i (:350-359)."""
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    for t in range(i, j, -1):
        P = (np.asarray(maps[t - 1], dtype=np.float64) @ P.T).T
        if normalize:
            P = _normalize_rows(P)
    return P


def binary_with_other(P):
    """Binary copies of the populations plus an 'Other' row for uncovered
    cells when there are any (population.py:39-69).  Returns (P, has_other)."""
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    missing = P.sum(axis=0) == 0
    B = (P > 0).astype(np.float64)
    if missing.any():
        return np.vstack([B, missing.astype(np.float64)]), True
    return B, False


def fates(maps, P, i):
    """(cells of days 0..i) x populations (:55-66)."""
    B, _ = binary_with_other(P)
    blocks = [B.T]
    cur = B
    for t in range(i, 0, -1):
        cur = pull_back(maps, cur, t, t - 1, normalize=False)
        blocks.insert(0, cur.T)
    X = np.concatenate(blocks)
    return X / X.sum(axis=1, keepdims=True)


def transition_table(maps, P_start, i_start, P_end, i_end):
    """start populations x end populations, total 1 (:86-100)."""
    S, _ = binary_with_other(P_start)
    E, _ = binary_with_other(P_end)
    t = i_end
    while t > 0 and t > i_start:
        E = pull_back(maps, E, t, t - 1, normalize=False)
        t -= 1
    p = S @ E.T
    return p / p.sum()


def trajectories(maps, P, i):
    """(cells of every day) x populations (:118-142)."""
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    P = _normalize_rows(P)
    blocks = [P.T]
    cur = P
    for t in range(i, 0, -1):
        cur = pull_back(maps, cur, t, t - 1)
        blocks.insert(0, cur.T)
    cur = P
    for t in range(i, len(maps)):
        cur = push_forward(maps, cur, t, t + 1)
        blocks.append(cur.T)
    return np.concatenate(blocks)


def census(P, cell_sets):
    """Share of each population's mass in each cell set (:605-613), for a set
    matrix covering every cell of the day in the same order."""
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    out = []
    for p in P:
        s = p.sum()
        q = p if np.isclose(s, 0) else p / s
        out.append(q @ cell_sets)
    return np.asarray(out)
