"""Measures over the cells of one day (wot/population.py:6-69) and the
time-point helpers of wot/tmap (util.py:52-72, chaining.py:43-110).

The model accepts any object with ``time``, ``p`` and ``name`` attributes, so
the reference's own Population instances work as well as this one."""
import numpy as np


class Population:
    """A measure ``p`` over the cells measured at ``time`` (population.py:6-23)."""

    def __init__(self, time, p, name=None):
        self.time = time
        self.p = np.asarray(p, dtype=np.float64)
        self.name = name

    def normalize(self):
        """Divide by the total mass (population.py:25-29)."""
        self.p = self.p / self.p.sum()

    def make_binary(self):
        """1 where the measure is positive, else 0 (population.py:31-37)."""
        self.p = (self.p > 0.0).astype(np.float64)

    @staticmethod
    def get_missing_population(*populations):
        """'Other': the cells no population covers, or None (population.py:39-47)."""
        covered = np.zeros(len(populations[0].p))
        for pop in populations:
            covered = covered + pop.p
        missing = covered == 0
        if not missing.any():
            return None
        return Population(populations[0].time, missing.astype(np.float64), 'Other')

    @staticmethod
    def copy(*populations, normalize=None, add_missing=False):
        """Copies, plus 'Other' when ``add_missing``; normalize True -> sum 1,
        False -> binary, None -> unchanged (population.py:49-69)."""
        out = [Population(pop.time, pop.p, pop.name) for pop in populations]
        if add_missing:
            other = Population.get_missing_population(*out)
            if other is not None:
                out.append(other)
        if normalize is not None:
            for pop in out:
                if normalize:
                    pop.normalize()
                else:
                    pop.make_binary()
        return out


def unique_timepoint(*populations):
    """The one day all populations share (wot/tmap/util.py:52-72)."""
    times = {pop.time for pop in populations}
    if len(times) > 1:
        raise ValueError("Several populations were given, but they are not from the same day")
    if not times:
        raise ValueError("No cells found at the given day")
    return times.pop()


def find_path(t0, t1, available_pairs, timepoints):
    """Path of day pairs from t0 to t1 with the most steps (wot/tmap/chaining.py:43-110).

    With ``available_pairs`` None: every consecutive pair between t0 and t1.
    Otherwise the longest chain of available pairs (pairs go forward in time,
    so ascending days order the graph); a day's predecessor changes only for
    a strictly longer chain, so among equally long chains the one reaching
    each day first in ascending order wins, as in the reference.  Only days
    reachable from t0 take part (the reference lets unreachable days seed a
    chain and then fails with a KeyError; here that case is a ValueError).
    """
    timepoints = list(timepoints)
    if t0 not in timepoints or t1 not in timepoints:
        raise ValueError("Unable to build path : t0 and t1 must be present in timepoints list")
    i0, i1 = timepoints.index(t0), timepoints.index(t1)
    if available_pairs is None:
        return [(timepoints[i], timepoints[i + 1]) for i in range(i0, i1)]
    succ = {t: [] for t in timepoints}
    for a, b in available_pairs:
        succ[a].append(b)
    best = {t0: (0, None)}  # day -> (steps from t0, previous day)
    for t in timepoints[i0:]:
        if t not in best:
            continue
        d = best[t][0]
        for u in sorted(succ[t]):
            if u not in best or d >= best[u][0]:
                best[u] = (d + 1, t)
    if t1 not in best or best[t1][0] == 0:
        raise ValueError("{} is not reachable from {} with the chosen day_pairs".format(t1, t0))
    path = []
    cur = t1
    while cur != t0:
        prev = best[cur][1]
        path.insert(0, (prev, cur))
        cur = prev
    return path
