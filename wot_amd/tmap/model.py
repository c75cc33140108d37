"""TransportMapModel with the maps resident in HBM
(wot/tmap/transport_map_model.py:12-615).

The reference keeps maps as host AnnData and multiplies populations through
them with numpy one map at a time.  Here every map is uploaded once (FP64 by
default, or FP32 to halve memory and traffic) and the multi-map walks of
push_forward / pull_back / fates / trajectories / transition_table /
ancestor_census run as single libwotk calls over the device-resident chain:
the populations cross the host boundary once on the way in and once on the
way out.  Method names, arguments, return types and error messages follow the
reference.
"""
import numpy as np
import pandas as pd

from .. import _anndata
from ..ot import device as _dev
from .population import Population, find_path, unique_timepoint
from .propagate import PULL_BACK, PUSH_FORWARD, DeviceMap, normalize_rows, propagate


class TransportMapModel:
    """Pre-computed transport maps over a time course (transport_map_model.py:12-38).

    tmaps : dict (t0, t1) [or (t0, t1, cv0, cv1)] -> AnnData-like map (X, obs,
        var), numpy array, torch tensor or DeviceMap
    meta : DataFrame indexed by cell id with a 'day' column, in the cell order
        of the maps
    dtype : 'float64' (default, the reference's precision) or 'float32' for the
        device copies of the maps
    transposed_maps : keep a transposed copy of every map in HBM (default), so
        that pull backs stream rows as coalesced as push forwards; False halves
        the memory and pulls read the maps' rows directly
    """

    def __init__(self, tmaps, meta, timepoints=None, day_pairs=None, cache=False, dtype='float64',
                 device='cuda', transposed_maps=True):
        self.tmaps = tmaps
        self.meta = meta
        self.cache = cache
        if timepoints is None:
            timepoints = sorted(meta['day'].unique())
        self.timepoints = list(timepoints)
        if day_pairs is None:
            day_pairs = [(self.timepoints[i], self.timepoints[i + 1]) for i in range(len(self.timepoints) - 1)]
        self.day_pairs = day_pairs
        self.dtype = dtype
        self.device = device
        self.transposed_maps = transposed_maps
        self._device_maps = {}

    # ------------------------------------------------------------------
    # maps
    # ------------------------------------------------------------------
    @classmethod
    def from_time_course(cls, result, days, meta, combo=0, dtype='float64', transposed_maps=True):
        """Model over the maps of a batched time-course solve
        (``wot_amd.ot.batch.solve_time_course``): every pair's map of
        combination ``combo`` is materialised on the device from its
        potentials and never leaves HBM.  ``days[p]`` is the day of the p-th
        coordinate block; ``meta`` lists the cells day by day in that order."""
        tmaps = {}
        for p in range(len(result.pairs)):
            t0, t1 = days[p], days[p + 1]
            n_i, n_j = result.shapes[p]
            dm = DeviceMap.empty(n_i, n_j, 'float64', obs_names=meta.index[meta['day'] == t0],
                                 var_names=meta.index[meta['day'] == t1])
            result.plan(p, combo, out=dm.X)
            if dtype == 'float32':
                f32 = DeviceMap.empty(n_i, n_j, 'float32', obs_names=dm.obs_names, var_names=dm.var_names)
                f32.X.copy_(dm.X)
                dm = f32
            tmaps[(t0, t1)] = dm
        return cls(tmaps, meta, timepoints=list(days), dtype=dtype, transposed_maps=transposed_maps)

    # ------------------------------------------------------------------
    # index files and map directories (transport_map_model.py:617-732)
    # ------------------------------------------------------------------
    def to_json(self, path):
        """Day pairs, map paths, time points and cell days as JSON (:617-628)."""
        import json
        meta = self.meta.to_dict(orient='list')
        meta['id'] = self.meta.index.values.tolist()
        paths = [self.tmaps[tuple(dp)] for dp in self.day_pairs]
        d = {'day_pairs': list(self.day_pairs), 'paths': paths, 'timepoints': list(self.timepoints), 'meta': meta}
        with open(path, 'w') as f:
            json.dump(d, f, ensure_ascii=False)

    @staticmethod
    def from_json(index_path, **kw):
        """Model from a to_json index (:630-649); maps stay paths until first use."""
        import json
        if str(index_path).startswith('gs://'):
            raise ValueError('gs:// index files need google-cloud-storage; download the index first')
        with open(index_path, 'r') as r:
            info = json.load(r)
        meta = pd.DataFrame(index=info['meta']['id'], data={'day': info['meta']['day']})
        day_pairs = info['day_pairs']
        tmaps = {tuple(day_pairs[i]): info['paths'][i] for i in range(len(day_pairs))}
        return TransportMapModel(tmaps=tmaps, meta=meta, timepoints=info['timepoints'], day_pairs=day_pairs, **kw)

    @staticmethod
    def from_directory(tmap_out, with_covariates=False, cache=False, **kw):
        """Model over the maps ``<prefix>_<t0>_<t1>.h5ad`` in a directory
        (:651-732): cell ids and days come from the maps' /obs and /var
        indices (the target ids of the last pair only), without reading X."""
        import os
        import re
        from .. import h5ad
        tmap_dir, tmap_prefix = os.path.split(tmap_out)
        tmap_dir = tmap_dir or '.'
        tmap_prefix = tmap_prefix or 'tmaps'
        day = r'([0-9]*\.?[0-9]+)'
        # the reference's suffix is a character class ('[\.h5ad|\.loom]'): one of those characters
        if with_covariates:
            pattern = re.compile(tmap_prefix + r'_{}_{}_cv([a-zA-Z0-9]+)_cv([a-zA-Z0-9]+)[\.h5ad|\.loom]'.format(day, day))
        else:
            pattern = re.compile(tmap_prefix + r'_{}_{}[\.h5ad|\.loom]'.format(day, day))
        tmaps = {}
        for f in os.listdir(tmap_dir):
            path = os.path.join(tmap_dir, f)
            if not os.path.isfile(path):
                continue
            m = pattern.match(f)
            if m is None:
                continue
            try:
                t1, t2 = float(m.group(1)), float(m.group(2))
            except ValueError:
                print('Unable to find day pair for ' + f)
                continue
            if with_covariates:
                tmaps[(t1, t2, m.group(3), m.group(4))] = path
            else:
                tmaps[(t1, t2)] = path
        if len(tmaps) == 0:
            raise ValueError('No transport maps found in ' + tmap_dir + ' with prefix ' + tmap_prefix)
        day_pairs, timepoints = set(), set()
        keys = sorted(tmaps.keys(), key=lambda x: x[0])
        parts = []
        for i, key in enumerate(keys):
            t0, t1 = key[0], key[1]
            day_pairs.add((t0, t1))
            timepoints.add(t0)
            timepoints.add(t1)
            if not with_covariates:
                path = tmaps[key]
                if path.endswith('.loom'):
                    raise ValueError('loom transport maps are not supported here: ' + path)
                parts.append(pd.DataFrame(index=h5ad.read_index(path, 'obs'), data={'day': t0}))
                if i == len(keys) - 1:
                    parts.append(pd.DataFrame(index=h5ad.read_index(path, 'var'), data={'day': t1}))
        meta = pd.concat(parts, copy=False) if parts else None
        return TransportMapModel(tmaps=tmaps, meta=meta, timepoints=sorted(timepoints), day_pairs=day_pairs,
                                 cache=cache, **kw)

    def device_map(self, t0, t1):
        """The map (t0, t1) on the device, uploaded on first use and kept."""
        key = (t0, t1)
        dm = self._device_maps.get(key)
        if dm is None:
            src = self.tmaps.get(key)
            if src is None:
                src = self.get_coupling(t0, t1)  # raises for a missing pair
            if isinstance(src, DeviceMap):
                dm = src
                if self.dtype == 'float32' and src.X.element_size() == 8:
                    dm = DeviceMap.upload(src.X, 'float32', self.device)
                    dm.obs_names, dm.var_names = src.obs_names, src.var_names
            else:
                dm = DeviceMap.upload(self._as_anndata(src, key), self.dtype, self.device)
            if self.transposed_maps:
                dm.make_transposed()
            self._device_maps[key] = dm
        return dm

    def get_coupling(self, t0, t1, covariate=None):
        """The map t0 -> t1 as AnnData (transport_map_model.py:145-190);
        non-consecutive days chain the maps of find_path."""
        if t0 not in self.timepoints or t1 not in self.timepoints:
            raise ValueError("Day pair {}, {} not found".format(t0, t1))
        atomic = (self.day_pairs is not None and (t0, t1) in self.day_pairs) \
            or self.timepoints.index(t1) == self.timepoints.index(t0) + 1
        if not atomic and covariate is not None:
            raise ValueError("Covariate-restricted transport maps can only be atomic")
        if atomic:
            key = (t0, t1) if covariate is None else (t0, t1, str(covariate[0]), str(covariate[1]))
            src = self.tmaps.get(key)
            if src is None:
                raise ValueError('No transport map found for {}', key)
            return self._as_anndata(src, key)
        return self._chained_coupling(find_path(t0, t1, self.day_pairs, self.timepoints))

    def _chained_coupling(self, path):
        """T_0 T_1 ... over the device maps of ``path`` (wot/tmap/chaining.py:4-40
        with glue_transport_maps, util.py:74-94): FP64 products (rocBLAS DGEMM),
        rows of each next map re-ordered to the previous map's columns by cell
        id as the reference does, no normalisation."""
        torch = _dev.torch_mod()
        for i in range(len(path) - 1):
            if path[i][1] != path[i + 1][0]:
                raise ValueError("Transport maps cannot be glued together : invalid path")
        for a, b in path:
            if a >= b:
                raise ValueError("({}, {}) is not a valid transport map : it goes backwards in time".format(a, b))
        first = self.device_map(*path[0])
        X = first.X.double()
        var = first.var_names
        for a, b in path[1:]:
            nxt = self.device_map(a, b)
            Y = nxt.X.double()
            if var is not None and nxt.obs_names is not None and not pd.Index(nxt.obs_names).equals(pd.Index(var)):
                idx = pd.Index(nxt.obs_names).get_indexer_for(pd.Index(var))
                Y = Y.index_select(0, torch.from_numpy(np.asarray(idx, dtype=np.int64)).to(Y.device))
            X = X @ Y
            var = nxt.var_names
        n0, n1 = X.shape
        obs = pd.DataFrame(index=first.obs_names if first.obs_names is not None else
                           pd.RangeIndex(n0).astype(str))
        var_df = pd.DataFrame(index=var if var is not None else pd.RangeIndex(n1).astype(str))
        return _anndata.AnnData(X.cpu().numpy(), obs, var_df)

    def _as_anndata(self, src, key):
        if isinstance(src, DeviceMap):
            X = src.X.double().cpu().numpy()
            obs = pd.DataFrame(index=src.obs_names if src.obs_names is not None else
                               pd.RangeIndex(X.shape[0]).astype(str))
            var = pd.DataFrame(index=src.var_names if src.var_names is not None else
                               pd.RangeIndex(X.shape[1]).astype(str))
            return _anndata.AnnData(X, obs, var)
        if isinstance(src, str):
            from .. import io as _io
            ds = _io.read_dataset(src)
            if self.cache:
                self.tmaps[key] = ds
            return ds
        if hasattr(src, 'X'):
            return src
        X = np.asarray(src.cpu().numpy() if hasattr(src, 'cpu') else src, dtype=np.float64)
        return _anndata.AnnData(X)

    def _chain(self, i, j):
        """Device maps walked from day index i to j (push: i < j, pull: i > j)."""
        tp = self.timepoints
        if i < j:
            return [self.device_map(tp[s], tp[s + 1]) for s in range(i, j)]
        return [self.device_map(tp[s - 1], tp[s]) for s in range(i, j, -1)]

    def _n_cells(self, t):
        i = self.timepoints.index(t)
        if i + 1 < len(self.timepoints):
            return self.device_map(t, self.timepoints[i + 1]).shape[0]
        return self.device_map(self.timepoints[i - 1], t).shape[1]

    # ------------------------------------------------------------------
    # push / pull (transport_map_model.py:192-365)
    # ------------------------------------------------------------------
    def can_push_forward(self, *populations):
        return self.timepoints.index(unique_timepoint(*populations)) < len(self.timepoints) - 1

    def can_pull_back(self, *populations):
        return self.timepoints.index(unique_timepoint(*populations)) > 0

    @staticmethod
    def _upload(torch, populations):
        P = np.vstack([pop.p for pop in populations]).T
        return torch.from_numpy(np.ascontiguousarray(P, dtype=np.float64)).cuda()

    def _walk(self, populations, i, j, normalize):
        """Device populations (cells x k) after walking day index i -> j."""
        torch = _dev.torch_mod()
        P = self._upload(torch, populations)
        if i == j:
            return P
        direction = PUSH_FORWARD if i < j else PULL_BACK
        return propagate(self._chain(i, j), P, direction, normalize=normalize)

    def _result(self, P, t, as_list):
        p = P.cpu().numpy()
        result = [Population(t, np.ascontiguousarray(p[:, k])) for k in range(p.shape[1])]
        if len(result) == 1 and not as_list:
            return result[0]
        return result

    def push_forward(self, *populations, to_time=None, normalize=True, as_list=False):
        """p <- p T through every map up to ``to_time`` (default: the next day),
        normalised to sum 1 after each map when ``normalize``."""
        i = self.timepoints.index(unique_timepoint(*populations))
        j = i + 1 if to_time is None else self.timepoints.index(to_time)
        if j >= len(self.timepoints):
            raise ValueError("No further timepoints. Unable to push forward")
        if i > j:
            raise ValueError("Destination timepoint is before source. Unable to push forward")
        return self._result(self._walk(populations, i, j, normalize), self.timepoints[j], as_list)

    def pull_back(self, *populations, to_time=None, normalize=True, as_list=False):
        """p <- (T p^T)^T back to ``to_time`` (default: the previous day)."""
        i = self.timepoints.index(unique_timepoint(*populations))
        j = i - 1 if to_time is None else self.timepoints.index(to_time)
        if i == 0:
            raise ValueError("No previous timepoints. Unable to pull back")
        if i < j:
            raise ValueError("Destination timepoint is after source. Unable to pull back")
        return self._result(self._walk(populations, i, j, normalize), self.timepoints[j], as_list)

    def ancestors(self, *populations, at_time=None, as_list=False):
        return self.pull_back(*populations, to_time=at_time, as_list=as_list)

    def descendants(self, *populations, at_time=None, as_list=False):
        return self.push_forward(*populations, to_time=at_time, as_list=as_list)

    # ------------------------------------------------------------------
    # whole-course products (transport_map_model.py:40-143)
    # ------------------------------------------------------------------
    def _day_blocks(self, days):
        """Row offsets of each day's block in a matrix of the given days stacked in order."""
        offs, n = {}, 0
        for t in days:
            offs[t] = n
            n += self._n_cells(t)
        return offs, n

    def fates(self, populations):
        """Probability that each cell of every day up to the populations' day
        ends in each population (binary populations plus 'Other', pulled back
        unnormalised; every cell's row normalised to sum 1)."""
        torch = _dev.torch_mod()
        start_day = unique_timepoint(*populations)
        populations = Population.copy(*populations, normalize=False, add_missing=True)
        names = [pop.name for pop in populations]
        i = self.timepoints.index(start_day)
        days = self.timepoints[:i + 1]
        offs, n = self._day_blocks(days)
        k = len(populations)
        X = torch.empty((n, k), dtype=torch.float64, device='cuda')
        P = self._upload(torch, populations)
        X[offs[start_day]:offs[start_day] + P.shape[0]].copy_(P)
        if i > 0:
            outs = [X[offs[days[s - 1]]:offs[days[s - 1]] + self._n_cells(days[s - 1])] for s in range(i, 0, -1)]
            propagate(self._chain(i, 0), P, PULL_BACK, normalize=False, outs=outs)
        normalize_rows(X)
        obs = self.meta.copy()
        obs = obs[obs['day'] <= start_day]
        return _anndata.AnnData(X.cpu().numpy(), obs, pd.DataFrame(index=names))

    def transition_table(self, start_populations, end_populations):
        """Mass moving from each start population to each end population
        (both binary plus 'Other'), normalised to sum 1."""
        torch = _dev.torch_mod()
        start_time = unique_timepoint(*start_populations)
        start_populations = Population.copy(*start_populations, normalize=False, add_missing=True)
        end_populations = Population.copy(*end_populations, normalize=False, add_missing=True)
        end_time = unique_timepoint(*end_populations)
        i = self.timepoints.index(end_time)
        # pulled back while the end day is after the start day
        j = min(i, self.timepoints.index(start_time))
        end_p = self._walk(end_populations, i, j, normalize=False)  # cells x k2 at day j
        start_p = torch.from_numpy(np.vstack([pop.p for pop in start_populations])).cuda()  # k1 x cells
        if start_p.shape[1] != end_p.shape[0]:
            raise ValueError('matmul: start populations have %d cells, end populations %d'
                             % (start_p.shape[1], end_p.shape[0]))
        # k1 x k2 = start_p @ end_p: a pull back of end_p through start_p as a map
        p = propagate([DeviceMap(start_p.contiguous())], end_p, PULL_BACK, normalize=False).cpu().numpy()
        p = p / p.sum()
        return _anndata.AnnData(p, pd.DataFrame(index=[q.name for q in start_populations]),
                                pd.DataFrame(index=[q.name for q in end_populations]))

    def trajectories(self, populations):
        """Each population's ancestors and descendants on every day (every
        step normalised), stacked day by day over all cells."""
        torch = _dev.torch_mod()
        unique_timepoint(*populations)
        populations = Population.copy(*populations, normalize=True, add_missing=False)
        names = [p.name for p in populations]
        t = populations[0].time
        i = self.timepoints.index(t)
        last = len(self.timepoints) - 1
        offs, n = self._day_blocks(self.timepoints)
        k = len(populations)
        X = torch.empty((n, k), dtype=torch.float64, device='cuda')
        P = self._upload(torch, populations)
        X[offs[t]:offs[t] + P.shape[0]].copy_(P)
        tp = self.timepoints

        def block(day):
            return X[offs[day]:offs[day] + self._n_cells(day)]

        if i > 0:
            propagate(self._chain(i, 0), P, PULL_BACK, normalize=True,
                      outs=[block(tp[s - 1]) for s in range(i, 0, -1)])
        if i < last:
            propagate(self._chain(i, last), P, PUSH_FORWARD, normalize=True,
                      outs=[block(tp[s + 1]) for s in range(i, last)])
        return _anndata.AnnData(X.cpu().numpy(), self.meta.copy(), pd.DataFrame(index=names))

    # ------------------------------------------------------------------
    # populations from cell ids, census (transport_map_model.py:455-615)
    # ------------------------------------------------------------------
    def population_from_ids(self, *ids, at_time):
        """Binary (unnormalised) populations over the cells of ``at_time``; None
        where no id is found there."""
        day = float(at_time)
        df = self.meta[self.meta['day'] == day]

        def one(id_list):
            idx = df.index.get_indexer_for(id_list)
            idx = idx[idx > -1]
            if len(idx) == 0:
                return None
            p = np.zeros(len(df), dtype=np.float64)
            p[idx] = 1.0
            return Population(day, p)

        return [one(x) for x in ids]

    def population_from_cell_sets(self, cell_sets, at_time):
        keys = list(cell_sets.keys())
        pops = self.population_from_ids(*[cell_sets[k] for k in keys], at_time=at_time)
        out = []
        for key, pop in zip(keys, pops):
            if pop is not None:
                pop.name = key
                out.append(pop)
        return out

    def cell_ids(self, population):
        return self.meta[self.meta['day'] == population.time].index.values

    def population_census(self, cell_set_matrix, *populations):
        """census[i] = share of each population's mass in cell set i (host:
        O(cells x sets) on populations already on the host)."""
        day = unique_timepoint(*populations)
        df = self.meta[self.meta['day'] == day]
        inter = cell_set_matrix.obs.index.intersection(df.index)
        if len(inter) == 0:
            return [[0] * cell_set_matrix.X.shape[1]] * len(populations)
        pop_idx = df.index.get_indexer_for(inter)
        csm_idx = cell_set_matrix.obs.index.get_indexer_for(inter)
        M = np.asarray(cell_set_matrix.X)[csm_idx, :]

        def norm(p):
            s = np.sum(p)
            return p if np.isclose(s, 0) else p / s

        return np.asarray([np.dot(norm(pop.p)[pop_idx], M) for pop in populations], dtype=np.float64)

    def ancestor_census(self, cset_matrix, *populations):
        """Census of the populations' ancestors and descendants on every day."""
        timepoints, census = [], []
        t = unique_timepoint(*populations)
        i = self.timepoints.index(t)
        last = len(self.timepoints) - 1
        timepoints.append(t)
        census.append(self.population_census(cset_matrix, *populations))
        torch = _dev.torch_mod()
        P = self._upload(torch, populations)
        tp = self.timepoints
        if i > 0:
            outs = [torch.empty((self._n_cells(tp[s - 1]), P.shape[1]), dtype=torch.float64, device='cuda')
                    for s in range(i, 0, -1)]
            propagate(self._chain(i, 0), P, PULL_BACK, normalize=True, outs=outs)
            for s, o in zip(range(i, 0, -1), outs):
                pops = self._result(o, tp[s - 1], True)
                timepoints.insert(0, tp[s - 1])
                census.insert(0, self.population_census(cset_matrix, *pops))
        if i < last:
            outs = [torch.empty((self._n_cells(tp[s + 1]), P.shape[1]), dtype=torch.float64, device='cuda')
                    for s in range(i, last)]
            propagate(self._chain(i, last), P, PUSH_FORWARD, normalize=True, outs=outs)
            for s, o in zip(range(i, last), outs):
                pops = self._result(o, tp[s + 1], True)
                timepoints.append(tp[s + 1])
                census.append(self.population_census(cset_matrix, *pops))
        census = np.asarray(census)
        if census.ndim == 3:
            census = np.asarray([census[:, c, :] for c in range(census.shape[1])])
        return timepoints, census
