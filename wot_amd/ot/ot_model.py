"""OTModel with the reference's interface (wot/ot/ot_model.py:19-326); every
transport map is solved on MI355X through libwotk.

Inputs, configuration keys, solver names, per-pair parameters, error types and
the returned AnnData (obs = source cell ids with learned growth g0..gN,
var = target cell ids, X = float64 plan) follow the reference.  The cost is
computed on the device from the (PCA) coordinates and never copied back.
"""
import itertools
import logging
import os

import numpy as np
import pandas as pd

from .. import _anndata
from ..io import filter_adata
from . import device as _dev
from . import optimal_transport as _ot
from .initializer import parse_configuration, parse_parameter_file
from .pca import compute_pca

logger = logging.getLogger('wot')

DEFAULT_CONFIG = {'local_pca': 30, 'growth_iters': 1, 'epsilon': 0.05, 'lambda1': 1, 'lambda2': 50,
                  'epsilon0': 1, 'tau': 10000, 'scaling_iter': 3000, 'inner_iter_max': 50, 'tolerance': 1e-8,
                  'max_iter': 1e7, 'batch_size': 5, 'extra_iter': 1000}


def _dense(x):
    import scipy.sparse
    return x.toarray() if scipy.sparse.issparse(x) else np.asarray(x)


class OTModel:
    """Computes transport maps (wot/ot/ot_model.py:19-114).

    Parameters follow the reference: ``matrix`` (AnnData, cells x features,
    ``obs[day_field]``, optional ``obs[growth_rate_field]``), and OT
    configuration keyword arguments (``epsilon``, ``lambda1``, ``lambda2``,
    ``growth_iters``, ``local_pca``, ``solver`` = 'duality_gap' (default) or
    'fixed_iters', ...).  ``cost_mode`` ('f64' default, 'bf16x3' or 'bf16x6') selects
    the device cost kernel."""

    def __init__(self, matrix, day_field='day', covariate_field='covariate', growth_rate_field='cell_growth_rate',
                 **kwargs):
        self.matrix = matrix
        self.day_field = day_field
        self.covariate_field = covariate_field
        self.cell_growth_rate_field = growth_rate_field
        self.day_pairs = parse_configuration(kwargs.pop('config', None))
        cell_filter = kwargs.pop('cell_filter', None)
        gene_filter = kwargs.pop('gene_filter', None)
        day_filter = kwargs.pop('cell_day_filter', None)
        ncounts = kwargs.pop('ncounts', None)
        ncells = kwargs.pop('ncells', None)
        self.cost_mode = kwargs.pop('cost_mode', 'f64')
        self.matrix = filter_adata(self.matrix, obs_filter=cell_filter, var_filter=gene_filter)
        if day_filter is not None:
            days = [float(day) for day in day_filter.split(',')] if isinstance(day_filter, str) else day_filter
            self.matrix = self.matrix[self.matrix.obs[self.day_field].isin(days).values].copy()

        cvs = set(self.matrix.obs[self.covariate_field]) if self.covariate_field in self.matrix.obs else [None]
        if ncells is not None:
            index_list = []
            for day in sorted(set(self.matrix.obs[self.day_field])):
                day_query = (self.matrix.obs[self.day_field] == day).values
                for cv in cvs:
                    if cv is None:
                        indices = np.where(day_query)[0]
                    else:
                        indices = np.where(day_query & (self.matrix.obs[self.covariate_field] == cv).values)[0]
                    if len(indices) > ncells:
                        np.random.shuffle(indices)
                        indices = indices[0:ncells]
                    index_list.append(indices)
            self.matrix = self.matrix[np.concatenate(index_list)]
        if ncounts is not None:
            X = _dense(self.matrix.X).astype(np.float64)
            for i in range(X.shape[0]):
                total = X[i].sum()
                if total > ncounts:
                    X[i] = np.random.multinomial(ncounts, X[i] / total, size=1)[0]
            self.matrix.X = X

        if self.matrix.X.shape[0] == 0:
            raise ValueError('No cells in matrix')

        self.ot_config = dict(DEFAULT_CONFIG)
        solver = kwargs.pop('solver', 'duality_gap')
        if solver == 'fixed_iters':
            self.solver = _ot.transport_stablev2
        elif solver == 'duality_gap':
            self.solver = _ot.optimal_transport_duality_gap
        else:
            raise ValueError('Unknown solver')

        parameters_from_file = kwargs.pop('parameters', None)
        for k in kwargs.keys():
            self.ot_config[k] = kwargs[k]
        if parameters_from_file is not None:
            for k, v in parse_parameter_file(parameters_from_file).items():
                self.ot_config[k] = v

        local_pca = self.ot_config['local_pca']
        if local_pca > self.matrix.X.shape[1]:
            logger.warning("local_pca set to {}, above gene count of {}. Disabling PCA"
                           .format(local_pca, self.matrix.X.shape[1]))
            self.ot_config['local_pca'] = 0
        if self.day_field not in self.matrix.obs.columns:
            raise ValueError("Days information not available for matrix")
        if any(self.matrix.obs[self.day_field].isnull()):
            self.matrix = self.matrix[(self.matrix.obs[self.day_field].isnull() == False).values]  # noqa: E712
        self.timepoints = sorted(set(self.matrix.obs[self.day_field]))

    def get_covariate_pairs(self):
        """All covariate pairs in the dataset (ot_model.py:116-122)."""
        if self.covariate_field not in self.matrix.obs.columns:
            raise ValueError("Covariate value not available in dataset")
        covariate = set(self.matrix.obs[self.covariate_field])
        return itertools.product(covariate, covariate)

    def compute_all_transport_maps(self, tmap_out='tmaps', overwrite=True, output_file_format='h5ad',
                                   with_covariates=False, cost_matrices=None, batch='auto'):
        """Computes and saves every consecutive (or configured) day pair
        (ot_model.py:124-201).  Existing outputs are skipped when
        ``overwrite`` is False; learned growth goes to ``<prefix>_g.txt``
        when ``growth_iters > 1``.  Each map is written by one background
        thread while the next pair is solved (at most one write in flight, so
        one extra host copy of a map); write errors surface here.

        ``batch``: with the fixed_iters solver, no covariates and no given
        cost matrices, every pair (each with its own epsilon / lambda1 /
        lambda2 from the day-pair configuration) is solved in one batched run
        on the device -- one persistent launch per half-iteration for all
        pairs, growth iterations through the plans' row sums -- when the
        pairs' FP32 costs fit in device memory ('auto', the default), always
        (True) or never (False).  Results match the pair-by-pair solves."""
        from concurrent.futures import ThreadPoolExecutor
        from ..io import check_file_extension, write_dataset
        tmap_dir, tmap_prefix = os.path.split(tmap_out) if tmap_out is not None else (None, None)
        tmap_prefix = tmap_prefix or "tmaps"
        tmap_dir = tmap_dir or '.'
        if not os.path.exists(tmap_dir):
            os.makedirs(tmap_dir)
        t = self.timepoints
        day_pairs = self.day_pairs
        if day_pairs is None or len(day_pairs) == 0:
            day_pairs = [(t[i], t[i + 1]) for i in range(len(t) - 1)]
        if with_covariates:
            day_pairs = [(*d, c) for d, c in itertools.product(day_pairs, self.get_covariate_pairs())]
        if not day_pairs:
            logger.info('No day pairs')
            return
        use_batch = batch is True or (batch == 'auto' and cost_matrices is None and not with_covariates)
        if cost_matrices is None:
            cost_matrices = [None for _ in day_pairs]
        full_learned_growth_df = None
        save_learned_growth = self.ot_config.get('growth_iters', 1) > 1
        todo = []
        for day_pair in day_pairs:
            path = tmap_prefix + ("_{}_{}".format(*day_pair) if not with_covariates
                                  else "_{}_{}_cv{}_cv{}".format(*day_pair))
            todo.append(check_file_extension(os.path.join(tmap_dir, path), output_file_format))
        batched = None
        if use_batch:
            batched = self._batched_maps(day_pairs, todo, overwrite, cost_matrices, with_covariates,
                                         required=batch is True)
        with ThreadPoolExecutor(1) as writer:
            pending = None
            for k, (day_pair, cost_matrix) in enumerate(zip(day_pairs, cost_matrices)):
                if batched is not None:
                    tmap, output_file = batched(k), todo[k]
                else:
                    tmap, output_file = self._next_map(day_pair, cost_matrix, with_covariates, tmap_dir,
                                                       tmap_prefix, output_file_format, overwrite,
                                                       check_file_extension)
                if tmap is None:
                    continue
                if pending is not None:
                    pending.result()
                pending = writer.submit(write_dataset, tmap, output_file, output_format=output_file_format)
                if save_learned_growth:
                    learned_growth_df = tmap.obs
                    full_learned_growth_df = learned_growth_df if full_learned_growth_df is None else pd.concat(
                        (full_learned_growth_df, learned_growth_df), copy=False)
            if pending is not None:
                pending.result()
        if full_learned_growth_df is not None:
            full_learned_growth_df.to_csv(os.path.join(tmap_dir, tmap_prefix + '_g.txt'), sep='\t',
                                          index_label='id')

    def _batched_maps(self, day_pairs, outputs, overwrite, cost_matrices, with_covariates, required=False):
        """Solve every pair of compute_all_transport_maps in one batched
        device run (libwotk wotk_batch_begin / advance, per-pair parameters).
        Returns a function k -> AnnData map of pair k (None: skipped, as the
        sequential loop skips it), or None when the pairs cannot be batched
        (then the caller solves them one by one); ``required`` turns that
        into a ValueError."""
        from . import batch as _batch

        def refuse(why):
            if required:
                raise ValueError('compute_all_transport_maps cannot batch these pairs: ' + why)
            logger.info('solving pairs one by one: %s', why)
            return None

        if self.solver is not _ot.transport_stablev2:
            return refuse("batching runs the 'fixed_iters' solver")
        if with_covariates or any(c is not None for c in cost_matrices):
            return refuse('covariates or given cost matrices')
        keys = ('growth_iters', 'scaling_iter', 'extra_iter', 'inner_iter_max', 'epsilon0', 'local_pca')
        coords, G, prm, slots = [], [], [], {}
        shared = None
        for k, (t0, t1) in enumerate(day_pairs):
            if os.path.exists(outputs[k]) and not overwrite:
                logger.info('Found existing tmap at ' + outputs[k] + '. ')
                continue
            local = self.day_pairs[(t0, t1)] if self.day_pairs is not None else {}
            if self.day_pairs is not None and (t0, t1) not in self.day_pairs:
                raise ValueError("Transport map ({},{}) is not present in day_pairs".format(t0, t1))
            cfg = {**self.ot_config, **local}
            sh = tuple(cfg.get(x) for x in keys)
            if shared is None:
                shared = sh
            elif sh != shared:
                return refuse('pairs differ in ' + ', '.join(x for x, a, b in zip(keys, sh, shared) if a != b))
            day = self.matrix.obs[self.day_field]
            p0 = self.matrix[(day == float(t0)).values, :]
            p1 = self.matrix[(day == float(t1)).values, :]
            if p0.shape[0] == 0 or p1.shape[0] == 0:
                logger.info('No cells at {}'.format(t0 if p0.shape[0] == 0 else t1))
                continue
            if cfg['local_pca'] is not None and cfg['local_pca'] > 0:
                x, y, pca, _ = compute_pca(p0.X, p1.X, cfg['local_pca'])
                ev = np.diag(pca.singular_values_)
                x, y = np.asarray(x).dot(ev), np.asarray(y).dot(ev)
            else:
                x, y = _dense(p0.X), _dense(p1.X)
            dt = t1 - t0
            if self.cell_growth_rate_field in p0.obs.columns:
                g = np.power(p0.obs[self.cell_growth_rate_field].values.astype(np.float64), dt)
            else:
                g = np.ones(p0.shape[0])
            slots[k] = (len(prm), p0.obs.index, p1.obs.index, dt)
            coords += [np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)]
            G.append(g)
            prm.append((cfg['epsilon'], cfg['lambda1'], cfg['lambda2']))
        if not prm:
            return lambda k: None
        torch = _dev.torch_mod()
        n = len(prm)
        need = sum(8.0 * c.shape[0] * coords[2 * i + 1].shape[0] for i, c in enumerate(coords[0::2]))
        free, _ = torch.cuda.mem_get_info()
        if need > 0.7 * free:
            return refuse('the %d costs need %.1f GB of the %.1f GB free' % (n, need / 1e9, free / 1e9))
        growth_iters, scaling_iter, extra_iter, inner_iter_max, epsilon0, _ = shared
        tc = _batch.prepare_time_course(coords, [G[i // 2] if i % 2 == 0 else None for i in range(2 * n)],
                                        pairs=[(2 * i, 2 * i + 1) for i in range(n)], cost_mode=self.cost_mode)
        eps = np.array([[p[0]] for p in prm], dtype=np.float64)
        l1 = np.array([[p[1]] for p in prm], dtype=np.float64)
        l2 = np.array([[p[2]] for p in prm], dtype=np.float64)
        learned = [[torch.from_numpy(g).cuda()] for g in G]
        Gk = tc.G
        res = None
        for it in range(int(growth_iters)):
            if it > 0:
                Gk = torch.cat([res.row_sums(i, 0) for i in range(n)])
                for i in range(n):
                    learned[i].append(res.row_sums(i, 0))
            with _batch.TimeCourseRun(tc, eps, l1, l2, scaling_iter=scaling_iter, epsilon0=epsilon0,
                                      extra_iter=extra_iter, inner_iter_max=inner_iter_max, G=Gk) as run:
                run.advance(run.n_iter)
                res = run.result()

        def get(k):
            if k not in slots:
                return None
            i, obs_ids, var_ids, dt = slots[k]
            R = res.plan(i, 0, device_out=True)
            rows = learned[i] + [R.sum(1)]
            obs = pd.DataFrame(index=obs_ids, data={'g' + str(j): np.power(r.cpu().numpy(), 1.0 / dt)
                                                    for j, r in enumerate(rows)})
            out = _anndata.AnnData(R.cpu().numpy(), obs, pd.DataFrame(index=var_ids))
            del R
            return out
        return get

    def _next_map(self, day_pair, cost_matrix, with_covariates, tmap_dir, tmap_prefix, output_file_format,
                  overwrite, check_file_extension):
        """(map, output path) of one pair of compute_all_transport_maps, or
        (None, path) when the file exists and overwrite is off."""
        path = tmap_prefix
        if not with_covariates:
            path += "_{}_{}".format(*day_pair)
        else:
            path += "_{}_{}_cv{}_cv{}".format(*day_pair)
        output_file = check_file_extension(os.path.join(tmap_dir, path), output_file_format)
        if os.path.exists(output_file) and not overwrite:
            logger.info('Found existing tmap at ' + output_file + '. ')
            return None, output_file
        return self.compute_transport_map(*day_pair, cost_matrix=cost_matrix), output_file

    def compute_transport_map(self, t0, t1, covariate=None, cost_matrix=None):
        """Transport map from t0 to t1 (ot_model.py:203-240)."""
        if self.day_pairs is not None:
            if (t0, t1) not in self.day_pairs:
                raise ValueError("Transport map ({},{}) is not present in day_pairs".format(t0, t1))
            local_config = self.day_pairs[(t0, t1)]
        else:
            local_config = {}
        if covariate is None:
            logger.info('Computing transport map from {} to {}'.format(t0, t1))
        else:
            logger.info('Computing transport map from {} {} to {} {}'.format(t0, covariate[0], t1, covariate[1]))
        config = {**self.ot_config, **local_config, 't0': t0, 't1': t1, 'covariate': covariate,
                  'C': cost_matrix}
        return self.compute_single_transport_map(config)

    @staticmethod
    def compute_default_cost_matrix(a, b, eigenvals=None, device_out=False, cost_mode='f64'):
        """sqeuclid(a, b) / median (ot_model.py:242-253), computed on the
        device.  Returns a float64 numpy matrix like the reference, or the
        device-resident :class:`DeviceCost` with ``device_out=True``."""
        a, b = _dense(a), _dense(b)
        if eigenvals is not None:
            a = a.dot(eigenvals)
            b = b.dot(eigenvals)
        cost = _dev.DeviceCost.from_coordinates(a, b, mode=cost_mode)
        if device_out:
            return cost
        # the FP64 cost of the reference: recompute it exactly on the device
        c64 = _dev.sqeuclid_f64(a, b)
        return (c64 / cost.median).cpu().numpy()

    def compute_single_transport_map(self, config):
        """One transport map (ot_model.py:255-326)."""
        t0 = config.pop('t0', None)
        t1 = config.pop('t1', None)
        if t0 is None or t1 is None:
            raise ValueError("config must have both t0 and t1, indicating target timepoints")
        ds = self.matrix
        covariate = config.pop('covariate', None)
        day = ds.obs[self.day_field]
        if covariate is None:
            p0_indices = (day == float(t0)).values
            p1_indices = (day == float(t1)).values
        else:
            cv = ds.obs[self.covariate_field]
            p0_indices = ((day == float(t0)) & (cv == covariate[0])).values
            p1_indices = ((day == float(t1)) & (cv == covariate[1])).values
        p0 = ds[p0_indices, :]
        p1 = ds[p1_indices, :]
        if p0.shape[0] == 0:
            logger.info('No cells at {}'.format(t0))
            return None
        if p1.shape[0] == 0:
            logger.info('No cells at {}'.format(t1))
            return None

        local_pca = config.pop('local_pca', None)
        eigenvals = None
        if local_pca is not None and local_pca > 0:
            p0_x, p1_x, pca, mean = compute_pca(p0.X, p1.X, local_pca)
            eigenvals = np.diag(pca.singular_values_)
        else:
            p0_x = p0.X
            p1_x = p1.X
        if config['C'] is None:
            config['C'] = OTModel.compute_default_cost_matrix(p0_x, p1_x, eigenvals, device_out=True,
                                                             cost_mode=self.cost_mode)
        C = config['C']
        delta_days = t1 - t0
        n_i = C.shape[0]
        if self.cell_growth_rate_field in p0.obs.columns:
            config['G'] = np.power(p0.obs[self.cell_growth_rate_field].values.astype(np.float64), delta_days)
        else:
            config['G'] = np.ones(n_i)
        tmap, learned_growth = _ot.compute_transport_matrix(solver=self.solver, **config)
        learned_growth.append(tmap.sum(axis=1))
        obs_growth = {}
        for i in range(len(learned_growth)):
            g = np.power(np.asarray(learned_growth[i]), 1.0 / delta_days)
            obs_growth['g' + str(i)] = g
        obs = pd.DataFrame(index=p0.obs.index, data=obs_growth)
        return _anndata.AnnData(tmap, obs, pd.DataFrame(index=p1.obs.index))
