"""Drop-in solver callables with the reference's signatures
(wot/ot/optimal_transport.py), running on MI355X through libwotk.

``transport_stablev2``, ``optimal_transport_duality_gap`` and
``compute_transport_matrix`` accept and return numpy arrays exactly like the
reference, so ``OTModel`` (or any caller) can swap them in.  Passing a
:class:`DeviceCost` as ``C`` keeps the cost on the device; ``device_out=True``
returns a torch tensor instead of copying the plan to the host.
"""
import logging

import numpy as np

from . import device as _dev

logger = logging.getLogger('wot')

_MAX_ITER_MSG = "Reached max_iter with duality gap still above threshold. Returning"


def _finish(res, device_out):
    if res['stats'].get('max_iter_reached'):
        logger.warning(_MAX_ITER_MSG)
    plan = res['plan']
    return plan if device_out else plan.cpu().numpy()


def transport_stablev2(C, lambda1, lambda2, epsilon, scaling_iter, G, tau, epsilon0, extra_iter,
                       inner_iter_max, device_out=False, stats=None, **ignored):
    """Fixed-iteration stabilised solver (wot/ot/optimal_transport.py:167-236).

    ``scaling_iter`` iterations with eps(n) = (eps0 - eps) e^-n + eps changed
    every ``inner_iter_max`` iterations, then ``extra_iter`` iterations at the
    last eps; returns R / J (float64, I x J)."""
    cost = _dev.ensure_cost(C)
    res = _dev.solve(cost, G, 'fixed_iters',
                     dict(lambda1=lambda1, lambda2=lambda2, epsilon=epsilon, scaling_iter=scaling_iter,
                          tau=tau, epsilon0=epsilon0, extra_iter=extra_iter, inner_iter_max=inner_iter_max))
    if stats is not None:
        stats.update(res['stats'])
    return _finish(res, device_out)


def optimal_transport_duality_gap(C, G, lambda1, lambda2, epsilon, batch_size, tolerance, tau, epsilon0,
                                  max_iter, device_out=False, stats=None, coords=None, terms=6, **ignored):
    """Duality-gap solver, OTModel's default (wot/ot/optimal_transport.py:67-164).

    Five epsilon scalings down to eps0*eps, dual-evolution checks (1e-6) then
    a relative duality gap <= ``tolerance``; returns R / J, or the
    un-normalised plan with a warning when ``max_iter`` is reached; a NaN gap
    raises RuntimeError.

    With ``C=None`` and ``coords=(x, y)`` the cost sqeuclid(x, y) / median is
    never stored (split-BF16 cross term on the matrix cores with ``terms`` = 3
    or 6 products, regenerated per sweep and per gap pass); the result is then
    a ``large.CoordsResult`` holding the potentials."""
    if C is None:
        if coords is None:
            raise ValueError('optimal_transport_duality_gap needs a cost matrix C or coords=(x, y)')
        from . import large
        res = large.optimal_transport_duality_gap_coords(coords[0], coords[1], G, lambda1=lambda1, lambda2=lambda2,
                                                         epsilon=epsilon, batch_size=batch_size, tolerance=tolerance,
                                                         tau=tau, epsilon0=epsilon0, max_iter=max_iter, terms=terms)
        if stats is not None:
            stats.update(res.stats)
        return res
    cost = _dev.ensure_cost(C)
    res = _dev.solve(cost, G, 'duality_gap',
                     dict(lambda1=lambda1, lambda2=lambda2, epsilon=epsilon, batch_size=batch_size,
                          tolerance=tolerance, tau=tau, epsilon0=epsilon0, max_iter=max_iter))
    if stats is not None:
        stats.update(res['stats'])
    return _finish(res, device_out)


_SOLVER_NAMES = {}


def _register(fn, name):
    _SOLVER_NAMES[fn] = name
    return fn


_register(transport_stablev2, 'fixed_iters')
_register(optimal_transport_duality_gap, 'duality_gap')


def compute_transport_matrix(solver, **params):
    """Growth-learning outer loop (wot/ot/optimal_transport.py:10-33).

    Iteration 0 solves with ``G``; iteration k > 0 with the previous plan's
    row sums.  Returns ``(tmap, learned_growth)``.  For this package's solvers
    the loop stays on the device: the cost is uploaded once and only the last
    plan is materialised (earlier iterations only produce their row sums)."""
    name = _SOLVER_NAMES.get(solver)
    n_outer = params['growth_iters']
    if name is None:
        # a foreign solver callable: the plain reference loop
        learned = []
        tmap = None
        for it in range(n_outer):
            rows = params['G'] if it == 0 else tmap.sum(axis=1)
            params['G'] = rows
            learned.append(rows)
            tmap = solver(**params)
        return tmap, learned
    device_out = params.get('device_out', False)
    cost = _dev.ensure_cost(params['C'])
    G = params['G']
    learned = []
    res = None
    for it in range(n_outer):
        if it > 0:
            G = res['row_sums']
        learned.append(G if (device_out or it == 0) else G.cpu().numpy())
        last = it == n_outer - 1
        res = _dev.solve(cost, G, name, params, want_plan=last, want_row_sums=not last)
        if res['stats'].get('max_iter_reached'):
            logger.warning(_MAX_ITER_MSG)
    if n_outer == 0:
        return None, learned
    plan = res['plan']
    return (plan if device_out else plan.cpu().numpy()), learned


def transport_stablev1_learnx(C, G, lambda1, lambda2, epsilon, scaling_iter, tau, epsilon0, extra_iter,
                              inner_iter_max, growth_iters=1, device_out=False, coords=None, terms=6, **ignored):
    """Growth-learning fixed-iteration transport (the ``learnx`` variant).

    The snapshot of wot this package follows no longer ships a
    ``transport_stablev1_learnx``; its contract here is: the growth outer loop
    of compute_transport_matrix (optimal_transport.py:10-33) around
    transport_stablev2 semantics (:167-236), learning the source growth from
    the plan's row sums ``growth_iters`` times.  Returns ``(tmap,
    learned_growth)`` where ``learned_growth`` has ``growth_iters`` entries
    (the G used by each solve).

    ``C`` may be a matrix or a :class:`DeviceCost`.  With ``C=None`` and
    ``coords=(x, y)`` (source and target PCA coordinates) the cost
    sqeuclid(x, y) / median (ot_model.py:242-253) is never stored: every sweep
    regenerates it on the matrix cores (split-BF16 cross term with ``terms`` =
    3 or 6 products), so a pair needs O(I + J) memory; ``tmap`` is then a
    ``large.CoordsResult`` holding the potentials (``tmap.dense()`` or
    ``tmap.plan_rows(rows)`` materialise R / J)."""
    if C is None:
        if coords is None:
            raise ValueError('transport_stablev1_learnx needs a cost matrix C or coords=(x, y)')
        from . import large
        x, y = coords
        return large.transport_map_coords(x, y, G, growth_iters=growth_iters, terms=terms, lambda1=lambda1,
                                          lambda2=lambda2, epsilon=epsilon, scaling_iter=scaling_iter, tau=tau,
                                          epsilon0=epsilon0, extra_iter=extra_iter, inner_iter_max=inner_iter_max)
    params = dict(C=C, G=G, lambda1=lambda1, lambda2=lambda2, epsilon=epsilon, scaling_iter=scaling_iter,
                  tau=tau, epsilon0=epsilon0, extra_iter=extra_iter, inner_iter_max=inner_iter_max,
                  growth_iters=growth_iters, device_out=device_out)
    return compute_transport_matrix(transport_stablev2, **params)


# ---- Chizat primal / dual on dense arrays (optimal_transport.py:37-62) -----
# API-compatible helpers evaluated on the device in FP64 with torch; the
# solvers' own gap checks use the fused libwotk reduction instead.
def _t(x):
    torch = _dev.torch_mod()
    return _dev.as_device_f64(torch, x, 'cuda')


def fdiv(l, x, p, dx):
    x, p, dx = _t(x), _t(p), _t(dx)
    return float(l * (dx * (x * (x / p).log() - x + p)).sum())


def fdivstar(l, u, p, dx):
    u, p, dx = _t(u), _t(p), _t(dx)
    return float(l * ((p * dx) * ((u / l).exp() - 1)).sum())


def primal(C, K, R, dx, dy, p, q, a, b, epsilon, lambda1, lambda2):
    torch = _dev.torch_mod()
    C, K, R, dx, dy, p, q = (_t(v) for v in (C, K, R, dx, dy, p, q))
    n_i, n_j = len(p), len(q)
    logr = torch.nan_to_num(R.log())
    return (fdiv(lambda1, R @ dy, p, dx) + fdiv(lambda2, R.T @ dx, q, dy)
            + float((epsilon * (R * logr - R + K).sum() + (R * C).sum()) / (n_i * n_j)))


def dual(C, K, R, dx, dy, p, q, a, b, epsilon, lambda1, lambda2):
    C, K, R, dx, dy, p, q, a, b = (_t(v) for v in (C, K, R, dx, dy, p, q, a, b))
    n_i, n_j = len(p), len(q)
    return (-fdivstar(lambda1, -epsilon * a.log(), p, dx) - fdivstar(lambda2, -epsilon * b.log(), q, dy)
            - float(epsilon * (R - K).sum() / (n_i * n_j)))
