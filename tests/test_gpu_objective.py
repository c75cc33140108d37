"""Objective reductions on the GPU (wot/ot/optimal_transport.py:37-62).

- The dense-array helpers ``fdiv``, ``fdivstar``, ``primal`` and ``dual`` of
  ``wot_amd.ot.optimal_transport`` (torch FP64 on the device) against the
  golden vectors written from the reference file (objective_{0,1,2}.npz; case
  0 holds the NaN that 0 * log 0 gives in the reference's fdiv).
- The fused FP64 gap pass behind ``wotk_primal_dual`` (``device.primal_dual``)
  against the oracle's primal and dual evaluated at the same potentials, with
  the plan, the dual variables and K-tilde built as the reference's gap check
  builds them (optimal_transport.py:147-156), on the FP32-rounded cost the
  device holds.
- The converged duality gap of the device solver against the oracle solver's
  on the same inputs.
"""
import os

import numpy as np
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.fixture(scope='module')
def mods():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a ROCm device'
    import wot_amd.ot as ot
    from wot_amd.ot import device as dev
    from wot_amd.ot import optimal_transport as otm
    return torch, ot, dev, otm


def _load(n):
    z = np.load(os.path.join(GOLD, 'objective_%d.npz' % n), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize('n', [0, 1, 2])
def test_dense_helpers_golden(mods, n):
    _, _, _, otm = mods
    z = _load(n)
    ni, nj = z['R'].shape
    dx, dy = np.ones(ni) / ni, np.ones(nj) / nj
    eps, l1, l2 = float(z['eps']), float(z['lambda1']), float(z['lambda2'])
    args = (z['C'], z['K'], z['R'], dx, dy, z['p'], z['q'], z['a'], z['b'], eps, l1, l2)
    pr, du = otm.primal(*args), otm.dual(*args)
    # the golden fdiv / fdivstar calls (tests/golden/make_golden.py)
    fd = otm.fdiv(l1, z['R'].dot(dy) + 0.1, z['p'], dx)
    fs = otm.fdivstar(l2, -eps * np.log(z['b']), z['q'], dy)
    np.testing.assert_allclose(pr, float(z['primal']), rtol=1e-12, equal_nan=True)
    np.testing.assert_allclose(du, float(z['dual']), rtol=1e-12, equal_nan=True)
    np.testing.assert_allclose(fd, float(z['fdiv']), rtol=1e-12)
    np.testing.assert_allclose(fs, float(z['fdivstar']), rtol=1e-12)
    if n == 0:
        assert np.isnan(pr), 'R = 0 must give NaN (0 * log 0 in fdiv), as the reference does'


def _oracle_at(C32, G, f, g, eps, l1, l2):
    """The reference's gap-check primal and dual at potentials (f, g):
    R = exp((f_i + g_j - C_ij)/eps), a = exp(f/eps), b = exp(g/eps),
    K-tilde = exp(-C/eps) (optimal_transport.py:147-156)."""
    C = C32.astype(np.float64)
    ni, nj = C.shape
    dx, dy = np.ones(ni) / ni, np.ones(nj) / nj
    p = np.asarray(G, dtype=np.float64)
    q = np.ones(nj) * np.average(p)
    with np.errstate(all='ignore'):
        R = np.exp((f[:, None] + g[None, :] - C) / eps)
        Kt = np.exp(-C / eps)
        a, b = np.exp(f / eps), np.exp(g / eps)
        args = (C, Kt, R, dx, dy, p, q, a, b, eps, l1, l2)
        return O.primal(*args), O.dual(*args)


def _case(ni, nj, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(ni, 5))
    y = rng.normal(size=(nj, 5)) + 0.3
    G = np.exp(rng.normal(0.0, 0.3, ni))
    return x, y, G


@pytest.mark.parametrize('ni,nj,eps,seed', [(64, 80, 0.05, 1), (300, 400, 0.1, 2), (7, 5, 0.2, 3)])
def test_primal_dual_kernel_vs_oracle(mods, ni, nj, eps, seed):
    torch, ot, dev, _ = mods
    x, y, G = _case(ni, nj, seed)
    cost = ot.DeviceCost.from_coordinates(x, y)
    C32 = cost.dense().cpu().numpy()
    l1, l2 = 1.0, 50.0
    # potentials from a converged device solve (the state the gap check sees)
    res = dev.solve(cost, G, 'duality_gap', dict(lambda1=l1, lambda2=l2, epsilon=eps, batch_size=5, tolerance=1e-8,
                                                 tau=1e4, epsilon0=1.0, max_iter=1e7),
                    want_plan=False, want_potentials=True)
    f, g = res['f'].cpu().numpy(), res['g'].cpu().numpy()
    eps_f = res['stats']['epsilon_final']
    pr, du = dev.primal_dual(cost, G, f, g, l1, l2, eps_f)
    pr_o, du_o = _oracle_at(C32, G, f, g, eps_f, l1, l2)
    assert abs(pr - pr_o) <= 1e-11 * abs(pr_o), (pr, pr_o)
    assert abs(du - du_o) <= 1e-11 * abs(pr_o), (du, du_o)
    # weak duality, and the converged gap the solver reported
    assert pr >= du
    np.testing.assert_allclose((pr - du) / abs(pr), res['stats']['gap'], rtol=1e-6, atol=1e-12)
    # away from convergence (perturbed potentials) the two still agree
    rng = np.random.default_rng(seed + 10)
    f2 = f + 0.05 * eps_f * rng.normal(size=ni)
    g2 = g + 0.05 * eps_f * rng.normal(size=nj)
    pr2, du2 = dev.primal_dual(cost, G, f2, g2, l1, l2, eps_f)
    pr2_o, du2_o = _oracle_at(C32, G, f2, g2, eps_f, l1, l2)
    assert abs(pr2 - pr2_o) <= 1e-11 * abs(pr2_o), (pr2, pr2_o)
    assert abs(du2 - du2_o) <= 1e-11 * abs(pr2_o), (du2, du2_o)
    assert pr2 - du2 > pr - du


def test_primal_dual_kernel_nan_row(mods):
    """A source row whose plan entries all underflow has row sum 0, and the
    reference's fdiv gives 0 * log 0 = NaN there; the kernel gives NaN too
    (the solver turns that into the reference's RuntimeError)."""
    torch, ot, dev, _ = mods
    x, y, G = _case(40, 30, 4)
    cost = ot.DeviceCost.from_coordinates(x, y)
    C32 = cost.dense().cpu().numpy()
    f = np.zeros(40)
    g = np.zeros(30)
    f[3] = -1.0e6
    pr, du = dev.primal_dual(cost, G, f, g, 1.0, 50.0, 0.05)
    pr_o, du_o = _oracle_at(C32, G, f, g, 0.05, 1.0, 50.0)
    assert np.isnan(pr_o) and np.isnan(pr)
    np.testing.assert_allclose(du, du_o, rtol=1e-11)


@pytest.mark.parametrize('ni,nj,seed', [(120, 150, 5), (500, 450, 6)])
def test_converged_gap_vs_oracle(mods, ni, nj, seed):
    """The relative duality gap at exit, iteration and stabilisation counts of
    the device solver equal the oracle solver's (FP32 cost on both)."""
    torch, ot, dev, _ = mods
    x, y, G = _case(ni, nj, seed)
    cost = ot.DeviceCost.from_coordinates(x, y)
    C32 = cost.dense().cpu().numpy().astype(np.float64)
    prm = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, batch_size=5, tolerance=1e-8, tau=1e4, epsilon0=1.0,
               max_iter=1e7)
    st = {}
    R = ot.optimal_transport_duality_gap(cost, G, stats=st, **prm)
    so = {}
    Ro = O.optimal_transport_duality_gap(C32, G, stats=so, **prm)
    assert st['iterations'] == so['iterations']
    assert st['stabilisations'] == so['stabilisations']
    assert 0.0 <= st['gap'] <= 1e-8 and 0.0 <= so['gap'] <= 1e-8
    np.testing.assert_allclose(st['gap'], so['gap'], rtol=1e-3, atol=1e-12)
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-5
