"""The no-allocation fixed-iteration entry point (wotk_transport_stablev2_noalloc,
state in a static per-GPU device array) against wotk_transport_stablev2 with
a caller workspace: identical plans, row sums and potentials on two calls in
a row, the second on a smaller problem (the array keeps the first call's
contents), and problems beyond the array's capacity rejected on the host."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRM = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=300, tau=1e4, epsilon0=1.0, extra_iter=100,
           inner_iter_max=50)


@pytest.fixture(scope='module')
def mods():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a ROCm device'
    import wot_amd.ot as ot
    from wot_amd import _native
    from wot_amd.ot import device as dev
    return torch, ot, _native, dev


def _outputs(torch, n_i, n_j):
    return (torch.empty((n_i, n_j), dtype=torch.float64, device='cuda'),
            torch.empty(n_i, dtype=torch.float64, device='cuda'),
            torch.empty(n_i, dtype=torch.float64, device='cuda'),
            torch.empty(n_j, dtype=torch.float64, device='cuda'))


def _call(torch, _native, dev, cost, G, noalloc):
    L = _native.lib()
    n_i, n_j = cost.shape
    plan, rows, f, g = _outputs(torch, n_i, n_j)
    st = _native.Stats()
    p = dev._ptr
    args = (p(cost.C), cost.ldc, p(cost.CT), cost.ldct, n_i, n_j, p(G), PRM['lambda1'], PRM['lambda2'],
            PRM['epsilon'], PRM['scaling_iter'], PRM['tau'], PRM['epsilon0'], PRM['extra_iter'],
            PRM['inner_iter_max'], p(plan), n_j, p(rows), p(f), p(g))
    if noalloc:
        rc = L.wotk_transport_stablev2_noalloc(*args, ctypes.byref(st), dev._stream(torch))
    else:
        ws = dev.workspace(torch, L.wotk_solver_workspace_size(n_i, n_j), cost.device)
        rc = L.wotk_transport_stablev2(*args, p(ws), ws.numel(), ctypes.byref(st), dev._stream(torch))
    _native.check(rc)
    torch.cuda.synchronize()
    return plan, rows, f, g, st.as_dict()


def test_noalloc_matches_workspace_entry_point_twice(mods):
    torch, ot, _native, dev = mods
    from wot_amd.synthetic import day_pair
    for (n_i, n_j, seed) in [(900, 1100, 3), (257, 130, 4)]:  # the second call is smaller
        dp = day_pair(n_i, n_j, dim=30, n_components=5, seed=seed)
        cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
        G = torch.from_numpy(dp.growth).cuda()
        ref = _call(torch, _native, dev, cost, G, noalloc=False)
        got = _call(torch, _native, dev, cost, G, noalloc=True)
        for a, b in zip(got[:4], ref[:4]):
            assert torch.equal(a, b), (n_i, n_j)
        assert got[4]['iterations'] == ref[4]['iterations'] and got[4]['epsilon_final'] == ref[4]['epsilon_final']
        # and both against the FP64 oracle (the workspace path's own gate)
        from oracle import ot_oracle as O
        Ro = O.transport_stablev2(O.default_cost_matrix(dp.x, dp.y), G=dp.growth, **PRM)
        R = got[0].cpu().numpy()
        assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-5


def test_noalloc_rejects_problems_beyond_capacity(mods):
    torch, ot, _native, dev = mods
    L = _native.lib()
    cap = L.wotk_scratch_capacity()
    n = 200000
    assert L.wotk_solver_workspace_size(n, n) > cap
    rc = L.wotk_transport_stablev2_noalloc(None, n, None, n, n, n, None, 1.0, 50.0, 0.05, 10, 1e4, 1.0, 0, 50,
                                           None, n, None, None, None, None, None)
    assert rc == _native.WOTK_EINVAL
    assert 'no-allocation entry point holds' in L.wotk_last_error().decode()
