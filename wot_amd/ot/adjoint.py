"""Differentiable OT objective: ``ot_objective(x, log_g, y, ...)`` returns the
converged regularised unbalanced OT objective (the reference's primal,
wot/ot/optimal_transport.py:45-53, at the optimal plan) as a torch scalar whose
backward runs the hand-written adjoint kernel (``wotk_ot_adjoint``).

The forward builds C = sqeuclid(x, y) / median on the device, G = exp(dt *
log_g) (ot_model.py:314-315 with g = exp(log_g)), solves to convergence and
evaluates the primal in FP64.  The backward uses the envelope theorem on the
converged potentials with a recomputed plan: dP/dC_ij = (R_ij - exp(-C_ij/eps))
/ (I J), chained through the median-normalised cost (including the median's
subgradient) to x, and through p and q = mean(p) to log g.
"""
import ctypes

from .. import _native
from . import device as _dev

DEFAULTS = dict(epsilon=0.05, lambda1=1.0, lambda2=50.0, epsilon0=1.0, tau=1e4, batch_size=5, tolerance=1e-8,
                max_iter=1e7, scaling_iter=3000, inner_iter_max=50, extra_iter=1000)


def _function():
    torch = _dev.torch_mod()

    class OTObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, log_g, y, solver, dt, params):
            X = x.detach().to(torch.float64).contiguous()
            Y = y.detach().to(torch.float64).contiguous()
            G = torch.exp(dt * log_g.detach().to(torch.float64)).contiguous()
            cost = _dev.DeviceCost.from_coordinates(X, Y)
            res = _dev.solve(cost, G, solver, params, want_plan=False, want_potentials=True)
            eps = res['stats']['epsilon_final']
            pri, _ = _dev.primal_dual(cost, G, res['f'], res['g'], params['lambda1'], params['lambda2'], eps)
            ctx.state = (cost, X, Y, G, res['f'], res['g'], eps, dt, params)
            ctx.dtypes = (x.dtype, log_g.dtype)
            ctx.stats = res['stats']
            return torch.tensor(pri, dtype=torch.float64, device=x.device)

        @staticmethod
        def backward(ctx, grad_out):
            cost, X, Y, G, f, g, eps, dt, params = ctx.state
            gx, glg = adjoint(cost, X, Y, G, f, g, params['lambda1'], params['lambda2'], eps, dt,
                              want_x=ctx.needs_input_grad[0], want_log_g=ctx.needs_input_grad[1])
            out_x = (grad_out * gx).to(ctx.dtypes[0]) if gx is not None else None
            out_g = (grad_out * glg).to(ctx.dtypes[1]) if glg is not None else None
            return out_x, out_g, None, None, None, None

    return OTObjective


_FN = None


def ot_objective(x, log_g, y, solver='duality_gap', dt=1.0, **params):
    """Converged OT objective between source cells ``x`` (I x d) with log growth
    ``log_g`` (I) and target cells ``y`` (J x d); differentiable in x and log_g.
    Keyword arguments override the OT configuration (DEFAULTS, the reference's
    OTModel defaults)."""
    global _FN
    if _FN is None:
        _FN = _function()
    prm = dict(DEFAULTS, **params)
    val = _FN.apply(x, log_g, y, solver, float(dt), prm)
    return val.to(x.dtype) if x.dtype != val.dtype else val


def adjoint(cost, X, Y, G, f, g, lambda1, lambda2, epsilon, dt=1.0, want_x=True, want_log_g=True):
    """Raw adjoint call: (dOT/dX (I x d float64) or None, dOT/dlog g (I) or None)."""
    torch = _dev.torch_mod()
    n_i, n_j = cost.shape
    d = X.shape[1]
    dev = cost.device
    L = _native.lib()
    ws = _dev.workspace(torch, L.wotk_adjoint_workspace_size(n_i, n_j, d), dev)
    gx = torch.empty((n_i, d), dtype=torch.float64, device=dev) if want_x else None
    glg = torch.empty(n_i, dtype=torch.float64, device=dev) if want_log_g else None
    p = _dev._ptr
    _native.check(L.wotk_ot_adjoint(p(cost.C), cost.ldc, p(cost.CT), cost.ldct, n_i, n_j, p(X), p(Y), d, p(G),
                                    float(dt), p(f), p(g), float(lambda1), float(lambda2), float(epsilon),
                                    p(cost.median), p(cost.median_index), p(gx), p(glg), p(ws), ws.numel(),
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return gx, glg
