"""Plain-GPU baseline: the reference solver restated literally in torch FP64
(dense stabilised kernel K, torch.mv for K b and K^T a, a full K rebuild at
every absorption / epsilon change), one day pair at a time.

This is the "plain GPU Sinkhorn" the fused HIP path is measured against; it
follows wot/ot/optimal_transport.py:167-236 statement for statement.
"""
import numpy as np


def transport_stablev2(C, lambda1, lambda2, epsilon, scaling_iter, G, tau, epsilon0, extra_iter, inner_iter_max,
                       device='cuda', **ignored):
    import torch
    C = torch.as_tensor(np.asarray(C, dtype=np.float64), device=device) if not isinstance(C, torch.Tensor) \
        else C.to(device=device, dtype=torch.float64)
    G = torch.as_tensor(np.asarray(G, dtype=np.float64), device=device) if not isinstance(G, torch.Tensor) \
        else G.to(device=device, dtype=torch.float64)
    n_i, n_j = C.shape
    eps_final = epsilon

    def reg(n):
        return (epsilon0 - eps_final) * np.exp(-n) + eps_final

    eps_i = epsilon0
    dx = torch.full((n_i,), 1.0 / n_i, dtype=torch.float64, device=device)
    dy = torch.full((n_j,), 1.0 / n_j, dtype=torch.float64, device=device)
    p = G
    q = torch.full((n_j,), float(G.mean()), dtype=torch.float64, device=device)
    u = torch.zeros(n_i, dtype=torch.float64, device=device)
    v = torch.zeros(n_j, dtype=torch.float64, device=device)
    b = torch.ones(n_j, dtype=torch.float64, device=device)
    K = torch.exp(-C / eps_i)
    al1 = lambda1 / (lambda1 + eps_i)
    al2 = lambda2 / (lambda2 + eps_i)
    since = 0
    n_eps = 0
    a = None

    def rebuild():
        return torch.exp((u[:, None] - C + v[None, :]) / eps_i)

    for _ in range(scaling_iter):
        a = (p / torch.mv(K, b * dy)) ** al1 * torch.exp(-u / (lambda1 + eps_i))
        b = (q / torch.mv(K.T, a * dx)) ** al2 * torch.exp(-v / (lambda2 + eps_i))
        since += 1
        if max(float(a.abs().max()), float(b.abs().max())) > tau:
            u = u + eps_i * torch.log(a)
            v = v + eps_i * torch.log(b)
            K = rebuild()
            a = torch.ones_like(a)
            b = torch.ones_like(b)
        if since == inner_iter_max:
            n_eps += 1
            since = 0
            u = u + eps_i * torch.log(a)
            v = v + eps_i * torch.log(b)
            eps_i = reg(n_eps)
            al1 = lambda1 / (lambda1 + eps_i)
            al2 = lambda2 / (lambda2 + eps_i)
            K = rebuild()
            a = torch.ones_like(a)
            b = torch.ones_like(b)
    for _ in range(extra_iter):
        a = (p / torch.mv(K, b * dy)) ** al1 * torch.exp(-u / (lambda1 + eps_i))
        b = (q / torch.mv(K.T, a * dx)) ** al2 * torch.exp(-v / (lambda2 + eps_i))
    R = (K.T * a).T * b
    return R / n_j


def optimal_transport_duality_gap(C, G, lambda1, lambda2, epsilon, batch_size, tolerance, tau, epsilon0, max_iter,
                                  device='cuda', stats=None, **ignored):
    """The reference's default solver (optimal_transport.py:67-164) restated in
    torch FP64: dense stabilised K and K~ = exp(-C/eps), torch.mv per half
    iteration, dual-variable evolution checks on the first five epsilon
    scalings and primal / dual (optimal_transport.py:45-62, with
    nan_to_num(log R)) every ``batch_size`` iterations of the last one."""
    import torch
    C = torch.as_tensor(np.asarray(C, dtype=np.float64), device=device) if not isinstance(C, torch.Tensor) \
        else C.to(device=device, dtype=torch.float64)
    G = torch.as_tensor(np.asarray(G, dtype=np.float64), device=device) if not isinstance(G, torch.Tensor) \
        else G.to(device=device, dtype=torch.float64)
    n_scal = 5
    factor = float(np.exp(-np.log(epsilon) / n_scal))
    n_i, n_j = C.shape
    dx = torch.full((n_i,), 1.0 / n_i, dtype=torch.float64, device=device)
    dy = torch.full((n_j,), 1.0 / n_j, dtype=torch.float64, device=device)
    p = G
    q = torch.full((n_j,), float(G.mean()), dtype=torch.float64, device=device)
    u = torch.zeros(n_i, dtype=torch.float64, device=device)
    v = torch.zeros(n_j, dtype=torch.float64, device=device)
    a = torch.ones(n_i, dtype=torch.float64, device=device)
    b = torch.ones(n_j, dtype=torch.float64, device=device)
    eps_i = epsilon0 * factor
    n_iter = n_stab = n_final_checks = 0
    gap = float('inf')
    R = None

    def fdiv(l, x, pp, d):
        return l * torch.sum(d * (x * torch.log(x / pp) - x + pp))

    def fdivstar(l, w, pp, d):
        return l * torch.sum((pp * d) * (torch.exp(w / l) - 1))

    def stab():
        return torch.exp((u[:, None] - C + v[None, :]) / eps_i)

    for e in range(n_scal + 1):
        gap = float('inf')
        u = u + eps_i * torch.log(a)
        v = v + eps_i * torch.log(b)
        eps_i = eps_i / factor
        Kt = torch.exp(-C / eps_i)
        al1 = lambda1 / (lambda1 + eps_i)
        al2 = lambda2 / (lambda2 + eps_i)
        K = stab()
        a = torch.ones(n_i, dtype=torch.float64, device=device)
        b = torch.ones(n_j, dtype=torch.float64, device=device)
        old_a, old_b = a, b
        thr = tolerance if e == n_scal else 1e-6
        while gap > thr:
            for _ in range(batch_size if e == n_scal else 5):
                n_iter += 1
                old_a, old_b = a, b
                a = (p / torch.mv(K, b * dy)) ** al1 * torch.exp(-u / (lambda1 + eps_i))
                b = (q / torch.mv(K.T, a * dx)) ** al2 * torch.exp(-v / (lambda2 + eps_i))
                if max(float(a.abs().max()), float(b.abs().max())) > tau:
                    u = u + eps_i * torch.log(a)
                    v = v + eps_i * torch.log(b)
                    K = stab()
                    a = torch.ones(n_i, dtype=torch.float64, device=device)
                    b = torch.ones(n_j, dtype=torch.float64, device=device)
                    n_stab += 1
                if n_iter >= max_iter:
                    if stats is not None:
                        stats.update(iterations=n_iter, stabilisations=n_stab, final_checks=n_final_checks,
                                     max_iter_reached=True)
                    return (K.T * a).T * b
            a_true = a * torch.exp(u / eps_i)
            b_true = b * torch.exp(v / eps_i)
            if e == n_scal:
                n_final_checks += 1
                R = (K.T * a).T * b
                logR = torch.nan_to_num(torch.log(R))
                pri = (fdiv(lambda1, R @ dy, p, dx) + fdiv(lambda2, R.T @ dx, q, dy)
                       + (eps_i * torch.sum(R * logR - R + Kt) + torch.sum(R * C)) / (n_i * n_j))
                del logR
                dua = (-fdivstar(lambda1, -eps_i * torch.log(a_true), p, dx)
                       - fdivstar(lambda2, -eps_i * torch.log(b_true), q, dy)
                       - eps_i * torch.sum(R - Kt) / (n_i * n_j))
                gap = float((pri - dua) / torch.abs(pri))
            else:
                gap = max(float(torch.linalg.norm(a_true - old_a * torch.exp(u / eps_i))
                                / (1 + torch.linalg.norm(a_true))),
                          float(torch.linalg.norm(b_true - old_b * torch.exp(v / eps_i))
                                / (1 + torch.linalg.norm(b_true))))
    if np.isnan(gap):
        raise RuntimeError("Overflow encountered in duality gap computation, please report this incident")
    if stats is not None:
        stats.update(iterations=n_iter, stabilisations=n_stab, final_checks=n_final_checks, gap=gap,
                     max_iter_reached=False)
    return R / n_j
