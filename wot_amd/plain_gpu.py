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
