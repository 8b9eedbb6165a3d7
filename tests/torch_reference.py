"""Float64 PyTorch reference of the GaNi upper bound (test infrastructure).

AH(theta) for a = exp(theta) computed with torch.fft and a CG loop that autograd
differentiates through (no envelope argument): an independent check of the hand-written
backward. Projection G1 = xi xi^T/|xi|^2 on the 'r' half spectrum, grid-mean inner
product and absolute stop rule as in ffthompy/general/solver.py:80-139.
"""
import numpy as np
import torch


def g1_half(N, Y):
    xis = []
    for a, n in enumerate(N):
        k = np.fft.fftfreq(n, d=1. / n) if a < len(N) - 1 else np.arange(n // 2 + 1)
        xis.append(torch.tensor(k / Y[a], dtype=torch.float64))
    grids = torch.meshgrid(*xis, indexing='ij')
    q = sum(g * g for g in grids)
    q[(0,) * len(N)] = 1.
    d = len(N)
    G = torch.empty((d, d) + tuple(grids[0].shape), dtype=torch.float64)
    for i in range(d):
        for j in range(d):
            G[i, j] = grids[i] * grids[j] / q
    return G


def upper_bound_torch(theta, Y=None, tol=1e-13, maxiter=2000):
    a = torch.exp(theta)
    fields = fields_torch(theta, Y, tol, maxiter)
    d = len(fields)
    n = float(np.prod(theta.shape))
    AH = torch.zeros((d, d), dtype=torch.float64)
    for i in range(d):
        for j in range(d):
            AH[i, j] = torch.sum(a * torch.sum(fields[i] * fields[j], dim=0)) / n
    return AH


def fields_torch(theta, Y=None, tol=1e-13, maxiter=2000):
    """Unit-load solution fields e_i = E_i + x_i (list of (d,)+N tensors), differentiable
    through the CG iterations."""
    N = tuple(theta.shape)
    d = len(N)
    Y = np.ones(d) if Y is None else Y
    n = float(np.prod(N))
    a = torch.exp(theta)
    G = g1_half(N, Y)
    dims = tuple(range(1, d + 1))

    def Afun(x):
        Yh = torch.fft.rfftn(a * x, dim=dims)
        Z = torch.einsum('ij...,j...->i...', G.to(Yh.dtype), Yh)
        return torch.fft.irfftn(Z, s=N, dim=dims)

    scal = lambda u, v: torch.sum(u * v) / n
    fields = []
    for iL in range(d):
        E = torch.zeros((d,) + N, dtype=torch.float64)
        E[iL] = 1.
        B = Afun(-E)
        x = torch.zeros_like(B)
        r = B
        p = r
        rr = scal(r, r)
        k = 0
        while float(rr.detach()) ** 0.5 > tol and k < maxiter:
            Ap = Afun(p)
            alp = rr / scal(p, Ap)
            x = x + alp * p
            r = r - alp * Ap
            rrn = scal(r, r)
            p = r + (rrn / rr) * p
            rr = rrn
            k += 1
        fields.append(x + E)
    return fields
