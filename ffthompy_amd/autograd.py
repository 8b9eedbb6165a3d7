"""Differentiable homogenized matrices and solution fields: torch.autograd.Function over the
HIP solve.

Forward: for a voxel field theta (torch float64, shape N), a(x) = exp(theta(x)), solve
the d unit-load GaNi problems on the GPU and return the primal matrix AH (the upper
bound, ffthompy/postprocess.py:53-70). Backward: the solution fields e_i minimise the
energy, so with the fields held fixed dAH_ij/da(x) = e_i(x).e_j(x)/n (envelope
theorem, also for i != j by polarisation); the hand-written kernel fh_energy_gradient
evaluates sum_ij W_ij e_i.e_j a(x)/n for W = dL/dAH in one pass over the stored
fields. No adjoint solve is needed for losses on AH.

SolutionFields returns the unit-load fields e_i = E_i + x_i themselves; a loss on them needs
one adjoint CG per load: lam_i solves G(a lam_i) = G dL/de_i (the same operator, fp64), and
dL/da(x) = -sum_i lam_i(x) . e_i(x) (kernel fh_field_gradient).
"""
import numpy as np
import torch

from . import _native as nat


class UpperBound(torch.autograd.Function):
    """AH(theta) with a = exp(theta), scalar conductivity on an N grid of cell Y."""

    @staticmethod
    def forward(ctx, theta, Y, tol, maxiter, precision):
        N = tuple(int(n) for n in theta.shape)
        d = len(N)
        a = torch.exp(theta.detach()).to('cpu', torch.float64).numpy()
        plan = nat.Plan(N, np.asarray(Y, dtype=np.float64), 'scalar', 'primal')
        plan.set_coeff('iso', a)
        fields, kits = [], []
        B = nat.DeviceArray((d,) + N)
        for iL in range(d):
            E = np.zeros(d)
            E[iL] = 1.
            EN = nat.DeviceArray.zeros((d,) + N)
            nat.add_constant(EN, -E)
            plan.matvec(EN, B)
            x = nat.DeviceArray.zeros((d,) + N)
            info = plan.cg_solve(B, x, tol=tol, maxiter=maxiter, mode=precision)
            if info['status'] != nat.FH_CG_OK:
                raise FloatingPointError('CG breakdown')
            kits.append(info['kit'])
            nat.add_constant(x, E)
            fields.append(x)
        AH = plan.energy_matrix(fields)
        ctx.plan, ctx.fields = plan, fields
        ctx.a_dev = nat.DeviceArray.from_host(a)
        ctx.theta_device = theta.device
        ctx.kits = kits
        return torch.from_numpy(AH).to(theta.device)

    @staticmethod
    def backward(ctx, gAH):
        W = gAH.detach().to('cpu', torch.float64).numpy()
        g = ctx.plan.energy_gradient(ctx.fields, W, chain=ctx.a_dev)
        return torch.from_numpy(g.to_host()).to(ctx.theta_device), None, None, None, None


def upper_bound(theta, Y=None, tol=1e-10, maxiter=5000, precision='fp64'):
    """Differentiable primal homogenized matrix AH(theta), a = exp(theta)."""
    if Y is None:
        Y = np.ones(theta.dim())
    return UpperBound.apply(theta, Y, tol, maxiter, precision)


def _solve_unit_loads(theta, Y, tol, maxiter, precision):
    N = tuple(int(n) for n in theta.shape)
    d = len(N)
    a = torch.exp(theta.detach()).to('cpu', torch.float64).numpy()
    plan = nat.Plan(N, np.asarray(Y, dtype=np.float64), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    fields = []
    B = nat.DeviceArray((d,) + N)
    for iL in range(d):
        E = np.zeros(d)
        E[iL] = 1.
        EN = nat.DeviceArray.zeros((d,) + N)
        nat.add_constant(EN, -E)
        plan.matvec(EN, B)
        x = nat.DeviceArray.zeros((d,) + N)
        info = plan.cg_solve(B, x, tol=tol, maxiter=maxiter, mode=precision)
        if info['status'] != nat.FH_CG_OK:
            raise FloatingPointError('CG breakdown')
        nat.add_constant(x, E)
        fields.append(x)
    return plan, a, fields


class SolutionFields(torch.autograd.Function):
    """e (d, d) + N: e[i] = E_i + x_i for a = exp(theta); backward by adjoint CG solves."""

    @staticmethod
    def forward(ctx, theta, Y, tol, maxiter):
        plan, a, fields = _solve_unit_loads(theta, Y, tol, maxiter, 'fp64')
        ctx.plan, ctx.fields, ctx.tol, ctx.maxiter = plan, fields, tol, maxiter
        ctx.a_dev = nat.DeviceArray.from_host(a)
        ctx.theta_device = theta.device
        return torch.from_numpy(np.stack([f.to_host() for f in fields])).to(theta.device)

    @staticmethod
    def backward(ctx, g):
        plan = ctx.plan
        g = g.detach().to('cpu', torch.float64).numpy()
        adj = []
        for i, f in enumerate(ctx.fields):
            gi = nat.DeviceArray.from_host(np.ascontiguousarray(g[i]))
            b = nat.DeviceArray(f.shape)
            plan.project(gi, b)
            lam = nat.DeviceArray.zeros(f.shape)
            info = plan.cg_solve(b, lam, tol=ctx.tol, maxiter=ctx.maxiter, mode='fp64')
            if info['status'] != nat.FH_CG_OK:
                raise FloatingPointError('adjoint CG breakdown')
            adj.append(lam)
        grad = plan.field_gradient(ctx.fields, adj, chain=ctx.a_dev)
        return torch.from_numpy(grad.to_host()).to(ctx.theta_device), None, None, None


def solution_fields(theta, Y=None, tol=1e-10, maxiter=5000):
    """Differentiable unit-load solution fields e (d, d)+N for a = exp(theta) (fp64)."""
    if Y is None:
        Y = np.ones(theta.dim())
    return SolutionFields.apply(theta, Y, tol, maxiter)
