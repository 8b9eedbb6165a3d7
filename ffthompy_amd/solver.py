"""Linear solvers on the GPU operator: linear_solver, CG and the residual callback.

Mirrors ffthompy/general/solver.py:8-139 and ffthompy/general/solver_pp.py:6-34. With a
GpuOperator the whole CG loop runs on the device (fh_cg_solve): same algorithm, same
absolute stop sqrt(<r,r>) <= tol on the grid-mean inner product, x0 honoured, info with
'kit', 'norm_res' and 'time'. Extra keys of ``par`` select the arithmetic:
'precision' = 'fp64' (default) or 'mixed' (fp32 CG with fp64 reliable updates),
'delta' = reliable-update threshold of the mixed mode.
"""
import time

import numpy as np

from . import _native as nat
from .tensors import GpuOperator, Tensor


class CallBack:
    """Residual log of a solve (solver_pp.py:6-24). The reference evaluates B - A x after
    every iteration with an extra matvec; the device solve instead records
    sqrt(<r,r>) of the CG residual after every iteration (the same quantity up to
    rounding), with no extra work. res_norm[0] is the initial residual."""

    def __init__(self, A=None, B=None, **kwargs):
        self.__dict__.update(kwargs)
        self.A, self.B = A, B
        self.iter = -1
        self.res_norm = []

    def record(self, log):
        self.res_norm = [float(v) for v in log]
        self.iter = len(self.res_norm) - 1

    def __call__(self, x):
        """Host-side evaluation ||B - A x|| (one extra matvec), as the reference does."""
        self.iter += 1
        self.res_norm.append((self.B - self.A(x)).norm())

    def __repr__(self):
        if not self.res_norm:
            return 'the results are not initialized yet'
        return '    iterations : %d\n    res_norm : %g\n' % (self.iter, self.res_norm[-1])


def linear_solver(Afun, B, ATfun=None, x0=None, par=None, solver=None, callback=None):
    """Solve Afun x = B (solver.py:8-60). Only 'CG' on a GpuOperator is provided."""
    par = {} if par is None else par
    if solver is None or solver.lower() != 'cg':
        raise NotImplementedError("This kind (%s) of linear solver is not implemented" % solver)
    if not isinstance(Afun, GpuOperator):
        raise TypeError('linear_solver runs on a GpuOperator (the HIP matvec)')
    t0 = time.perf_counter()
    x, info = CG(Afun, B, x0=x0, par=par, callback=callback)
    info['time'] = time.perf_counter() - t0
    return x, info


def CG(Afun, B, x0=None, par=None, callback=None):
    """Conjugate gradients on the device (solver.py:80-139)."""
    par = {} if par is None else par
    par.setdefault('tol', 1e-6)
    par.setdefault('maxiter', int(1e3))
    mode = par.get('precision', 'fp64')
    B = B.to_device()
    x = B.zeros_like(name='x') if x0 is None else x0.to_device().copy(name='x')
    hist_len = int(par['maxiter']) + 1 if callback is not None else 0
    out = Afun.plan.cg_solve(B.val, x.val, tol=float(par['tol']), maxiter=int(par['maxiter']), mode=mode,
                             delta=float(par.get('delta', 0.01)), hist_len=hist_len)
    if out['status'] == nat.FH_CG_BREAKDOWN:
        raise FloatingPointError('CG breakdown: <p, A p> <= 0 (coefficient not positive definite?)')
    info = {'kit': out['kit'], 'norm_res': out['norm_res']}
    if mode == 'mixed':
        info['reliable_updates'] = Afun.plan.last_updates()
    if callback is not None:
        if isinstance(callback, CallBack):
            callback.record(out['res_log'])
        else:
            callback(x)
    return x, info
