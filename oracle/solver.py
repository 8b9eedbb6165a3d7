"""Conjugate gradients over fields, the linear_solver wrapper and the residual callback.

Restates ffthompy/general/solver.py:8-139,287-298 and
ffthompy/general/solver_pp.py:6-34. Test infrastructure only.
"""
import time
import numpy as np


class CallBack:
    """Logs the true residual ||B - A x|| after every iteration, which costs one
    extra matvec per call (solver_pp.py:6-24)."""

    def __init__(self, A, B):
        self.A = A
        self.B = B
        self.iter = -1
        self.res_norm = []

    def __call__(self, x):
        self.iter += 1
        self.res_norm.append((self.B - self.A(x)).norm())


def CG(Afun, B, x0, par=None, callback=None):
    """Unpreconditioned CG with the absolute stop sqrt(<r,r>) <= tol, where <,> is the
    grid-mean inner product (solver.py:80-139)."""
    par = {} if par is None else par
    par.setdefault('tol', 1e-6)
    par.setdefault('maxiter', int(1e3))
    scal = par.get('scal', lambda X, Y: X * Y)   # get_scal for Tensor operands, solver.py:287-298
    res = {}
    x = x0
    R = B - Afun(x0)
    P = R
    rr = scal(R, R)
    res['kit'] = 0
    res['norm_res'] = np.double(rr) ** 0.5
    log = [res['norm_res']]
    while res['norm_res'] > par['tol'] and res['kit'] < par['maxiter']:
        res['kit'] += 1
        AP = Afun(P)
        alp = float(rr / scal(P, AP))
        x = x + alp * P
        R = R - alp * AP
        rrn = scal(R, R)
        bet = rrn / rr
        rr = rrn
        P = R + bet * P
        res['norm_res'] = np.double(rr) ** 0.5
        log.append(res['norm_res'])
        if callback is not None:
            callback(x)
    if res['kit'] == 0:
        res['norm_res'] = 0
    res['norm_res_log'] = log
    return x, res


def richardson(Afun, B, x0, par=None, callback=None):
    """x <- x + (1/alpha)(B - A x) (solver.py:63-77)."""
    omega = 1. / par['alpha']
    res = {'norm_res': 1e15, 'kit': 0}
    x = x0
    while res['norm_res'] > par['tol'] and res['kit'] < par['maxiter']:
        res['kit'] += 1
        r = B - Afun(x)
        x = x + omega * r
        res['norm_res'] = (r * r) ** 0.5
        if callback is not None:
            callback(x)
    return x, res


def linear_solver(Afun, B, ATfun=None, x0=None, par=None, solver=None, callback=None):
    """Dispatch wrapper: calls callback(x0) first, then the solver; stores wall time
    in info['time'] (solver.py:8-60)."""
    t0 = time.perf_counter()
    if x0 is None:
        x0 = B.zeros_like()
    if callback is not None:
        callback(x0)
    if solver.lower() == 'cg':
        x, info = CG(Afun, B, x0=x0, par=par, callback=callback)
    elif solver.lower() in ('iterative', 'richardson'):
        x, info = richardson(Afun, B, x0, par=par, callback=callback)
    else:
        raise NotImplementedError("This kind (%s) of linear solver is not implemented" % solver)
    info['time'] = time.perf_counter() - t0
    return x, info
