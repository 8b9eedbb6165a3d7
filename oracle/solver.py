"""Conjugate gradients over fields, the linear_solver wrapper and the residual callback.

Restates ffthompy/general/solver.py:8-139,206-298 and
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


def cheby2TERM(A, B, x0, par=None, callback=None):
    """Two-term Chebyshev iteration over the eigenvalue range par['eigrange'] (solver.py:206-285):
    returns at once when |B - A x0| / |B| < tol; otherwise v = r - p v, x = x + w v,
    r = B - A x with the reference's (p, w) sequence, until |r| / |r0| <= tol or par['maxit']."""
    par = {} if par is None else par
    par.setdefault('tol', 1e-06)
    par.setdefault('maxit', 1e7)
    if 'eigrange' not in par:
        raise NotImplementedError("It is necessary to calculate eigenvalues.")
    e0, e1 = par['eigrange']
    res = {'kit': 0}
    bn = (B * B) ** 0.5
    x = x0
    r = B - A(x)
    r0 = np.double(r * r) ** 0.5
    res['norm_res'] = r0 / bn
    if res['norm_res'] < par['tol']:
        return x, res
    d = (e1 + e0) / 2.0
    c = (e1 - e0) / 2.0
    v = 0 * x0
    w = 0.
    while res['norm_res'] > par['tol'] and res['kit'] < par['maxit']:
        res['kit'] += 1
        if res['kit'] == 1:
            p = 0.
            w = 1 / d
        elif res['kit'] == 2:
            p = -(1 / 2) * (c / d) * (c / d)
            w = 1 / (d - c * c / 2 / d)
        else:
            p = -(c * c / 4) * w * w
            w = 1 / (d - c * c * w / 4)
        v = r - p * v
        x = x + w * v
        r = B - A(x)
        res['norm_res'] = (1.0 / r0) * r.norm()
        if callback is not None:
            callback(x)
    if res['kit'] == 0:
        res['norm_res'] = 0
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
    elif solver.lower() in ('chebyshev', 'cheby'):
        x, info = cheby2TERM(Afun, B, x0, par=par, callback=callback)
    else:
        raise NotImplementedError("This kind (%s) of linear solver is not implemented" % solver)
    info['time'] = time.perf_counter() - t0
    return x, info
