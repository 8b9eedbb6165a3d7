"""Linear solvers on the GPU operator: linear_solver, CG, Richardson, Chebyshev, SciPy's CG and
the residual callback.

Mirrors ffthompy/general/solver.py:8-285 and ffthompy/general/solver_pp.py:6-34. With a
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
    every iteration with an extra matvec; the device solve by default records
    sqrt(<r,r>) of the CG residual after every iteration (the same quantity up to
    rounding), with no extra work. res_norm[0] is the initial residual. With
    par['true_residual_every'] = k (fp64 CG) the device evaluates ||B - A x|| itself at x0, every
    k-th iteration and the last (one extra fp64 matvec each): res_norm holds those values,
    res_kit their iterations, res_norm_recursive the recursive log."""

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
    """Solve Afun x = B (solver.py:8-60): 'CG', 'richardson'/'iterative' and 'chebyshev'/'cheby' run
    on the device; 'scipy_cg' ('scipy.sparse.linalg.cg') runs SciPy's CG on the host with the
    GPU matvec as its LinearOperator, as the reference does. 'BiCG' is not provided (the
    reference's BiCG needs matrix operations that Tensor lacks, solver.py:174-184)."""
    par = {} if par is None else par
    if not isinstance(Afun, GpuOperator):
        raise TypeError('linear_solver runs on a GpuOperator (the HIP matvec)')
    if solver is None:
        raise NotImplementedError("This kind (%s) of linear solver is not implemented" % solver)
    t0 = time.perf_counter()
    kind = solver.lower()
    if callback is not None and not isinstance(callback, CallBack) and not kind.startswith('scipy'):
        # the reference calls callback(x0) before the solver (solver.py:15-17); a CallBack's log
        # starts with the initial residual instead
        callback(B.zeros_like(name='x0') if x0 is None else x0)
    if kind == 'cg':
        x, info = CG(Afun, B, x0=x0, par=par, callback=callback)
    elif kind in ('iterative', 'richardson'):
        x, info = richardson(Afun, B, x0, par=par, callback=callback)
    elif kind in ('chebyshev', 'cheby'):
        x, info = cheby2TERM(Afun, B, x0=x0, par=par, callback=callback)
    elif solver in ('scipy.sparse.linalg.cg', 'scipy_cg'):
        x, info = scipy_cg(Afun, B, x0=x0, par=par, callback=callback)
    else:
        raise NotImplementedError("This kind (%s) of linear solver is not implemented" % solver)
    info['time'] = time.perf_counter() - t0
    return x, info


class _IterateCallback:
    """callback(x) on the device iterate after every iteration of a device solve (the reference's
    callback(xCG), solver.py:134-136, and callback(x) of richardson / cheby2TERM, solver.py:75-76,
    277-278): registered on the plan for one solve; x is handed over as a fresh device Tensor (a
    snapshot, as the reference's functional iterates are). An exception raised by the callback stops
    the solve and is re-raised here."""

    def __init__(self, Afun, B, callback):
        self.plan, self.B, self.callback = Afun.plan, B, callback

    def __enter__(self):
        if self.callback is None or isinstance(self.callback, CallBack):
            return self
        B, cb = self.B, self.callback
        self.plan.set_iterate_callback(lambda kit, x: cb(B._like(x.copy(), name='x')), every=1)
        return self

    def __exit__(self, etype, exc, tb):
        if self.callback is None or isinstance(self.callback, CallBack):
            return False
        self.plan.set_iterate_callback(None)
        err = self.plan.pop_callback_error()
        if err is not None and etype is nat.FhError:
            raise err
        return False


def _iterate(method, Afun, B, x0, coefs, tol, maxiter, callback):
    B, x = _start(B, x0)
    hist_len = int(maxiter) + 1 if isinstance(callback, CallBack) else 0
    with _IterateCallback(Afun, B, callback):
        out = Afun.plan.iter_solve(method, B.val, x.val, coefs, float(tol), int(maxiter), hist_len=hist_len)
    info = {'kit': out['kit'], 'norm_res': out['norm_res']}
    if isinstance(callback, CallBack):
        callback.record(out['res_log'])
    return x, info


def richardson(Afun, B, x0=None, par=None, callback=None):
    """x <- x + (B - A x)/alpha until sqrt(<r,r>) <= tol (solver.py:63-77), on the device.
    par: 'alpha', 'tol', 'maxiter'. CallBack records the norm of the residual taken before
    each step (the quantity the reference tests)."""
    return _iterate('richardson', Afun, B, x0, [par['alpha']], par['tol'], par['maxiter'], callback)


def cheby2TERM(Afun, B, x0=None, par=None, callback=None):
    """Two-term Chebyshev iteration (solver.py:206-285) on the device. par: 'eigrange' (required),
    'tol' (default 1e-6), 'maxit' (default 1e7); norm_res is relative to the initial residual."""
    par = {} if par is None else par
    par.setdefault('tol', 1e-06)
    par.setdefault('maxit', 1e7)
    if 'eigrange' not in par:
        raise NotImplementedError("It is necessary to calculate eigenvalues.")
    maxit = int(min(par['maxit'], 2 ** 31 - 1))
    return _iterate('chebyshev', Afun, B, x0, par['eigrange'], par['tol'], maxit, callback)


def scipy_cg(Afun, B, x0=None, par=None, callback=None):
    """scipy.sparse.linalg.cg with the GPU operator (solver.py:28-44): SciPy iterates on the host,
    every matvec round-trips through the device. Its stop rule is SciPy's (relative residual)."""
    import scipy.sparse.linalg as spslin
    shape = B.val.shape
    n = int(np.prod(shape))
    dx = nat.DeviceArray(shape)
    dy = nat.DeviceArray(shape)

    def mv(v):
        h = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(shape))
        nat.call('fh_memcpy_htod', dx.ptr, h.ctypes.data, h.nbytes)
        Afun.plan.matvec(dx, dy)
        return dy.to_host().reshape(-1)

    op = spslin.LinearOperator((n, n), matvec=mv, dtype=np.float64)
    b = B.to_host().val.reshape(-1)
    x0v = np.zeros(n) if x0 is None else x0.to_host().val.reshape(-1)
    cb = callback
    if isinstance(callback, CallBack):
        # SciPy hands over the iterate as a flat vector: the CallBack gets it as a device Tensor and
        # logs ||B - A x|| with its own A and B (this operator and right-hand side if unset)
        if callback.A is None:
            callback.A = Afun
        if callback.B is None:
            callback.B = B.to_device()
        Bd = B.to_device()

        def cb(xk):
            callback(Bd._like(nat.DeviceArray.from_host(np.asarray(xk, dtype=np.float64).reshape(shape)), name='x'))
    xcol, code = spslin.cg(op, b, x0=x0v, rtol=par['tol'], maxiter=par['maxiter'], M=None, callback=cb)
    x = B.zeros_like(name='x').to_host()
    x.val = xcol.reshape(shape)
    return x.to_device(), {'info': code}


def _start(B, x0):
    """Device copies of the right-hand side and the initial guess; x0 must match B."""
    B = B.to_device()
    if x0 is None:
        return B, B.zeros_like(name='x')
    if tuple(x0.val.shape) != tuple(B.val.shape) or np.dtype(x0.val.dtype) != np.dtype(B.val.dtype):
        raise ValueError('x0 (shape %s, %s) does not match B (shape %s, %s)'
                         % (tuple(x0.val.shape), x0.val.dtype, tuple(B.val.shape), B.val.dtype))
    return B, x0.to_device().copy(name='x')


def CG(Afun, B, x0=None, par=None, callback=None):
    """Conjugate gradients on the device (solver.py:80-139)."""
    par = {} if par is None else par
    par.setdefault('tol', 1e-6)
    par.setdefault('maxiter', int(1e3))
    mode = par.get('precision', 'fp64')
    nat.solve_mode(mode)
    B, x = _start(B, x0)
    hist_len = int(par['maxiter']) + 1 if isinstance(callback, CallBack) else 0
    every = int(par.get('true_residual_every', 0)) if isinstance(callback, CallBack) else 0
    Afun.plan.set_residual_log(every)
    try:
        with _IterateCallback(Afun, B, callback):
            out = Afun.plan.cg_solve(B.val, x.val, tol=float(par['tol']), maxiter=int(par['maxiter']), mode=mode,
                                     delta=float(par.get('delta', 1e-4)), hist_len=hist_len)
    finally:
        if every:
            Afun.plan.set_residual_log(0)
    if out['status'] == nat.FH_CG_BREAKDOWN:
        raise FloatingPointError('CG breakdown: <p, A p> <= 0 (coefficient not positive definite?)')
    info = {'kit': out['kit'], 'norm_res': out['norm_res']}
    if mode == 'mixed':
        info['reliable_updates'] = Afun.plan.last_updates()
    if isinstance(callback, CallBack):
        callback.record(out['res_log'])
        if every:
            # the reference's CallBack quantity ||B - A x|| (fp64 operator), iterations in res_kit
            callback.res_norm_recursive = callback.res_norm
            kits, norms = Afun.plan.residual_log()
            callback.res_kit = [int(k) for k in kits]
            callback.res_norm = [float(v) for v in norms]
    return x, info
