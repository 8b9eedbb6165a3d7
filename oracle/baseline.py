"""Reference-style CPU path for the cpu_baseline leg of bench.py (test infrastructure).

Operation for operation what FFTHomPy does per CG iteration of the scalar GaNi solve:
the materialised 'r'-form projection hG1 (ffthompy/projections.py:9-112) and the
full (d,d)+N coefficient (ffthompy/materials.py:115-123) applied with einsum, FFTs
with numpy.fft.rfftn/irfftn (ffthompy/tensors/fft.py:39-43), an allocating Tensor-style
CG (ffthompy/general/solver.py:80-139) and the CallBack true-residual matvec
(ffthompy/general/solver_pp.py:13-24). numpy.fft and these einsums run on one core.
"""
import time

import numpy as np

from . import grid


def materialised_G1_half(N, Y):
    """hG1 in the 'r' form, (d,d)+N[:-1]+(N[-1]//2+1,), values xi_m xi_n/|xi|^2
    (zero at xi=0); equal to projections.scalar(...)[1].val for odd N."""
    xi = grid.frequencies(N, Y, fft_form='r')
    d = len(N)
    shp = tuple(len(x) for x in xi)
    X = [np.broadcast_to(xi[m].reshape([-1 if a == m else 1 for a in range(d)]), shp) for m in range(d)]
    q = sum(x * x for x in X)
    q = np.array(q)
    q[(0,) * d] = 1.
    G = np.empty((d, d) + shp)
    for m in range(d):
        for n in range(d):
            G[m, n] = X[m] * X[n] / q
    return G


class RefStyleOperator:
    """Afun(x) = irfftn(hG1 . rfftn(A . x)) with fresh arrays at every step. workers=None: numpy.fft
    (the reference's, one core); workers=k: scipy.fft with k threads (the tuned CPU variant)."""

    def __init__(self, a, Y, workers=None):
        self.N = a.shape
        d = len(self.N)
        self.A = np.einsum('ij,...->ij...', np.eye(d), a)
        self.G = materialised_G1_half(self.N, Y)
        self.axes = tuple(range(1, d + 1))
        self.workers = workers

    def __call__(self, x):
        y = np.einsum('ij...,j...->i...', self.A, x)
        if self.workers is None:
            Y = np.fft.rfftn(y, s=self.N, axes=self.axes)
        else:
            import scipy.fft
            Y = scipy.fft.rfftn(y, s=self.N, axes=self.axes, workers=self.workers)
        Y = np.einsum('ij...,j...->i...', self.G, Y)
        if self.workers is None:
            return np.fft.irfftn(Y, s=self.N, axes=self.axes) + 0.
        import scipy.fft
        return scipy.fft.irfftn(Y, s=self.N, axes=self.axes, workers=self.workers) + 0.


def time_cg_iterations(a, Y, iters=2, callback=True, workers=None):
    """Run `iters` reference-style CG iterations on the load E = e_1 and return
    seconds per iteration (with the CallBack matvec when callback=True; FFTs as RefStyleOperator)."""
    Afun = RefStyleOperator(a, Y, workers=workers)
    d = a.ndim
    n = float(np.prod(a.shape))
    EN = np.zeros((d,) + a.shape)
    EN[0] = 1.
    B = Afun(-EN)
    scal = lambda u, v: np.sum(u * v) / n
    x = np.zeros_like(B)
    R = B - Afun(x)
    P = R
    rr = scal(R, R)
    t0 = time.perf_counter()
    for _ in range(iters):
        AP = Afun(P)
        alp = float(rr / scal(P, AP))
        x = x + alp * P
        R = R - alp * AP
        rrn = scal(R, R)
        P = R + (rrn / rr) * P
        rr = rrn
        if callback:
            res = B - Afun(x)
            _ = scal(res, res) ** 0.5
    return (time.perf_counter() - t0) / iters
