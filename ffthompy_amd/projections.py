"""Fourier projections G0, G1, G2 as operators that are never materialised.

Mirrors the interface of ffthompy/projections.py:9-112 (scalar) and 114-267
(elasticity): the returned objects act on 'r'-form spectra held on the device
(Tensor with Fourier=True) through fh_apply_projection_hat, which evaluates G(xi) from
the frequency index. For elasticity the reference returns the parts (G1h, G1s, G2h,
G2s) whose sums the driver uses (ffthompy/applications.py:121-122); here the sums
G1 = G1h + G1s and G2 = G2h + G2s are returned directly.
"""
import numpy as np

from . import _native as nat
from .tensors import Tensor


class FourierProjection:
    """G(xi) acting on an 'r'-form spectrum (the einsum 'ij...,j...->i...' of
    ffthompy/tensors/objects.py:223-232)."""

    def __init__(self, N, Y, physics, which, name):
        self.N = tuple(int(n) for n in N)
        self.Y = np.asarray(Y, dtype=np.float64)
        self.physics = physics
        self.which = which
        self.name = name
        self.plan = nat.Plan(self.N, self.Y, physics, 'primal' if which == 1 else 'dual')
        self.ncomp = self.plan.ncomp

    def __call__(self, X):
        assert X.Fourier and X.fft_form == 'r'
        X = X.to_device()
        out = nat.DeviceArray(X.val.shape, X.val.dtype)
        self.plan.apply_projection_hat(X.val, out)
        return X._like(out, name='{0}({1})'.format(self.name, X.name[:10]))

    def __repr__(self):
        return 'FourierProjection(%s, N=%s, %s)' % (self.name, self.N, self.physics)


class MeanProjection:
    """G0: keeps the zero-frequency coefficient only (projections.py:62)."""

    def __init__(self, N, name='hG0'):
        self.N = tuple(int(n) for n in N)
        self.name = name

    def __call__(self, X):
        host = X.to_host().val
        out = np.zeros_like(host)
        out[(Ellipsis,) + (0,) * len(self.N)] = host[(Ellipsis,) + (0,) * len(self.N)]
        return X._like(nat.DeviceArray.from_host(out))


def scalar(N, Y, NyqNul=True, tensor=True, fft_form='r'):
    """(G0, G1, G2) of the scalar problem. Nyquist modes of even grids are nulled
    (NyqNul=True, the reference default; False is not supported)."""
    if not NyqNul or fft_form != 'r':
        raise NotImplementedError("only NyqNul=True and fft_form 'r'")
    return (MeanProjection(N), FourierProjection(N, Y, 'scalar', 1, 'hG1'),
            FourierProjection(N, Y, 'scalar', 2, 'hG2'))


def elasticity(N, Y, NyqNul=True, tensor=True, fft_form='r'):
    """(G0, G1, G2) of small-strain elasticity in Mandel notation."""
    if not NyqNul or fft_form != 'r':
        raise NotImplementedError("only NyqNul=True and fft_form 'r'")
    return (MeanProjection(N), FourierProjection(N, Y, 'elasticity', 1, 'hG1'),
            FourierProjection(N, Y, 'elasticity', 2, 'hG2'))


class Operator:
    """Composition of linear maps applied right to left (ffthompy/tensors/operators.py:108-144)."""

    def __init__(self, name='Operator', mat=None):
        self.name = name
        self.mat_rev = [list(reversed(s)) for s in mat]

    def __call__(self, x):
        res = None
        for summand in self.mat_rev:
            prod = x
            for op in summand:
                prod = op(prod)
            res = prod if res is None else res + prod
        res.name = '{0}({1})'.format(self.name[:6], x.name[:10])
        return res
