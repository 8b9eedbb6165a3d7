"""Fourier projections G0, G1, G2 as operators that are never materialised.

Mirrors the interface of ffthompy/projections.py:9-112 (scalar) and 114-267
(elasticity): the returned objects act on spectra held on the device (Tensor with
Fourier=True, any fft_form 'r', 0 or 'c') through fh_project_spectrum, which evaluates
G(xi) from the frequency index. NyqNul=True (the reference's default) sets the Nyquist
modes of even axes to zero, as the reference's G built on the odd grid and zero-padded;
NyqNul=False keeps G at the frequency -N/2 there. For elasticity the reference returns
the parts (G1h, G1s, G2h, G2s) whose sums the driver uses
(ffthompy/applications.py:121-122); here the sums G1 = G1h + G1s and G2 = G2h + G2s are
returned directly.
"""
import ctypes

import numpy as np

from . import _native as nat
from .tensors import Tensor


def _check_form(fft_form):
    nat.form_code(fft_form)
    return fft_form


class FourierProjection:
    """G(xi) acting on a spectrum of the projection's fft_form (the einsum 'ij...,j...->i...' of
    ffthompy/tensors/objects.py:223-232)."""

    def __init__(self, N, Y, physics, which, name, NyqNul=True, fft_form='r'):
        self.N = tuple(int(n) for n in N)
        self.Y = np.asarray(Y, dtype=np.float64)
        self.physics = physics
        self.which = which
        self.name = name
        self.NyqNul = bool(NyqNul)
        self.fft_form = _check_form(fft_form)
        d = len(self.N)
        if d not in (2, 3) or physics not in ('scalar', 'elasticity'):
            raise nat.FhError('projections are defined for scalar or elasticity problems in 2-D or 3-D',
                              nat.FH_ERR_ARG)
        self.ncomp = d if physics == 'scalar' else d * (d + 1) // 2

    def __call__(self, X):
        assert X.Fourier and nat.form_code(X.fft_form) == nat.form_code(self.fft_form)
        X = X.to_device()
        out = nat.project_spectrum(X.val, self.N, self.Y, self.physics, self.which, self.fft_form, self.NyqNul)
        return X._like(out, name='{0}({1})'.format(self.name, X.name[:10]))

    def __repr__(self):
        return 'FourierProjection(%s, N=%s, %s, NyqNul=%s, fft_form=%r)' % (self.name, self.N, self.physics,
                                                                           self.NyqNul, self.fft_form)


class MeanProjection:
    """G0: keeps the zero-frequency coefficient only (projections.py:62), copied on the device."""

    def __init__(self, N, name='hG0', fft_form='r'):
        self.N = tuple(int(n) for n in N)
        self.name = name
        self.fft_form = _check_form(fft_form)

    def __call__(self, X):
        assert X.Fourier
        X = X.to_device()
        out = nat.DeviceArray.zeros(X.val.shape, X.val.dtype)
        nspec = int(np.prod(X.val.shape[X.order:]))
        ic = X.mean_index()
        off = int(np.ravel_multi_index(ic, X.val.shape[X.order:]))
        isz = X.val.dtype.itemsize
        for c in range(X._count()):
            b = (c * nspec + off) * isz
            nat.call('fh_memcpy_dtod', ctypes.c_void_p(out.ptr.value + b), ctypes.c_void_p(X.val.ptr.value + b),
                     ctypes.c_size_t(isz))
        return X._like(out, name='{0}({1})'.format(self.name, X.name[:10]))


def scalar(N, Y, NyqNul=True, tensor=True, fft_form='r'):
    """(G0, G1, G2) of the scalar problem (projections.py:9-112) for spectra of fft_form."""
    return (MeanProjection(N, fft_form=fft_form),
            FourierProjection(N, Y, 'scalar', 1, 'hG1', NyqNul, fft_form),
            FourierProjection(N, Y, 'scalar', 2, 'hG2', NyqNul, fft_form))


def elasticity(N, Y, NyqNul=True, tensor=True, fft_form='r'):
    """(G0, G1, G2) of small-strain elasticity in Mandel notation (projections.py:114-267)."""
    return (MeanProjection(N, fft_form=fft_form),
            FourierProjection(N, Y, 'elasticity', 1, 'hG1', NyqNul, fft_form),
            FourierProjection(N, Y, 'elasticity', 2, 'hG2', NyqNul, fft_form))


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
