"""Potential formulation of the scalar homogenisation problem on the GPU: the drop-in for
homog_GaNi_full_potential / homog_Ga_full_potential / get_preconditioner of
ffthompy/tensorsLowRank/homogenisation.py:43-128, 296-300.

The unknown is the rfftn spectrum of a scalar potential u (one complex half spectrum instead of
d real fields), the field is e = grad u, and CG runs on P (-div) F A F^-1 grad P with
P = 1/|2 pi xi| (fh_potential_solve). As in the reference the macroscopic load is E = e_1, CG
starts from zero, and AH is the scalar <A (e + E), e + E> (GaNi coefficient, or the Ga
coefficient after interpolation onto the quadrature grid).

The Ga variant solves on the quadrature grid of the Ga coefficient (2N-1 per axis, or the next
length with prime factors <= 13) with grad, div and P restricted to the frequencies of the grid N,
which is the reference's enlarge / project between N and 2N-1 with the same inner products.
"""
import time

import numpy as np

from . import _native as nat
from .tensors import Tensor, classify_coefficient


class Struct(dict):
    """Attribute dictionary for results (the reference's ffthompy.Struct)."""

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError:
            raise AttributeError(key)

    def __setattr__(self, key, val):
        self[key] = val


def _par(pars, key, default=None):
    if isinstance(pars, dict):
        return pars.get(key, default)
    return getattr(pars, key, default)


def _solver_par(pars):
    """tol, maxiter and the arithmetic: solver['precision'] 'fp64' (default, the reference's) or
    'mixed' (fp32 iterations with fp64 reliable updates, solver['delta'] their trigger)."""
    s = _par(pars, 'solver', {}) or {}
    mode = s.get('precision', 'fp64')
    nat.solve_mode(mode)
    return (float(s.get('tol', 1e-6)), int(s.get('maxiter', 1000)),
            {'mode': mode, 'delta': float(s.get('delta', 1e-4))})


def _e1(d):
    E = np.zeros(d)
    E[0] = 1.
    return E


def get_preconditioner(N, pars):
    """P = 1/|2 pi xi| on the 'r' half spectrum, 1 at the mean (homogenisation.py:296-300);
    a host Tensor (Fourier, multype 00). The GPU solve evaluates P inside its kernels."""
    N = tuple(int(n) for n in N)
    Y = np.asarray(_par(pars, 'Y', np.ones(len(N))), dtype=np.float64)
    d = len(N)
    shape = N[:-1] + (N[-1] // 2 + 1,)
    k2 = np.zeros(shape)
    for j in range(d):
        k = np.fft.ifftshift(np.arange(np.fix(-N[j] / 2.), np.fix(N[j] / 2. + 0.5)))
        if j == d - 1:
            k = k[:shape[-1]]
        sh = [1] * d
        sh[j] = -1
        k2 = k2 + (2 * np.pi * k / Y[j]).reshape(sh) ** 2
    k2[(0,) * d] = 1.
    return Tensor(name='P', val=1. / k2 ** 0.5, order=0, N=N, Y=Y, Fourier=True, multype=0)


def _fourier_tensor(name, val, N, Y):
    return Tensor(name=name, val=val, order=0, N=N, Y=Y, Fourier=True)


def homog_GaNi_full_potential(Agani, Aga, pars):
    """GaNi potential solve (homogenisation.py:84-128). Agani: the nodal coefficient Tensor
    (d,d)+N (Material.get_A_GaNi); Aga: None (GaNi value) or the Ga coefficient of
    Material.get_A_Ga(2N-1) for the Ga value of the same solution. Returns
    Struct(AH, Fu, info, time, pars)."""
    N = tuple(int(n) for n in Agani.N)
    Y = np.asarray(getattr(Agani, 'Y', None) if getattr(Agani, 'Y', None) is not None
                   else _par(pars, 'Y', np.ones(len(N))), dtype=np.float64)
    d = len(N)
    tol, maxiter, prec = _solver_par(pars)
    plan = nat.Plan(N, Y, 'scalar', 'primal')
    plan.set_coeff(*classify_coefficient(Agani.val, 'scalar'))
    E = _e1(d)
    X = nat.DeviceArray((d,) + N)
    t0 = time.perf_counter()
    Fu, info = plan.potential_solve(E, tol, maxiter, X=X, **prec)
    elapsed = time.perf_counter() - t0
    nat.add_constant(X, E)
    # the Gram matrix of d copies of one field has every entry <A XEN, XEN>
    if Aga is None:
        AH = plan.energy_matrix([X] * d)[0, 0]
    elif all(n % 2 for n in N):
        # odd N: grad Fu is Hermitian, so enlarging it is interpolating the real field X
        ga = Aga.load(nat.Ga(plan))
        AH = ga.energy_matrix([X] * d)[0, 0]
    else:
        # even N: the Nyquist entries of grad Fu (frequency -N/2) are not Hermitian; as the
        # reference (homogenisation.py:123-126) the gradient spectrum itself is enlarged to 2N-1
        # (Tensor.project, Nyquist entries split in halves) and transformed there
        ga = Aga.load(nat.Ga(plan))
        M = ga.M
        XM = _enlarged_gradient(Fu, N, M, Y)
        nat.add_constant(XM, E)
        AH = ga.energy_matrix([XM] * d, on_grid=True)[0, 0]
    return Struct(AH=AH, Fu=_fourier_tensor('Fu', Fu, N, Y), info=info, time=elapsed, pars=pars)


def _enlarged_gradient(Fu, N, M, Y):
    """F^-1 on grid M of grad(Fu) enlarged from N to M (Tensor.enlarge of the 'r' spectrum,
    ffthompy/tensors/objects.py:428-467), on the device."""
    d = len(N)
    G = nat.spectrum_grad(Fu, N, Y)
    full = nat.spectrum_half_to_full(G, N, d, 1. / float(np.prod(N)), centred=True)
    big = nat.spectrum_resize(full, N, M, d, enlarge=True)
    half = nat.spectrum_full_to_half(big, M, d, float(np.prod(M)), centred=True)
    X = nat.DeviceArray((d,) + tuple(M))
    Mh = tuple(M[:-1]) + (M[-1] // 2 + 1,)
    nat.Plan(M, Y, 'scalar', 'primal').irfftn(half.reshape((d,) + Mh), X, d)
    return X


def _crop_spectrum(val, M, N):
    """'r' spectrum on grid M of a trigonometric polynomial with frequencies of grid N (odd) ->
    its 'r' spectrum on N (the reference's Tensor.decrease of an 'r' spectrum, values times
    prod(N)/prod(M)), on the device."""
    full = nat.spectrum_half_to_full(val, M, 1, 1. / float(np.prod(M)), centred=True)
    small = nat.spectrum_resize(full, M, N, 1, enlarge=False)
    half = nat.spectrum_full_to_half(small, N, 1, float(np.prod(N)), centred=True)
    return half.reshape(tuple(N[:-1]) + (N[-1] // 2 + 1,))


def homog_Ga_full_potential(Aga, pars):
    """Ga potential solve (homogenisation.py:43-82): unknown on N = (Nbar+1)/2, coefficient the
    exactly integrated Aga of Material.get_A_Ga(Nbar). Returns Struct(AH, e, Fu, info, time):
    e = F^-1 grad Fu on the quadrature grid (the reference's Nbar grid when 2N-1 has no prime
    factor above 13), Fu on the grid N."""
    Nbar = tuple(int(n) for n in Aga.N)
    N = tuple((n + 1) // 2 for n in Nbar)
    if any(n % 2 == 0 for n in N):
        # an even N has a Nyquist mode, which the band |k| <= (N-1)/2 of the device solve leaves out
        raise NotImplementedError('homog_Ga_full_potential needs an odd grid N = (Nbar+1)/2 (got %s)' % (N,))
    Y = np.asarray(Aga.Y, dtype=np.float64)
    d = len(N)
    tol, maxiter, prec = _solver_par(pars)
    coarse = nat.Plan(N, Y, 'scalar', 'primal')
    ga = Aga.load(nat.Ga(coarse))
    M = ga.M
    plan = nat.Plan(M, Y, 'scalar', 'primal')
    plan.set_band([(n - 1) // 2 for n in N], 1.0)
    ga.set_plan_coeff(plan)
    E = _e1(d)
    X = nat.DeviceArray((d,) + M)
    t0 = time.perf_counter()
    FuM, info = plan.potential_solve(E, tol, maxiter, X=X, **prec)
    elapsed = time.perf_counter() - t0
    XEN = X.copy()
    nat.add_constant(XEN, E)
    AH = ga.energy_matrix([XEN] * d, on_grid=True)[0, 0]
    Fu = _crop_spectrum(FuM, M, N)
    return Struct(AH=AH, e=Tensor(name='e', val=X, order=1, Y=Y), Fu=_fourier_tensor('Fu', Fu, N, Y), info=info,
                  time=elapsed, pars=pars)
