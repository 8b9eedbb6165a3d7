"""Potential formulation of the scalar homogenisation problem, restated from
ffthompy/tensorsLowRank/homogenisation.py:43-128 (homog_Ga_full_potential,
homog_GaNi_full_potential) and :296-300 (get_preconditioner), with grad/div of
ffthompy/tensors/operators.py:227-290, 371-386.

The unknown is the 'r' spectrum of a scalar potential u; e = grad u; CG runs on
P (-div) F A F^-1 grad P with P = 1/|2 pi xi| (1 at xi = 0), x0 = 0, E = e_1.

Test infrastructure only (see oracle/__init__.py). Pinned in tests/test_oracle_potential.py:
the homogenised value equals the one of the gradient (projection) formulation, GaNi and Ga;
for a constant coefficient the operator is the closed form xi^T A xi / |xi|^2 (a for A = a I);
symmetry and positivity of the operator in the inner product of 'r' spectra; mean of Fu = 0.
"""
import numpy as np

from . import grid
from .fields import DFT, Field, half_length
from .solver import CG


def grad_symbol(N, Y):
    """2 pi i xi_j on the 'r' half spectrum, shape (d,) + half_length(N) (operators.py:371-386)."""
    N = tuple(int(n) for n in N)
    d = len(N)
    xis = grid.frequencies(N, Y, fft_form='r')
    shape = half_length(N)
    g = np.zeros((d,) + shape, dtype=np.complex128)
    for j in range(d):
        sh = [1] * d
        sh[j] = -1
        g[j] = np.broadcast_to((xis[j] * 2 * np.pi).reshape(sh), shape) * 1j
    return g


def preconditioner(N, Y):
    """P = 1/|2 pi xi| with P = 1 at the mean (homogenisation.py:296-300)."""
    g = grad_symbol(N, Y)
    k2 = np.einsum('i...,i...', g, np.conj(g)).real
    k2[(0,) * len(N)] = 1.
    return 1. / k2 ** 0.5


def grad(X):
    """Gradient of a scalar 'r' spectrum (operators.py:227-260)."""
    assert X.Fourier and X.order == 0
    return X.copy(val=grad_symbol(X.N, X.Y) * X.val[np.newaxis], order=1)


def div(X):
    """Divergence of a vector 'r' spectrum (operators.py:262-290)."""
    assert X.Fourier and X.order == 1
    return X.copy(val=np.einsum('i...,i...->...', grad_symbol(X.N, X.Y), X.val), order=0)


def _scaled(P, X):
    return X.copy(val=P * X.val)


def _macro(N, Y):
    d = len(N)
    E = np.zeros(d)
    E[0] = 1.
    EN = Field(shape=(d,), N=N, Y=Y)
    EN.set_mean(E)
    return EN


def operator_gani(Agani):
    """x -> P (-div) F A F^-1 grad (P x) on the grid of Agani (homogenisation.py:97-102, 108)."""
    N, Y = Agani.N, Agani.Y
    F, iF = DFT(N), DFT(N, inverse=True)
    P = preconditioner(N, Y)

    def op(X):
        FAX = F(Agani(iF(grad(_scaled(P, X)))))
        return _scaled(P, -div(FAX))
    return op


def operator_ga(Aga):
    """x -> P (-div) project_N F2 A F2^-1 enlarge_Nbar grad (P x), x on N = (Nbar+1)/2
    (homogenisation.py:57-67)."""
    Nbar, Y = Aga.N, Aga.Y
    N = tuple((n + 1) // 2 for n in Nbar)
    F2, iF2 = DFT(Nbar), DFT(Nbar, inverse=True)
    P = preconditioner(N, Y)

    def op(X):
        FAX = F2(Aga(iF2(grad(_scaled(P, X)).enlarge(Nbar))))
        FAX = FAX.project(N)
        return _scaled(P, -div(FAX))
    return op


def homog_GaNi_full_potential(Agani, Aga, tol=1e-8, maxiter=50):
    """homogenisation.py:84-128. Agani: (d,d)+N coefficient Field (multype 21); Aga: the Ga
    coefficient on 2N-1 or None. Returns dict(AH, Fu, info)."""
    N, Y = Agani.N, Agani.Y
    F, iF = DFT(N), DFT(N, inverse=True)
    P = preconditioner(N, Y)
    EN = _macro(N, Y)
    op = operator_gani(Agani)
    B = div(F(Agani(EN)))
    PB = _scaled(P, B)
    x0 = Field(shape=(), N=N, Y=Y, Fourier=True)
    iPU, info = CG(op, PB, x0=x0, par={'tol': tol, 'maxiter': maxiter})
    Fu = _scaled(P, iPU)
    if Aga is None:
        XEN = iF(grad(Fu)) + EN
        AH = Agani(XEN) * XEN
    else:
        Nbar = tuple(2 * n - 1 for n in N)
        XEN = DFT(Nbar, inverse=True)(grad(Fu).project(Nbar)) + EN.project(Nbar)
        AH = Aga(XEN) * XEN
    return {'AH': AH, 'Fu': Fu, 'info': info}


def homog_Ga_full_potential(Aga, tol=1e-8, maxiter=50):
    """homogenisation.py:43-82: solve on the grid N = (Nbar+1)/2 with the Ga coefficient on Nbar.
    Returns dict(AH, e, Fu, info)."""
    Nbar, Y = Aga.N, Aga.Y
    N = tuple((n + 1) // 2 for n in Nbar)
    F2, iF2 = DFT(Nbar), DFT(Nbar, inverse=True)
    P = preconditioner(N, Y)
    EN = _macro(Nbar, Y)

    B = div(F2(Aga(EN)).decrease(N))
    x0 = Field(shape=(), N=N, Y=Y, Fourier=True)
    iPU, info = CG(operator_ga(Aga), _scaled(P, B), x0=x0, par={'tol': tol, 'maxiter': maxiter})
    Fu = _scaled(P, iPU)
    X = iF2(grad(Fu).project(Nbar))
    AH = Aga(X + EN) * (X + EN)
    return {'AH': AH, 'e': X, 'Fu': Fu, 'info': info}
