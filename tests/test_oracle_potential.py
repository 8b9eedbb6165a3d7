"""Oracle of the potential formulation (oracle/potential.py) pinned without the GPU:

* The homogenised values equal the reference's golden results of the gradient (projection)
  formulation of the same Galerkin problems: GaNi value and Ga post-processing of the GaNi
  solution (scalar_2d prob1, scalar_3d prob1), Ga solves (scalar_2d prob2-5, scalar_3d prob2-3,
  from_file prob2), to the goldens' 1e-9.
* For a constant coefficient the operator is the multiplier xi^T A xi / |xi|^2 (0 at xi = 0).
* The operator is symmetric and positive in the inner product of 'r' spectra for a random
  coefficient; the potential has zero mean.
* The host helpers of ffthompy_amd (the preconditioner, the spectrum crop, the inner product and
  mean of 'r' spectra) against the oracle's.
"""
import json
import os

import numpy as np
import pytest

from oracle import examples
from oracle import potential as opot
from oracle.fields import DFT, Field
from oracle.materials import Material as OMaterial

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), 'golden', 'reference_results.json')))['problems']
TOL = 1e-9


def _mat(maker, name):
    return OMaterial(maker()[name])


@pytest.mark.parametrize('maker,name,N,key', [
    (examples.scalar_2d_materials, 'square', (5, 5), 'scalar_2d_prob1'),
    (examples.scalar_3d_materials, 'cube', (5, 5, 5), 'scalar_3d_prob1'),
])
def test_gani_potential_equals_golden_gani_and_ga_postprocessing(maker, name, N, key):
    mat = _mat(maker, name)
    Agani = mat.get_A_GaNi(N, 'primal')
    r = opot.homog_GaNi_full_potential(Agani, None, tol=1e-10, maxiter=200)
    assert abs(r['AH'] - GOLD[key]['diag']['AH_GaNi_primal']) < TOL
    assert abs(r['Fu'].mean()) < 1e-15
    Nbar = tuple(2 * n - 1 for n in N)
    for order, P, gkey in ((None, None, 'AH_Ga_primal'), (0, N, 'AH_Ga_o0_P5_primal')):
        if gkey not in GOLD[key]['diag']:
            continue
        Aga = mat.get_A_Ga(Nbar, 'primal', order=order, P=P)
        r = opot.homog_GaNi_full_potential(Agani, Aga, tol=1e-10, maxiter=200)
        assert abs(r['AH'] - GOLD[key]['diag'][gkey]) < TOL, (order, r['AH'])


@pytest.mark.parametrize('maker,name,N,key', [
    (examples.scalar_2d_materials, 'square', (5, 5), 'scalar_2d_prob2'),
    (examples.scalar_2d_materials, 'square_Ga', (5, 5), 'scalar_2d_prob3'),
    (examples.scalar_2d_materials, 'pyramid', (5, 5), 'scalar_2d_prob4'),
    (examples.scalar_2d_materials, 'ball', (5, 5), 'scalar_2d_prob5'),
    (examples.scalar_3d_materials, 'cube', (5, 5, 5), 'scalar_3d_prob2'),
    (examples.scalar_3d_materials, 'cube_Ga', (5, 5, 5), 'scalar_3d_prob3'),
    (examples.from_file_materials, 'file', (5, 5), 'from_file_prob2'),
])
def test_ga_potential_equals_golden_ga_solve(maker, name, N, key):
    mat = _mat(maker, name)
    Aga = mat.get_A_Ga(tuple(2 * n - 1 for n in N), 'primal')
    r = opot.homog_Ga_full_potential(Aga, tol=1e-10, maxiter=200)
    assert abs(r['AH'] - GOLD[key]['diag']['AH_Ga_primal']) < TOL, r['AH']
    assert abs(r['Fu'].mean()) < 1e-15


def _random_spectrum(N, rng):
    x = Field(val=rng.standard_normal(N), order=0)
    return DFT(N)(x)


@pytest.mark.parametrize('N', [(5, 7), (5, 3, 7)])
def test_constant_coefficient_operator_is_the_multiplier(N):
    rng = np.random.default_rng(3)
    d = len(N)
    B = rng.standard_normal((d, d))
    A0 = B @ B.T + d * np.eye(d)
    Agani = Field(val=np.einsum('ij,...->ij...', A0, np.ones(N)), order=2, multype=21)
    X = _random_spectrum(N, rng)
    got = opot.operator_gani(Agani)(X).val
    q = opot.grad_symbol(N, np.ones(d)).imag
    qq = np.einsum('i...,j...->ij...', q, q)
    k2 = np.einsum('ii...->...', qq)
    k2[(0,) * d] = 1.
    mult = np.einsum('ij,ij...->...', A0, qq) / k2
    want = mult * X.val
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    assert got[(0,) * d] == 0.


@pytest.mark.parametrize('N', [(7, 5), (5, 5, 3)])
def test_operator_symmetric_positive(N):
    rng = np.random.default_rng(5)
    d = len(N)
    a = np.exp(rng.standard_normal(N))
    Agani = Field(val=np.einsum('ij,...->ij...', np.eye(d), a), order=2, multype=21)
    op = opot.operator_gani(Agani)
    X, Z = _random_spectrum(N, rng), _random_spectrum(N, rng)
    X.val[(0,) * d] = 0.
    xy, yx = op(X) * Z, X * op(Z)
    assert abs(xy - yx) < 1e-12 * abs(xy)
    assert op(X) * X > 0


def test_host_preconditioner_matches_oracle():
    from ffthompy_amd.potential import get_preconditioner
    for N, Y in (((5, 7), (1., 2.)), ((3, 5, 9), (1., 1., 0.5))):
        P = get_preconditioner(N, {'Y': np.asarray(Y)})
        assert np.abs(P.val - opot.preconditioner(N, np.asarray(Y))).max() < 1e-15


def test_host_inner_product_of_r_spectra_matches_oracle():
    from ffthompy_amd.tensors import Tensor
    rng = np.random.default_rng(21)
    for N in ((5, 7), (4, 6), (3, 5, 8)):
        X, Z = _random_spectrum(N, rng), _random_spectrum(N, rng)
        tx = Tensor(val=X.val, order=0, N=N, Fourier=True)
        tz = Tensor(val=Z.val, order=0, N=N, Fourier=True)
        assert abs(tx * tz - X * Z) < 1e-13 * abs(X * X)
        assert abs(tx.mean() - X.mean()) < 1e-15


def test_ga_potential_rejects_even_solution_grid():
    from ffthompy_amd.materials import GaCoefficient
    from ffthompy_amd.potential import homog_Ga_full_potential
    A = GaCoefficient((7, 7), np.ones(2), 'A', 'iso', np.ones(1), inclusions=(['all'], np.zeros((1, 2)),
                                                                              np.zeros((1, 2))))
    with pytest.raises(NotImplementedError):
        homog_Ga_full_potential(A, {'solver': {'tol': 1e-8, 'maxiter': 10}})
