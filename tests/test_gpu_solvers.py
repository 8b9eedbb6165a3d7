"""Solvers other than CG on the GPU operator: Richardson and the two-term Chebyshev iteration on
the device (fh_iter_solve), SciPy's CG with the GPU matvec; against the oracle's restatements of
ffthompy/general/solver.py (same iterates in fp64: iteration counts equal, solutions to 1e-12)
and, as ffthompy/general/unittest_solver.py:35-76 does, against the CG solution."""
import numpy as np
import pytest

from oracle import projections
from oracle.fields import DFT, Field, Operator
from oracle.solver import linear_solver as o_linear_solver

pytestmark = pytest.mark.gpu
ffa = pytest.importorskip('ffthompy_amd')
from ffthompy_amd import GpuOperator, Tensor, linear_solver, CallBack  # noqa: E402


def _problem(N, seed):
    rng = np.random.default_rng(seed)
    dim = len(N)
    a = 1. + 10. * rng.random(N)
    A = np.einsum('ij,...->ij...', np.eye(dim), a)
    _, G1, _ = projections.scalar(N, np.ones(dim))
    GN = Operator([[DFT(N, inverse=True), G1, DFT(N)]])
    oAfun = Operator([[GN, Field(val=A, order=2, multype=21)]])
    E = np.zeros(dim)
    E[0] = 1.
    oE = Field(shape=(dim,), N=N)
    oE.set_mean(E)
    oB = oAfun(-oE)
    Afun = GpuOperator(N, np.ones(dim), 'scalar', 'primal', A=A)
    EN = Tensor(name='EN', N=N, shape=(dim,), Y=np.ones(dim))
    EN.set_mean(E)
    B = Afun(-EN)
    assert np.abs(B.to_host().val - oB.val).max() < 1e-12
    return Afun, B, oAfun, oB


@pytest.mark.parametrize('N', [(5, 5), (15, 15, 15), (27, 27, 27)])
@pytest.mark.parametrize('solver', ['richardson', 'chebyshev'])
def test_stationary_solver_matches_oracle(N, solver):
    print('case solver', solver, N, flush=True)
    Afun, B, oAfun, oB = _problem(N, 5)
    par = {'tol': 1e-10, 'maxiter': 3000, 'alpha': 0.5 * (1. + 11.), 'eigrange': [1., 11.]}
    x, info = linear_solver(Afun, B, x0=None, par=dict(par), solver=solver)
    ox, oinfo = o_linear_solver(oAfun, oB, x0=oB.zeros_like(), par=dict(par), solver=solver)
    assert info['kit'] == oinfo['kit'], (info['kit'], oinfo['kit'])
    assert abs(info['norm_res'] - oinfo['norm_res']) <= 1e-6 * oinfo['norm_res']
    assert np.abs(x.to_host().val - ox.val).max() < 1e-12
    X, _ = linear_solver(Afun, B, x0=None, par=dict(par), solver='CG')
    assert np.abs(x.to_host().val - X.to_host().val).max() < 1e-8


def test_solvers_converge_to_the_cg_solution_with_callback():
    """unittest_solver.py:35-76: CG, scipy_cg, richardson and chebyshev agree to 1e-8."""
    Afun, B, _, _ = _problem((5, 5), 7)
    par = {'tol': 1e-10, 'maxiter': int(1e3), 'alpha': 0.5 * (1. + 10.), 'eigrange': [1., 10.]}
    X, _ = linear_solver(Afun, B, x0=None, par=dict(par), solver='CG')
    for solver in ('scipy_cg', 'richardson', 'chebyshev'):
        cb = CallBack(A=Afun, B=B)
        x, info = linear_solver(Afun, B, x0=None, par=dict(par), solver=solver, callback=cb)
        assert np.linalg.norm(X.to_host().val - x.to_host().val) < 1e-8, solver
        if solver != 'scipy_cg':
            assert len(cb.res_norm) == info['kit'] + 1 and cb.res_norm[-1] == pytest.approx(info['norm_res'])


def test_chebyshev_early_return_and_zero_iterations():
    Afun, B, oAfun, oB = _problem((5, 5), 9)
    par = {'tol': 1e-10, 'maxiter': 3000, 'eigrange': [1., 11.]}
    X, _ = linear_solver(Afun, B, x0=None, par=dict(par), solver='CG')
    x, info = linear_solver(Afun, B, x0=X, par=dict(par), solver='chebyshev')
    assert info['kit'] == 0 and info['norm_res'] < 1e-10
    x, info = linear_solver(Afun, B, x0=None, par=dict(par, maxit=0), solver='chebyshev')
    assert info['kit'] == 0 and info['norm_res'] == 0.
    x, info = linear_solver(Afun, B, x0=None, par=dict(par, alpha=6., maxiter=0), solver='richardson')
    assert info['kit'] == 0 and info['norm_res'] == 1e15
