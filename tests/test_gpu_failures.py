"""Failure handling, each condition forced from the host (every kernel still exits normally):

* maxiter exhaustion: the solve returns after maxiter iterations with its residual above tol and
  status OK (the reference's CG exits the same way, ffthompy/general/solver.py:123,137-139), for
  fp64, mixed and the potential solve;
* breakdown: a negative coefficient makes <p, A p> < 0; the device reports FH_CG_BREAKDOWN after the
  first iteration and linear_solver raises FloatingPointError (the reference does not check);
* non-finite coefficients and grid lengths with a prime factor above 13 are rejected with an error
  before any device work on them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd import GpuOperator, Tensor, linear_solver  # noqa: E402

N = (15, 15, 15)


def _rhs(plan):
    EN = nat.DeviceArray.zeros((3,) + N)
    nat.add_constant(EN, -np.eye(3)[0])
    B = nat.DeviceArray((3,) + N)
    plan.matvec(EN, B)
    return B


def _plan(a):
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    return plan


@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
def test_maxiter_exhaustion_returns_with_status_ok(mode):
    print('case maxiter', mode, flush=True)
    a = np.exp(np.random.default_rng(41).standard_normal(N))
    plan = _plan(a)
    B = _rhs(plan)
    x = nat.DeviceArray.zeros((3,) + N)
    info = plan.cg_solve(B, x, tol=1e-14, maxiter=3, mode=mode, delta=1e-4)
    assert info['kit'] == 3 and info['status'] == nat.FH_CG_OK and info['norm_res'] > 1e-14


def test_potential_maxiter_exhaustion():
    print('case potential maxiter', flush=True)
    a = np.exp(np.random.default_rng(43).standard_normal(N))
    plan = _plan(a)
    for mode in ('fp64', 'mixed'):
        _, info = plan.potential_solve(np.eye(3)[0], 1e-14, 4, mode=mode)
        assert info['kit'] == 4 and info['status'] == nat.FH_CG_OK and info['norm_res'] > 1e-14


@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
def test_negative_coefficient_reports_breakdown(mode):
    print('case breakdown', mode, flush=True)
    rng = np.random.default_rng(47)
    a = -np.exp(rng.standard_normal(N))
    plan = _plan(a)
    B = _rhs(plan)
    x = nat.DeviceArray.zeros((3,) + N)
    info = plan.cg_solve(B, x, tol=1e-8, maxiter=50, mode=mode, delta=1e-4)
    assert info['status'] == nat.FH_CG_BREAKDOWN and info['kit'] == 0
    _, pinfo = plan.potential_solve(np.eye(3)[0], 1e-8, 50, mode=mode)
    assert pinfo['status'] == nat.FH_CG_BREAKDOWN


def test_linear_solver_raises_on_breakdown():
    a = -np.exp(np.random.default_rng(53).standard_normal(N))  # (a constant one gives B = 0: no iteration)
    A = np.einsum('ij,...->ij...', np.eye(3), a)
    Afun = GpuOperator(N, np.ones(3), 'scalar', 'primal', A=Tensor(name='A', val=A, order=2, multype=21))
    EN = Tensor(name='EN', shape=(3,), N=N)
    EN.set_mean(np.eye(3)[0])
    B = Afun(-EN)
    with pytest.raises(FloatingPointError):
        linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': 1e-8, 'maxiter': 50}, solver='CG')


def test_non_finite_coefficient_rejected():
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    a = np.ones(N)
    a[3, 4, 5] = np.nan
    with pytest.raises(nat.FhError):
        plan.set_coeff('iso', a)
    a[3, 4, 5] = np.inf
    with pytest.raises(nat.FhError):
        plan.set_coeff('iso', a)
    dev = nat.DeviceArray.from_host(a)
    with pytest.raises(nat.FhError):
        plan.set_coeff('iso', dev)


@pytest.mark.parametrize('grid', [(17, 15, 15), (15, 34), (19, 19)])
def test_grid_with_large_prime_factor_is_planned(grid):
    """Any length is accepted (Bluestein beyond radix 13); its values are tested against numpy
    and the oracle in test_gpu_bluestein.py."""
    pl = nat.Plan(grid, (1.,) * len(grid), 'scalar', 'primal')
    assert pl.ncomp == len(grid)


@pytest.mark.parametrize('grid', [(5, 100003), (3, 65537, 3)])
def test_row_too_long_for_a_workgroup_rejected(grid):
    """A prime length whose Bluestein convolution does not fit one workgroup's LDS is refused at
    plan creation with FH_ERR_UNSUPPORTED (nothing launched)."""
    with pytest.raises(nat.FhError) as e:
        nat.Plan(grid, (1.,) * len(grid), 'scalar', 'primal')
    assert e.value.code == nat.FH_ERR_UNSUPPORTED
