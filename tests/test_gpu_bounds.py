"""Upper and lower bound matrices of the headline workload's family (random spheres, contrast
100, CG tol 1e-8, primal with a and dual with 1/a, loads e1-e3) at 45^3 on the GPU against the
oracle's fp64 solves: fp64 within 1e-10 and fp32-with-fp64-refinement within 1e-8 (relative),
lower <= upper within the rounding slack 64 eps max|AH|."""
import numpy as np
import pytest

from oracle import projections
from oracle.fields import DFT, Field, Operator
from oracle.homogenize import assembly_matrix
from oracle.solver import linear_solver as o_linear_solver

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd.microstructure import spheres_conductivity  # noqa: E402

N = (45, 45, 45)


def _oracle_bounds(a):
    out = {}
    _, hG1, hG2 = projections.scalar(N, np.ones(3), NyqNul=True)
    for pd, coef, hG in (('primal', a, hG1), ('dual', 1. / a, hG2)):
        GN = Operator([[DFT(N, inverse=True), hG, DFT(N)]])
        A = Field(val=np.einsum('ij,...->ij...', np.eye(3), coef), order=2, multype=21)
        Afun = Operator([[GN, A]])
        sols = []
        for i in range(3):
            EN = Field(shape=(3,), N=N)
            EN.set_mean(np.eye(3)[i])
            B = Afun(-EN)
            X, _ = o_linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': 1e-8, 'maxiter': 1000}, solver='CG')
            sols.append(X + EN)
        AH = assembly_matrix(A, sols)
        out[pd] = AH if pd == 'primal' else np.linalg.inv(AH)
    return out


@pytest.fixture(scope='module')
def case():
    a = spheres_conductivity(N, count=30, contrast=100.)[0]
    return a, _oracle_bounds(a)


@pytest.mark.parametrize('mode,tol', [('fp64', 1e-10), ('mixed', 1e-8)])
def test_sphere_bounds_match_oracle(case, mode, tol):
    a, ref = case
    print('case bounds', mode, flush=True)
    got = {}
    for pd, coef in (('primal', a), ('dual', 1. / a)):
        plan = nat.Plan(N, (1., 1., 1.), 'scalar', pd)
        plan.set_coeff('iso', coef)
        sols = []
        for i in range(3):
            E = np.eye(3)[i]
            EN = nat.DeviceArray.zeros((3,) + N)
            nat.add_constant(EN, -E)
            B = nat.DeviceArray((3,) + N)
            plan.matvec(EN, B)
            x = nat.DeviceArray.zeros((3,) + N)
            info = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode=mode, delta=1e-4)
            assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= 1e-8
            nat.add_constant(x, E)
            sols.append(x)
        AH = plan.energy_matrix(sols)
        got[pd] = AH if pd == 'primal' else np.linalg.inv(AH)
    for pd in ('primal', 'dual'):
        err = np.abs(got[pd] - ref[pd]).max() / np.abs(ref[pd]).max()
        assert err < tol, (pd, err)
    # GaNi primal and dual are equal analytically and the CG iterates only overestimate both
    # energies, so lower <= upper up to the rounding of the double-double energies: slack 64 eps |AH|
    slack = 64 * np.finfo(np.float64).eps * np.abs(got['primal']).max()
    gap = np.linalg.eigvalsh(0.5 * ((got['primal'] - got['dual']) + (got['primal'] - got['dual']).T)).min()
    assert gap >= -slack, (gap, slack)
