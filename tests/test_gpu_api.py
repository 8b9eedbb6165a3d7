"""The drop-in API on the GPU: Problem / applications / linear_solver / postprocess run the
reference's GaNi example problems end to end, against the golden results, and the
Tensor / DFT / projection objects against the oracle."""
import json
import os

import numpy as np
import pytest

from oracle import examples, projections as oproj
from oracle.fields import rfftn as o_rfftn

pytestmark = pytest.mark.gpu
ffa = pytest.importorskip('ffthompy_amd')
from ffthompy_amd import Problem, Tensor, DFT, GpuOperator, linear_solver, CallBack  # noqa: E402
from ffthompy_amd import projections  # noqa: E402

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), 'golden', 'reference_results.json')))
GANI = [c for c in examples.problems() if c[2]['solve']['kind'] == 'GaNi']


@pytest.mark.parametrize('case', GANI, ids=['%s_%s' % (c[0], c[1]) for c in GANI])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_problem_api_gani_goldens(case, precision):
    f, name, conf, mats = case
    print('case api', f, name, precision, flush=True)
    conf = dict(conf)
    conf['postprocess'] = [{'kind': 'GaNi'}]
    conf['solver'] = dict(conf['solver'], precision=precision)
    pb = Problem(conf, type('conf', (), {'materials': mats()}))
    pb.calculate()
    gold = GOLD['problems']['%s_%s' % (f, name)]
    for pd in pb.solve['primaldual']:
        AH = pb.output['mat_' + pd]['AH_GaNi_' + pd]
        kits = [r['info']['kit'] for r in pb.output['res_' + pd]]
        if precision == 'fp64':
            assert kits == gold['kit'][pd]
            cb = pb.output['res_' + pd][0]['cb']
            assert len(cb.res_norm) == kits[0] + 1 and cb.res_norm[-1] <= conf['solver']['tol']
        key = 'AH_GaNi_' + pd
        if key in gold.get('diag', {}):
            assert np.abs(AH - gold['diag'][key] * np.eye(AH.shape[0])).max() < 1e-9
        for ij, v in gold.get('entries', {}).get(key, {}).items():
            i, j = map(int, ij.split(','))
            assert abs(AH[i, j] - v) < 1e-9


def test_tensor_arithmetic_and_dft():
    rng = np.random.default_rng(40)
    N = (9, 7, 11)
    a, b = rng.standard_normal((3,) + N), rng.standard_normal((3,) + N)
    ta = Tensor(val=ffa._native.DeviceArray.from_host(a), order=1)
    tb = Tensor(val=ffa._native.DeviceArray.from_host(b), order=1)
    assert abs(ta * tb - np.sum(a * b) / np.prod(N)) < 1e-12
    assert abs(ta.norm() - np.sqrt(np.sum(a * a) / np.prod(N))) < 1e-12
    assert np.abs((ta + tb).to_host().val - (a + b)).max() < 1e-14
    assert np.abs((ta - tb).to_host().val - (a - b)).max() < 1e-14
    assert np.abs((2.5 * ta).to_host().val - 2.5 * a).max() < 1e-14
    assert np.abs(ta.mean() - a.mean(axis=(1, 2, 3))).max() < 1e-14
    E = np.array([1., -2., 0.5])
    tc = ta.copy()
    tc.set_mean(E)
    assert np.abs(tc.mean() - E).max() < 1e-13
    F = DFT(N=N)
    Fa = F(ta)
    assert Fa.Fourier and np.abs(Fa.to_host().val - o_rfftn(a, N)).max() < 1e-11
    back = DFT(N=N, inverse=True)(Fa)
    assert np.abs(back.to_host().val - a).max() < 1e-13


@pytest.mark.parametrize('physics', ['scalar', 'elasticity'])
def test_projection_objects(physics):
    N, Y = (5, 7, 9), np.array([1., 1.5, 2.])
    G0, G1, G2 = projections.scalar(N, Y) if physics == 'scalar' else projections.elasticity(N, Y)
    if physics == 'scalar':
        _, o1, o2 = oproj.scalar(N, Y)
        M1, M2 = o1.val, o2.val
    else:
        _, a, b, c, d = oproj.elasticity(N, Y)
        M1, M2 = a.val + b.val, c.val + d.val
    nc = G1.ncomp
    rng = np.random.default_rng(41)
    Nh = N[:-1] + (N[-1] // 2 + 1,)
    X = rng.standard_normal((nc,) + Nh) + 1j * rng.standard_normal((nc,) + Nh)
    tX = Tensor(val=ffa._native.DeviceArray.from_host(X), order=1, N=N, Fourier=True)
    for G, M in ((G1, M1), (G2, M2)):
        out = G(tX).to_host().val
        assert np.abs(out - np.einsum('ij...,j...->i...', M, X)).max() < 1e-13


def test_linear_solver_x0_and_callback():
    """A non-zero x0 is honoured, and the callback log matches the solve."""
    rng = np.random.default_rng(42)
    N = (15, 15, 15)
    a = 1. + 9. * rng.random(N)
    Afun = GpuOperator(N, np.ones(3), 'scalar', 'primal', coef=('iso', a))
    EN = Tensor(shape=(3,), N=N)
    EN.set_mean(np.array([0., 1., 0.]))
    B = Afun(-EN)
    cb = CallBack(A=Afun, B=B)
    X, info = linear_solver(Afun, B, x0=None, par={'tol': 1e-9, 'maxiter': 500}, solver='CG', callback=cb)
    assert info['norm_res'] <= 1e-9 and len(cb.res_norm) == info['kit'] + 1
    X2, info2 = linear_solver(Afun, B, x0=X, par={'tol': 1e-9, 'maxiter': 500}, solver='CG')
    assert info2['kit'] == 0
    assert (B - Afun(X)).norm() <= 1e-9 * 1.01
