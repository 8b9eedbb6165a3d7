"""Grid lengths with a prime factor above 13 (the reference takes any length through pocketfft,
ffthompy/tensors/fft.py:39-43): the runtime-radix kernels transform them with Bluestein's chirp
convolution on a 7-smooth length >= 2N-1.

Tolerances: FFTs against numpy.fft to 1e-12 (float64) and 5e-5 (float32, relative to the largest
entry: the convolution length adds a few rounding steps to the radix path's 2e-5); the operator
against the oracle's materialised Operator to 1e-12; CG iteration counts exactly and bound
matrices to 1e-10 relative against the oracle's solve.
"""
import numpy as np
import pytest

from oracle import examples, projections
from oracle.fields import DFT, Field, Operator, rfftn as o_rfftn, irfftn as o_irfftn
from oracle.homogenize import Problem as OracleProblem

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd import Problem  # noqa: E402


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


SHAPES = [(17, 17), (19, 23), (101, 101), (97, 6), (238, 5), (5, 131), (17, 17, 17), (34, 19, 10),
          (23, 45, 17), (1009, 3)]


@pytest.mark.parametrize('N', SHAPES, ids=[str(s) for s in SHAPES])
def test_prime_length_fft_matches_numpy(N):
    print('case bluestein fft', N, flush=True)
    rng = np.random.default_rng(20)
    nc = 2
    plan = nat.Plan(N, np.ones(len(N)))
    x = rng.standard_normal((nc,) + N)
    Nh = N[:-1] + (N[-1] // 2 + 1,)
    ref = o_rfftn(x, N)
    for dt, ct, tol in ((np.float64, np.complex128, 1e-12), (np.float32, np.complex64, 5e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dX = nat.DeviceArray((nc,) + Nh, ct)
        plan.rfftn(dx, dX, nc)
        assert rel(dX.to_host(), ref) < tol, (dt, rel(dX.to_host(), ref))
        Z = (rng.standard_normal((nc,) + Nh) + 1j * rng.standard_normal((nc,) + Nh)).astype(ct)
        dZ = nat.DeviceArray.from_host(Z)
        dz = nat.DeviceArray((nc,) + N, dt)
        plan.irfftn(dZ, dz, nc)
        zref = o_irfftn(Z.astype(np.complex128), N)
        assert rel(dz.to_host(), zref) < tol, (dt, rel(dz.to_host(), zref))


def _proj(physics, N, Y):
    if physics == 'scalar':
        _, G1, G2 = projections.scalar(N, Y)
        return G1.val, G2.val
    _, G1h, G1s, G2h, G2s = projections.elasticity(N, Y)
    return G1h.val + G1s.val, G2h.val + G2s.val


MV = [('scalar', 'iso', (17, 17, 17)), ('scalar', 'full', (101, 101)), ('elasticity', 'lame', (19, 23)),
      ('elasticity', 'iso', (17, 9, 19)), ('scalar', 'iso', (34, 19, 10))]


@pytest.mark.parametrize('physics,kind,N', MV, ids=['%s-%s-%s' % c for c in MV])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_prime_length_matvec_matches_oracle(physics, kind, N, pd):
    print('case bluestein matvec', physics, kind, N, pd, flush=True)
    rng = np.random.default_rng(21)
    Y = np.linspace(1., 1.4, len(N))
    G1, G2 = _proj(physics, N, Y)
    plan = nat.Plan(N, Y, physics, pd)
    nc = plan.ncomp
    if kind == 'iso':
        a = 1. + rng.random(N)
        A = np.einsum('ij,...->ij...', np.eye(nc), a)
        vals = a
    elif kind == 'lame':
        lam, mu = 1. + rng.random(N), 0.5 + rng.random(N)
        m = np.zeros(nc)
        m[:len(N)] = 1.
        A = np.einsum('i,j,...->ij...', m, m, lam) + 2. * np.einsum('ij,...->ij...', np.eye(nc), mu)
        vals = np.stack([lam, mu])
    else:
        B = rng.standard_normal((nc, nc) + N) * 0.2
        A = np.einsum('ik...,jk...->ij...', B, B) + 2. * np.einsum('ij,...->ij...', np.eye(nc), np.ones(N))
        vals = A
    plan.set_coeff(kind, vals)
    hG = Field(val=G1 if pd == 'primal' else G2, order=2, N=N, multype=21, Fourier=True)
    Afun = Operator([[DFT(N, inverse=True), hG, DFT(N), Field(val=A, order=2, multype=21)]])
    x = rng.standard_normal((nc,) + N)
    ref = Afun(Field(val=x, order=1)).val
    for dt, tol in ((np.float64, 1e-12), (np.float32, 5e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dy = nat.DeviceArray((nc,) + N, dt)
        plan.matvec(dx, dy)
        assert rel(dy.to_host(), ref) < tol, (dt, rel(dy.to_host(), ref))


@pytest.mark.parametrize('N', [(17, 17, 17), (101, 101), (19, 23)])
def test_prime_length_gani_solve_matches_oracle(N):
    """GaNi solves of the cube / square example material on prime grids: iteration counts equal,
    bounds 1e-10, in fp64 and mixed precision (bounds 1e-8)."""
    print('case bluestein solve', N, flush=True)
    dim = len(N)
    mats = examples.scalar_3d_materials if dim == 3 else examples.scalar_2d_materials
    for prec in ('fp64', 'mixed'):
        conf = {'name': 'prime', 'physics': 'scalar', 'material': 'cube' if dim == 3 else 'square',
                'solve': {'kind': 'GaNi', 'N': np.array(N), 'primaldual': ['primal', 'dual']},
                'postprocess': [{'kind': 'GaNi'}],
                'solver': {'kind': 'CG', 'tol': 1e-8, 'maxiter': 1e3, 'precision': prec}}
        pb = Problem(dict(conf), type('conf', (), {'materials': mats()}))
        pb.calculate()
        ref = OracleProblem(dict(conf, solver={'kind': 'CG', 'tol': 1e-8, 'maxiter': 1e3}), mats())
        ref.calculate()
        tol = 1e-10 if prec == 'fp64' else 1e-8
        for pd in ('primal', 'dual'):
            AH = pb.output['mat_' + pd]['AH_GaNi_' + pd]
            AR = ref.output['mat_' + pd]['AH_GaNi_' + pd]
            assert np.abs(AH - AR).max() < tol * np.abs(AR).max(), (prec, pd)
            if prec == 'fp64':
                kits = [r['info']['kit'] for r in pb.output['res_' + pd]]
                assert kits == [r['info']['kit'] for r in ref.output['res_' + pd]], pd


@pytest.mark.parametrize('N', [(17, 17), (17, 17, 17)])
def test_ga_solve_on_prime_coarse_grid_matches_oracle(N):
    """A Ga solve whose solution grid N = 17 has no radix schedule (2N-1 = 33 has one): the coarse
    plan now exists, and bounds and iteration counts equal the oracle's."""
    print('case bluestein ga solve', N, flush=True)
    dim = len(N)
    mats = examples.scalar_3d_materials if dim == 3 else examples.scalar_2d_materials
    conf = {'name': 'ga17', 'physics': 'scalar', 'material': 'cube' if dim == 3 else 'square',
            'solve': {'kind': 'Ga', 'N': np.array(N), 'primaldual': ['primal', 'dual']},
            'postprocess': [{'kind': 'Ga'}, {'kind': 'GaNi'}],
            'solver': {'kind': 'CG', 'tol': 1e-8, 'maxiter': 1e3}}
    pb = Problem(dict(conf), type('conf', (), {'materials': mats()}))
    pb.calculate()
    ref = OracleProblem(dict(conf), mats())
    ref.calculate()
    for pd in ('primal', 'dual'):
        for key in ('AH_Ga_' + pd, 'AH_GaNi_' + pd):
            AH, AR = pb.output['mat_' + pd][key], ref.output['mat_' + pd][key]
            assert np.abs(AH - AR).max() < 1e-10 * np.abs(AR).max(), (pd, key)
        kits = [r['info']['kit'] for r in pb.output['res_' + pd]]
        assert kits == [r['info']['kit'] for r in ref.output['res_' + pd]], pd
