"""Pins of the oracle beyond the goldens: closed forms, library routines, brute force.

Mirrors the reference's own unit tests (ffthompy/tensors/unittest_operators.py,
ffthompy/general/unittest_solver.py, ffthompy/tensors/unittest_tensors.py,
ffthompy/unittest_materials.py) as property tests on the restatement.
"""
import itertools

import numpy as np
import pytest

from oracle import fields, grid, projections
from oracle.fields import DFT, Field, Operator
from oracle.homogenize import Problem
from oracle.matcoef import ElasticTensor
from oracle.materials import Material
from oracle.solver import linear_solver


@pytest.mark.parametrize('dim', [2, 3])
@pytest.mark.parametrize('n', [5, 6, 7])
def test_scalar_projection_equals_closed_form(dim, n):
    """ffthompy/general/unittest_solver.py:21-33: materialised G == xi xi^T/|xi|^2.
    Even n: the materialised form nulls Nyquist modes, the closed form does not,
    so even grids are compared away from Nyquist frequencies only."""
    N = (n,) * dim
    Y = np.ones(dim)
    G0, G1, G2 = projections.scalar(N, Y, NyqNul=(n % 2 == 1))
    C0, C1, C2 = projections.scalar_closed_form(N, Y)
    for a, b in ((G0.val, C0), (G1.val, C1), (G2.val, C2)):
        assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize('dim', [2, 3])
def test_projection_idempotent_orthogonal(dim):
    """ffthompy/tensors/unittest_operators.py:177-203."""
    N = (5,) * dim
    Y = np.ones(dim)
    for P, Q in itertools.product(projections.scalar(N, Y), repeat=2):
        prod = np.einsum('ij...,jk...->ik...', P.val, Q.val)
        ref = P.val if P is Q else 0.
        assert np.abs(prod - ref).max() < 1e-13
    for P, Q in itertools.product(projections.elasticity(N, Y), repeat=2):
        prod = np.einsum('ij...,jk...->ik...', P.val, Q.val)
        ref = P.val if P is Q else 0.
        assert np.abs(prod - ref).max() < 1e-13


@pytest.mark.parametrize('fft_form', ['r', 0])
def test_elastic_projection_is_mandel_gamma(fft_form):
    """G1h+G1s equals the Mandel form of the 4th-order small-strain Gamma
    (ffthompy/tensors/projection.py:33-51), and G2h+G2s = I[xi!=0] - G1."""
    N = (5, 7, 3)
    Y = np.array([1., 2., 0.5])
    G0, G1h, G1s, G2h, G2s = projections.elasticity(N, Y, fft_form=fft_form)
    M = projections.mandel_matrix_of(projections.gamma_small_strain(N, Y, fft_form=fft_form))
    assert np.abs((G1h.val + G1s.val) - M).max() < 1e-13
    I0 = np.zeros_like(M)
    for a in range(6):
        I0[a, a] = 1.
        I0[(a, a) + (0, 0, 0)] = 0.
    assert np.abs((G2h.val + G2s.val) - (I0 - M)).max() < 1e-13


def _nyquist_mask(N, fft_form):
    """True where some even axis sits at its Nyquist frequency -N/2 (layout of fft_form)."""
    xi = grid.frequencies(N, np.ones(len(N)), fft_form=fft_form)
    m = np.zeros(tuple(len(x) for x in xi), dtype=bool)
    for a, (x, n) in enumerate(zip(xi, N)):
        if n % 2 == 0:
            shp = [1] * len(N)
            shp[a] = len(x)
            m |= (x == -n // 2).reshape(shp)
    return m


@pytest.mark.parametrize('N', [(6, 8, 4), (5, 7, 9), (8, 6), (4, 7)])
@pytest.mark.parametrize('fft_form', ['r', 0, 'c'])
def test_projections_of_every_form_equal_closed_forms(N, fft_form):
    """Scalar G1 = xi xi^T/|xi|^2 and elastic G1h+G1s = Mandel(Gamma) (ffthompy/tensors/projection.py)
    on the frequencies of each layout, -N/2 at the Nyquist index: exactly with NyqNul=False, and
    with NyqNul=True away from the Nyquist modes, which are zero."""
    d = len(N)
    Y = np.array([1., 2., 0.5])[:d]
    mask = _nyquist_mask(N, fft_form)
    C0, C1, C2 = projections.scalar_closed_form(N, Y, fft_form=fft_form)
    Gam = projections.mandel_matrix_of(projections.gamma_small_strain(N, Y, fft_form=fft_form))
    for nyq in (False, True):
        _, G1, G2 = projections.scalar(N, Y, NyqNul=nyq, fft_form=fft_form)
        _, a, b, c, e = projections.elasticity(N, Y, NyqNul=nyq, fft_form=fft_form)
        for got, ref in ((G1.val, C1), (G2.val, C2), (a.val + b.val, Gam)):
            if nyq:
                assert np.abs(got[..., mask]).max(initial=0.) == 0.
                got, ref = got[..., ~mask], ref[..., ~mask]
            assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize('N', [(5, 5), (5, 4), (3, 5, 7), (4, 4, 4)])
def test_fft_forms_round_trip_and_dft_matrix(N):
    """ffthompy/tensors/unittest_operators.py:23-100 and matvecs/unittest_matvec.py:14-28."""
    rng = np.random.default_rng(0)
    u = Field(val=rng.random(N), order=0)
    for form in ('r', 0, 'c'):
        v = u.copy(fft_form=form)
        Fu = DFT(N, fft_form=form)(v)
        back = DFT(N, inverse=True, fft_form=form)(Fu)
        assert np.abs(back.val - u.val).max() < 1e-13
        for other in ('r', 0, 'c'):
            Fc = Fu.copy().set_fft_form(other)
            assert abs(Fc.norm() - Fu.norm()) < 1e-13
            assert np.abs(Fc.copy().set_fft_form(form).val - Fu.val).max() < 1e-13
    Fc = DFT(N, fft_form='c')(u.copy(fft_form='c'))
    Mdft = DFT(N, fft_form='c').matrix()
    assert np.abs(Mdft @ u.val.ravel() - Fc.val.ravel()).max() < 1e-13


@pytest.mark.parametrize('dim,n', [(2, 4), (2, 5), (3, 4), (3, 5)])
def test_project_to_double_grid(dim, n):
    """ffthompy/tensors/unittest_tensors.py:11-47: interpolation onto 2N keeps the
    mean and, for odd n, the norm; nodal values are reproduced."""
    rng = np.random.default_rng(1)
    N = (n,) * dim
    M = tuple(2 * np.array(N))
    u = Field(val=rng.random(N), order=0)
    uM = u.fourier(copy=True).project(M).fourier()
    if n % 2:
        assert abs(u.norm() - uM.norm()) < 1e-12
    else:
        assert u.norm() >= uM.norm()
    assert abs(u.mean() - uM.mean()) < 1e-13
    assert np.abs(u.val - uM.val[tuple(slice(0, m, 2) for m in M)]).max() < 1e-12


def _dense_operator(Afun, shape, N):
    n = int(np.prod(shape + N))
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.
        cols.append(Afun(Field(val=e.reshape(shape + N), order=len(shape))).val.ravel())
    return np.array(cols).T


def test_cg_matches_dense_solve():
    """CG on G A x = b agrees with a dense least-squares solve on range(G)
    (brute force at N=5x5, random coefficient 1+10U as in unittest_solver.py:35-76)."""
    rng = np.random.default_rng(2)
    dim, N = 2, (5, 5)
    _, G1, _ = projections.scalar(N, np.ones(dim))
    F, iF = DFT(N), DFT(N, inverse=True)
    GN = Operator([[iF, G1, F]])
    A = Field(val=np.einsum('ij,...->ij...', np.eye(dim), 1. + 10. * rng.random(N)), order=2, multype=21)
    Afun = Operator([[GN, A]])
    E = Field(shape=(dim,), N=N)
    E.set_mean(np.array([1., 0.]))
    B = Afun(-E)
    x, info = linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': 1e-12, 'maxiter': 1000}, solver='CG')
    # The CG iterates live in range(G): solve Q^T A Q y = Q^T b with Q an orthonormal
    # basis of range(G) (eigenvectors of the dense projection with eigenvalue 1).
    Gd = _dense_operator(GN, (dim,), N)
    w, V = np.linalg.eigh(0.5 * (Gd + Gd.T))
    Q = V[:, w > 0.5]
    Ad = np.kron(np.eye(dim), np.diag(A.val[0, 0].ravel()))
    y = np.linalg.solve(Q.T @ Ad @ Q, Q.T @ B.val.ravel())
    assert np.abs(x.val.ravel() - Q @ y).max() < 1e-9
    assert info['norm_res'] <= 1e-12


def test_richardson_matches_cg():
    """ffthompy/general/unittest_solver.py:35-76 (CG vs richardson)."""
    rng = np.random.default_rng(3)
    dim, N = 2, (5, 5)
    _, G1, _ = projections.scalar(N, np.ones(dim))
    GN = Operator([[DFT(N, inverse=True), G1, DFT(N)]])
    A = Field(val=np.einsum('ij,...->ij...', np.eye(dim), 1. + 10. * rng.random(N)), order=2, multype=21)
    Afun = Operator([[GN, A]])
    E = Field(shape=(dim,), N=N)
    E.set_mean(np.array([1., 0.]))
    B = Afun(-E)
    par = {'tol': 1e-10, 'maxiter': 10000, 'alpha': 0.5 * (1. + 11.)}
    X, _ = linear_solver(Afun, B, x0=B.zeros_like(), par=dict(par), solver='CG')
    x, _ = linear_solver(Afun, B, x0=B.zeros_like(), par=dict(par), solver='richardson')
    assert np.linalg.norm(X.val - x.val) < 1e-8


def _diagonal_operator(rng, N, e0, e1):
    lam = e0 + (e1 - e0) * rng.random((1,) + N)
    lam[0, 0, 0], lam[0, 0, 1] = e0, e1
    return lam, (lambda x: x.copy(val=lam * x.val))


def test_chebyshev_residual_is_the_chebyshev_polynomial():
    """Textbook: after k steps of the two-term Chebyshev iteration on [e0, e1] the residual is
    T_k((d - A)/c) / T_k(d/c) r0, d, c the centre and half-width (pins the (p, w) sequence)."""
    from oracle.solver import cheby2TERM
    rng = np.random.default_rng(0)
    N, e0, e1 = (7, 7), 1., 10.
    lam, A = _diagonal_operator(rng, N, e0, e1)
    B = Field(val=rng.standard_normal((1,) + N), order=1)
    d, c = (e0 + e1) / 2, (e1 - e0) / 2
    for k in range(1, 9):
        x, res = cheby2TERM(A, B, B.zeros_like(), par={'eigrange': [e0, e1], 'tol': 0., 'maxit': k})
        Tk = np.polynomial.chebyshev.Chebyshev.basis(k)
        pred = Tk((d - lam) / c) / Tk(d / c) * B.val
        assert res['kit'] == k
        assert np.abs((B.val - lam * x.val) - pred).max() < 1e-13


def test_richardson_error_is_a_power_of_the_iteration_matrix():
    """x_k - x = (I - A/alpha)^k (x_0 - x) on a diagonal operator."""
    from oracle.solver import richardson
    rng = np.random.default_rng(1)
    N = (6, 5)
    lam, A = _diagonal_operator(rng, N, 1., 4.)
    B = Field(val=rng.standard_normal((1,) + N), order=1)
    xs = B.val / lam
    for k in (1, 2, 5):
        x, res = richardson(A, B, B.zeros_like(), par={'alpha': 3., 'tol': 0., 'maxiter': k})
        assert res['kit'] == k
        assert np.abs((x.val - xs) - (1. - lam / 3.) ** k * (0. - xs)).max() < 1e-13


def test_chebyshev_matches_cg():
    """ffthompy/general/unittest_solver.py:35-76 (CG vs chebyshev, eigenvalue range of a)."""
    rng = np.random.default_rng(4)
    dim, N = 2, (5, 5)
    _, G1, _ = projections.scalar(N, np.ones(dim))
    GN = Operator([[DFT(N, inverse=True), G1, DFT(N)]])
    A = Field(val=np.einsum('ij,...->ij...', np.eye(dim), 1. + 10. * rng.random(N)), order=2, multype=21)
    Afun = Operator([[GN, A]])
    E = Field(shape=(dim,), N=N)
    E.set_mean(np.array([1., 0.]))
    B = Afun(-E)
    par = {'tol': 1e-10, 'maxiter': 10000, 'eigrange': [1., 11.]}
    X, _ = linear_solver(Afun, B, x0=B.zeros_like(), par=dict(par), solver='CG')
    x, info = linear_solver(Afun, B, x0=B.zeros_like(), par=dict(par), solver='chebyshev')
    assert info['norm_res'] <= 1e-10
    assert np.linalg.norm(X.val - x.val) < 1e-8


def _scalar_problem(material, N, pd=('primal', 'dual')):
    return Problem({'physics': 'scalar', 'material': material,
                    'solve': {'kind': 'GaNi', 'N': np.array(N), 'primaldual': list(pd)},
                    'postprocess': [{'kind': 'GaNi'}],
                    'solver': {'kind': 'CG', 'tol': 1e-12, 'maxiter': 1e3}})


def test_constant_coefficient_gives_exact_ah():
    """A = C everywhere: B = -G C E = 0, so kit = 0 and AH = C."""
    C = np.array([[3., 1.], [1., 2.]])
    mat = {'inclusions': ['all'], 'positions': [''], 'params': [''], 'vals': [C], 'Y': np.ones(2)}
    pb = _scalar_problem(mat, (7, 7))
    pb.calculate()
    assert all(r['info']['kit'] == 0 for r in pb.output['res_primal'])
    assert np.abs(pb.output['mat_primal']['AH_GaNi_primal'] - C).max() < 1e-14
    assert np.abs(pb.output['mat_dual']['AH_GaNi_dual'] - C).max() < 1e-13


def test_laminate_gives_arithmetic_and_harmonic_means():
    """A grid-aligned laminate: AH = arithmetic mean along the layers and harmonic
    mean across them (examples/scalar/scalar_2d.py laminate material)."""
    N = (9, 9)
    vals = np.where(np.arange(9) < 4, 11., 1.)

    def fun(coord):
        a = np.broadcast_to(vals[:, None], N)
        return np.multiply.outer(np.eye(2), a)

    pb = _scalar_problem({'fun': fun, 'Y': np.ones(2)}, N)
    pb.calculate()
    AH = pb.output['mat_primal']['AH_GaNi_primal']
    assert abs(AH[1, 1] - vals.mean()) < 1e-10
    assert abs(AH[0, 0] - 1. / np.mean(1. / vals)) < 1e-10
    assert np.abs(pb.output['mat_dual']['AH_GaNi_dual'] - AH).max() < 1e-10


def test_voigt_reuss_bounds_and_primal_dual():
    rng = np.random.default_rng(4)
    N = (7, 7, 7)
    a = np.exp(rng.standard_normal(N))

    def fun(coord):
        return np.multiply.outer(np.eye(3), a)

    pb = _scalar_problem({'fun': fun, 'Y': np.ones(3)}, N)
    pb.calculate()
    AH = pb.output['mat_primal']['AH_GaNi_primal']
    AHd = pb.output['mat_dual']['AH_GaNi_dual']
    ev = np.linalg.eigvalsh(AH)
    assert ev.min() >= 1. / np.mean(1. / a) - 1e-12 and ev.max() <= a.mean() + 1e-12
    assert np.abs(AH - AH.T).max() < 1e-12
    assert np.abs(AH - AHd).max() < 1e-9 * np.abs(AH).max()


def test_elastic_constants_mandel():
    """ffthompy/mechanics/unittest_matcoef.py: C = 3K vol + 2 mu dev; Mandel form is
    lambda m m^T + 2 mu I with m = (1,1,1,0,0,0)."""
    K, mu = 10., 5.
    C = ElasticTensor(bulk=K, mu=mu).mandel
    lam = K - 2. * mu / 3.
    m = np.array([1., 1., 1., 0., 0., 0.])
    assert np.abs(C - (lam * np.outer(m, m) + 2 * mu * np.eye(6))).max() < 1e-13
    S = ElasticTensor(bulk=K, mu=mu, stiffness=False).mandel
    assert np.abs(C @ S - np.eye(6)).max() < 1e-13


@pytest.mark.parametrize('dim', [2, 3])
@pytest.mark.parametrize('kind', ['square', 'pyramid'])
def test_exact_integration_equals_grid_integration(dim, kind):
    """ffthompy/unittest_materials.py:17-59: order=None (exact) == order 0/1 on the grid."""
    N = (5,) * dim
    base = {'positions': [np.zeros(dim), ''], 'vals': [10 * np.eye(dim), 1. * np.eye(dim)],
            'Y': np.ones(dim)}
    if kind == 'square':
        c0 = dict(base, inclusions=['square', 'otherwise'], params=[0.6 * np.ones(dim), ''], order=None)
        c1 = dict(base, inclusions=['square', 'otherwise'], params=[0.6 * np.ones(dim), ''], order=0,
                  P=np.array(N))
    else:
        c0 = dict(base, inclusions=['pyramid', 'all'], params=[0.8 * np.ones(dim), ''], order=None)
        c1 = dict(base, inclusions=['pyramid', 'all'], params=[0.8 * np.ones(dim), ''], order=1,
                  P=np.array(N))
    c0['positions'] = [np.zeros(dim), '']
    c1['positions'] = [np.zeros(dim), '']
    A0 = Material(c0).get_A_Ga(N)
    A1 = Material(c1).get_A_Ga(N)
    assert np.linalg.norm(A0.val[0, 0] - A1.val[0, 0]) < 1e-13


def test_pointwise_inverse():
    rng = np.random.default_rng(5)
    A = rng.random((3, 3, 4, 5)) + 3 * np.eye(3)[:, :, None, None]
    inv = grid.pointwise_inverse(A)
    prod = np.einsum('ij...,jk...->ik...', A, inv)
    assert np.abs(prod - np.eye(3)[:, :, None, None]).max() < 1e-13


def test_rfft_matches_library():
    rng = np.random.default_rng(6)
    x = rng.random((3, 5, 7, 9))
    X = fields.rfftn(x, (5, 7, 9))
    assert X.shape == (3, 5, 7, 5)
    assert np.abs(fields.irfftn(X, (5, 7, 9)) - x).max() < 1e-13
    for c in range(3):
        assert np.abs(X[c] - np.fft.fftn(x[c])[..., :5]).max() < 1e-12
