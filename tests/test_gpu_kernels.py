"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (float64 unless stated): FFT and operator results agree with NumPy /
the oracle to 1e-12 relative to the largest entry; float32 kernels to 2e-5
relative; CG iteration counts exactly; homogenized matrices to the reference's own
1e-9 absolute (run_unittests.py:64).
"""
import numpy as np
import pytest

from oracle import projections
from oracle.fields import DFT, Field, Operator, rfftn as o_rfftn, irfftn as o_irfftn
from oracle.homogenize import assembly_matrix
from oracle.solver import linear_solver as o_linear_solver

pytestmark = pytest.mark.gpu

nat = pytest.importorskip('ffthompy_amd._native')


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


FFT_SHAPES = [(5, 5), (9, 15), (4, 6), (26, 21), (45, 45), (5, 5, 5), (9, 7, 11), (4, 6, 8),
              (13, 10, 9), (15, 15, 15), (21, 33, 25), (63, 45, 27), (99, 81, 105)]


@pytest.mark.parametrize('N', FFT_SHAPES, ids=[str(s) for s in FFT_SHAPES])
def test_rfftn_irfftn_match_numpy(N):
    print('case rfftn', N, flush=True)
    rng = np.random.default_rng(10)
    nc = 2
    plan = nat.Plan(N, np.ones(len(N)))
    x = rng.standard_normal((nc,) + N)
    Nh = N[:-1] + (N[-1] // 2 + 1,)
    ref = o_rfftn(x, N)
    for dt, ct, tol in ((np.float64, np.complex128, 1e-12), (np.float32, np.complex64, 2e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dX = nat.DeviceArray((nc,) + Nh, ct)
        plan.rfftn(dx, dX, nc)
        X = dX.to_host()
        assert rel(X, ref) < tol, (dt, rel(X, ref))
        # inverse on an arbitrary (non-Hermitian) spectrum, as numpy.fft.irfftn defines it
        Z = (rng.standard_normal((nc,) + Nh) + 1j * rng.standard_normal((nc,) + Nh)).astype(ct)
        dZ = nat.DeviceArray.from_host(Z)
        dz = nat.DeviceArray((nc,) + N, dt)
        plan.irfftn(dZ, dz, nc)
        zref = o_irfftn(Z.astype(np.complex128), N)
        assert rel(dz.to_host(), zref) < tol, (dt, rel(dz.to_host(), zref))


def _projections(physics, N, Y):
    if physics == 'scalar':
        _, G1, G2 = projections.scalar(N, Y)
        return G1.val, G2.val
    _, G1h, G1s, G2h, G2s = projections.elasticity(N, Y)
    return G1h.val + G1s.val, G2h.val + G2s.val


PROJ_CASES = [('scalar', (5, 7)), ('scalar', (6, 8)), ('scalar', (5, 7, 9)), ('scalar', (4, 6, 5)),
              ('elasticity', (5, 7)), ('elasticity', (5, 7, 9)), ('elasticity', (9, 5, 7))]


@pytest.mark.parametrize('physics,N', PROJ_CASES, ids=['%s-%s' % c for c in PROJ_CASES])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_projection_matches_materialised(physics, N, pd):
    print('case projection', physics, N, pd, flush=True)
    rng = np.random.default_rng(11)
    Y = np.linspace(1., 2., len(N))
    G1, G2 = _projections(physics, N, Y)
    G = G1 if pd == 'primal' else G2
    plan = nat.Plan(N, Y, physics, pd)
    nc = plan.ncomp
    Nh = N[:-1] + (N[-1] // 2 + 1,)
    X = rng.standard_normal((nc,) + Nh) + 1j * rng.standard_normal((nc,) + Nh)
    ref = np.einsum('ij...,j...->i...', G, X)
    dX = nat.DeviceArray.from_host(X)
    dY = nat.DeviceArray((nc,) + Nh, np.complex128)
    plan.apply_projection_hat(dX, dY)
    assert rel(dY.to_host(), ref) < 1e-13


def _random_coef(rng, kind, nc, N, physics):
    n = int(np.prod(N))
    if kind == 'iso':
        a = 1. + 10. * rng.random(N)
        A = np.einsum('ij,...->ij...', np.eye(nc), a)
        return a, A
    if kind == 'lame':
        lam = 1. + rng.random(N)
        mu = 0.5 + rng.random(N)
        m = np.zeros(nc)
        m[:3 if nc == 6 else 2] = 1.
        A = np.einsum('i,j,...->ij...', m, m, lam) + 2 * np.einsum('ij,...->ij...', np.eye(nc), mu)
        return np.stack([lam, mu]), A
    M = rng.standard_normal((nc, nc) + N) * 0.3
    A = np.einsum('ik...,jk...->ij...', M, M) + 2. * np.eye(nc).reshape((nc, nc) + (1,) * len(N))
    return A.reshape((nc * nc,) + N), A


MV_CASES = [('scalar', 'iso', (5, 5)), ('scalar', 'full', (9, 15)), ('scalar', 'iso', (5, 5, 5)),
            ('scalar', 'full', (9, 7, 11)), ('scalar', 'iso', (4, 6, 8)), ('elasticity', 'lame', (5, 7, 9)),
            ('elasticity', 'full', (5, 5, 5)), ('elasticity', 'lame', (7, 9)), ('scalar', 'iso', (45, 45, 45))]


@pytest.mark.parametrize('physics,kind,N', MV_CASES, ids=['%s-%s-%s' % c for c in MV_CASES])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_matvec_matches_oracle_operator(physics, kind, N, pd):
    print('case matvec', physics, kind, N, pd, flush=True)
    rng = np.random.default_rng(12)
    Y = np.linspace(1., 1.5, len(N))
    G1, G2 = _projections(physics, N, Y)
    plan = nat.Plan(N, Y, physics, pd)
    nc = plan.ncomp
    vals, A = _random_coef(rng, kind, nc, N, physics)
    plan.set_coeff(kind, vals)
    hG = Field(val=G1 if pd == 'primal' else G2, order=2, N=N, multype=21, Fourier=True)
    Afun = Operator([[DFT(N, inverse=True), hG, DFT(N), Field(val=A, order=2, multype=21)]])
    x = rng.standard_normal((nc,) + N)
    ref = Afun(Field(val=x, order=1)).val
    for dt, tol in ((np.float64, 1e-12), (np.float32, 2e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dy = nat.DeviceArray((nc,) + N, dt)
        plan.matvec(dx, dy)
        assert rel(dy.to_host(), ref) < tol, (dt, rel(dy.to_host(), ref))
    # bare projection
    refp = Operator([[DFT(N, inverse=True), hG, DFT(N)]])(Field(val=x, order=1)).val
    dx = nat.DeviceArray.from_host(x)
    dy = nat.DeviceArray((nc,) + N)
    plan.project(dx, dy)
    assert rel(dy.to_host(), refp) < 1e-12


def test_energy_matrix_matches_assembly():
    print('case energy', flush=True)
    rng = np.random.default_rng(13)
    N = (9, 7, 11)
    for physics, kind in (('scalar', 'iso'), ('scalar', 'full'), ('elasticity', 'lame'), ('elasticity', 'full')):
        plan = nat.Plan(N, np.ones(3), physics)
        nc = plan.ncomp
        vals, A = _random_coef(rng, kind, nc, N, physics)
        plan.set_coeff(kind, vals)
        sols = [rng.standard_normal((nc,) + N) for _ in range(nc)]
        ref = assembly_matrix(Field(val=A, order=2, multype=21), [Field(val=s, order=1) for s in sols])
        AH = plan.energy_matrix([nat.DeviceArray.from_host(s) for s in sols])
        assert rel(AH, ref) < 1e-13, (physics, kind)


def _oracle_solve(N, Y, physics, pd, A, E, tol):
    G1, G2 = _projections(physics, N, Y)
    hG = Field(val=G1 if pd == 'primal' else G2, order=2, N=N, multype=21, Fourier=True)
    Afun = Operator([[DFT(N, inverse=True), hG, DFT(N), Field(val=A, order=2, multype=21)]])
    EN = Field(shape=(len(E),), N=N)
    EN.set_mean(E)
    B = Afun(-EN)
    x, info = o_linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': tol, 'maxiter': 1000}, solver='CG')
    return B.val, x.val, info


@pytest.mark.parametrize('physics,kind,N', [('scalar', 'iso', (9, 9, 9)), ('scalar', 'full', (15, 11)),
                                            ('elasticity', 'lame', (7, 9, 5))])
def test_cg_fp64_matches_oracle(physics, kind, N):
    print('case cg', physics, kind, N, flush=True)
    rng = np.random.default_rng(14)
    Y = np.ones(len(N))
    plan = nat.Plan(N, Y, physics, 'primal')
    nc = plan.ncomp
    vals, A = _random_coef(rng, kind, nc, N, physics)
    plan.set_coeff(kind, vals)
    E = np.zeros(nc)
    E[0] = 1.
    B, xref, info = _oracle_solve(N, Y, physics, 'primal', A, E, 1e-10)
    dB = nat.DeviceArray.from_host(B)
    dx = nat.DeviceArray.zeros((nc,) + N)
    out = plan.cg_solve(dB, dx, tol=1e-10, maxiter=1000, hist_len=1001)
    assert out['status'] == nat.FH_CG_OK
    assert out['kit'] == info['kit'], (out['kit'], info['kit'])
    assert np.abs(out['res_log'] - np.array(info['norm_res_log'])).max() < 1e-9 * info['norm_res_log'][0]
    assert rel(dx.to_host(), xref) < 1e-8


# Shapes that put every specialised length (5, 9, 15, 25, 27, 45, 63, 75, 81, 99, 105, 125, 135, 225,
# 243) on the contiguous axis, the middle axis and the outer (projection) axis.
FIXED_SHAPES = [(243, 9, 243), (5, 243, 27), (27, 45, 63), (81, 99, 105), (135, 15, 9), (9, 5, 15),
                (45, 63, 81), (99, 105, 135), (63, 81), (105, 135), (243, 45), (15, 99),
                (25, 75, 125), (125, 225, 25), (75, 25, 225), (225, 75), (125, 25)]


@pytest.mark.parametrize('N', FIXED_SHAPES, ids=[str(s) for s in FIXED_SHAPES])
def test_fixed_length_fft_matches_numpy(N):
    print('case fixed rfftn', N, flush=True)
    rng = np.random.default_rng(20)
    nc = 1
    plan = nat.Plan(N, np.ones(len(N)))
    x = rng.standard_normal((nc,) + N)
    Nh = N[:-1] + (N[-1] // 2 + 1,)
    ref = o_rfftn(x, N)
    for dt, ct, tol in ((np.float64, np.complex128, 1e-12), (np.float32, np.complex64, 2e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dX = nat.DeviceArray((nc,) + Nh, ct)
        plan.rfftn(dx, dX, nc)
        assert rel(dX.to_host(), ref) < tol, (dt, rel(dX.to_host(), ref))
        dz = nat.DeviceArray((nc,) + N, dt)
        plan.irfftn(dX, dz, nc)
        assert rel(dz.to_host(), x) < tol


FIXED_MV = [('scalar', (243, 9, 45)), ('scalar', (63, 81)), ('scalar', (105, 135)), ('elasticity', (27, 45, 15)),
            ('elasticity', (81, 9, 5)), ('scalar', (99, 27, 63)), ('elasticity', (15, 45)),
            ('scalar', (25, 75, 125)), ('scalar', (225, 125)), ('elasticity', (75, 25, 25))]


@pytest.mark.parametrize('physics,N', FIXED_MV, ids=['%s-%s' % c for c in FIXED_MV])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_fixed_matvec_matches_oracle_and_generic(physics, N, pd):
    print('case fixed matvec', physics, N, pd, flush=True)
    rng = np.random.default_rng(21)
    Y = np.linspace(1., 1.3, len(N))
    G1, G2 = _projections(physics, N, Y)
    plan = nat.Plan(N, Y, physics, pd)
    nc = plan.ncomp
    kind = 'iso' if physics == 'scalar' else 'lame'
    vals, A = _random_coef(rng, kind, nc, N, physics)
    plan.set_coeff(kind, vals)
    hG = Field(val=G1 if pd == 'primal' else G2, order=2, N=N, multype=21, Fourier=True)
    x = rng.standard_normal((nc,) + N)
    ref = Operator([[DFT(N, inverse=True), hG, DFT(N), Field(val=A, order=2, multype=21)]])(Field(val=x, order=1)).val
    for dt, tol in ((np.float64, 1e-12), (np.float32, 2e-5)):
        dx = nat.DeviceArray.from_host(x.astype(dt))
        dy = nat.DeviceArray((nc,) + N, dt)
        plan.use_generic(False)
        plan.matvec(dx, dy)
        yf = dy.to_host()
        plan.use_generic(True)
        plan.matvec(dx, dy)
        yg = dy.to_host()
        assert rel(yf, ref) < tol, (dt, rel(yf, ref))
        assert rel(yg, ref) < tol, (dt, rel(yg, ref))


MIXED_CASES = [('scalar', 'iso', (45, 45, 45), 100.), ('scalar', 'iso', (63, 81), 100.),
               ('elasticity', 'lame', (27, 27, 27), 10.), ('scalar', 'full', (21, 25, 27), 10.)]


@pytest.mark.parametrize('physics,kind,N,contrast', MIXED_CASES, ids=['%s-%s-%s' % c[:3] for c in MIXED_CASES])
def test_mixed_precision_solve_matches_fp64(physics, kind, N, contrast):
    """fp32 CG with fp64 reliable updates, against the plain fp64 solve. Each fp32 correction is
    projected in fp64 before it is added, so both solutions lie in the range of G and differ by
    what the tolerance leaves: the field is checked to 1e-6 and the energy (AH entry) to 1e-12
    (with the corrections folded unprojected and x projected once at the end the field differed
    at first order in the fp32 drift, 1e-5 here). The fp64 solve also matches the oracle's CG
    iteration count."""
    print('case mixed', physics, kind, N, flush=True)
    rng = np.random.default_rng(30)
    Y = np.ones(len(N))
    plan = nat.Plan(N, Y, physics, 'primal')
    nc = plan.ncomp
    if kind == 'iso':
        a = np.where(rng.random(N) < 0.3, contrast, 1.)
        vals, A = a, np.einsum('ij,...->ij...', np.eye(nc), a)
    elif kind == 'lame':
        ph = rng.random(N) < 0.3
        lam = np.where(ph, contrast, 1.) * 0.5
        mu = np.where(ph, contrast, 1.) * 0.4
        m = np.zeros(nc)
        m[:3 if nc == 6 else 2] = 1.
        vals = np.stack([lam, mu])
        A = np.einsum('i,j,...->ij...', m, m, lam) + 2 * np.einsum('ij,...->ij...', np.eye(nc), mu)
    else:
        vals, A = _random_coef(rng, 'full', nc, N, physics)
    plan.set_coeff(kind, vals)
    E = np.zeros(nc)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((nc,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((nc,) + N)
    plan.matvec(EN, B)
    sols = {}
    infos = {}
    for mode in ('fp64', 'mixed'):
        x = nat.DeviceArray.zeros((nc,) + N)
        infos[mode] = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode=mode, delta=0.01)
        assert infos[mode]['status'] == nat.FH_CG_OK
        nat.add_constant(x, E)
        sols[mode] = x
    x64, xmx = sols['fp64'].to_host(), sols['mixed'].to_host()
    assert rel(xmx, x64) < 1e-6, rel(xmx, x64)
    e64 = np.einsum('i...,i...', np.einsum('ij...,j...->i...', A, x64), x64).mean()
    emx = np.einsum('i...,i...', np.einsum('ij...,j...->i...', A, xmx), xmx).mean()
    assert abs(emx - e64) / abs(e64) < 1e-12, (emx, e64)
    assert infos['mixed']['norm_res'] <= 1e-8
    if int(np.prod(N)) <= 30000:
        Bh, xref, info = _oracle_solve(N, Y, physics, 'primal', A, E, 1e-8)
        assert infos['fp64']['kit'] == info['kit']


@pytest.mark.parametrize('N', [(45, 45, 45), (63, 81)])
def test_fused_mixed_iteration_matches_operator_form(N):
    """The fused single-precision iteration (denominator <p, a p>, solution step deferred
    into the next forward pass, residual updated in the inverse pass) against the
    operator-form iteration: same iteration count within 2 and the same energy to 1e-10."""
    print('case fused mixed', N, flush=True)
    rng = np.random.default_rng(31)
    a = np.where(rng.random(N) < 0.3, 100., 1.)
    d = len(N)
    plan = nat.Plan(N, np.ones(d), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    E = np.zeros(d)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((d,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((d,) + N)
    plan.matvec(EN, B)
    out = {}
    for fused in (True, False):
        plan.set_cg_variant(fused)
        x = nat.DeviceArray.zeros((d,) + N)
        info = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode='mixed', delta=0.01)
        assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= 1e-8
        nat.add_constant(x, E)
        xh = x.to_host()
        out[fused] = (info['kit'], float(np.mean(a * np.sum(xh * xh, axis=0))))
    assert abs(out[True][0] - out[False][0]) <= 2, out
    assert abs(out[True][1] - out[False][1]) / out[False][1] < 1e-10, out


RELIABLE_CASES = [((45, 45, 45), 'scalar', 1e-4, 'zero'), ((45, 45, 45), 'scalar', 1e-2, 'zero'),
                  ((45, 45, 45), 'scalar', 3e-5, 'random'), ((45, 27, 45), 'elasticity', 3e-5, 'zero'),
                  ((63, 81), 'scalar', 1e-4, 'zero'), ((63, 81), 'elasticity', 1e-3, 'random')]


@pytest.mark.parametrize('N,physics,delta,x0', RELIABLE_CASES)
def test_mixed_solution_residual_matches_host(N, physics, delta, x0):
    """The mixed solve returns a solution whose residual ||B - G A x||, recomputed from an fp64
    matvec, is the norm it reports (its last reliable update, formed in the fp64 operator's inverse
    row pass) and within the tolerance; the solution equals the fp64 solve's to the tolerance's
    accuracy. Regression: folding the fp32 corrections unprojected and projecting x once at the end
    left a solution 1e-5 away from the fp64 one with a residual 4e-4 while 8.7e-9 was reported
    (45^3, contrast 100), because the iteration had offset the fp32 drift outside the range of G
    with an error inside it. A nonzero x0 keeps its part outside the range of G in both solves."""
    print('case mixed-solution residual', N, physics, delta, x0, flush=True)
    rng = np.random.default_rng(41)
    d = len(N)
    plan = nat.Plan(N, np.ones(d), physics, 'primal')
    if physics == 'scalar':
        plan.set_coeff('iso', np.where(rng.random(N) < 0.3, 100., 1.))
    else:
        plan.set_coeff('lame', np.stack([1. + 5. * rng.random(N), 0.5 + 5. * rng.random(N)]))
    nc = plan.ncomp
    E = np.zeros(nc)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((nc,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((nc,) + N)
    plan.matvec(EN, B)
    start = np.zeros((nc,) + N) if x0 == 'zero' else 0.1 * rng.standard_normal((nc,) + N)
    x = nat.DeviceArray.from_host(start)
    info = plan.cg_solve(B, x, tol=1e-8, maxiter=3000, mode='mixed', delta=delta)
    assert info['status'] == nat.FH_CG_OK and 0 < info['kit'] < 3000
    x64 = nat.DeviceArray.from_host(start)
    i64 = plan.cg_solve(B, x64, tol=1e-8, maxiter=3000, mode='fp64')
    y = nat.DeviceArray((nc,) + N)
    plan.matvec(x, y)
    r = B.to_host() - y.to_host()
    host = float(np.sqrt(np.sum(r * r) / np.prod(N)))
    print('kit %d (fp64 %d) updates %d norm_res %.4e host %.4e' % (info['kit'], i64['kit'], plan.last_updates(),
                                                                   info['norm_res'], host), flush=True)
    assert host <= 1e-8 * (1 + 1e-6), host
    assert abs(host - info['norm_res']) <= 1e-6 * info['norm_res'], (host, info['norm_res'])
    xm, xd = x.to_host(), x64.to_host()
    assert np.abs(xm - xd).max() <= 1e-6 * np.abs(xd).max()
    # the restarts cost a few iterations over the fp64 solve
    assert info['kit'] <= i64['kit'] + 10, (info['kit'], i64['kit'])


@pytest.mark.parametrize('N,physics', [((27, 27, 27), 'scalar'), ((45, 45), 'scalar'), ((81, 81, 81), 'elasticity'),
                                       ((45, 45, 45), 'elasticity'), ((63, 81), 'elasticity'), ((5, 9, 15), 'scalar')])
def test_deferred_alpha_beta_match_kernels(N, physics):
    """Small grids: alpha and beta formed inside the passes (every forward-pass workgroup sums the
    inverse pass's partials in one fixed order, block 0 of the first strided pass finalises)
    against the fused iteration with its one-workgroup alpha/beta kernels: the sums differ in
    their order only, so the iteration counts agree within 1 and the energies to 1e-10; both
    converge to the tolerance, and a plan switches between the two without stale state."""
    print('case deferred alpha/beta', N, physics, flush=True)
    rng = np.random.default_rng(33)
    ph = rng.random(N) < 0.3
    plan = nat.Plan(N, np.ones(len(N)), physics, 'primal')
    nc = plan.ncomp
    if physics == 'scalar':
        a = np.where(ph, 100., 1.)
        plan.set_coeff('iso', a)
    else:
        lam, mu = np.where(ph, 5.0, 0.5), np.where(ph, 4.0, 0.4)
        plan.set_coeff('lame', np.stack([lam, mu]))
    E = np.zeros(nc)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((nc,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((nc,) + N)
    plan.matvec(EN, B)
    out = {}
    for defer in (True, False, True):
        plan.set_cg_variant(True, defer=defer)
        x = nat.DeviceArray.zeros((nc,) + N)
        info = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode='mixed', delta=1e-4)
        assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= 1e-8
        nat.add_constant(x, E)
        AH = plan.energy_matrix([x] * nc)[0, 0]
        if defer in out:
            assert out[defer] == (info['kit'], AH), 'deferred path not reproducible'
        out[defer] = (info['kit'], AH)
    assert abs(out[True][0] - out[False][0]) <= 1, out
    assert abs(out[True][1] - out[False][1]) < 1e-10 * abs(out[False][1]), out


@pytest.mark.parametrize('N', [(27, 27, 27), (45, 45, 45), (63, 81)])
def test_fused_mixed_iteration_lame(N):
    """Fused single-precision iteration with a Lame coefficient (staged row pass computes
    C:e = lambda tr(e) m + 2 mu e and <p, C:p>) against the operator-form iteration."""
    print('case fused mixed lame', N, flush=True)
    rng = np.random.default_rng(32)
    ph = rng.random(N) < 0.3
    lam = np.where(ph, 5.0, 0.5)
    mu = np.where(ph, 4.0, 0.4)
    plan = nat.Plan(N, np.ones(len(N)), 'elasticity', 'primal')
    nc = plan.ncomp
    plan.set_coeff('lame', np.stack([lam, mu]))
    E = np.zeros(nc)
    E[nc - 1] = 1.
    EN = nat.DeviceArray.zeros((nc,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((nc,) + N)
    plan.matvec(EN, B)
    m = np.zeros(nc)
    m[:3 if nc == 6 else 2] = 1.
    out = {}
    for fused in (True, False):
        plan.set_cg_variant(fused)
        x = nat.DeviceArray.zeros((nc,) + N)
        info = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode='mixed', delta=0.01)
        assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= 1e-8
        nat.add_constant(x, E)
        xh = x.to_host()
        tr = np.einsum('i,i...->...', m, xh)
        out[fused] = (info['kit'], float(np.mean(lam * tr * tr + 2 * mu * np.sum(xh * xh, axis=0))))
    assert abs(out[True][0] - out[False][0]) <= 2, out
    assert abs(out[True][1] - out[False][1]) / out[False][1] < 1e-10, out


@pytest.mark.parametrize('case', ['iso-243', 'iso-2d', 'lame-45'])
def test_graph_replayed_iterations_equal_direct_launches(case):
    """Mixed CG with batches of fused iterations replayed from a captured HIP graph against the
    same solve with every kernel launched directly: bitwise equal solutions and iteration
    counts, for solves before and after the direct one (replays with eager work between)."""
    print('case graph', case, flush=True)
    rng = np.random.default_rng(33)
    if case == 'iso-243':
        N, physics = (243, 243, 243), 'scalar'
    elif case == 'iso-2d':
        N, physics = (135, 81), 'scalar'
    else:
        N, physics = (45, 45, 45), 'elasticity'
    plan = nat.Plan(N, np.ones(len(N)), physics, 'primal')
    nc = plan.ncomp
    ph = rng.random(N) < 0.3
    if physics == 'scalar':
        plan.set_coeff('iso', np.where(ph, 100., 1.))
    else:
        plan.set_coeff('lame', np.stack([np.where(ph, 5.0, 0.5), np.where(ph, 4.0, 0.4)]))
    E = np.zeros(nc)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((nc,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((nc,) + N)
    plan.matvec(EN, B)
    res = []
    for graphs in (True, False, True):
        plan.use_graphs(graphs)
        x = nat.DeviceArray.zeros((nc,) + N)
        info = plan.cg_solve(B, x, tol=1e-8, maxiter=3000, mode='mixed', delta=1e-4)
        assert info['status'] == nat.FH_CG_OK
        nb = plan.last_graph_batches()
        assert (nb >= 2) if graphs else (nb == 0), (graphs, nb)
        res.append((info['kit'], x.to_host()))
        y = nat.DeviceArray((nc,) + N)
        plan.matvec(x, y)  # eager work between the solves
    for kit, xh in res[1:]:
        assert kit == res[0][0]
        assert np.array_equal(xh, res[0][1])


@pytest.mark.parametrize('N', [(75, 75, 75), (125, 225), (25, 125, 75)])
def test_fused_mixed_cg_on_added_lengths(N):
    """Fused fp32 iteration with fp64 reliable updates on the 25/75/125/225 schedules: the energy
    of the solution equals the fp64 solve's (1e-8), final true residual <= tol."""
    print('case fused mixed lengths', N, flush=True)
    d = len(N)
    rng = np.random.default_rng(73)
    a = np.where(rng.random(N) < 0.3, 30., 1.)
    plan = nat.Plan(N, np.ones(d), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    E = np.eye(d)[0]
    EN = nat.DeviceArray.zeros((d,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((d,) + N)
    plan.matvec(EN, B)
    ah = {}
    for mode in ('fp64', 'mixed'):
        x = nat.DeviceArray.zeros((d,) + N)
        info = plan.cg_solve(B, x, tol=1e-9, maxiter=3000, mode=mode, delta=1e-4)
        assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= 1e-9
        nat.add_constant(x, E)
        ah[mode] = plan.energy_matrix([x] * d)[0, 0]
    assert abs(ah['mixed'] - ah['fp64']) < 1e-8 * ah['fp64']
