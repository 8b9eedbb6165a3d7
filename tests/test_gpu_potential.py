"""Potential formulation on the GPU (fh_potential_solve through ffthompy_amd.potential) against
the oracle (oracle/potential.py), fp64 both:

* GaNi: CG iteration counts equal, AH within 1e-12 (relative), Fu within 1e-10 of its largest
  entry; Aga = None and Ga post-processing (exact and order-0 coefficients); 2-D and 3-D; an
  anisotropic (full) coefficient; a rough random 27^3 coefficient (residual histories, count at a
  tolerance with margin, AH).
* Ga: on quadrature grids 2N-1 = 9 (as the reference) and 29 (the device solves on 30): counts
  equal, AH within 1e-10, Fu on the grid N.
* At 45^3 (random spheres, contrast 10) the potential value equals the gradient formulation's
  AH_11 from the device CG (both to tol 1e-10): size-independent check of the same Galerkin
  problem.
* Even grids (GaNi, with and without Ga post-processing), fp64 and mixed, against the oracle.
"""
import numpy as np
import pytest

from oracle import examples
from oracle import potential as opot
from oracle.fields import Field
from oracle.materials import Material as OMaterial

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd import potential as gpot  # noqa: E402
from ffthompy_amd.materials import Material  # noqa: E402
from ffthompy_amd.microstructure import spheres_conductivity  # noqa: E402

ANISO = {2: np.array([[5.5, -4.5], [-4.5, 5.5]]),
         3: np.array([[4.25, -3.25, -1.76776695], [-3.25, 4.25, 1.76776695], [-1.76776695, 1.76776695, 7.5]])}
PARS = {'solver': {'tol': 1e-10, 'maxiter': 300}}


def _check_fu(got, ref, tol):
    g = got.val.to_host() if hasattr(got.val, 'to_host') else np.asarray(got.val)
    assert g.shape == ref.val.shape
    assert np.abs(g - ref.val).max() <= tol * np.abs(ref.val).max()


def _materials(maker, name):
    conf = maker()[name]
    return Material(dict(conf)), OMaterial(dict(conf))


@pytest.mark.parametrize('maker,name,N,ga', [
    (examples.scalar_2d_materials, 'square', (5, 5), None),
    (examples.scalar_2d_materials, 'square', (45, 45), None),
    (examples.scalar_2d_materials, 'pyramid', (15, 15), 'exact'),
    (examples.scalar_2d_materials, 'square_Ga', (5, 5), 0),
    (examples.scalar_3d_materials, 'cube', (5, 5, 5), 'exact'),
    (examples.scalar_3d_materials, 'cube', (15, 15, 15), None),
])
def test_gani_potential_matches_oracle(maker, name, N, ga):
    print('case gani potential', name, N, ga, flush=True)
    mat, omat = _materials(maker, name)
    Agani, oAgani = mat.get_A_GaNi(N, 'primal'), omat.get_A_GaNi(N, 'primal')
    Nbar = tuple(2 * n - 1 for n in N)
    order = None if ga == 'exact' else ga
    Aga = None if ga is None else mat.get_A_Ga(Nbar, 'primal', order=order, P=N if order == 0 else None)
    oAga = None if ga is None else omat.get_A_Ga(Nbar, 'primal', order=order, P=N if order == 0 else None)
    res = gpot.homog_GaNi_full_potential(Agani, Aga, PARS)
    ref = opot.homog_GaNi_full_potential(oAgani, oAga, tol=1e-10, maxiter=300)
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-12 * abs(ref['AH'])
    _check_fu(res.Fu, ref['Fu'], 1e-10)
    assert abs(res.Fu.mean()) < 1e-14


@pytest.mark.parametrize('maker,name,N', [
    (examples.scalar_2d_materials, 'square', (5, 5)),
    (examples.scalar_2d_materials, 'ball', (5, 5)),
    (examples.scalar_2d_materials, 'pyramid', (15, 15)),
    (examples.scalar_3d_materials, 'cube_Ga', (5, 5, 5)),
    (examples.from_file_materials, 'file', (5, 5)),
])
def test_ga_potential_matches_oracle(maker, name, N):
    print('case ga potential', name, N, flush=True)
    mat, omat = _materials(maker, name)
    Nbar = tuple(2 * n - 1 for n in N)
    res = gpot.homog_Ga_full_potential(mat.get_A_Ga(Nbar, 'primal'), PARS)
    ref = opot.homog_Ga_full_potential(omat.get_A_Ga(Nbar, 'primal'), tol=1e-10, maxiter=300)
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-10 * abs(ref['AH'])
    _check_fu(res.Fu, ref['Fu'], 1e-9)


@pytest.mark.parametrize('d', [2, 3])
def test_anisotropic_coefficient_matches_oracle(d):
    """The reference's materials with an anisotropic constant added to both coefficients
    (square inclusion, order-0 Ga coefficient on P = N)."""
    print('case aniso potential', d, flush=True)
    N = (5,) * d
    conf = {'inclusions': ['square', 'otherwise'], 'positions': [np.zeros(d), ''], 'params': [0.6 * np.ones(d), ''],
            'vals': [10 * np.eye(d), 1. * np.eye(d)], 'Y': np.ones(d), 'P': N, 'order': 0}
    mat, omat = Material(dict(conf)), OMaterial(dict(conf))
    Nbar = tuple(2 * n - 1 for n in N)
    Agani = mat.get_A_GaNi(N, 'primal').add_mean(ANISO[d])
    oAgani = omat.get_A_GaNi(N, 'primal').add_mean(ANISO[d])
    Aga = mat.get_A_Ga(Nbar, 'primal').add_mean(ANISO[d])
    oAga = omat.get_A_Ga(Nbar, 'primal').add_mean(ANISO[d])
    res = gpot.homog_GaNi_full_potential(Agani, Aga, PARS)
    ref = opot.homog_GaNi_full_potential(oAgani, oAga, tol=1e-10, maxiter=300)
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-12 * abs(ref['AH'])
    res = gpot.homog_Ga_full_potential(Aga, PARS)
    ref = opot.homog_Ga_full_potential(oAga, tol=1e-10, maxiter=300)
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-10 * abs(ref['AH'])


def test_random_coefficient_27_matches_oracle():
    """A rough coefficient (exp of white noise): the residual histories agree while the residual
    is above 1e-7, the iteration count at a tolerance placed midway (geometrically) between the
    oracle's residuals of iterations 40 and 41 is equal, and AH at tol 1e-10 agrees to 1e-12 (the
    late iterations of the two solves round differently, so their counts at 1e-10 may differ)."""
    print('case potential random 27^3', flush=True)
    N = (27, 27, 27)
    rng = np.random.default_rng(11)
    a = np.exp(rng.standard_normal(N))
    A = np.einsum('ij,...->ij...', np.eye(3), a)
    ref = opot.homog_GaNi_full_potential(Field(val=A, order=2, multype=21), None, tol=1e-10, maxiter=300)
    log = np.asarray(ref['info']['norm_res_log'])
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    _, info = plan.potential_solve(np.eye(3)[0], 1e-10, 300, hist_len=400)
    got = info['res_log']
    m = int(np.sum(log > 1e-7))
    assert np.abs(got[:m] - log[:m]).max() <= 1e-4 * log[:m].min() + 1e-9 * log[0]
    tol = float(np.sqrt(log[40] * log[41]))
    _, info = plan.potential_solve(np.eye(3)[0], tol, 300)
    assert info['kit'] == 41
    Agani = gpot.Tensor(name='A', val=A, order=2, multype=21)
    res = gpot.homog_GaNi_full_potential(Agani, None, PARS)
    assert abs(res.AH - ref['AH']) < 1e-12 * abs(ref['AH'])
    _check_fu(res.Fu, ref['Fu'], 1e-8)


def test_potential_equals_gradient_formulation_45():
    print('case potential vs gradient 45^3', flush=True)
    N = (45, 45, 45)
    a = spheres_conductivity(N, count=30, contrast=10.)[0]
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    E = np.eye(3)[0]
    EN = nat.DeviceArray.zeros((3,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((3,) + N)
    plan.matvec(EN, B)
    x = nat.DeviceArray.zeros((3,) + N)
    info = plan.cg_solve(B, x, tol=1e-10, maxiter=2000, mode='fp64')
    assert info['status'] == nat.FH_CG_OK
    nat.add_constant(x, E)
    ah_grad = plan.energy_matrix([x, x, x])[0, 0]
    Agani = gpot.Tensor(name='A', val=np.einsum('ij,...->ij...', np.eye(3), a), order=2, multype=21)
    res = gpot.homog_GaNi_full_potential(Agani, None, {'solver': {'tol': 1e-10, 'maxiter': 2000}})
    assert res.info['status'] == 0
    assert abs(res.AH - ah_grad) < 1e-9 * ah_grad, (res.AH, ah_grad)


@pytest.mark.parametrize('N,ga', [((6, 8), None), ((8, 8), 'exact'), ((44, 46), None), ((6, 5, 4), None),
                                  ((16, 16, 16), None), ((6, 6, 6), 'exact')])
def test_even_grid_gani_potential_matches_oracle(N, ga):
    """Even grids as the reference runs them (homogenisation.py:86-128): the Nyquist index of an
    even axis carries the frequency -N/2 in grad, div and P, the inverse real transform drops the
    imaginary parts of the last axis' self-conjugate planes, and the inner product counts the
    Nyquist plane of an even last axis once. The result differs from the gradient formulation's
    (whose projection nulls the Nyquist modes); the oracle restates the reference's steps."""
    print('case even-grid potential', N, ga, flush=True)
    d = len(N)
    rng = np.random.default_rng(23)
    a = 1. + 9. * rng.random(N)
    A = np.einsum('ij,...->ij...', np.eye(d), a)
    Agani = gpot.Tensor(name='A', val=A, order=2, multype=21)
    oAgani = Field(val=A, order=2, N=N, Y=np.ones(d), multype=21)
    Aga = oAga = None
    if ga:
        mat, omat = _materials(examples.scalar_2d_materials if d == 2 else examples.scalar_3d_materials,
                               'square' if d == 2 else 'cube')
        Nbar = tuple(2 * n - 1 for n in N)
        Aga, oAga = mat.get_A_Ga(Nbar, 'primal'), omat.get_A_Ga(Nbar, 'primal')
        Agani, oAgani = mat.get_A_GaNi(N, 'primal'), omat.get_A_GaNi(N, 'primal')
    ref = opot.homog_GaNi_full_potential(oAgani, oAga, tol=1e-10, maxiter=500)
    for prec in ('fp64', 'mixed'):
        res = gpot.homog_GaNi_full_potential(Agani, Aga, {'solver': {'tol': 1e-10, 'maxiter': 500,
                                                                      'precision': prec}})
        assert res.info['status'] == 0
        if prec == 'fp64':
            assert res.info['kit'] == ref['info']['kit']
            assert abs(res.AH - ref['AH']) < 1e-12 * abs(ref['AH'])
            _check_fu(res.Fu, ref['Fu'], 1e-10)
        else:
            assert abs(res.AH - ref['AH']) < 1e-8 * abs(ref['AH'])


@pytest.mark.parametrize('N', [(27, 27, 27), (45, 45), (15, 27, 9)])
def test_fused_passes_equal_generic_kernels(N):
    """The fixed-length kernels (gradient expansion and divergence inside the outermost strided
    passes, C2R-a-R2C in one row kernel) against the runtime-radix one-way passes with separate
    elementwise kernels: same iteration count, Fu and X within 1e-12 of their largest entries."""
    print('case potential fused vs generic', N, flush=True)
    d = len(N)
    rng = np.random.default_rng(17)
    a = np.exp(0.5 * rng.standard_normal(N))
    out = []
    for generic in (False, True):
        plan = nat.Plan(N, (1.,) * d, 'scalar', 'primal')
        plan.set_coeff('iso', a)
        plan.use_generic(generic)
        X = nat.DeviceArray((d,) + N)
        Fu, info = plan.potential_solve(np.eye(d)[0], 1e-9, 500, X=X)
        out.append((info['kit'], Fu.to_host(), X.to_host()))
    assert out[0][0] == out[1][0]
    for i in (1, 2):
        assert np.abs(out[0][i] - out[1][i]).max() <= 1e-12 * np.abs(out[1][i]).max()


@pytest.mark.parametrize('N,Y', [((15, 9, 21), (1., 2., 0.5)), ((27, 15), (0.5, 3.))])
def test_rectangular_cell_matches_oracle(N, Y):
    """Unequal axes and cell sizes (the frequencies xi = k / Y per axis of the view); the tolerance
    sits midway (geometrically) between two of the oracle's residuals, five iterations before its
    1e-10 stop, so the count is compared with a margin."""
    print('case potential cell', N, Y, flush=True)
    d = len(N)
    rng = np.random.default_rng(23)
    a = np.exp(0.7 * rng.standard_normal(N))
    A = np.einsum('ij,...->ij...', np.eye(d), a)
    oA = Field(val=A, order=2, Y=np.asarray(Y), multype=21)
    log = opot.homog_GaNi_full_potential(oA, None, tol=1e-10, maxiter=300)['info']['norm_res_log']
    k = len(log) - 6
    tol = float(np.sqrt(log[k] * log[k + 1]))
    res = gpot.homog_GaNi_full_potential(gpot.Tensor(name='A', val=A, order=2, Y=np.asarray(Y), multype=21), None,
                                         {'solver': {'tol': tol, 'maxiter': 300}})
    ref = opot.homog_GaNi_full_potential(oA, None, tol=tol, maxiter=300)
    assert ref['info']['kit'] == k + 1
    assert res.info['kit'] == k + 1
    assert abs(res.AH - ref['AH']) < 1e-12 * abs(ref['AH'])
    _check_fu(res.Fu, ref['Fu'], 1e-9)


@pytest.mark.parametrize('case', ['gani45x2', 'gani27_rough', 'gani45_spheres', 'aniso2', 'ga15x2', 'ga5x3'])
def test_mixed_precision_matches_fp64(case):
    """fp32 iterations with fp64 reliable updates (delta 1e-4) against the fp64 solve, both to
    tol 1e-10: AH within 1e-8, Fu within 2e-6 of its largest entry, final true residual <= tol.
    Covers the fixed-length kernels (45, 27, 9), the runtime-radix passes (30) and the full
    coefficient (one-way row passes)."""
    print('case potential mixed', case, flush=True)
    rng = np.random.default_rng(29)
    mixed = {'solver': {'tol': 1e-10, 'maxiter': 2000, 'precision': 'mixed', 'delta': 1e-4}}
    fp64 = {'solver': {'tol': 1e-10, 'maxiter': 2000}}
    if case.startswith('ga'):
        N = (15, 15) if case == 'ga15x2' else (5, 5, 5)
        maker = examples.scalar_2d_materials if len(N) == 2 else examples.scalar_3d_materials
        mat = Material(dict(maker()['pyramid' if len(N) == 2 else 'cube_Ga']))
        Aga = mat.get_A_Ga(tuple(2 * n - 1 for n in N), 'primal')
        r64 = gpot.homog_Ga_full_potential(Aga, fp64)
        r32 = gpot.homog_Ga_full_potential(Aga, mixed)
    else:
        if case == 'aniso2':
            N = (45, 45)
            conf = examples.scalar_2d_materials()['square']
            A = Material(dict(conf)).get_A_GaNi(N, 'primal').add_mean(ANISO[2]).val
        else:
            N = {'gani45x2': (45, 45), 'gani27_rough': (27, 27, 27), 'gani45_spheres': (45, 45, 45)}[case]
            if case == 'gani45_spheres':
                a = spheres_conductivity(N, count=30, contrast=100.)[0]
            elif case == 'gani27_rough':
                a = np.exp(rng.standard_normal(N))
            else:
                a = np.exp(0.5 * rng.standard_normal(N))
            A = np.einsum('ij,...->ij...', np.eye(len(N)), a)
        Agani = gpot.Tensor(name='A', val=A, order=2, multype=21)
        r64 = gpot.homog_GaNi_full_potential(Agani, None, fp64)
        r32 = gpot.homog_GaNi_full_potential(Agani, None, mixed)
    assert r32.info['status'] == 0 and r32.info['norm_res'] <= 1e-10
    assert abs(r32.AH - r64.AH) < 1e-8 * abs(r64.AH), (r32.AH, r64.AH)
    f64 = r64.Fu.val.to_host() if hasattr(r64.Fu.val, 'to_host') else r64.Fu.val
    f32 = r32.Fu.val.to_host() if hasattr(r32.Fu.val, 'to_host') else r32.Fu.val
    # both residuals are <= 1e-10 in the grid-mean norm, i.e. 1e-10 * prod(N) in the unnormalised
    # spectrum; times the operator's condition number this allows ~1e-6 of the largest entry
    assert np.abs(f32 - f64).max() <= 2e-6 * np.abs(f64).max()


def _lowrank_example_material(material, dim, N):
    """The material dictionaries of the reference's low-rank examples (square with order 0 on P = 5,
    pyramid with order 1 on P = N; material 3 adds an anisotropic constant)."""
    if material in (0, 3):
        return {'inclusions': ['square', 'otherwise'], 'positions': [np.zeros(dim), ''],
                'params': [0.6 * np.ones(dim), ''], 'vals': [10 * np.eye(dim), 1. * np.eye(dim)],
                'Y': np.ones(dim), 'P': (5,) * dim, 'order': 0}
    return {'inclusions': ['pyramid', 'all'], 'positions': [np.zeros(dim), ''], 'params': [0.8 * np.ones(dim), ''],
            'vals': [10 * np.eye(dim), 1. * np.eye(dim)], 'Y': np.ones(dim), 'P': N, 'order': 1}


@pytest.mark.parametrize('dim,material', [(2, 0), (2, 3), (3, 1), (3, 0)])
def test_reference_unittest_cases_match_oracle(dim, material):
    """The cases of the reference's potential test (N = 5, tol 1e-8, maxiter 50): materials 0 and 3
    through homog_Ga_full_potential, material 1 through homog_GaNi_full_potential with A_Ga;
    mean of Fu zero, AH and iteration counts as the oracle's."""
    print('case potential unittest', dim, material, flush=True)
    N = (5,) * dim
    conf = _lowrank_example_material(material, dim, N)
    mat, omat = Material(dict(conf)), OMaterial(dict(conf))
    Nbar = tuple(2 * n - 1 for n in N)
    pars = {'solver': {'tol': 1e-8, 'maxiter': 50}}
    Aga, oAga = mat.get_A_Ga(Nbar, 'primal'), omat.get_A_Ga(Nbar, 'primal')
    if material == 3:
        Aga.add_mean(ANISO[dim])
        oAga.add_mean(ANISO[dim])
    if material in (0, 3):
        res = gpot.homog_Ga_full_potential(Aga, pars)
        ref = opot.homog_Ga_full_potential(oAga, tol=1e-8, maxiter=50)
    else:
        Agani, oAgani = mat.get_A_GaNi(N, 'primal'), omat.get_A_GaNi(N, 'primal')
        res = gpot.homog_GaNi_full_potential(Agani, Aga, pars)
        ref = opot.homog_GaNi_full_potential(oAgani, oAga, tol=1e-8, maxiter=50)
    assert abs(res.Fu.mean()) < 1e-12
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-10 * abs(ref['AH'])


def test_ga_potential_rectangular_cell_matches_oracle():
    """Ga variant on a cell of sizes (1, 2) and grids 7 x 5 (quadrature grids 13 x 9), centred inclusion.
    (An off-centre inclusion gets the reference's phase convention in its Ga coefficient, which can make
    the coefficient negative somewhere; CG then meets <p, A p> <= 0, which the device reports as a
    breakdown where the reference carries on.)"""
    print('case ga potential cell', flush=True)
    d = 2
    conf = {'inclusions': ['square', 'otherwise'], 'positions': [np.zeros(d), ''],
            'params': [np.array([0.5, 1.1]), ''], 'vals': [7 * np.eye(d), 1. * np.eye(d)],
            'Y': np.array([1., 2.]), 'order': None}
    mat, omat = Material(dict(conf)), OMaterial(dict(conf))
    Nbar = (13, 9)
    res = gpot.homog_Ga_full_potential(mat.get_A_Ga(Nbar, 'primal'), PARS)
    ref = opot.homog_Ga_full_potential(omat.get_A_Ga(Nbar, 'primal'), tol=1e-10, maxiter=300)
    assert res.info['kit'] == ref['info']['kit']
    assert abs(res.AH - ref['AH']) < 1e-10 * abs(ref['AH'])
    _check_fu(res.Fu, ref['Fu'], 1e-9)


def test_potential_on_new_fixed_lengths_equals_generic():
    """The potential kernels on the 25/75/125 schedules against the runtime-radix passes."""
    print('case potential new lengths', flush=True)
    N = (25, 75, 125)
    a = np.exp(0.5 * np.random.default_rng(71).standard_normal(N))
    out = []
    for generic in (False, True):
        plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
        plan.set_coeff('iso', a)
        plan.use_generic(generic)
        Fu, info = plan.potential_solve(np.eye(3)[0], 1e-9, 500)
        out.append((info['kit'], Fu.to_host()))
    assert out[0][0] == out[1][0]
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-12 * np.abs(out[1][1]).max()


def test_device_crop_of_spectra_matches_oracle():
    """The crop of a Ga solve's potential spectrum from the quadrature grid M back to N (the
    reference's Tensor.decrease of an 'r' spectrum) against the oracle's project."""
    from ffthompy_amd.potential import _crop_spectrum
    rng = np.random.default_rng(9)
    for N, M in (((5, 7), (10, 15)), ((5, 5, 3), (9, 12, 5))):
        x = rng.standard_normal(N)
        X = opot.DFT(N)(Field(val=x, order=0, N=N, Y=np.ones(len(N))))
        big = X.project(M)
        got = _crop_spectrum(nat.DeviceArray.from_host(big.val), M, N).to_host()
        assert np.abs(got - X.val).max() < 1e-12 * np.abs(X.val).max()
