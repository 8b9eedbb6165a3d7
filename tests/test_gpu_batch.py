"""Batched 2-D scalar homogenization (one persistent kernel over many samples of uneven
grids) against the oracle's per-sample GaNi solve (ffthompy/applications.py:24-90,
ffthompy/postprocess.py:53-70)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd.batch import homogenize_batch  # noqa: E402
from ffthompy_amd.microstructure import gaussian_random_field, uq_samples, BASE_SEED  # noqa: E402
from oracle.homogenize import Problem  # noqa: E402


def oracle_bounds(a, tol=1e-8, Y=None):
    """Upper and lower GaNi bounds of the oracle for origin-0 conductivity a (2-D)."""
    Y = np.ones(2) if Y is None else np.asarray(Y, dtype=np.float64)
    conf = {'physics': 'scalar',
            'material': {'fun': lambda c: np.einsum('ij,...->ij...', np.eye(2), np.fft.fftshift(a)), 'Y': Y},
            'solve': {'kind': 'GaNi', 'N': np.array(a.shape), 'primaldual': ['primal', 'dual']},
            'postprocess': [{'kind': 'GaNi'}],
            'solver': {'kind': 'CG', 'tol': tol, 'maxiter': 1e4}}
    pb = Problem(conf)
    pb.calculate()
    return pb.output['mat_primal']['AH_GaNi_primal'], pb.output['mat_dual']['AH_GaNi_dual']


def rel(A, B):
    return np.abs(A - B).max() / np.abs(B).max()


def grf_samples(shapes, corr=0.1, std=1.0, seed=BASE_SEED + 11):
    return [np.exp(gaussian_random_field(s, corr, std, seed + i)) for i, s in enumerate(shapes)]


SMALL = [(27, 27), (27, 45), (45, 27), (63, 81), (81, 63), (45, 135), (105, 27), (135, 63)]


@pytest.mark.parametrize('precision,tolerance', [('fp64', 1e-10), ('mixed', 1e-10)])
def test_batch_matches_oracle(precision, tolerance):
    print('case batch vs oracle', precision, SMALL, flush=True)
    fields = grf_samples(SMALL)
    out = homogenize_batch(fields, precision=precision, tol=1e-8)
    assert (out['status'] == 0).all()
    for s, a in enumerate(fields):
        up, lo = oracle_bounds(a)
        assert rel(out['upper'][s], up) < tolerance, (s, precision, rel(out['upper'][s], up))
        assert rel(out['lower'][s], lo) < tolerance, (s, precision, rel(out['lower'][s], lo))
        assert np.linalg.eigvalsh(out['upper'][s] - out['lower'][s]).min() >= -1e-12


def test_batch_nonunit_cell_and_repeat():
    """Cells of other sizes, and a second solve on the same batch (graph replay)."""
    print('case batch cells', flush=True)
    shapes = [(27, 45), (63, 63)]
    Y = np.array([[1.0, 2.0], [0.5, 1.5]])
    fields = grf_samples(shapes, seed=BASE_SEED + 3)
    b = nat.Batch(shapes, Y)
    b.set_coeff(fields)
    r1 = b.solve(tol=1e-8, mode='fp64')
    r2 = b.solve(tol=1e-8, mode='fp64')
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    for s, a in enumerate(fields):
        up, lo = oracle_bounds(a, Y=Y[s])
        assert rel(r1[0][s], up) < 1e-10
        assert rel(r1[1][s], lo) < 1e-10


def test_batch_matches_single_plan_solve():
    """The same sample through the per-problem HIP path (fh_cg_solve, fp64): equal
    iteration counts and bounds to rounding."""
    print('case batch vs plan', flush=True)
    a = grf_samples([(81, 105)], seed=BASE_SEED + 7)[0]
    out = homogenize_batch([a], precision='fp64', tol=1e-8)
    plan = nat.Plan((81, 105), (1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    fields, kits = [], []
    for iL in range(2):
        E = np.zeros(2)
        E[iL] = 1.
        EN = nat.DeviceArray.zeros((2, 81, 105))
        nat.add_constant(EN, -E)
        B = nat.DeviceArray((2, 81, 105))
        plan.matvec(EN, B)
        x = nat.DeviceArray.zeros((2, 81, 105))
        info = plan.cg_solve(B, x, tol=1e-8, maxiter=10000, mode='fp64')
        kits.append(info['kit'])
        nat.add_constant(x, E)
        fields.append(x)
    AH = plan.energy_matrix(fields)
    assert rel(out['upper'][0], AH) < 1e-12
    assert abs(int(out['kit'][0, 0, 0]) - kits[0]) <= 1 and abs(int(out['kit'][0, 0, 1]) - kits[1]) <= 1


def test_uq_workload_512():
    """512 samples, grids from {63,81,105,135}^2, exp of a GRF (correlation 0.1, log-std 1):
    mixed precision to tol 1e-8; every sample converged, lower <= upper, and eight samples
    (the largest grids included) against the oracle within 1e-10 (each reliable update projects the
    fp32 correction, so the mixed solutions stay in the range of G and their bounds agree with the
    fp64 ones to rounding; with unprojected corrections they differed by 6.6e-11)."""
    print('case batch uq 512', flush=True)
    dims, fields = uq_samples()
    out = homogenize_batch(fields, precision='mixed', tol=1e-8)
    print('uq 512: %.1f ms, iterations %d..%d, fp64 updates %d..%d' % (
        out['time'] * 1e3, out['kit'].min(), out['kit'].max(), out['updates'].min(), out['updates'].max()),
        flush=True)
    assert (out['status'] == 0).all()
    assert out['kit'].max() < 10000
    gap = np.array([np.linalg.eigvalsh(u - l).min() for u, l in zip(out['upper'], out['lower'])])
    assert (gap >= -1e-12).all()
    big = int(np.argmax([d[0] * d[1] for d in dims]))
    pick = sorted(set([big] + list(np.random.default_rng(5).choice(len(dims), 7, replace=False))))
    for s in pick:
        up, lo = oracle_bounds(fields[s])
        assert rel(out['upper'][s], up) < 1e-10, (s, dims[s], rel(out['upper'][s], up))
        assert rel(out['lower'][s], lo) < 1e-10, (s, dims[s], rel(out['lower'][s], lo))


def test_schedules_agree_and_autotune_picks_fastest():
    """Both radix-schedule variants and both task orders give the same bounds; the
    autotuner's choice is the fastest of its own timings."""
    print('case batch schedules', flush=True)
    shapes = [(63, 81), (135, 45), (105, 105), (27, 63), (81, 81), (45, 135)]
    fields = grf_samples(shapes, seed=BASE_SEED + 21)
    res = {}
    for var in (0, 1):
        for order in (0, 1):
            res[(var, order)] = homogenize_batch(fields, precision='fp64', tol=1e-8, schedule=(var, order))
    ref = res[(0, 0)]
    for key, r in res.items():
        assert np.abs(r['upper'] - ref['upper']).max() < 1e-12 * np.abs(ref['upper']).max(), key
        assert np.abs(r['lower'] - ref['lower']).max() < 1e-12 * np.abs(ref['lower']).max(), key
    out = homogenize_batch(fields, precision='mixed', tol=1e-8)
    t = out['tune_ms']
    var, order = out['schedule']
    assert t.shape == (2, 2) and (t > 0).all()
    assert t[order, var] == t.min()
    assert rel(out['upper'], ref['upper']) < 1e-8
