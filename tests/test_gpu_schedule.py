"""Schedule variants of the fixed-length passes and the plan autotuner (fh_plan_set_schedule,
fh_plan_autotune): every variant computes the same transforms per pencil, so the matvec must be
equal to the last bit across variants; the autotuner must pick one of them, and a mixed CG after it
converges alike (the row variant only re-blocks the CG's partial sums). The matvec itself is
pinned to the oracle at variant 0 by test_gpu_kernels.py."""
import numpy as np
import pytest

from ffthompy_amd import _native as nat

pytestmark = pytest.mark.gpu

CASES = [((45, 45, 45), np.float32, 'scalar'), ((45, 45, 45), np.float64, 'scalar'),
         ((81, 81, 81), np.float32, 'scalar'), ((243, 27, 243), np.float32, 'scalar'),
         ((27, 45, 243), np.float64, 'scalar'),
         ((63, 135), np.float32, 'scalar'), ((99, 99), np.float64, 'scalar'),
         ((81, 81, 81), np.float32, 'elasticity'), ((45, 27, 45), np.float64, 'elasticity'),
         ((81, 45), np.float32, 'elasticity')]


@pytest.mark.parametrize('N,dt,physics', CASES, ids=['%s-%s-%s' % (c[0], np.dtype(c[1]).name, c[2]) for c in CASES])
def test_schedule_variants_are_bitwise_equal(N, dt, physics):
    rng = np.random.default_rng(11)
    pl = nat.Plan(N, np.ones(len(N)), physics, 'primal')
    if physics == 'scalar':
        pl.set_coeff('iso', 1. + 9. * rng.random(N))
    else:
        pl.set_coeff('lame', np.stack([1. + rng.random(N), 0.5 + rng.random(N)]))
    nc = pl.ncomp
    x = nat.DeviceArray.from_host(rng.standard_normal((nc,) + N).astype(dt))
    y = nat.DeviceArray((nc,) + N, dt)
    ref = None
    for pv in range(4):
        for cv in range(2):
            for rv in range(3):
                pl.set_schedule(dt, pv, cv, rv)
                assert pl.schedule(dt) == (pv, cv, rv)
                pl.matvec(x, y)
                got = y.to_host()
                if ref is None:
                    ref = got
                else:
                    assert np.array_equal(got, ref), (pv, cv, rv)
    print('%s %s %s: 24 schedules equal' % (N, np.dtype(dt).name, physics), flush=True)


def test_schedule_rejects_bad_variants():
    pl = nat.Plan((15, 15, 15), np.ones(3), 'scalar', 'primal')
    with pytest.raises(nat.FhError):
        pl.set_schedule(np.float32, 4, 0, 0)
    with pytest.raises(nat.FhError):
        pl.set_schedule(np.float64, 0, 2, 0)
    with pytest.raises(nat.FhError):
        pl.set_schedule(np.float64, 0, 0, 3)


@pytest.mark.parametrize('N', [(81, 81, 81), (105, 135)])
def test_autotune_picks_a_variant_and_keeps_results(N):
    rng = np.random.default_rng(12)
    a = 1. + 99. * (rng.random(N) > 0.7)
    pl = nat.Plan(N, np.ones(len(N)), 'scalar', 'primal')
    pl.set_coeff('iso', a)
    d = len(N)
    E = np.zeros(d)
    E[0] = 1.
    EN = nat.DeviceArray.zeros((d,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((d,) + N)
    pl.matvec(EN, B)
    x0 = nat.DeviceArray.zeros((d,) + N)
    info0 = pl.cg_solve(B, x0, tol=1e-8, maxiter=500, mode='mixed', delta=1e-4)
    for dt in (np.float32, np.float64):
        res = pl.autotune(dt, nrep=3)
        print(N, np.dtype(dt).name, res, flush=True)
        assert 0 <= res['proj_variant'] < 4 and 0 <= res['cols_variant'] < 2 and 0 <= res['rows_variant'] < 3
        assert all(t > 0 for t in res['proj_ms']) and all(t > 0 for t in res['rows_ms'])
        assert pl.schedule(dt) == (res['proj_variant'], res['cols_variant'], res['rows_variant'])
        assert res['proj_ms'][res['proj_variant']] == min(res['proj_ms'])
    x1 = nat.DeviceArray.zeros((d,) + N)
    info1 = pl.cg_solve(B, x1, tol=1e-8, maxiter=500, mode='mixed', delta=1e-4)
    # the row variant changes how the CG's partial sums are blocked (rounding only): the solve
    # converges alike, within the solver tolerance
    assert abs(info1['kit'] - info0['kit']) <= 2 and info1['norm_res'] <= 1e-8
    a0 = x0.to_host()
    assert np.abs(x1.to_host() - a0).max() < 1e-6 * np.abs(a0).max()
