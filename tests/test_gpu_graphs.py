"""Captured iteration graphs and plan changes: a plan reused after set_primaldual / set_coeff /
use_generic captures its mixed-CG graph anew, so a mixed gradient solve (and a potential solve) after
each change equals the same solve on a fresh plan with those settings (the captured kernels and their
arguments belong to the settings at capture time).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')


def _coef(N, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random(N) < 0.3, 50., 1.)


def _gradient_mixed(plan, N, E):
    EN = nat.DeviceArray.zeros((3,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((3,) + N)
    plan.matvec(EN, B)
    x = nat.DeviceArray.zeros((3,) + N)
    info = plan.cg_solve(B, x, tol=1e-8, maxiter=2000, mode='mixed', delta=1e-4)
    return info['kit'], x.to_host()


def test_reused_plan_captures_anew_after_changes():
    print('case graphs after plan changes', flush=True)
    N = (45, 45, 45)
    a = _coef(N, 67)
    E = np.eye(3)[0]
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    _gradient_mixed(plan, N, E)
    assert plan.last_graph_batches() >= 1
    # dual projection and inverse coefficient on the same plan (same work buffers)
    plan.set_primaldual('dual')
    plan.set_coeff('iso', 1. / a)
    got = _gradient_mixed(plan, N, E)
    fresh = nat.Plan(N, (1., 1., 1.), 'scalar', 'dual')
    fresh.set_coeff('iso', 1. / a)
    want = _gradient_mixed(fresh, N, E)
    assert got[0] == want[0] and np.array_equal(got[1], want[1])
    # another coefficient and the runtime-radix kernels for the potential solve
    plan.set_primaldual('primal')
    plan.set_coeff('iso', a + 1.)
    plan.use_generic(True)
    Fu, info = plan.potential_solve(E, 1e-9, 1000, mode='mixed')
    fresh = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    fresh.set_coeff('iso', a + 1.)
    fresh.use_generic(True)
    Fw, winfo = fresh.potential_solve(E, 1e-9, 1000, mode='mixed')
    assert info['kit'] == winfo['kit'] and np.array_equal(Fu.to_host(), Fw.to_host())
