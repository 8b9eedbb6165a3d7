"""The reference's CallBack on the device: ||B - A x|| evaluated in fp64 at x0 and after the
iterations of the CG (fh_plan_set_residual_log), against the oracle's CallBack, which evaluates
B - A(x) on the host after every iteration (ffthompy/general/solver_pp.py:13-24).

* every = 1: one entry per iteration (and x0), equal to the oracle's within 1e-12 of the initial
  residual plus 1e-6 relative (both are fp64 evaluations of nearly the same iterates).
* every = 3: the entries at iterations 0, 3, 6, ... and the last, bitwise those of every = 1 (the
  log does not touch the CG's arithmetic); every = 100 > the iteration count: x0 and the last.
* The mixed-precision solve rejects a plan with the log on.
"""
import numpy as np
import pytest

from oracle import projections
from oracle.fields import DFT, Field, Operator
from oracle.solver import CallBack as OCallBack
from oracle.solver import linear_solver as o_linear_solver

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd import GpuOperator, Tensor, linear_solver  # noqa: E402
from ffthompy_amd.solver import CallBack  # noqa: E402


def _case(N, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    d = len(N)
    a = np.exp(spread * rng.standard_normal(N))
    return a, np.einsum('ij,...->ij...', np.eye(d), a)


def _oracle_log(N, A, tol):
    """The oracle's CallBack log at a tolerance placed midway (geometrically) between the two
    recursive residuals around `tol`, so that both solves stop at the same iteration with margin."""
    d = len(N)
    _, hG1, _ = projections.scalar(N, np.ones(d), NyqNul=True)
    Afun = Operator([[Operator([[DFT(N, inverse=True), hG1, DFT(N)]]), Field(val=A, order=2, multype=21)]])
    EN = Field(shape=(d,), N=N)
    EN.set_mean(np.eye(d)[0])
    B = Afun(-EN)
    _, info = o_linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': tol, 'maxiter': 500}, solver='CG')
    log = info['norm_res_log']
    k = info['kit']
    tol_mid = float(np.sqrt(log[k - 1] * log[k]))
    cb = OCallBack(Afun, B)
    _, info = o_linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': tol_mid, 'maxiter': 500}, solver='CG',
                              callback=cb)
    assert info['kit'] == k
    return np.asarray(cb.res_norm), k, tol_mid


@pytest.mark.parametrize('N', [(27, 27, 27), (45, 45)])
def test_true_residual_log_matches_oracle_callback(N):
    print('case residual log', N, flush=True)
    d = len(N)
    # a moderate contrast keeps the two fp64 CGs' iterates within rounding of each other to the
    # end (with a rough coefficient they drift apart by more than the late true residuals)
    a, A = _case(N, 31, spread=0.3)
    ref, kit_ref, tol = _oracle_log(N, A, 1e-8)
    Afun = GpuOperator(N, np.ones(d), 'scalar', 'primal', A=Tensor(name='A', val=A, order=2, multype=21))
    EN = Tensor(name='EN', shape=(d,), N=N)
    EN.set_mean(np.eye(d)[0])
    B = Afun(-EN)
    logs = {}
    for every in (1, 3, 100):
        cb = CallBack(A=Afun, B=B)
        x, info = linear_solver(Afun, B, x0=B.zeros_like(), par={'tol': tol, 'maxiter': 500,
                                                                'true_residual_every': every},
                                solver='CG', callback=cb)
        assert info['kit'] == kit_ref
        logs[every] = (np.asarray(cb.res_kit), np.asarray(cb.res_norm))
        assert len(cb.res_norm_recursive) == kit_ref + 1
    k1, v1 = logs[1]
    assert list(k1) == list(range(kit_ref + 1))
    assert np.all(np.abs(v1 - ref) <= 1e-12 * ref[0] + 1e-6 * ref)
    k3, v3 = logs[3]
    want = sorted(set(list(range(0, kit_ref + 1, 3)) + [kit_ref]))
    assert list(k3) == want
    assert np.array_equal(v3, v1[k3])
    k100, v100 = logs[100]
    assert list(k100) == [0, kit_ref] and np.array_equal(v100, v1[k100])


def test_mixed_solve_rejects_residual_log():
    N = (15, 15, 15)
    a, A = _case(N, 33)
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    b = nat.DeviceArray.zeros((3,) + N)
    x = nat.DeviceArray.zeros((3,) + N)
    plan.set_residual_log(1)
    with pytest.raises(nat.FhError):
        plan.cg_solve(b, x, tol=1e-8, maxiter=10, mode='mixed', delta=1e-4)
    plan.set_residual_log(0)


def _oracle_iterates(N, A, tol, method='CG', par=None):
    """The oracle's iterates x_k handed to a plain callback (x0 first, as linear_solver does)."""
    d = len(N)
    _, hG1, _ = projections.scalar(N, np.ones(d), NyqNul=True)
    Afun = Operator([[Operator([[DFT(N, inverse=True), hG1, DFT(N)]]), Field(val=A, order=2, multype=21)]])
    EN = Field(shape=(d,), N=N)
    EN.set_mean(np.eye(d)[0])
    B = Afun(-EN)
    xs = []
    p = dict(par or {}, tol=tol, maxiter=500)
    _, info = o_linear_solver(Afun, B, x0=B.zeros_like(), par=p, solver=method, callback=lambda x: xs.append(x.val))
    return xs, info['kit']


@pytest.mark.parametrize('method', ['CG', 'richardson'])
def test_host_callback_every_iteration_matches_oracle_iterates(method):
    """A plain callable gets x0 and then every iterate, as the reference's linear_solver and CG /
    richardson call it (solver.py:15-17, 75-76, 134-136); each iterate equals the oracle's to 1e-10
    relative and the callback's last argument is the returned solution."""
    N = (9, 9, 9)
    print('case host callback', method, flush=True)
    a, A = _case(N, 41, spread=0.3)
    par = {'alpha': float(a.max())} if method == 'richardson' else {}
    tol = 1e-8 if method == 'CG' else 1e-6
    ref, kit_ref = _oracle_iterates(N, A, tol, method, par)
    Afun = GpuOperator(N, np.ones(3), 'scalar', 'primal', A=Tensor(name='A', val=A, order=2, multype=21))
    EN = Tensor(name='EN', shape=(3,), N=N)
    EN.set_mean(np.eye(3)[0])
    B = Afun(-EN)
    got = []
    x, info = linear_solver(Afun, B, x0=None, par=dict(par, tol=tol, maxiter=500), solver=method,
                            callback=lambda t: got.append(t.to_host().val))
    assert info['kit'] == kit_ref and len(got) == kit_ref + 1 == len(ref)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert np.abs(g - r).max() <= 1e-10 * max(np.abs(r).max(), 1e-300) + 1e-14, k
    assert np.array_equal(got[-1], x.to_host().val)


def test_host_callback_exception_stops_the_solve():
    N = (9, 9, 9)
    a, A = _case(N, 42)
    Afun = GpuOperator(N, np.ones(3), 'scalar', 'primal', A=Tensor(name='A', val=A, order=2, multype=21))
    EN = Tensor(name='EN', shape=(3,), N=N)
    EN.set_mean(np.eye(3)[0])
    B = Afun(-EN)
    seen = []

    def cb(x):
        seen.append(1)
        if len(seen) == 4:
            raise KeyError('stop here')
    with pytest.raises(KeyError):
        linear_solver(Afun, B, par={'tol': 1e-12, 'maxiter': 500}, solver='CG', callback=cb)
    assert len(seen) == 4
    # the plan is usable afterwards, with the callback cleared
    x, info = linear_solver(Afun, B, par={'tol': 1e-8, 'maxiter': 500}, solver='CG')
    assert info['kit'] > 3


def test_host_callback_needs_fp64_and_every_k_schedule():
    N = (15, 15, 15)
    a, A = _case(N, 43)
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    EN = nat.DeviceArray.zeros((3,) + N)
    nat.add_constant(EN, -np.eye(3)[0])
    b = nat.DeviceArray((3,) + N)
    plan.matvec(EN, b)
    kits = []
    plan.set_iterate_callback(lambda k, x: kits.append(k), every=5)
    x = nat.DeviceArray.zeros((3,) + N)
    with pytest.raises(nat.FhError):
        plan.cg_solve(b, x, tol=1e-8, maxiter=500, mode='mixed', delta=1e-4)
    info = plan.cg_solve(b, x, tol=1e-8, maxiter=500, mode='fp64')
    plan.set_iterate_callback(None)
    want = sorted(set(list(range(5, info['kit'] + 1, 5)) + [info['kit']]))
    assert kits == want


def test_scipy_cg_logs_the_callback():
    """scipy_cg with a CallBack: SciPy's iterates reach the CallBack, which logs ||B - A x||."""
    N = (9, 9)
    a, A = _case(N, 44)
    Afun = GpuOperator(N, np.ones(2), 'scalar', 'primal', A=Tensor(name='A', val=A, order=2, multype=21))
    EN = Tensor(name='EN', shape=(2,), N=N)
    EN.set_mean(np.eye(2)[0])
    B = Afun(-EN)
    cb = CallBack(A=Afun, B=B)
    x, info = linear_solver(Afun, B, par={'tol': 1e-10, 'maxiter': 500}, solver='scipy_cg', callback=cb)
    assert info['info'] == 0 and len(cb.res_norm) >= 3
    assert cb.res_norm[-1] < 1e-6 * cb.res_norm[0]
