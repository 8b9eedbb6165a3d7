"""Property tests (hypothesis) of the device operator over random grids: odd and even lengths
with prime factors 2, 3, 5, 7, 11, 13 (fixed-length and runtime-radix kernels mixed), random
cell sizes and random positive coefficients.

* fp64 matvec equals the oracle's operator F^-1 G F A (1e-12 relative), primal and dual;
* on odd grids the potential solve converges (no breakdown) and its field gives the gradient
  formulation's AH_11 (both to tol 1e-11; 1e-9 relative).
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from oracle import projections
from oracle.fields import DFT, Field, Operator

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')

lengths = st.sampled_from([3, 5, 6, 7, 9, 10, 11, 12, 13, 14, 15, 21, 22, 25, 27])
odd_lengths = st.sampled_from([3, 5, 7, 9, 11, 13, 15, 21, 25, 27])
grids = st.one_of(st.tuples(lengths, lengths), st.tuples(lengths, lengths, lengths))
odd_grids = st.one_of(st.tuples(odd_lengths, odd_lengths), st.tuples(odd_lengths, odd_lengths, odd_lengths))
seeds = st.integers(min_value=0, max_value=2 ** 31 - 1)
SETTINGS = dict(max_examples=20, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@settings(**SETTINGS)
@given(grids, seeds, st.sampled_from(['primal', 'dual']))
def test_matvec_matches_oracle_on_random_grids(N, seed, pd):
    print('case property matvec', N, pd, flush=True)
    rng = np.random.default_rng(seed)
    d = len(N)
    Y = 0.5 + rng.random(d)
    a = np.exp(rng.standard_normal(N))
    _, G1, G2 = projections.scalar(N, Y, NyqNul=True)
    hG = G1 if pd == 'primal' else G2
    A = np.einsum('ij,...->ij...', np.eye(d), a)
    Afun = Operator([[DFT(N, inverse=True), hG, DFT(N), Field(val=A, order=2, multype=21)]])
    x = rng.standard_normal((d,) + N)
    ref = Afun(Field(val=x, order=1)).val
    plan = nat.Plan(N, Y, 'scalar', pd)
    plan.set_coeff('iso', a)
    dx = nat.DeviceArray.from_host(x)
    dy = nat.DeviceArray((d,) + N)
    plan.matvec(dx, dy)
    assert _rel(dy.to_host(), ref) < 1e-12


@settings(**SETTINGS)
@given(odd_grids, seeds)
def test_potential_solution_equals_gradient_formulation(N, seed):
    print('case property potential', N, flush=True)
    rng = np.random.default_rng(seed)
    d = len(N)
    a = np.exp(0.5 * rng.standard_normal(N))
    plan = nat.Plan(N, (1.,) * d, 'scalar', 'primal')
    plan.set_coeff('iso', a)
    E = np.eye(d)[0]
    X = nat.DeviceArray((d,) + N)
    _, pinfo = plan.potential_solve(E, 1e-11, 1000, X=X)
    assert pinfo['status'] == nat.FH_CG_OK
    nat.add_constant(X, E)
    ahp = plan.energy_matrix([X] * d)[0, 0]
    EN = nat.DeviceArray.zeros((d,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((d,) + N)
    plan.matvec(EN, B)
    x = nat.DeviceArray.zeros((d,) + N)
    info = plan.cg_solve(B, x, tol=1e-11, maxiter=1000, mode='fp64')
    assert info['status'] == nat.FH_CG_OK
    nat.add_constant(x, E)
    ahg = plan.energy_matrix([x] * d)[0, 0]
    assert abs(ahp - ahg) < 1e-9 * ahg
