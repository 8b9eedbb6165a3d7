"""Property tests (hypothesis) of the invariants the reference's unit tests assert
(ffthompy/tensors/unittest_operators.py, unittest_tensors.py), on the oracle, over random grids:

* the scalar projections G1, G2 are idempotent, mutually orthogonal and sum to the identity off the
  mean (unittest_operators.py:177-203);
* a gradient field is left alone by G1 and removed by G2 (unittest_operators.py:102-175);
* FFT forms round-trip (unittest_operators.py:23-100); spectral enlargement to a finer grid keeps
  the mean and the norm and decreasing it back restores the field (unittest_tensors.py:11-47);
* div grad = -|2 pi xi|^2, so P (-div grad) P is the identity off the mean (the potential operator
  for a unit coefficient).
"""
import numpy as np
from hypothesis import given, settings, strategies as st

from oracle import potential as opot
from oracle import projections
from oracle.fields import DFT, Field

odd = st.integers(min_value=1, max_value=5).map(lambda k: 2 * k + 1)
grids = st.one_of(st.tuples(odd, odd), st.tuples(odd, odd, st.sampled_from([3, 5])))
seeds = st.integers(min_value=0, max_value=2 ** 31 - 1)


def _field(N, shape, rng):
    return Field(val=rng.standard_normal(tuple(shape) + tuple(N)), order=len(shape))


@settings(max_examples=25, deadline=None, derandomize=True)
@given(grids, seeds)
def test_scalar_projections_are_complementary_orthogonal_projections(N, seed):
    rng = np.random.default_rng(seed)
    d = len(N)
    _, G1, G2 = projections.scalar(N, np.ones(d), NyqNul=True)
    g1 = np.moveaxis(G1.val, (0, 1), (-2, -1))
    g2 = np.moveaxis(G2.val, (0, 1), (-2, -1))
    assert np.abs(g1 @ g1 - g1).max() < 1e-13
    assert np.abs(g2 @ g2 - g2).max() < 1e-13
    assert np.abs(g1 @ g2).max() < 1e-13
    eye = np.broadcast_to(np.eye(d), g1.shape).copy()
    eye[(0,) * d] = 0.
    assert np.abs(g1 + g2 - eye).max() < 1e-13


@settings(max_examples=25, deadline=None, derandomize=True)
@given(grids, seeds)
def test_gradient_fields_are_fixed_by_g1_and_removed_by_g2(N, seed):
    rng = np.random.default_rng(seed)
    d = len(N)
    u = DFT(N)(_field(N, (), rng))
    e = opot.grad(u)
    _, G1, G2 = projections.scalar(N, np.ones(d), NyqNul=True)
    v1 = np.einsum('ij...,j...->i...', G1.val, e.val)
    v2 = np.einsum('ij...,j...->i...', G2.val, e.val)
    assert np.abs(v1 - e.val).max() < 1e-11 * np.abs(e.val).max()
    assert np.abs(v2).max() < 1e-11 * np.abs(e.val).max()


@settings(max_examples=25, deadline=None, derandomize=True)
@given(grids, seeds)
def test_fft_round_trip_and_enlarge_decrease(N, seed):
    rng = np.random.default_rng(seed)
    x = _field(N, (2,), rng)
    for form in ('r', 0, 'c'):
        y = DFT(N, inverse=True, fft_form=form)(DFT(N, fft_form=form)(x.copy(fft_form=form)))
        assert np.abs(y.val - x.val).max() < 1e-12
    M = tuple(2 * n + 1 for n in N)
    X = DFT(N)(x)
    big = X.project(M)
    xb = DFT(M, inverse=True)(big)
    assert np.allclose(xb.mean(), x.mean(), atol=1e-13)
    assert abs(xb.norm() - x.norm()) < 1e-12 * x.norm()
    back = big.project(N)
    assert np.abs(back.val - X.val).max() < 1e-12 * np.abs(X.val).max()


@settings(max_examples=25, deadline=None, derandomize=True)
@given(grids, seeds)
def test_preconditioned_laplacian_is_identity_off_the_mean(N, seed):
    rng = np.random.default_rng(seed)
    d = len(N)
    u = DFT(N)(_field(N, (), rng))
    lap = opot.div(opot.grad(u)).val
    q2 = np.sum(np.abs(opot.grad_symbol(N, np.ones(d))) ** 2, axis=0)
    assert np.abs(lap + q2 * u.val).max() < 1e-11 * np.abs(q2 * u.val).max()
    A = Field(val=np.einsum('ij,...->ij...', np.eye(d), np.ones(N)), order=2, multype=21)
    got = opot.operator_gani(A)(u).val
    want = u.val.copy()
    want[(0,) * d] = 0.
    assert np.abs(got - want).max() < 1e-12 * np.abs(u.val).max()
