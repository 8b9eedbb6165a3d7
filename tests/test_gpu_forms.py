"""Spectrum forms, shifts, grid changes, pointwise inverses and projections of every form on the
GPU against the float64 oracle (oracle/fields.py, oracle/projections.py, oracle/grid.py):

* DFT in the forms 'r', 0 and 'c' (ffthompy/tensors/fft.py:18-43, operators.py:14-64), forward on
  real fields and inverse on arbitrary (non-Hermitian) full spectra, whose inverse is the real
  part of ifftn;
* Tensor.set_fft_form between every pair of forms (objects.py:135-167), Tensor.shift
  (objects.py:169-186, a pure permutation: bit-exact), Tensor.inv (objects.py:247-251);
* Tensor.enlarge / decrease / project (objects.py:428-511) on even and odd grids, including the
  Nyquist split of even axes and the reference's side effect of leaving the tensor in form 'c';
* projections.scalar / elasticity for NyqNul True and False in every form (projections.py:9-267).
float64 results within 1e-12 relative of the oracle (1e-13 for pure data movement), float32 1e-5.
"""
import numpy as np
import pytest

from ffthompy_amd import _native as nat
from ffthompy_amd import projections
from ffthompy_amd.tensors import DFT, Tensor
from oracle import grid as ogrid
from oracle import projections as oproj
from oracle.fields import DFT as ODFT
from oracle.fields import Field

pytestmark = pytest.mark.gpu

FORMS = ['r', 0, 'c']
GRIDS = [(5, 7, 9), (6, 8, 10), (4, 7), (9, 6), (8, 8, 5)]


def _dev(val, order, N, Fourier=False, fft_form='r', dtype=None):
    v = np.ascontiguousarray(val if dtype is None else val.astype(dtype))
    return Tensor(val=nat.DeviceArray.from_host(v), order=order, N=N, Fourier=Fourier, fft_form=fft_form)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _spectrum(N, form, rng, shape=(2,)):
    """Spectrum in `form` of a random real field (oracle), and the field."""
    x = rng.standard_normal(shape + tuple(N))
    X = ODFT(N, fft_form=form)(Field(val=x, order=len(shape), fft_form=form))
    return X, x


def _fid(N, form):
    return '%s-%s' % ('x'.join(map(str, N)), form)


@pytest.mark.parametrize('N,form', [(N, f) for N in GRIDS for f in FORMS], ids=[_fid(N, f) for N in GRIDS
                                                                                for f in FORMS])
def test_dft_forms_match_oracle(N, form):
    print('case dft', N, form, flush=True)
    rng = np.random.default_rng(7)
    X, x = _spectrum(N, form, rng)
    G = DFT(N=N, fft_form=form)(_dev(x, 1, N, fft_form=form))
    assert G.Fourier and G.fft_form == form
    assert _rel(G.to_host().val, X.val) < 1e-12
    back = DFT(N=N, inverse=True, fft_form=form)(G)
    assert _rel(back.to_host().val, x) < 1e-12
    if form != 'r':
        # arbitrary complex spectrum: the inverse is the real part of ifftn (times prod(N))
        Z = rng.standard_normal((2,) + N) + 1j * rng.standard_normal((2,) + N)
        ref = ODFT(N, inverse=True, fft_form=form)(Field(val=Z, order=1, Fourier=True, fft_form=form))
        got = DFT(N=N, inverse=True, fft_form=form)(_dev(Z, 1, N, Fourier=True, fft_form=form))
        assert _rel(got.to_host().val, ref.val) < 1e-12
    # float32
    G32 = DFT(N=N, fft_form=form)(_dev(x, 1, N, fft_form=form, dtype=np.float32))
    assert G32.val.dtype == np.complex64 and _rel(G32.to_host().val, X.val) < 1e-5


@pytest.mark.parametrize('N', GRIDS, ids=['x'.join(map(str, N)) for N in GRIDS])
def test_set_fft_form_every_pair(N):
    print('case set_fft_form', N, flush=True)
    rng = np.random.default_rng(8)
    for src in FORMS:
        X, _ = _spectrum(N, src, rng)
        cases = [X]
        if src != 'r':  # also a non-Hermitian full spectrum
            Z = rng.standard_normal((2,) + N) + 1j * rng.standard_normal((2,) + N)
            cases.append(Field(val=Z, order=1, Fourier=True, fft_form=src))
        for S in cases:
            for dst in FORMS:
                ref = S.copy().set_fft_form(dst)
                T = _dev(S.val, 1, N, Fourier=True, fft_form=src)
                got = T.set_fft_form(dst, copy=True)
                assert got.fft_form == dst and T.fft_form == src
                assert got.val.shape == ref.val.shape, (src, dst)
                assert _rel(got.to_host().val, ref.val) < 1e-13, (src, dst)
                assert np.allclose(got.mean(), ref.mean(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('N', GRIDS, ids=['x'.join(map(str, N)) for N in GRIDS])
def test_shift_is_fftshift(N):
    rng = np.random.default_rng(9)
    axes = tuple(range(1, 1 + len(N)))
    for dt in (np.float64, np.float32, np.complex128):
        x = rng.standard_normal((3,) + N).astype(dt)
        T = _dev(x, 1, N)
        T.shift()
        assert T.origin == 'c' and np.array_equal(T.to_host().val, np.fft.fftshift(x, axes))
        T.shift()
        assert T.origin == 0 and np.array_equal(T.to_host().val, x)
        assert T.shift(origin=0) is T
    F = Field(val=x.real.copy(), order=1).shift()
    T = _dev(x.real.copy(), 1, N).shift()
    assert np.array_equal(T.to_host().val, F.val)


@pytest.mark.parametrize('k', [1, 2, 3, 6])
def test_inv_matches_pointwise_inverse(k):
    rng = np.random.default_rng(10 + k)
    N = (7, 6, 5)
    A = rng.standard_normal((k, k) + N) * 0.3
    for i in range(k):
        A[i, i] += 2. + rng.random(N)
    ref = ogrid.pointwise_inverse(A)
    got = _dev(A, 2, N).inv()
    assert _rel(got.to_host().val, ref) < 1e-12
    g32 = _dev(A, 2, N, dtype=np.float32).inv()
    assert _rel(g32.to_host().val, ref) < 1e-5
    # an inverse times the matrix is the identity at every voxel
    prod = np.einsum('ij...,jk...->ik...', got.to_host().val, A)
    assert np.abs(prod - np.eye(k).reshape((k, k) + (1,) * 3)).max() < 1e-12


RESIZE = [((5, 7, 9), (9, 11, 13)), ((6, 8, 10), (7, 12, 11)), ((4, 7), (9, 8)), ((6, 4), (8, 6)),
          ((9, 6), (5, 4)), ((8, 8, 5), (6, 5, 3)), ((8, 8, 5), (7, 8, 5))]


@pytest.mark.parametrize('N,M', RESIZE, ids=['%s-%s' % ('x'.join(map(str, a)), 'x'.join(map(str, b)))
                                             for a, b in RESIZE])
def test_enlarge_decrease_match_oracle(N, M):
    print('case resize', N, M, flush=True)
    rng = np.random.default_rng(12)
    grow = all(m >= n for m, n in zip(M, N))
    for form in FORMS:
        X, _ = _spectrum(N, form, rng)
        cases = [X]
        if form != 'r':
            Z = rng.standard_normal((2,) + N) + 1j * rng.standard_normal((2,) + N)
            cases.append(Field(val=Z, order=1, Fourier=True, fft_form=form))
        for S in cases:
            So = S.copy()
            ref = So.enlarge(M) if grow else So.decrease(M)
            T = _dev(S.val, 1, N, Fourier=True, fft_form=form)
            got = T.enlarge(M) if grow else T.decrease(M)
            assert got.N == tuple(M) and got.fft_form == form
            assert _rel(got.to_host().val, ref.val) < 1e-13, form
            # the reference leaves the resized tensor itself in form 'c'
            assert T.fft_form == So.fft_form == 'c'
            assert _rel(T.to_host().val, So.val) < 1e-13


@pytest.mark.parametrize('N,M', [((5, 7, 9), (11, 15, 19)), ((6, 8), (12, 16)), ((10, 9), (4, 5)),
                                 ((8, 6, 7), (5, 3, 4))])
def test_project_real_fields_match_oracle(N, M):
    print('case project', N, M, flush=True)
    rng = np.random.default_rng(13)
    x = rng.standard_normal((3,) + N)
    ref = Field(val=x.copy(), order=1).project(M)
    got = _dev(x, 1, N).project(M)
    assert not got.Fourier and got.N == tuple(M)
    assert _rel(got.to_host().val, ref.val) < 1e-12
    if all(m > n for m, n in zip(M, N)) and all(m % n == 0 for m, n in zip(M, N)):
        # interpolation reproduces the nodal values
        sub = got.to_host().val[(slice(None),) + tuple(slice(0, m, m // n) for m, n in zip(M, N))]
        assert np.abs(sub - x).max() < 1e-12
    with pytest.raises(NotImplementedError):
        _dev(x, 1, N).project(tuple(n + 1 if i == 0 else n - 1 for i, n in enumerate(N)))


def _oracle_G(physics, N, Y, nyq, form):
    if physics == 'scalar':
        G0, G1, G2 = oproj.scalar(N, Y, NyqNul=nyq, fft_form=form)
        return G0.val, G1.val, G2.val
    G0, a, b, c, d = oproj.elasticity(N, Y, NyqNul=nyq, fft_form=form)
    return G0.val, a.val + b.val, c.val + d.val


PROJ = [(N, ph, nyq, f) for N in [(6, 8, 10), (5, 7, 9), (8, 6), (7, 9)] for ph in ('scalar', 'elasticity')
        for nyq in (True, False) for f in FORMS]


@pytest.mark.parametrize('N,physics,nyq,form', PROJ,
                         ids=['%s-%s-%s-%s' % ('x'.join(map(str, a)), b, 'nyqnul' if c else 'keepnyq', d)
                              for a, b, c, d in PROJ])
def test_projections_every_form_match_oracle(N, physics, nyq, form):
    # elasticity with NyqNul=True on even grids: the reference builds xi on N but its arrays on the
    # odd grid and raises; the oracle (as this library) nulls the Nyquist modes instead
    d = len(N)
    Y = np.array([1., 1.5, 2.])[:d]
    M0, M1, M2 = _oracle_G(physics, N, Y, nyq, form)
    nc = M1.shape[0]
    rng = np.random.default_rng(14)
    lay = N[:-1] + (N[-1] // 2 + 1,) if form == 'r' else N
    X = rng.standard_normal((nc,) + lay) + 1j * rng.standard_normal((nc,) + lay)
    G0, G1, G2 = (projections.scalar if physics == 'scalar' else projections.elasticity)(N, Y, NyqNul=nyq,
                                                                                        fft_form=form)
    T = _dev(X, 1, N, Fourier=True, fft_form=form)
    for G, M in ((G0, M0), (G1, M1), (G2, M2)):
        ref = np.einsum('ij...,j...->i...', M, X)
        assert _rel(G(T).to_host().val, ref) < 1e-13, G
    T32 = _dev(X, 1, N, Fourier=True, fft_form=form, dtype=np.complex64)
    assert _rel(G1(T32).to_host().val, np.einsum('ij...,j...->i...', M1, X)) < 1e-5


PRODUCTS = [(21, (3, 3), (3,)), (42, (3, 3, 3, 3), (3, 3)), (0, (2, 3), (2, 3)), ('grad', (3,), ()),
            ('div', (3,), (3,)), (21, (6, 6), (6,))]


@pytest.mark.parametrize('multype,xs,ys', PRODUCTS, ids=['%s-%s' % (p[0], 'x'.join(map(str, p[1]))) for p in PRODUCTS])
def test_tensor_products_match_einsum(multype, xs, ys):
    """Tensor.__mul__ with the reference's multypes (objects.py:223-245) on device fields, real
    fp64/fp32 and complex spectra, against numpy.einsum."""
    sub = {21: 'ij...,j...->i...', 42: 'ijkl...,kl...->ij...', 0: '...,...->...', 'grad': 'i...,...->i...',
           'div': 'i...,i...->...'}[multype]
    N = (6, 5, 7)
    rng = np.random.default_rng(15)
    X = rng.standard_normal(xs + N)
    Y = rng.standard_normal(ys + N)
    ref = np.einsum(sub, X, Y)
    for dt, tol in ((np.float64, 1e-14), (np.float32, 1e-5)):
        tx = Tensor(val=nat.DeviceArray.from_host(X.astype(dt)), order=len(xs), multype=multype)
        ty = Tensor(val=nat.DeviceArray.from_host(Y.astype(dt)), order=len(ys))
        got = tx(ty)
        assert got.order == ref.ndim - len(N) and got.N == N
        assert _rel(got.to_host().val, ref) < tol
    # complex spectra in form 0
    Xc = X + 1j * rng.standard_normal(X.shape)
    Yc = Y + 1j * rng.standard_normal(Y.shape)
    tx = Tensor(val=nat.DeviceArray.from_host(Xc), order=len(xs), multype=multype, Fourier=True, fft_form=0)
    ty = Tensor(val=nat.DeviceArray.from_host(Yc), order=len(ys), Fourier=True, fft_form=0)
    assert _rel(tx(ty).to_host().val, np.einsum(sub, Xc, Yc)) < 1e-14
    # a coefficient held on the host acts on a device field
    th = Tensor(val=X, order=len(xs), multype=multype)
    assert _rel(th(Tensor(val=nat.DeviceArray.from_host(Y), order=len(ys))).to_host().val, ref) < 1e-14


@pytest.mark.parametrize('N,physics', [((5, 7, 9), 'scalar'), ((6, 8), 'scalar'), ((5, 5, 7), 'elasticity')])
def test_operator_composition_matches_oracle(N, physics):
    """The reference driver's composition Afun = Operator([[Operator([[iF, hG1, F]]), A]])
    (ffthompy/applications.py:54-58) built from the drop-in's DFT, projection and coefficient Tensor,
    against the oracle's materialised operator."""
    from ffthompy_amd.projections import Operator
    from oracle.fields import Operator as OOperator
    d = len(N)
    rng = np.random.default_rng(16)
    Y = np.ones(d)
    nc = d if physics == 'scalar' else d * (d + 1) // 2
    a = 1. + rng.random(N)
    A = np.einsum('ij,...->ij...', np.eye(nc), a) + 0.1 * rng.standard_normal((nc, nc) + N)
    x = rng.standard_normal((nc,) + N)
    if physics == 'scalar':
        _, oG1, _ = oproj.scalar(N, Y)
        _, G1, _ = projections.scalar(N, Y)
        MG = oG1
    else:
        _, a1, b1, _, _ = oproj.elasticity(N, Y)
        _, G1, _ = projections.elasticity(N, Y)
        MG = a1 + b1
    oAfun = OOperator([[OOperator([[ODFT(N, inverse=True), MG, ODFT(N)]]), Field(val=A, order=2, multype=21)]])
    ref = oAfun(Field(val=x, order=1)).val
    Afun = Operator(mat=[[Operator(mat=[[DFT(N=N, inverse=True), G1, DFT(N=N)]]), Tensor(val=A, order=2, multype=21)]])
    got = Afun(_dev(x, 1, N)).to_host().val
    assert _rel(got, ref) < 1e-12
