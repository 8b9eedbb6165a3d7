"""Device-backed fields and the operator algebra of the reference, on the HIP library.

Mirrors ffthompy/tensors/objects.py (Tensor) and ffthompy/tensors/operators.py (DFT,
Operator) for what the GaNi solve uses: a Tensor's ``val`` is a DeviceArray (or a
NumPy array for coefficients held on the host), arithmetic runs as GPU kernels, and
``*`` between fields is the grid inner product sum(x*y)/prod(N)
(ffthompy/tensors/objects.py:618-636).
"""
import threading

import numpy as np

from . import _native as nat


class Tensor:
    """Tensor-valued trigonometric polynomial, layout shape+N (objects.py:92-133)."""

    def __init__(self, name='', val=None, order=None, shape=None, N=None, Y=None, multype='scal',
                 Fourier=False, fft_form='r', origin=0, dtype=np.float64):
        self.name = name
        self.Fourier = Fourier
        self.fft_form = fft_form
        self.origin = origin
        self.multype = multype
        if val is not None:
            self.val = val
            self.order = int(order)
            self.shape = tuple(val.shape[:self.order])
            self.N = tuple(int(n) for n in (N if (Fourier and fft_form == 'r' and N is not None)
                                            else val.shape[self.order:]))
        elif shape is not None and N is not None:
            self.N = tuple(int(n) for n in N)
            self.shape = tuple(int(s) for s in shape)
            self.order = len(self.shape)
            if Fourier:
                nf = self.N[:-1] + (self.N[-1] // 2 + 1,) if fft_form == 'r' else self.N
                cdt = np.complex128 if np.dtype(dtype) == np.float64 else np.complex64
                self.val = nat.DeviceArray.zeros(self.shape + nf, cdt)
            else:
                self.val = nat.DeviceArray.zeros(self.shape + self.N, dtype)
        else:
            raise ValueError('Initialization of Tensor.')
        self.dim = len(self.N)
        self.Y = np.ones(self.dim) if Y is None else np.asarray(Y, dtype=np.float64)

    # -- bookkeeping (objects.py:39-45)
    @property
    def on_device(self):
        return isinstance(self.val, nat.DeviceArray)

    def _like(self, val, **kw):
        d = dict(name=self.name, val=val, order=self.order, N=self.N, Y=self.Y, multype=self.multype,
                 Fourier=self.Fourier, fft_form=self.fft_form, origin=self.origin)
        d.update(kw)
        return Tensor(**d)

    def copy(self, **kw):
        val = kw.pop('val', None)
        if val is None:
            val = self.val.copy()
        return self._like(val, **kw)

    def zeros_like(self, name=None):
        if self.on_device:
            return self._like(nat.DeviceArray.zeros(self.val.shape, self.val.dtype), name=name or self.name)
        return self._like(np.zeros_like(self.val), name=name or self.name)

    def empty_like(self, name=None):
        if self.on_device:
            return self._like(nat.DeviceArray(self.val.shape, self.val.dtype), name=name or self.name)
        return self._like(np.empty_like(self.val), name=name or self.name)

    def to_host(self):
        """Copy of this tensor with a NumPy ``val``."""
        return self._like(self.val.to_host() if self.on_device else np.array(self.val))

    def to_device(self):
        return self if self.on_device else self._like(nat.DeviceArray.from_host(self.val))

    def pN(self):
        return int(np.prod(self.N))

    # -- arithmetic (objects.py:188-245), fields on the device
    def _check(self, other):
        assert self.Fourier == other.Fourier
        assert self.val.shape == other.val.shape

    def __add__(self, x):
        if isinstance(x, Tensor):
            self._check(x)
            out = self.copy(name='({0}+{1})'.format(self.name[:10], x.name[:10]))
            nat.axpby(1., x.val, 1., out.val)
            return out
        if np.size(x) == 1:
            out = self.copy()
            E = np.full(int(np.prod(self.shape)) if self.order else 1, float(x))
            nat.add_constant(out.val, E)
            return out
        raise ValueError('Tensor.__add__')

    def __sub__(self, x):
        return self.__add__(-x)

    def __neg__(self):
        out = self.copy(name='-' + self.name[:10])
        nat.axpby(0., out.val, -1., out.val)
        return out

    def __rmul__(self, c):
        if np.size(c) != 1:
            raise ValueError('Tensor.__rmul__')
        out = self.copy()
        nat.axpby(0., out.val, float(c), out.val)
        return out

    def __call__(self, x):
        return self.__mul__(x)

    def __mul__(self, Y, multype=None):
        """The reference's products (objects.py:223-245): 'scal' the grid inner product; 21
        'ij...,j...->i...', 42 'ijkl...,kl...->ij...', 00 elementwise, 'grad' 'i...,...->i...',
        'div' 'i...,i...->...' as pointwise kernels on the device (fh_pointwise_product)."""
        if multype is None:
            multype = self.multype
        if multype in ('scal', 'scalar'):
            return scalar_product(self, Y)
        X = self
        assert X.Fourier == Y.Fourier
        assert nat.form_code(X.fft_form) == nat.form_code(Y.fft_form)
        grid = X.val.shape[X.order:]
        assert tuple(grid) == tuple(Y.val.shape[Y.order:])
        nvox = int(np.prod(grid))
        xs, ys = tuple(X.shape), tuple(Y.shape)
        if multype in (21, '21'):
            assert len(xs) == 2 and ys == xs[1:]
            A, B, mode, out_shape = xs[0], xs[1], 0, xs[:1]
        elif multype in (42, '42'):
            assert len(xs) == 4 and ys == xs[2:]
            A, B, mode, out_shape = xs[0] * xs[1], xs[2] * xs[3], 0, xs[:2]
        elif multype in (0, '00', 'elementwise', 'hadamard'):
            assert xs == ys
            A, B, mode, out_shape = int(np.prod(xs)) if xs else 1, 1, 1, xs
        elif multype == 'grad':
            assert len(xs) == 1 and ys == ()
            A, B, mode, out_shape = xs[0], 1, 0, xs
        elif multype == 'div':
            assert len(xs) == 1 and ys == xs
            A, B, mode, out_shape = 1, xs[0], 0, ()
        else:
            raise NotImplementedError('multype %r' % (multype,))
        Xd, Yd = X.to_device().val, Y.to_device().val
        if Xd.dtype != Yd.dtype:
            raise ValueError('Tensor product of %s and %s values' % (Xd.dtype, Yd.dtype))
        val = nat.pointwise_product(Xd, Yd, A, B, nvox, mode).reshape(out_shape + tuple(grid))
        return Y._like(val, name='{0}({1})'.format(X.name, Y.name), order=len(out_shape))

    def norm(self):
        return scalar_product(self, self) ** 0.5

    @property
    def fft_coef(self):
        """prod(N) for the unnormalised 'r' form, 1 for forms 0 and 'c' (objects.py:47-66)."""
        return float(np.prod(self.N)) if self.fft_form == 'r' else 1.

    def mean_index(self):
        """Index of the zero frequency: 0 for forms 'r' and 0, N//2 for 'c' (ffthompy/trigpol.py:220-224)."""
        return tuple(n // 2 for n in self.N) if self.fft_form == 'c' else (0,) * len(self.N)

    def mean(self):
        """Mean per component (objects.py:259-272); of a spectrum: Re(val[mean_index]) / fft_coef."""
        m = np.zeros(self.shape)
        if self.Fourier or not self.on_device:
            val = self.val.to_host() if self.on_device else np.asarray(self.val)
            for di in (np.ndindex(*self.shape) if self.order else [()]):
                v = val[di]
                m[di] = np.real(v[self.mean_index()]) / self.fft_coef if self.Fourier else np.mean(v)
            return m
        n = self.pN()
        ones = _ones(n, self.val.dtype)
        flat = self.val.reshape((int(np.prod(self.shape)) if self.order else 1, n))
        for i, di in enumerate(np.ndindex(*self.shape) if self.order else [()]):
            comp = _component(flat, i, n)
            m[di] = nat.dot(comp, ones) / n
        return m

    def add_mean(self, mean):
        mean = np.asarray(mean, dtype=np.float64)
        if not self.on_device:  # a coefficient held on the host: add to every voxel
            self.val = self.val + mean.reshape(mean.shape + (1,) * len(self.N))
            return self
        nat.add_constant(self.val, mean.ravel())
        return self

    def set_mean(self, mean):
        """objects.py:289-300"""
        mean = np.asarray(mean, dtype=np.float64)
        self.add_mean(-self.mean())
        return self.add_mean(mean)

    # -- Fourier forms, origin, inverse and grid changes on device data
    #    (objects.py:135-186, 247-251, 406-511); host-held values move to the device first
    @property
    def axes(self):
        """Axes of the grid in ``val`` (objects.py:403-405)."""
        return tuple(range(self.order, self.order + self.dim))

    def _count(self):
        return int(np.prod(self.shape)) if self.order else 1

    def _dev(self):
        if not self.on_device:
            self.val = nat.DeviceArray.from_host(np.ascontiguousarray(self.val))
        return self.val

    def _layout(self, fft_form):
        if self.Fourier and fft_form == 'r':
            return self.shape + self.N[:-1] + (self.N[-1] // 2 + 1,)
        return self.shape + self.N

    def set_fft_form(self, fft_form='r', copy=False):
        """Change the layout of a spectrum between the forms 'r', 0 and 'c' (objects.py:135-167):
        'r' -> 0 / 'c' completes the Hermitian half and divides by prod(N) (then fftshift for 'c');
        0 / 'c' -> 'r' keeps the first N[-1]//2+1 entries of the last axis times prod(N)."""
        R = self.copy() if copy else self
        src, dst = nat.form_code(R.fft_form), nat.form_code(fft_form)
        if src == dst:
            return R
        if R.Fourier:
            val, N, cnt, pN = R._dev(), R.N, R._count(), float(np.prod(R.N))
            if src == nat.FH_FORM_R:
                out = nat.spectrum_half_to_full(val, N, cnt, 1. / pN, centred=dst == nat.FH_FORM_CENTRED)
            elif dst == nat.FH_FORM_R:
                out = nat.spectrum_full_to_half(val, N, cnt, pN, centred=src == nat.FH_FORM_CENTRED)
            else:
                out = nat.fftshift(val, N, cnt, inverse=src == nat.FH_FORM_CENTRED)
            R.val = out.reshape(R._layout(fft_form))
        R.fft_form = fft_form
        return R

    def shift(self, origin=None):
        """Move the real-space origin between index 0 and the centre (numpy fftshift / ifftshift
        over the grid axes, objects.py:169-186)."""
        assert not self.Fourier
        if origin == self.origin:
            return self
        if origin is not None:
            raise ValueError()
        inverse = self.origin == 'c'
        out = nat.fftshift(self._dev(), self.N, self._count(), inverse=inverse)
        self.val = out.reshape(self.val.shape)
        self.origin = 0 if inverse else 'c'
        return self

    def inv(self):
        """Pointwise inverse of a square second-order field (objects.py:247-251), on the GPU."""
        assert self.Fourier is False
        assert self.order == 2
        assert self.shape[0] == self.shape[1]
        src = self.to_device()
        return self.copy(name='inv({})'.format(self.name), val=nat.pointwise_inverse(src.val))

    def fourier(self, Fourier=None, copy=False):
        """Forward or inverse DFT of the tensor's fft_form (objects.py:406-426)."""
        assert self.origin == 0
        if self.Fourier == Fourier:
            return self.copy() if copy else self
        F = DFT(inverse=self.Fourier, N=self.N, fft_form=self.fft_form, Y=self.Y)
        res = F(self.to_device() if copy else self._like(self._dev()))
        if copy:
            return self.copy(val=res.val, Fourier=not self.Fourier)
        self.val = res.val
        self.Fourier = not self.Fourier
        return self

    def _resize(self, M, enlarge):
        M = tuple(int(m) for m in M)
        fft_form = self.fft_form
        self.set_fft_form('c')  # as the reference, the tensor itself is left in form 'c'
        val = nat.spectrum_resize(self._dev(), self.N, M, self._count(), enlarge)
        R = self.copy(val=val.reshape(self.shape + M), N=M, fft_form='c')
        return R.set_fft_form(fft_form)

    def enlarge(self, M):
        """Zero-pad the spectrum to grid M; the Nyquist entries of even axes are split in halves
        at -N/2 and +N/2 (objects.py:428-467)."""
        assert self.Fourier
        if np.allclose(self.N, M):
            return self
        return self._resize(M, True)

    def decrease(self, M):
        """Drop the frequencies outside grid M (objects.py:469-486)."""
        assert self.Fourier
        if np.allclose(self.N, M):
            return self
        return self._resize(M, False)

    def project(self, M):
        """Trigonometric interpolation onto grid M: enlarge when M > N on every axis, decrease
        when M < N on every axis (objects.py:488-511)."""
        if np.allclose(self.N, M):
            return self
        Fourier = self.Fourier
        Y = self.copy() if Fourier else self.fourier(copy=True)
        if np.all(np.greater(M, self.N)):
            Y = Y.enlarge(M)
        elif np.all(np.less(M, self.N)):
            Y = Y.decrease(M)
        else:
            raise NotImplementedError()
        if not Fourier:
            Y = Y.fourier()
        return Y

    def __repr__(self):
        return 'Tensor(name=%s, shape=%s, N=%s, Fourier=%s, on_device=%s)' % (
            self.name, self.shape, self.N, self.Fourier, self.on_device)


_ONES = {}


def _ones(n, dtype):
    key = (n, np.dtype(dtype).str)
    if key not in _ONES:
        _ONES.clear()
        _ONES[key] = nat.DeviceArray.from_host(np.ones(n, dtype=dtype))
    return _ONES[key]


class _ComponentView(nat.DeviceArray):
    """Non-owning view of component i of a component-major device field."""

    def __init__(self, owner, offset_bytes, n, dtype):  # noqa: no super().__init__ (no allocation)
        import ctypes
        self.shape = (n,)
        self.dtype = np.dtype(dtype)
        self.nbytes = n * self.dtype.itemsize
        self._ptr = ctypes.c_void_p(owner.ptr.value + offset_bytes)
        self._owner = owner


def _component(flat, i, n):
    return _ComponentView(flat, i * n * flat.dtype.itemsize, n, flat.dtype)


def scalar_product(y, x):
    """Grid inner product sum(y*x)/prod(N) (objects.py:618-636, real space)."""
    assert isinstance(x, Tensor) and y.val.shape == x.val.shape
    if y.Fourier:
        # 'r' spectra (results such as a potential's Fu, read on the host): plane k2 = 0, and the
        # Nyquist plane of an even last axis, once, the others twice, / prod(N)^2
        Yv, Xv = np.asarray(_host(y)), np.asarray(_host(x))
        w = np.full(Yv.shape[-1], 2.)
        w[0] = 1.
        if y.N[-1] % 2 == 0:
            w[-1] = 1.
        return float(np.sum(w * (Yv * np.conj(Xv)).real) / float(np.prod(y.N)) ** 2)
    if not (y.on_device and x.on_device):
        return float(np.sum(np.asarray(_host(y)) * np.asarray(_host(x))) / np.prod(y.N))
    return nat.dot(y.val, x.val) / float(np.prod(y.N))


def _host(t):
    return t.val.to_host() if t.on_device else t.val


_PLANS = {}


def _fft_plan(N, Y):
    """A plan of grid N for the transforms, kept per host thread (a plan is not re-entrant; a few
    grids are kept)."""
    key = (threading.get_ident(), tuple(int(n) for n in N))
    pl = _PLANS.get(key)
    if pl is None:
        if len(_PLANS) >= 8:
            _PLANS.pop(next(iter(_PLANS)))
        pl = _PLANS[key] = nat.Plan(key[1], np.ones(len(key[1])) if Y is None else Y)
    return pl


class DFT:
    """(Inverse) DFT over the grid axes on the GPU (ffthompy/tensors/operators.py:14-64,
    ffthompy/tensors/fft.py:18-43), fft_form
      'r': rfftn unnormalised / irfftn normalised;
      0  : fftn / prod(N) / real(ifftn) * prod(N);
      'c': form 0 with fftshift of the spectrum / ifftshift before the inverse.
    Forms 0 and 'c' run the real transform and complete or halve the spectrum on the device; the
    inverse takes the Hermitian part of the spectrum, which gives the real part of ifftn."""

    def __init__(self, inverse=False, N=None, fft_form='r', name=None, Y=None):
        self.form = nat.form_code(fft_form)
        self.N = tuple(int(n) for n in N)
        self.inverse = inverse
        self.fft_form = fft_form
        self.name = name or ('iDFT' if inverse else 'DFT')
        self._plan = _fft_plan(self.N, Y)

    def __call__(self, x):
        assert x.Fourier == self.inverse
        ncomp = int(np.prod(x.shape)) if x.order else 1
        x = x.to_device()
        N, pN = self.N, float(np.prod(self.N))
        nh = N[:-1] + (N[-1] // 2 + 1,)
        centred = self.form == nat.FH_FORM_CENTRED
        if self.inverse:
            real = np.float64 if x.val.dtype == np.complex128 else np.float32
            half = x.val
            if self.form != nat.FH_FORM_R:
                half = nat.spectrum_full_to_half(x.val, N, ncomp, pN, centred=centred, hermitian=True)
            out = nat.DeviceArray(x.shape + N, real)
            self._plan.irfftn(half.reshape(x.shape + nh), out, ncomp)
            return x._like(out, name='iF({0})'.format(x.name[:10]), Fourier=False)
        assert nat.form_code(x.fft_form) == self.form
        cdt = np.complex128 if x.val.dtype == np.float64 else np.complex64
        half = nat.DeviceArray(x.shape + nh, cdt)
        self._plan.rfftn(x.val, half, ncomp)
        if self.form == nat.FH_FORM_R:
            return x._like(half, name='F({0})'.format(x.name[:10]), Fourier=True)
        full = nat.spectrum_half_to_full(half, N, ncomp, 1. / pN, centred=centred)
        return x._like(full.reshape(x.shape + N), name='F({0})'.format(x.name[:10]), Fourier=True)


class GpuOperator:
    """Afun = F^-1 G(xi) F A(x), the matrix-vector product of the GaNi solve
    (Operator([[GN, A]]), ffthompy/applications.py:58; Operator.__call__,
    ffthompy/tensors/operators.py:136-144) as one call into the HIP library.

    A is given as the reference's coefficient Tensor (shape (d,d)+N or (D,D)+N, origin 0)
    or directly as (kind, values) with kind 'iso', 'lame' or 'full'.
    """

    def __init__(self, N, Y, physics='scalar', primaldual='primal', A=None, coef=None, name='FiGFA',
                 _defer=False):
        self.name = name
        self.N = tuple(int(n) for n in N)
        self.Y = np.asarray(Y, dtype=np.float64)
        self.physics = physics
        self.primaldual = primaldual
        self.plan = nat.Plan(self.N, self.Y, physics, primaldual)
        self.ncomp = self.plan.ncomp
        self.Nsolve = None  # Ga solve: the grid N whose projection is enlarged onto self.N
        if A is not None:
            kind, vals = classify_coefficient(A.val if isinstance(A, Tensor) else A, physics)
            self.plan.set_coeff(kind, vals)
        elif coef is not None:
            self.plan.set_coeff(*coef)
        elif not _defer:
            raise ValueError('GpuOperator needs a coefficient')

    @classmethod
    def ga(cls, N, Y, physics, primaldual, A_ga, grid=None, name='FiGFA'):
        """Operator of the Ga solve (ffthompy/applications.py:28-31, 48-51): the projection of
        the grid N enlarged onto the Ga grid and the exactly integrated coefficient A_ga
        (materials.GaCoefficient for Nbar = 2N-1). It runs on `grid` (default: 2N-1 per axis, or
        the next length with prime factors <= 13): the solutions have frequencies below N/2,
        so every grid of at least 2N-1 points gives the same operator, inner products and
        energies."""
        N = tuple(int(n) for n in N)
        grid = tuple(nat.smooth_length(2 * n - 1) for n in N) if grid is None else tuple(grid)
        op = cls(grid, Y, physics, primaldual, name=name, _defer=True)
        op.Nsolve = N
        op.coarse_plan = nat.Plan(N, Y, physics, primaldual)
        # the reference's enlarged hG carries the factor prod(2N-1)/prod(N) (fh_plan_set_band)
        op.plan.set_band([(n - 1) // 2 for n in N], float(np.prod([2 * n - 1 for n in N])) / float(np.prod(N)))
        ga = A_ga.load(nat.Ga(op.coarse_plan, grid=grid))
        ga.set_plan_coeff(op.plan)
        return op

    def __call__(self, x):
        x = x.to_device()
        out = nat.DeviceArray(x.val.shape, x.val.dtype)
        self.plan.matvec(x.val, out)
        return x._like(out, name='{0}({1})'.format(self.name[:6], x.name[:10]))

    def project(self, x):
        """G x without the coefficient (the projection operator G1N / G2N)."""
        x = x.to_device()
        out = nat.DeviceArray(x.val.shape, x.val.dtype)
        self.plan.project(x.val, out)
        return x._like(out)

    def __repr__(self):
        return 'GpuOperator(%s, N=%s, %s, %s, coefficient %s)' % (self.name, self.N, self.physics,
                                                                  self.primaldual, self.plan.coef_kind)


def classify_coefficient(A, physics):
    """Compact storage of a (k,k)+N coefficient: 'iso' (a I), 'lame' (Mandel
    lambda m m^T + 2 mu I) or 'full'."""
    A = np.asarray(A, dtype=np.float64)
    k = A.shape[0]
    off = A.copy()
    diag = np.array([A[i, i] for i in range(k)])
    for i in range(k):
        off[i, i] = 0.
    if np.all(off == 0.) and np.all(diag == diag[0]):
        return 'iso', np.ascontiguousarray(diag[0])
    if physics == 'elasticity' and k in (3, 6):
        dn = 3 if k == 6 else 2
        lam = A[0, 1]
        mu = 0.5 * A[dn, dn]
        m = np.zeros(k)
        m[:dn] = 1.
        rec = np.einsum('i,j,...->ij...', m, m, lam) + 2. * np.einsum('ij,...->ij...', np.eye(k), mu)
        if np.allclose(rec, A, rtol=1e-14, atol=1e-14 * max(np.abs(A).max(), 1.)):
            return 'lame', np.ascontiguousarray(np.stack([lam, mu]))
    return 'full', np.ascontiguousarray(A.reshape((k * k,) + A.shape[2:]))
