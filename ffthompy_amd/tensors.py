"""Device-backed fields and the operator algebra of the reference, on the HIP library.

Mirrors ffthompy/tensors/objects.py (Tensor) and ffthompy/tensors/operators.py (DFT,
Operator) for what the GaNi solve uses: a Tensor's ``val`` is a DeviceArray (or a
NumPy array for coefficients held on the host), arithmetic runs as GPU kernels, and
``*`` between fields is the grid inner product sum(x*y)/prod(N)
(ffthompy/tensors/objects.py:618-636).
"""
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
                nf = self.N[:-1] + (self.N[-1] // 2 + 1,)
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

    def __mul__(self, Y):
        if self.multype in ('scal', 'scalar'):
            return scalar_product(self, Y)
        raise NotImplementedError('only the grid inner product is defined between device fields')

    def norm(self):
        return scalar_product(self, self) ** 0.5

    def mean(self):
        """Mean per component (objects.py:259-272); of an 'r' spectrum: Re(val[0]) / prod(N)."""
        m = np.zeros(self.shape)
        if self.Fourier or not self.on_device:
            val = self.val.to_host() if self.on_device else np.asarray(self.val)
            for di in (np.ndindex(*self.shape) if self.order else [()]):
                v = val[di]
                m[di] = np.real(v[(0,) * v.ndim]) / float(np.prod(self.N)) if self.Fourier else np.mean(v)
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


class DFT:
    """(Inverse) DFT over the grid axes, 'r' form: rfftn unnormalised, irfftn normalised
    (ffthompy/tensors/operators.py:14-64, ffthompy/tensors/fft.py:39-43), on the GPU."""

    def __init__(self, inverse=False, N=None, fft_form='r', name=None, Y=None):
        if fft_form != 'r':
            raise NotImplementedError("the GPU DFT implements fft_form 'r'")
        self.N = tuple(int(n) for n in N)
        self.inverse = inverse
        self.fft_form = fft_form
        self.name = name or ('iDFT' if inverse else 'DFT')
        self._plan = nat.Plan(self.N, np.ones(len(self.N)) if Y is None else Y)

    def __call__(self, x):
        assert x.Fourier == self.inverse
        ncomp = int(np.prod(x.shape)) if x.order else 1
        x = x.to_device()
        if self.inverse:
            real = np.float64 if x.val.dtype == np.complex128 else np.float32
            out = nat.DeviceArray(x.shape + self.N, real)
            self._plan.irfftn(x.val, out, ncomp)
            return x._like(out, name='iF({0})'.format(x.name[:10]), Fourier=False)
        cdt = np.complex128 if x.val.dtype == np.float64 else np.complex64
        out = nat.DeviceArray(x.shape + self.N[:-1] + (self.N[-1] // 2 + 1,), cdt)
        self._plan.rfftn(x.val, out, ncomp)
        return x._like(out, name='F({0})'.format(x.name[:10]), Fourier=True)


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
