"""Fields of trigonometric polynomials, FFT forms, DFT and operator composition.

Restates the parts of ffthompy/tensors/{objects,operators,fft}.py that the
homogenization path uses. Test infrastructure only (see oracle/__init__.py).
"""
import itertools
import numpy as np

from . import grid


# ---- FFT wrappers: ffthompy/tensors/fft.py:4-43 (axes passed explicitly) ----

def _grid_axes(ndim, N):
    return tuple(range(ndim - len(N), ndim))


# None: numpy.fft (pocketfft, one thread), the reference's library. An integer k: scipy.fft with k
# threads (pocketfft as well), for large offline fixtures; the results agree to rounding.
FFT_WORKERS = None


def rfftn(x, N):
    """Unnormalised real FFT over the last len(N) axes (fft.py:39-40)."""
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.rfftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N), workers=FFT_WORKERS)
    return np.fft.rfftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N))


def irfftn(x, N):
    """Normalised inverse real FFT (fft.py:42-43)."""
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.irfftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N), workers=FFT_WORKERS)
    return np.fft.irfftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N))


def fftn_n(x, N):
    """fft_form 0: fftn / n (fft.py:33-34)."""
    return np.fft.fftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N)) / np.prod(N)


def ifftn_n(x, N):
    """fft_form 0 inverse: real(ifftn) * n (fft.py:36-37)."""
    return np.fft.ifftn(x, s=tuple(N), axes=_grid_axes(x.ndim, N)).real * np.prod(N)


def fftnc(x, N):
    """fft_form 'c': fftshift(fftn) / n (fft.py:18-22)."""
    ax = _grid_axes(x.ndim, N)
    return np.fft.fftshift(np.fft.fftn(x, s=tuple(N), axes=ax), ax) / np.prod(N)


def icfftn(x, N):
    """fft_form 'c' inverse (fft.py:24-30)."""
    ax = _grid_axes(x.ndim, N)
    return np.fft.ifftn(np.fft.ifftshift(x, ax), s=tuple(N), axes=ax).real * np.prod(N)


def cfftnc(x, N):
    """Real and Fourier centred FFT (fft.py:4-9)."""
    ax = _grid_axes(x.ndim, N)
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, ax), s=tuple(N), axes=ax), ax) / np.prod(N)


def icfftnc(x, N):
    """Real and Fourier centred inverse FFT (fft.py:11-16)."""
    ax = _grid_axes(x.ndim, N)
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, ax), s=tuple(N), axes=ax), ax).real * np.prod(N)


_FFT = {'r': (rfftn, irfftn), 0: (fftn_n, ifftn_n), 'c': (fftnc, icfftn)}


def half_length(N):
    """Shape of the 'r' spectrum: last axis N//2+1 (objects.py:79-83)."""
    N = list(N)
    N[-1] = int(np.fix(N[-1] / 2) + 1)
    return tuple(N)


class Field:
    """Tensor-valued trigonometric polynomial: ``val`` has layout shape+N.

    Mirrors the attributes and arithmetic of ffthompy/tensors/objects.py:92-133.
    ``multype`` decides what ``*`` means: 'scal' is the grid inner product
    (objects.py:618-636), 21 the pointwise matrix-vector product, 42 the
    pointwise 4th-order contraction (objects.py:223-245).
    """

    def __init__(self, val=None, order=None, shape=None, N=None, Y=None, name='',
                 multype='scal', Fourier=False, fft_form=grid.FFT_FORM_DEFAULT, origin=0):
        self.name = name
        self.Fourier = Fourier
        self.origin = origin
        self.multype = multype
        if val is not None:
            self.val = val
            self.order = int(order)
            self.shape = tuple(val.shape[:self.order])
            if fft_form == 'r' and Fourier:
                self.N = tuple(int(n) for n in N)
            else:
                self.N = tuple(val.shape[self.order:])
        else:
            self.N = tuple(int(n) for n in N)
            self.shape = tuple(int(s) for s in shape)
            self.order = len(self.shape)
            if Fourier:
                nf = half_length(self.N) if fft_form == 'r' else self.N
                self.val = np.zeros(self.shape + nf, dtype=np.complex128)
            else:
                self.val = np.zeros(self.shape + self.N)
        self.fft_form = fft_form
        self.dim = len(self.N)
        self.Y = np.ones(self.dim) if Y is None else np.asarray(Y, dtype=np.float64)

    # -- bookkeeping
    @property
    def axes(self):
        return tuple(range(self.order, self.order + self.dim))

    @property
    def fft_coef(self):
        return float(np.prod(self.N)) if self.fft_form == 'r' else 1.

    def copy(self, **kw):
        d = dict(val=self.val.copy(), order=self.order, N=self.N, Y=self.Y, name=self.name,
                 multype=self.multype, Fourier=self.Fourier, fft_form=self.fft_form,
                 origin=self.origin)
        d.update(kw)
        return Field(**d)

    def zeros_like(self, name=None):
        return self.copy(val=np.zeros_like(self.val), name=name or self.name)

    # -- arithmetic (objects.py:188-245)
    def __neg__(self):
        return self.copy(val=-self.val)

    def __add__(self, x):
        if isinstance(x, Field):
            assert self.Fourier == x.Fourier and self.val.shape == x.val.shape
            return self.copy(val=self.val + x.val)
        return self.copy(val=self.val + x)

    def __sub__(self, x):
        return self.__add__(-x)

    def __rmul__(self, c):
        return self.copy(val=c * self.val)

    def __call__(self, x):
        return self.__mul__(x)

    def __mul__(self, y):
        assert self.Fourier == y.Fourier and self.fft_form == y.fft_form
        if self.multype == 'scal':
            return scalar_product(self, y)
        if self.multype == 21:
            return y.copy(val=np.einsum('ij...,j...->i...', self.val, y.val), order=1)
        if self.multype == 42:
            return y.copy(val=np.einsum('ijkl...,kl...->ij...', self.val, y.val), order=2)
        raise ValueError(self.multype)

    def norm(self):
        return scalar_product(self, self) ** 0.5

    def mean(self):
        """objects.py:259-272"""
        m = np.zeros(self.shape)
        if self.Fourier:
            ind = grid.mean_index(self.N, self.fft_form)
            for di in np.ndindex(*self.shape):
                m[di] = np.real(self.val[di][ind]) / self.fft_coef
        else:
            for di in np.ndindex(*self.shape):
                m[di] = np.mean(self.val[di])
        return m

    def add_mean(self, mean):
        if self.Fourier:
            ind = grid.mean_index(self.N, self.fft_form)
            for di in np.ndindex(*self.shape):
                self.val[di + ind] = mean[di] * self.fft_coef
        else:
            for di in np.ndindex(*self.shape):
                self.val[di] += mean[di]
        return self

    def set_mean(self, mean):
        """objects.py:289-300"""
        mean = np.asarray(mean, dtype=np.float64)
        self.add_mean(-self.mean())
        return self.add_mean(mean)

    def shift(self):
        """Toggle the real-space origin between centred and 0 (objects.py:169-186)."""
        assert not self.Fourier
        if self.origin == 0:
            self.val = np.fft.fftshift(self.val, self.axes)
            self.origin = 'c'
        else:
            self.val = np.fft.ifftshift(self.val, self.axes)
            self.origin = 0
        return self

    def inv(self):
        assert not self.Fourier and self.order == 2
        return self.copy(val=grid.pointwise_inverse(self.val), name='inv(%s)' % self.name)

    # -- Fourier forms (objects.py:136-167, 402-511)
    def set_fft_form(self, fft_form):
        if self.fft_form == fft_form:
            return self
        orig = self.fft_form
        if self.Fourier:
            if orig == 'r':
                tail = np.flip(self.val[..., 1:].conj(), axis=-1)
                for ax in self.axes[:-1]:
                    head, rest = np.split(tail, [1], axis=ax)
                    tail = np.concatenate((head, np.flip(rest, axis=ax)), axis=ax)
                if self.N[-1] % 2 == 0:
                    tail = tail[..., 1:]
                self.val = np.concatenate((self.val, tail), axis=-1) / np.prod(self.N)
                if fft_form == 'c':
                    self.val = np.fft.fftshift(self.val, axes=self.axes)
            elif orig == 'c':
                self.val = np.fft.ifftshift(self.val, axes=self.axes)
                if fft_form == 'r':
                    self.val = self.val[..., :half_length(self.N)[-1]] * np.prod(self.N)
            elif orig == 0:
                if fft_form == 'c':
                    self.val = np.fft.fftshift(self.val, axes=self.axes)
                else:
                    self.val = self.val[..., :half_length(self.N)[-1]] * np.prod(self.N)
        self.fft_form = fft_form
        return self

    def fourier(self, copy=False):
        assert self.origin == 0
        fwd, inv = _FFT[self.fft_form]
        obj = self.copy() if copy else self
        obj.val = inv(obj.val, obj.N) if obj.Fourier else fwd(obj.val, obj.N)
        obj.Fourier = not obj.Fourier
        return obj

    def enlarge(self, M):
        """Zero-pad the spectrum to grid M, splitting even-N Nyquist modes (objects.py:428-467)."""
        assert self.Fourier
        M = tuple(int(m) for m in M)
        if np.allclose(self.N, M):
            return self
        fft_form = self.fft_form
        self.set_fft_form('c')
        val = self.val
        for ii, ax in enumerate(self.axes):
            if self.N[ii] % 2 == 0:
                n0, rest = np.split(val, [1], axis=ax)
                n2 = np.copy(n0)
                for axc in self.axes:
                    if axc == ax:
                        continue
                    if n2.shape[axc] % 2 == 0:
                        a, b = np.split(n2, [1], axis=axc)
                        n2 = np.concatenate((a, np.flip(b, axis=axc)), axis=axc)
                    else:
                        n2 = np.flip(n2, axis=axc)
                val = np.concatenate((0.5 * n0, rest, 0.5 * n2.conj()), axis=ax)
        Mf = np.asarray(M, dtype=np.float64)
        Nf = np.asarray(val.shape[self.order:], dtype=np.float64)
        ibeg = np.ceil((Mf - Nf) / 2).astype(np.int64)
        iend = np.ceil((Mf + Nf) / 2).astype(np.int64)
        newval = np.zeros(self.shape + M, dtype=self.val.dtype)
        newval[(slice(None),) * self.order + tuple(slice(b, e) for b, e in zip(ibeg, iend))] = val
        R = self.copy(val=newval, N=M, fft_form='c')
        return R.set_fft_form(fft_form)

    def decrease(self, M):
        """Crop the spectrum to grid M (objects.py:469-486)."""
        assert self.Fourier
        M = tuple(int(m) for m in M)
        if np.allclose(self.N, M):
            return self
        fft_form = self.fft_form
        self.set_fft_form('c')
        val = np.zeros(self.shape + M, dtype=self.val.dtype)
        for di in np.ndindex(*self.shape):
            val[di] = grid.decrease_array(self.val[di], M)
        R = self.copy(val=val, N=M, fft_form='c')
        return R.set_fft_form(fft_form)

    def project(self, M):
        """Trigonometric interpolation onto grid M (objects.py:488-511)."""
        M = tuple(int(m) for m in M)
        if np.allclose(self.N, M):
            return self
        was_fourier = self.Fourier
        Z = self.copy() if was_fourier else self.fourier(copy=True)
        if all(m > n for m, n in zip(M, self.N)):
            Z = Z.enlarge(M)
        elif all(m < n for m, n in zip(M, self.N)):
            Z = Z.decrease(M)
        else:
            raise NotImplementedError()
        if not was_fourier:
            Z = Z.fourier()
        return Z


def scalar_product(y, x):
    """Grid inner product (objects.py:618-636): real space sum(x*y)/n;
    'r' spectra weight plane 0 once, the others twice (Nyquist once), / n^2."""
    assert y.val.shape == x.val.shape and y.fft_form == x.fft_form
    if y.Fourier:
        if x.fft_form == 'r':
            n2 = float(np.prod(y.N)) ** 2
            if x.N[-1] % 2 == 1:
                return (np.sum(y.val[..., 0] * np.conj(x.val[..., 0])).real
                        + 2 * np.sum(y.val[..., 1:] * np.conj(x.val[..., 1:])).real) / n2
            return (np.sum(y.val[..., 0] * np.conj(x.val[..., 0])).real
                    + np.sum(y.val[..., -1] * np.conj(x.val[..., -1])).real
                    + 2 * np.sum(y.val[..., 1:-1] * np.conj(x.val[..., 1:-1])).real) / n2
        return np.sum(y.val * np.conj(x.val)).real
    return np.sum(y.val * x.val) / np.prod(y.N)


class DFT:
    """(Inverse) DFT over the grid axes in a given FFT form (operators.py:14-64)."""

    def __init__(self, N, inverse=False, fft_form=grid.FFT_FORM_DEFAULT, name=None):
        self.N = tuple(int(n) for n in N)
        self.inverse = inverse
        self.fft_form = fft_form
        self.name = name or ('iDFT' if inverse else 'DFT')

    def __call__(self, x):
        assert x.Fourier == self.inverse
        fwd, inv = _FFT[self.fft_form]
        if self.inverse:
            return x.copy(val=inv(x.val, self.N).real, Fourier=False)
        assert x.fft_form == self.fft_form
        return x.copy(val=fwd(x.val, self.N), Fourier=True)

    def matrix(self):
        """Dense centred DFT matrix, 'c' form (operators.py:66-97)."""
        N = np.asarray(self.N)
        kin = grid.index_set(N, fft_form=0)
        kout = grid.index_set(N, fft_form='c')
        n = int(np.prod(N))
        M = np.zeros((n, n), dtype=np.complex128)
        sgn = 1. if self.inverse else -1.
        scale = 1. if self.inverse else 1. / n
        for i, k in enumerate(itertools.product(*kout)):
            for j, l in enumerate(itertools.product(*kin)):
                M[i, j] = np.exp(sgn * 2j * np.pi * np.sum(np.asarray(k) * np.asarray(l) / N)) * scale
        return M


class Operator:
    """Composition of linear maps applied right to left (operators.py:108-144)."""

    def __init__(self, mat, name='Operator'):
        self.name = name
        self.mat_rev = [list(reversed(summand)) for summand in mat]

    def __call__(self, x):
        res = None
        for summand in self.mat_rev:
            prod = x
            for op in summand:
                prod = op(prod)
            res = prod if res is None else prod + res
        return res + 0.
